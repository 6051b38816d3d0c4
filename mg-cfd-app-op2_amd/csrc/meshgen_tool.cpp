// mgcfd_meshgen -- writes a seeded synthetic multigrid mesh as an input deck
// (input.dat, io.h:28-205 grammar) plus one HDF5 file per level with the
// datasets euler3d.cpp:243-336 reads.
//
//   mgcfd_meshgen --box N [--levels L] [--seed S] -o DIR
//   mgcfd_meshgen --knn NODES EDGES [--levels L] [--ratios r1,r2,...] [--seed S] -o DIR
#include <sys/stat.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "mgcfd_mesh.h"
#include "op_hdf5.h"

static int usage()
{
    std::fprintf(stderr,
                 "usage: mgcfd_meshgen --box N [--levels L] [--seed S] -o DIR\n"
                 "       mgcfd_meshgen --knn NODES EDGES [--levels L] [--ratios r1,r2,..] [--seed S] -o DIR\n");
    return 2;
}

static bool put(const std::string& f, const char* ds, const char* type, const void* p, long long r, long long c,
                bool first)
{
    if (mgcfd_h5_write(f.c_str(), ds, type, p, r, c, first ? 1 : 0)) {
        std::fprintf(stderr, "mgcfd_meshgen: %s\n", mgcfd_h5_last_error());
        return false;
    }
    return true;
}

int main(int argc, char** argv)
{
    int box = 0, nodes = 0, edges = 0, levels = 4;
    unsigned long long seed = 1;
    std::string out, ratios_s;
    for (int i = 1; i < argc; i++) {
        std::string a = argv[i];
        if (a == "--box" && i + 1 < argc) box = std::atoi(argv[++i]);
        else if (a == "--knn" && i + 2 < argc) { nodes = std::atoi(argv[++i]); edges = std::atoi(argv[++i]); }
        else if (a == "--levels" && i + 1 < argc) levels = std::atoi(argv[++i]);
        else if (a == "--seed" && i + 1 < argc) seed = std::strtoull(argv[++i], nullptr, 10);
        else if (a == "--ratios" && i + 1 < argc) ratios_s = argv[++i];
        else if (a == "-o" && i + 1 < argc) out = argv[++i];
        else return usage();
    }
    if (out.empty() || (box <= 0 && nodes <= 0)) return usage();
    mkdir(out.c_str(), 0755);
    mgcfd_mesh* m = nullptr;
    int rc;
    if (box > 0) {
        rc = mgcfd_mesh_box(box, levels, seed, &m);
    } else {
        std::vector<double> r{1.0};
        size_t p = 0;
        while (!ratios_s.empty() && p < ratios_s.size()) {
            size_t q = ratios_s.find(',', p);
            r.push_back(std::atof(ratios_s.substr(p, q == std::string::npos ? std::string::npos : q - p).c_str()));
            if (q == std::string::npos) break;
            p = q + 1;
        }
        rc = mgcfd_mesh_knn(nodes, edges, levels, ratios_s.empty() ? nullptr : r.data(), seed, &m);
    }
    if (rc) {
        std::fprintf(stderr, "mgcfd_meshgen: %s\n", mgcfd_mesh_last_error());
        return 1;
    }
    for (int l = 0; l < m->levels; l++) {
        const mgcfd_mesh_level& L = m->L[l];
        const std::string f = out + "/level" + std::to_string(l) + ".h5";
        bool ok = put(f, "node_coordinates", "double", L.coords, L.n_nodes, 3, true) &&
                  put(f, "edge-->node", "int", L.edge_to_node, L.n_edges, 2, false) &&
                  put(f, "edge_weights", "double", L.edge_weights, L.n_edges, 3, false) &&
                  put(f, "bnd_node-->node", "int", L.bnd_to_node, L.n_bnd, 1, false) &&
                  put(f, "bnd_node-->group", "int", L.bnd_group, L.n_bnd, 1, false) &&
                  put(f, "bnd_node_weights", "double", L.bnd_weights, L.n_bnd, 3, false) &&
                  (!L.node_to_mg_node || put(f, "node-->mg_node", "int", L.node_to_mg_node, L.n_nodes, 1, false));
        if (!ok) return 1;
        std::printf("level %d: %d nodes, %d edges, %d boundary nodes -> %s\n", l, L.n_nodes, L.n_edges, L.n_bnd,
                    f.c_str());
    }
    std::ofstream deck(out + "/input.dat");
    deck << "# seeded synthetic mesh (mgcfd_meshgen)\n";
    deck << "size=" << m->L[0].n_nodes << "\n";
    deck << "num_levels=" << m->levels << "\n";
    deck << "base_array_index=0\n";
    deck << "mesh_name=m6wing\n";
    deck << "[levels]\n";
    for (int l = 0; l < m->levels; l++) deck << l << "=level" << l << ".h5\n";
    mgcfd_mesh_free(m);
    return 0;
}
