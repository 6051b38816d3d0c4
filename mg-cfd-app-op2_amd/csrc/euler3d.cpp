// euler3d_mi355x -- multigrid Euler solver driver on the MI355X runtime.
//
// Same command line, input deck, HDF5 layout and loop sequence as the
// reference's euler3d (euler3d.cpp:69-827, config.h:107-547, io.h:28-205),
// written against op_seq.h / op_hdf5.h so that every op_par_loop runs on the
// GPU through libmgcfd.  Additions, all off by default:
//   --mach M --aoa DEG     freestream (defaults: const.h:32-33, Mach 1.2, 0 deg)
//   --device-loop          whole multigrid cycles on the device (hipGraph),
//                          no host round trip per smoothing step
//   --flux-variant NAME    gather_exact (default) | gather_fast | atomic | nodal_exact |
//                          part_coloured | part_atomic | mixed_coloured | mixed_atomic
//   --validate-strict      -v with a 1e-12 relative bound per variable
//   --print-rms            print the L0 RMS residual of every cycle
#include <getopt.h>
#include <unistd.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <limits>
#include <sstream>
#include <string>
#include <vector>

#include "op_hdf5.h"
#include "op_seq.h"

OP_KERNEL(initialize_variables_kernel);
OP_KERNEL(zero_5d_array_kernel);
OP_KERNEL(zero_1d_array_kernel);
OP_KERNEL(calculate_cell_volumes);
OP_KERNEL(dampen_ewt);
OP_KERNEL(copy_double_kernel);
OP_KERNEL(calculate_dt_kernel);
OP_KERNEL(get_min_dt_kernel);
OP_KERNEL(compute_step_factor_kernel);
OP_KERNEL(compute_flux_edge_kernel);
OP_KERNEL(compute_bnd_node_flux_kernel);
OP_KERNEL(time_step_kernel);
OP_KERNEL(unstructured_stream_kernel);
OP_KERNEL(residual_kernel);
OP_KERNEL(calc_rms_kernel);
OP_KERNEL(count_bad_vals);
OP_KERNEL(up_pre_kernel);
OP_KERNEL(up_kernel);
OP_KERNEL(up_post_kernel);
OP_KERNEL(down_kernel);
OP_KERNEL(down_v2_kernel_pre);
OP_KERNEL(down_v2_kernel);
OP_KERNEL(down_v2_kernel_post);
OP_KERNEL(identify_differences);
OP_KERNEL(count_non_zeros);

namespace {

constexpr int NVAR = 5, NDIM = 3, RK = 3;
constexpr double GAMMA = 1.4;

struct Config {
    std::string config_file, input_file, input_dir, output_prefix;
    bool legacy = false, validate = false, strict = false, mem_bound = false, renumber = false;
    int cycles = 25, flow_interval = 0;
    bool out_variables = false, out_fluxes = false, out_step_factors = false;
    double mach = 1.2, aoa = 0.0;
    bool device_loop = false, print_rms = false, down_v2 = false;
    std::string flux_variant = "gather_exact";
};

std::string trim(const std::string& s)
{
    size_t a = s.find_first_not_of(" \t\r\n"), b = s.find_last_not_of(" \t\r\n");
    return a == std::string::npos ? "" : s.substr(a, b - a + 1);
}

// key=value settings shared by the -c config file and the command line
// (config.h:167-262).
void set_param(Config& c, const std::string& k, const std::string& v)
{
    if (k == "input_file") c.input_file = v;
    else if (k == "input_file_directory") c.input_dir = v;
    else if (k == "output_file_prefix") c.output_prefix = v;
    else if (k == "validate_result") c.validate = (v == "Y");
    else if (k == "measure_mem_bound") c.mem_bound = (v == "Y");
    else if (k == "cycles") c.cycles = std::atoi(v.c_str());
    else if (k == "partitioner" || k == "partitioner-method" || k == "papi-config-file") {
        // accepted for compatibility; one GPU has nothing to partition
    } else if (k == "renumber") c.renumber = (v == "Y");
    else if (k == "output_step_factors") c.out_step_factors = (v == "Y");
    else if (k == "output_fluxes") c.out_fluxes = (v == "Y");
    else if (k == "output_variables") c.out_variables = (v == "Y");
    else if (k == "output_flow_interval") c.flow_interval = std::atoi(v.c_str());
    else std::printf("WARNING: Unknown key '%s' encountered during parsing of config file.\n", k.c_str());
}

bool read_config(Config& c)
{
    std::ifstream f(c.config_file);
    if (!f.is_open()) {
        std::fprintf(stderr, "ERROR: \"%s\" does not exist.\n", c.config_file.c_str());
        return false;
    }
    std::string line;
    while (std::getline(f, line)) {
        if (line.empty() || line[0] == '#') continue;
        size_t eq = line.find('=');
        if (eq == std::string::npos) continue;
        set_param(c, trim(line.substr(0, eq)), trim(line.substr(eq + 1)));
    }
    return true;
}

void print_help()
{
    std::fprintf(stderr,
                 "MG-CFD (MI355X) instructions\n\n"
                 "Usage: euler3d_mi355x [OPTIONS] -i input.dat\n"
                 "       euler3d_mi355x [OPTIONS] -c conf.conf\n\n"
                 "-h, --help                     print help\n"
                 "-i, --input-file=FILEPATH      multigrid input grid (.dat file)\n"
                 "-c, --config-filepath=FILEPATH config file\n"
                 "-d, --input-directory=DIRPATH  directory path to input files\n"
                 "-o, --output-file-prefix=STR   string to prepend to output filenames\n"
                 "-g, --num-cycles=INT           number of multigrid V-cycles\n"
                 "-m, -r, -n, -p                 accepted; no effect on one GPU\n"
                 "-l, --legacy-mode              read the legacy '.renumbered' datasets and 'areas'\n"
                 "-b, --measure-mem-bound        also run unstructured_stream\n"
                 "-v, --validate                 check final state against solution files\n"
                 "-I, --output-flow-interval=INT write L0 variables every INT cycles\n"
                 "--output-variables --output-fluxes --output-step-factors\n"
                 "--mach M --aoa DEG --device-loop --flux-variant NAME --validate-strict --print-rms\n"
                 "--down-v2                      prolong by inverse distance (mg.h:90-204) over the\n"
                 "                               edge-->mg_nodes map built from edge-->node and node-->mg_node\n");
}

bool parse_args(Config& c, int argc, char** argv)
{
    enum { OPT_OUTV = 1000, OPT_OUTF, OPT_OUTS, OPT_MACH, OPT_AOA, OPT_DEV, OPT_FV, OPT_STRICT, OPT_RMS, OPT_DV2 };
    static struct option opts[] = {
        {"help", no_argument, nullptr, 'h'},
        {"config-filepath", required_argument, nullptr, 'c'},
        {"legacy-mode", no_argument, nullptr, 'l'},
        {"input-file", required_argument, nullptr, 'i'},
        {"input-directory", required_argument, nullptr, 'd'},
        {"papi-config-file", required_argument, nullptr, 'p'},
        {"output-file-prefix", required_argument, nullptr, 'o'},
        {"num-cycles", required_argument, nullptr, 'g'},
        {"partitioner", required_argument, nullptr, 'm'},
        {"partitioner-method", required_argument, nullptr, 'r'},
        {"renumber", no_argument, nullptr, 'n'},
        {"validate", no_argument, nullptr, 'v'},
        {"measure-mem-bound", no_argument, nullptr, 'b'},
        {"output-variables", no_argument, nullptr, OPT_OUTV},
        {"output-fluxes", no_argument, nullptr, OPT_OUTF},
        {"output-step-factors", no_argument, nullptr, OPT_OUTS},
        {"output-flow-interval", required_argument, nullptr, 'I'},
        {"mach", required_argument, nullptr, OPT_MACH},
        {"aoa", required_argument, nullptr, OPT_AOA},
        {"device-loop", no_argument, nullptr, OPT_DEV},
        {"flux-variant", required_argument, nullptr, OPT_FV},
        {"validate-strict", no_argument, nullptr, OPT_STRICT},
        {"print-rms", no_argument, nullptr, OPT_RMS},
        {"down-v2", no_argument, nullptr, OPT_DV2},
        {nullptr, 0, nullptr, 0},
    };
    int o;
    while ((o = getopt_long(argc, argv, "hc:li:d:p:o:g:m:r:vbI:", opts, nullptr)) != -1) {
        switch (o) {
            case 'h': print_help(); return false;
            case 'i': c.input_file = optarg; break;
            case 'c': c.config_file = optarg; if (!read_config(c)) std::exit(-1); break;
            case 'd': c.input_dir = optarg; break;
            case 'p': case 'm': case 'r': break;
            case 'o': c.output_prefix = optarg; break;
            case 'g': c.cycles = std::atoi(optarg); break;
            case 'n': c.renumber = true; break;
            case 'l': c.legacy = true; break;
            case 'b': c.mem_bound = true; break;
            case 'v': c.validate = true; break;
            case 'I': c.flow_interval = std::atoi(optarg); break;
            case OPT_OUTV: c.out_variables = true; break;
            case OPT_OUTF: c.out_fluxes = true; break;
            case OPT_OUTS: c.out_step_factors = true; break;
            case OPT_MACH: c.mach = std::atof(optarg); break;
            case OPT_AOA: c.aoa = std::atof(optarg); break;
            case OPT_DEV: c.device_loop = true; break;
            case OPT_FV: c.flux_variant = optarg; break;
            case OPT_STRICT: c.validate = c.strict = true; break;
            case OPT_RMS: c.print_rms = true; break;
            case OPT_DV2: c.down_v2 = true; break;
            default: std::printf("Unknown command line parameter '%c'\n", o);
        }
    }
    return true;
}

// input.dat deck (io.h:28-205): size, num_levels, base_array_index,
// mesh_name, then [levels] and an optional [mg_mapping] section.
bool read_input_dat(const std::string& path, int& levels, int& base_index, std::vector<std::string>& files)
{
    std::ifstream f(path);
    if (!f.is_open()) {
        std::fprintf(stderr, "Error: Could not open input file '%s'\n", path.c_str());
        return false;
    }
    bool have_size = false, have_levels = false, have_name = false, have_files = false;
    std::string line;
    while (std::getline(f, line)) {
        if (!line.empty() && line[0] == '#') continue;
        std::string t = trim(line);
        if (!t.empty() && t[0] == '[') {
            const bool lv = (t == "[levels]"), mg = (t == "[mg_mapping]");
            if (!lv && !mg) continue;
            if (!have_levels) {
                std::fprintf(stderr, "Error parsing %s: Need to know number of levels before parsing level filenames\n",
                             path.c_str());
                return false;
            }
            const int count = lv ? levels : levels - 1;
            if (lv) { files.assign(levels, ""); have_files = true; }
            for (int i = 0; i < count; i++) {
                if (!std::getline(f, line)) {
                    std::fprintf(stderr, "Error parsing %s: Have reached EOF before reading all %s filenames\n",
                                 path.c_str(), lv ? "mesh" : "MG");
                    return false;
                }
                size_t eq = line.find('=');
                if (eq == std::string::npos) {
                    std::fprintf(stderr, "Error parsing '%s': Was expecting a key-value pair following %s\n",
                                 path.c_str(), t.c_str());
                    return false;
                }
                const int idx = std::atoi(trim(line.substr(0, eq)).c_str());
                if (lv && idx >= 0 && idx < levels) files[idx] = trim(line.substr(eq + 1));
            }
            continue;
        }
        size_t eq = line.find('=');
        if (eq == std::string::npos) continue;
        const std::string k = trim(line.substr(0, eq)), v = trim(line.substr(eq + 1));
        if (k == "size") have_size = true;
        else if (k == "num_levels") { levels = std::atoi(v.c_str()); have_levels = true; }
        else if (k == "base_array_index") base_index = std::atoi(v.c_str());
        else if (k == "mesh_name") {
            if (v != "la_cascade" && v != "rotor37" && v != "fvcorr" && v != "m6wing") {
                std::fprintf(stderr, "Error parsing %s: Unknown mesh_name '%s'\n", path.c_str(), v.c_str());
                return false;
            }
            have_name = true;
        }
    }
    const char* missing = !have_size ? "size" : !have_levels ? "number of levels" : !have_name ? "mesh name"
                        : !have_files ? "mesh filenames" : nullptr;
    if (missing) {
        std::fprintf(stderr, "Error parsing '%s': %s not present\n", path.c_str(), missing);
        return false;
    }
    return true;
}

int flux_variant_id(const std::string& n)
{
    if (n == "gather_exact") return MGCFD_FLUX_GATHER_EXACT;
    if (n == "gather_fast") return MGCFD_FLUX_GATHER_FAST;
    if (n == "atomic") return MGCFD_FLUX_ATOMIC;
    if (n == "nodal_exact") return MGCFD_FLUX_NODAL_EXACT;
    if (n == "part_coloured") return MGCFD_FLUX_PART_COLOURED;
    if (n == "part_atomic") return MGCFD_FLUX_PART_ATOMIC;
    if (n == "mixed_coloured") return MGCFD_FLUX_MIXED_COLOURED;
    if (n == "mixed_atomic") return MGCFD_FLUX_MIXED_ATOMIC;
    return -1;
}

// Freestream state and flux contributions (euler3d.cpp:156-189) with the
// Mach number and angle of attack taken at run time.
struct Freestream {
    double var[NVAR];
    double fmx[NDIM], fmy[NDIM], fmz[NDIM], fde[NDIM];
};

Freestream freestream(double mach, double aoa_deg)
{
    Freestream f{};
    const double aoa = double(3.1415926535897931 / 180.0) * aoa_deg;
    f.var[0] = 1.4;
    const double p = 1.0;
    const double c = std::sqrt(GAMMA * p / f.var[0]);
    const double speed = mach * c;
    const double u[3] = {speed * std::cos(aoa), speed * std::sin(aoa), 0.0};
    for (int k = 0; k < 3; k++) f.var[1 + k] = f.var[0] * u[k];
    f.var[4] = f.var[0] * (0.5 * (speed * speed)) + (p / (GAMMA - 1.0));
    const double* m = f.var + 1;
    f.fmx[0] = u[0] * m[0] + p; f.fmx[1] = u[0] * m[1]; f.fmx[2] = u[0] * m[2];
    f.fmy[0] = f.fmx[1]; f.fmy[1] = u[1] * m[1] + p; f.fmy[2] = u[1] * m[2];
    f.fmz[0] = f.fmx[2]; f.fmz[1] = f.fmy[2]; f.fmz[2] = u[2] * m[2] + p;
    const double de_p = f.var[4] + p;
    for (int k = 0; k < 3; k++) f.fde[k] = u[k] * de_p;
    return f;
}

std::string join(const std::string& dir, const std::string& f) { return dir.empty() ? f : dir + "/" + f; }

}  // namespace

// Globals of euler3d.cpp:46-52 (declared as OP2 constants below).
double smoothing_coefficient = double(0.2f);
double ff_variable[NVAR];
double ff_flux_contribution_momentum_x[NDIM];
double ff_flux_contribution_momentum_y[NDIM];
double ff_flux_contribution_momentum_z[NDIM];
double ff_flux_contribution_density_energy[NDIM];
int mesh_name = 0;

int main(int argc, char** argv)
{
    Config conf;
    if (!parse_args(conf, argc, argv)) return 1;
    if (conf.input_file.empty()) {
        op_printf("ERROR: input_file not set\n");
        return 1;
    }
    if (conf.down_v2 && conf.device_loop) {
        op_printf("ERROR: --down-v2 runs loop by loop; the device-resident loop prolongs with down_kernel\n");
        return 1;
    }
    const int variant = flux_variant_id(conf.flux_variant);
    if (variant < 0) {
        std::fprintf(stderr, "ERROR: unknown --flux-variant '%s'\n", conf.flux_variant.c_str());
        return 1;
    }

    int levels = 0, base_index = 1;
    std::vector<std::string> layers;
    if (!read_input_dat(join(conf.input_dir, conf.input_file), levels, base_index, layers)) return 1;
    for (auto& l : layers) l = join(conf.input_dir, l);

    std::vector<std::string> args(argv, argv + argc);
    if (base_index >= 1 && base_index <= 9) args.push_back("OP_MAPS_BASE_INDEX=" + std::to_string(base_index));
    std::vector<char*> cargs;
    for (auto& a : args) cargs.push_back(&a[0]);
    op_check_(op_init((int)cargs.size(), cargs.data(), 0), "op_init");
    op_check_(mgcfd_set_flux_variant(variant), "flux variant");

    {
        Freestream f = freestream(conf.mach, conf.aoa);
        std::memcpy(ff_variable, f.var, sizeof f.var);
        std::memcpy(ff_flux_contribution_momentum_x, f.fmx, sizeof f.fmx);
        std::memcpy(ff_flux_contribution_momentum_y, f.fmy, sizeof f.fmy);
        std::memcpy(ff_flux_contribution_momentum_z, f.fmz, sizeof f.fmz);
        std::memcpy(ff_flux_contribution_density_energy, f.fde, sizeof f.fde);
    }
    op_decl_const(1, "double", &smoothing_coefficient);
    op_decl_const(NVAR, "double", ff_variable);
    op_decl_const(NDIM, "double", ff_flux_contribution_momentum_x);
    op_decl_const(NDIM, "double", ff_flux_contribution_momentum_y);
    op_decl_const(NDIM, "double", ff_flux_contribution_momentum_z);
    op_decl_const(NDIM, "double", ff_flux_contribution_density_energy);
    op_decl_const(1, "int", &mesh_name);

    std::vector<op_set> nodes(levels), edges(levels), bnd(levels);
    std::vector<op_map> e2n(levels), b2n(levels), n2mg(std::max(levels - 1, 0)), e2mg(levels, nullptr);
    std::vector<op_dat> ew(levels), bw(levels), bg(levels), xyz(levels), correct(levels, nullptr);
    std::vector<op_dat> var(levels), old(levels), res(levels), vol(levels, nullptr), sf(levels), flux(levels),
        dummy(levels, nullptr), scratch(levels, nullptr), rpro(levels, nullptr), rwsum(levels, nullptr);
    const std::string rn = conf.legacy ? ".renumbered" : "";
    char name[128];

    op_printf("-----------------------------------------------------\n");
    op_printf("Loading from HDF5 files ...\n");
    for (int i = 0; i < levels; i++) {
        op_printf("Loading level %d / %d\n", i + 1, levels);
        const char* L = layers[i].c_str();
        std::snprintf(name, sizeof name, "op_nodes_L%d", i);
        nodes[i] = op_decl_set_hdf5_infer_size(L, name, ("node_coordinates" + rn).c_str());
        std::snprintf(name, sizeof name, "op_edges_L%d", i);
        edges[i] = op_decl_set_hdf5_infer_size(L, name, ("edge-->node" + rn).c_str());
        std::snprintf(name, sizeof name, "op_bnd_nodes_L%d", i);
        bnd[i] = op_decl_set_hdf5_infer_size(L, name, ("bnd_node-->node" + rn).c_str());
        e2n[i] = op_decl_map_hdf5(edges[i], nodes[i], 2, L, ("edge-->node" + rn).c_str());
        b2n[i] = op_decl_map_hdf5(bnd[i], nodes[i], 1, L, ("bnd_node-->node" + rn).c_str());
        bg[i] = op_decl_dat_hdf5(bnd[i], 1, "int", L, "bnd_node-->group");
        if (i > 0) {
            n2mg[i - 1] = op_decl_map_hdf5(nodes[i - 1], nodes[i], 1, layers[i - 1].c_str(),
                                           ("node-->mg_node" + rn).c_str());
            // the map euler3d.cpp:286-292 describes and leaves undeclared
            if (conf.down_v2) {
                e2mg[i - 1] = mgcfd_decl_map_compose(e2n[i - 1], n2mg[i - 1], "edge-->mg_nodes");
                if (!e2mg[i - 1]) op_check_(-1, "mgcfd_decl_map_compose");
            }
        }
        if (conf.legacy) {
            vol[i] = op_decl_dat_hdf5(nodes[i], 1, "double", L, "areas");
            std::snprintf(name, sizeof name, "p_volumes_L%d", i);
            vol[i]->name = strdup(name);
        }
        ew[i] = op_decl_dat_hdf5(edges[i], NDIM, "double", L,
                                 conf.legacy ? "edge_weights.recalculated" : "edge_weights");
        bw[i] = op_decl_dat_hdf5(bnd[i], NDIM, "double", L, "bnd_node_weights");
        xyz[i] = op_decl_dat_hdf5(nodes[i], NDIM, "double", L, ("node_coordinates" + rn).c_str());
        if (conf.validate) {
            std::string sol = join(conf.input_dir, "solution.variables.L" + std::to_string(i) + ".cycles=" +
                                                       std::to_string(conf.cycles) + ".h5");
            if (access(sol.c_str(), R_OK) != -1) {
                correct[i] = op_decl_dat_hdf5(nodes[i], NVAR, "double", sol.c_str(),
                                              ("p_variables_result_L" + std::to_string(i)).c_str());
            } else {
                op_printf("Cannot find level %d solution file: %s\n", i, sol.c_str());
            }
        }
    }
    op_printf("-----------------------------------------------------\n");
    op_printf("Partitioning ...\n");
    for (int i = 0; i < levels; i++)
        op_check_(op_partition("PARMETIS", "KWAY", nodes[i], e2n[i], xyz[i]), "op_partition");
    op_printf("PARTITIONING COMPLETE\n");

    for (int i = 0; i < levels; i++) {
        auto temp = [&](int dim, const char* type, const char* fmt) {
            std::snprintf(name, sizeof name, fmt, i);
            op_dat d = op_decl_dat_temp_char(nodes[i], dim, type, std::strcmp(type, "int") ? 8 : 8, name);
            if (!d) op_check_(-1, "op_decl_dat_temp");
            return d;
        };
        var[i] = temp(NVAR, "double", "p_variables_L%d");
        old[i] = temp(NVAR, "double", "p_old_variables_L%d");
        res[i] = temp(NVAR, "double", "p_residuals_L%d");
        if (!conf.legacy) vol[i] = temp(1, "double", "p_volumes_L%d");
        sf[i] = temp(1, "double", "p_step_factors_L%d");
        flux[i] = temp(NVAR, "double", "p_fluxes_L%d");
        if (conf.mem_bound) dummy[i] = temp(NVAR, "double", "p_dummy_fluxes_L%d");
        if (i > 0) scratch[i] = temp(1, "int", "p_up_scratch_L%d");   // euler3d.cpp:405
        if (e2mg[i]) {
            rpro[i] = temp(NVAR, "double", "p_residuals_prolonged_L%d");
            rwsum[i] = temp(1, "double", "p_residuals_prolonged_wsum_L%d");
        }
    }

    // euler3d.cpp:412-441
    for (int i = 0; i < levels; i++) {
        op_par_loop(initialize_variables_kernel, "initialize_variables_kernel", nodes[i],
                    op_arg_dat(var[i], -1, OP_ID, NVAR, "double", OP_WRITE));
        op_par_loop(zero_5d_array_kernel, "zero_5d_array_kernel", nodes[i],
                    op_arg_dat(flux[i], -1, OP_ID, NVAR, "double", OP_WRITE));
        if (conf.mem_bound)
            op_par_loop(zero_5d_array_kernel, "zero_5d_array_kernel", nodes[i],
                        op_arg_dat(dummy[i], -1, OP_ID, NVAR, "double", OP_WRITE));
        if (!conf.legacy) {
            op_par_loop(zero_1d_array_kernel, "zero_1d_array_kernel", nodes[i],
                        op_arg_dat(vol[i], -1, OP_ID, 1, "double", OP_WRITE));
            op_par_loop(calculate_cell_volumes, "calculate_cell_volumes", edges[i],
                        op_arg_dat(xyz[i], 0, e2n[i], NDIM, "double", OP_READ),
                        op_arg_dat(xyz[i], 1, e2n[i], NDIM, "double", OP_READ),
                        op_arg_dat(ew[i], -1, OP_ID, NDIM, "double", OP_RW),
                        op_arg_dat(vol[i], 0, e2n[i], 1, "double", OP_INC),
                        op_arg_dat(vol[i], 1, e2n[i], 1, "double", OP_INC));
        }
    }
    for (int l = 0; l < levels; l++) {
        op_par_loop(dampen_ewt, "dampen_ewt", edges[l], op_arg_dat(ew[l], -1, OP_ID, NDIM, "double", OP_RW));
        op_par_loop(dampen_ewt, "dampen_ewt", bnd[l], op_arg_dat(bw[l], -1, OP_ID, NDIM, "double", OP_RW));
    }

    op_printf("-----------------------------------------------------\n");
    op_printf("Compute beginning\n");
    double cpu1, wall1, cpu2, wall2;
    op_timers(&cpu1, &wall1);

    auto write_flow = [&](int cycle) {
        const char* old_name = var[0]->name;
        var[0]->name = "p_variables";
        std::string f = conf.output_prefix + "variables.L0.cycle=" + std::to_string(cycle) + ".h5";
        op_fetch_data_hdf5_file(var[0], f.c_str());
        var[0]->name = old_name;
    };

    if (conf.device_loop) {
        std::vector<mgcfd_level_bind> binds(levels);
        for (int l = 0; l < levels; l++) {
            mgcfd_level_bind& b = binds[l];
            b = mgcfd_level_bind{};
            b.nodes = nodes[l]; b.edges = edges[l]; b.bnd_nodes = bnd[l];
            b.edge_to_node = e2n[l]; b.bnd_to_node = b2n[l];
            b.node_to_mg_node = l + 1 < levels ? n2mg[l] : nullptr;
            b.variables = var[l]; b.old_variables = old[l]; b.residuals = res[l]; b.fluxes = flux[l];
            b.step_factors = sf[l]; b.volumes = vol[l]; b.coords = xyz[l]; b.edge_weights = ew[l];
            b.bnd_weights = bw[l]; b.bnd_groups = bg[l]; b.up_scratch = scratch[l];
        }
        op_check_(mgcfd_mg_bind(levels, binds.data()), "mgcfd_mg_bind");
        const int chunk = conf.flow_interval > 0 ? conf.flow_interval : conf.cycles;
        std::vector<double> rms(std::max(chunk, 1));
        for (int done = 0; done < conf.cycles;) {
            const int n = std::min(chunk, conf.cycles - done);
            int status = 0;
            op_check_(mgcfd_mg_run(n, 1, rms.data(), &status), "mgcfd_mg_run");
            for (int k = 0; k < n; k++) {
                op_printf("Performing MG cycle %d / %d\n", done + k + 1, conf.cycles);
                if (conf.print_rms) op_printf("  rms = %.17g\n", rms[k]);
            }
            if (status == 1) {
                op_printf("Fatal error during 'step factor' calculation, min_dt < 0\n");
                op_exit();
                return 1;
            }
            if (status == 2) {
                op_printf("Bad variable values detected, aborting\n");
                op_exit();
                return 1;
            }
            done += n;
            if (conf.flow_interval > 0 && done % conf.flow_interval == 0) write_flow(done);
        }
    } else {
        // euler3d.cpp:452-641, loop by loop
        int level = 0, i = 0;
        bool up = true;
        while (i < conf.cycles) {
            if (level == 0) op_printf("Performing MG cycle %d / %d", i + 1, conf.cycles);
            op_par_loop(copy_double_kernel, "copy_double_kernel", nodes[level],
                        op_arg_dat(var[level], -1, OP_ID, NVAR, "double", OP_READ),
                        op_arg_dat(old[level], -1, OP_ID, NVAR, "double", OP_WRITE));
            op_par_loop(calculate_dt_kernel, "calculate_dt_kernel", nodes[level],
                        op_arg_dat(var[level], -1, OP_ID, NVAR, "double", OP_READ),
                        op_arg_dat(vol[level], -1, OP_ID, 1, "double", OP_READ),
                        op_arg_dat(sf[level], -1, OP_ID, 1, "double", OP_WRITE));
            double min_dt = std::numeric_limits<double>::max();
            op_par_loop(get_min_dt_kernel, "get_min_dt_kernel", nodes[level],
                        op_arg_dat(sf[level], -1, OP_ID, 1, "double", OP_READ),
                        op_arg_gbl(&min_dt, 1, "double", OP_MIN));
            if (min_dt < 0.0f) {
                op_printf("Fatal error during 'step factor' calculation, min_dt = %.5e\n", min_dt);
                op_exit();
                return 1;
            }
            op_par_loop(compute_step_factor_kernel, "compute_step_factor_kernel", nodes[level],
                        op_arg_dat(var[level], -1, OP_ID, NVAR, "double", OP_READ),
                        op_arg_dat(vol[level], -1, OP_ID, 1, "double", OP_READ),
                        op_arg_gbl(&min_dt, 1, "double", OP_READ),
                        op_arg_dat(sf[level], -1, OP_ID, 1, "double", OP_WRITE));
            for (int rk = 0; rk < RK; rk++) {
                op_par_loop(compute_flux_edge_kernel, "compute_flux_edge_kernel", edges[level],
                            op_arg_dat(var[level], 0, e2n[level], NVAR, "double", OP_READ),
                            op_arg_dat(var[level], 1, e2n[level], NVAR, "double", OP_READ),
                            op_arg_dat(ew[level], -1, OP_ID, NDIM, "double", OP_READ),
                            op_arg_dat(flux[level], 0, e2n[level], NVAR, "double", OP_INC),
                            op_arg_dat(flux[level], 1, e2n[level], NVAR, "double", OP_INC));
                op_par_loop(compute_bnd_node_flux_kernel, "compute_bnd_node_flux_kernel", bnd[level],
                            op_arg_dat(bg[level], -1, OP_ID, 1, "int", OP_READ),
                            op_arg_dat(bw[level], -1, OP_ID, NDIM, "double", OP_READ),
                            op_arg_dat(var[level], 0, b2n[level], NVAR, "double", OP_READ),
                            op_arg_dat(flux[level], 0, b2n[level], NVAR, "double", OP_INC));
                op_par_loop(time_step_kernel, "time_step_kernel", nodes[level], op_arg_gbl(&rk, 1, "int", OP_READ),
                            op_arg_dat(sf[level], -1, OP_ID, 1, "double", OP_READ),
                            op_arg_dat(flux[level], -1, OP_ID, NVAR, "double", OP_RW),
                            op_arg_dat(old[level], -1, OP_ID, NVAR, "double", OP_READ),
                            op_arg_dat(var[level], -1, OP_ID, NVAR, "double", OP_WRITE));
                if (conf.mem_bound)
                    op_par_loop(unstructured_stream_kernel, "unstructured_stream_kernel", edges[level],
                                op_arg_dat(var[level], 0, e2n[level], NVAR, "double", OP_READ),
                                op_arg_dat(var[level], 1, e2n[level], NVAR, "double", OP_READ),
                                op_arg_dat(ew[level], -1, OP_ID, NDIM, "double", OP_READ),
                                op_arg_dat(dummy[level], 0, e2n[level], NVAR, "double", OP_INC),
                                op_arg_dat(dummy[level], 1, e2n[level], NVAR, "double", OP_INC));
            }
            op_par_loop(residual_kernel, "residual_kernel", nodes[level],
                        op_arg_dat(old[level], -1, OP_ID, NVAR, "double", OP_READ),
                        op_arg_dat(var[level], -1, OP_ID, NVAR, "double", OP_READ),
                        op_arg_dat(res[level], -1, OP_ID, NVAR, "double", OP_WRITE));
            if (level == 0) {
                double rms = 0.0;
                op_par_loop(calc_rms_kernel, "calc_rms_kernel", nodes[level],
                            op_arg_dat(res[level], -1, OP_ID, NVAR, "double", OP_READ),
                            op_arg_gbl(&rms, 1, "double", OP_INC));
                rms = std::sqrt(rms / double(op_get_size(nodes[level])));
                int bad = 0;
                op_par_loop(count_bad_vals, "count_bad_vals", nodes[level],
                            op_arg_dat(var[level], -1, OP_ID, NVAR, "double", OP_READ),
                            op_arg_gbl(&bad, 1, "int", OP_INC));
                if (bad > 0) {
                    op_printf("Bad variable values detected, aborting\n");
                    op_exit();
                    return 1;
                }
                if (conf.print_rms) op_printf("  rms = %.17g", rms);
                op_printf("\n");
            }
            if (conf.flow_interval > 0 && ((i + 1) % conf.flow_interval) == 0 && level == 0) write_flow(i + 1);
            if (levels <= 1) {
                i++;
            } else if (up) {
                level++;
                op_par_loop(up_pre_kernel, "up_pre_kernel", nodes[level - 1],
                            op_arg_dat(var[level], 0, n2mg[level - 1], NVAR, "double", OP_WRITE),
                            op_arg_dat(scratch[level], 0, n2mg[level - 1], 1, "int", OP_WRITE));
                op_par_loop(up_kernel, "up_kernel", nodes[level - 1],
                            op_arg_dat(var[level - 1], -1, OP_ID, NVAR, "double", OP_READ),
                            op_arg_dat(var[level], 0, n2mg[level - 1], NVAR, "double", OP_INC),
                            op_arg_dat(scratch[level], 0, n2mg[level - 1], 1, "int", OP_INC));
                op_par_loop(up_post_kernel, "up_post_kernel", nodes[level],
                            op_arg_dat(var[level], -1, OP_ID, NVAR, "double", OP_RW),
                            op_arg_dat(scratch[level], -1, OP_ID, 1, "int", OP_READ));
                if (level == levels - 1) up = false;
            } else {
                level--;
                if (e2mg[level]) {      // euler3d.cpp:603-623
                    op_map en = e2n[level], em = e2mg[level];
                    op_par_loop(down_v2_kernel_pre, "down_v2_kernel_pre", nodes[level],
                                op_arg_dat(rwsum[level], -1, OP_ID, 1, "double", OP_WRITE),
                                op_arg_dat(rpro[level], -1, OP_ID, NVAR, "double", OP_WRITE));
                    op_par_loop(down_v2_kernel, "down_v2_kernel", edges[level],
                                op_arg_dat(xyz[level], 0, en, NDIM, "double", OP_READ),
                                op_arg_dat(xyz[level], 1, en, NDIM, "double", OP_READ),
                                op_arg_dat(xyz[level + 1], 0, em, NDIM, "double", OP_READ),
                                op_arg_dat(xyz[level + 1], 1, em, NDIM, "double", OP_READ),
                                op_arg_dat(res[level + 1], 0, em, NVAR, "double", OP_READ),
                                op_arg_dat(res[level + 1], 1, em, NVAR, "double", OP_READ),
                                op_arg_dat(rpro[level], 0, en, NVAR, "double", OP_INC),
                                op_arg_dat(rpro[level], 1, en, NVAR, "double", OP_INC),
                                op_arg_dat(rwsum[level], 0, en, 1, "double", OP_INC),
                                op_arg_dat(rwsum[level], 1, en, 1, "double", OP_INC));
                    op_par_loop(down_v2_kernel_post, "down_v2_kernel_post", nodes[level],
                                op_arg_dat(rpro[level], -1, OP_ID, NVAR, "double", OP_READ),
                                op_arg_dat(rwsum[level], -1, OP_ID, 1, "double", OP_READ),
                                op_arg_dat(res[level], -1, OP_ID, NVAR, "double", OP_READ),
                                op_arg_dat(var[level], -1, OP_ID, NVAR, "double", OP_INC));
                } else {                // euler3d.cpp:625-631
                    op_par_loop(down_kernel, "down_kernel", nodes[level],
                                op_arg_dat(var[level], -1, OP_ID, NVAR, "double", OP_INC),
                                op_arg_dat(res[level], -1, OP_ID, NVAR, "double", OP_READ),
                                op_arg_dat(xyz[level], -1, OP_ID, NDIM, "double", OP_READ),
                                op_arg_dat(res[level + 1], 0, n2mg[level], NVAR, "double", OP_READ),
                                op_arg_dat(xyz[level + 1], 0, n2mg[level], NDIM, "double", OP_READ));
                }
                if (level == 0) { up = true; i++; }
            }
        }
    }
    op_printf("\n");
    op_printf("Compute complete\n");
    op_timers(&cpu2, &wall2);
    op_printf("Max total runtime = %f\n", wall2 - wall1);

    // euler3d.cpp:658-718 (+ strict bound)
    if (conf.validate) {
        op_printf("-----------------------------------------------------\n");
        bool failed = false;
        op_printf("Looking for NaN and infinity values ...");
        for (int l = 0; l < levels && !failed; l++) {
            int bad = 0;
            op_par_loop(count_bad_vals, "count_bad_vals", nodes[l],
                        op_arg_dat(var[l], -1, OP_ID, NVAR, "double", OP_READ), op_arg_gbl(&bad, 1, "int", OP_INC));
            if (bad > 0) {
                failed = true;
                op_printf("\nValue check of MG level %d failed: %d bad values detected\n", l, bad);
            }
        }
        if (!failed) {
            op_printf(" None found\n");
            op_printf("Validating result against solution ...");
            bool vfail = false;
            for (int l = 0; l < levels; l++) {
                if (!correct[l]) {
                    op_printf("\n- Do not have solution for level %d, cannot validate\n", l);
                    vfail = true;
                    continue;
                }
                std::snprintf(name, sizeof name, "p_var_diff_L%d", l);
                op_dat diff = op_decl_dat_temp_char(nodes[l], NVAR, "double", 8, name);
                op_par_loop(identify_differences, "identify_differences", nodes[l],
                            op_arg_dat(var[l], -1, OP_ID, NVAR, "double", OP_READ),
                            op_arg_dat(correct[l], -1, OP_ID, NVAR, "double", OP_READ),
                            op_arg_dat(diff, -1, OP_ID, NVAR, "double", OP_WRITE));
                int count = 0;
                op_par_loop(count_non_zeros, "count_non_zeros", nodes[l],
                            op_arg_dat(diff, -1, OP_ID, NVAR, "double", OP_READ), op_arg_gbl(&count, 1, "int", OP_INC));
                const int threshold = op_get_size(nodes[l]) / 5000;
                if (count > threshold) {
                    vfail = true;
                    op_printf("\nValidation of MG level %d failed: %d incorrect values in 'variables' array\n", l, count);
                    break;
                }
                if (conf.strict) {
                    const int n = op_get_size(nodes[l]);
                    std::vector<double> g((size_t)n * NVAR), r((size_t)n * NVAR);
                    op_check_(op_fetch_data_char(var[l], (char*)g.data()), "fetch");
                    op_check_(op_fetch_data_char(correct[l], (char*)r.data()), "fetch");
                    double worst = 0.0;
                    for (int v = 0; v < NVAR; v++) {
                        double dmax = 0.0, smax = 0.0;
                        for (int k = 0; k < n; k++) {
                            dmax = std::max(dmax, std::fabs(g[(size_t)k * NVAR + v] - r[(size_t)k * NVAR + v]));
                            smax = std::max(smax, std::fabs(r[(size_t)k * NVAR + v]));
                        }
                        if (smax > 0) worst = std::max(worst, dmax / smax);
                    }
                    op_printf("\n  level %d: max relative difference per variable %.3e", l, worst);
                    if (!(worst <= 1e-12)) {
                        vfail = true;
                        op_printf(" (above 1e-12)");
                    }
                }
            }
            if (vfail) {
                op_printf("\nValidation failed\n");
            } else {
                op_printf(" Result correct\n");
                op_printf("Validation passed\n");
            }
        }
    }

    // euler3d.cpp:720-779
    if (conf.out_variables || conf.out_fluxes || conf.out_step_factors) {
        op_printf("-----------------------------------------------------\n");
        op_printf("Writing out data...\n");
        for (int l = 0; l < levels; l++) {
            const std::string suffix = ".L" + std::to_string(l) + ".cycles=" + std::to_string(conf.cycles) + ".h5";
            auto dump = [&](op_dat d, const char* what) {
                const char* old_name = d->name;
                std::string nm = std::string("p_") + what + "_result_L" + std::to_string(l);
                d->name = nm.c_str();
                op_fetch_data_hdf5_file(d, (conf.output_prefix + what + suffix).c_str());
                d->name = old_name;
            };
            if (conf.out_step_factors) dump(sf[l], "step_factors");
            if (conf.out_fluxes) dump(flux[l], "fluxes");
            if (conf.out_variables) dump(var[l], "variables");
        }
    }
    op_printf("-----------------------------------------------------\n");
    op_printf("Winding down OP2\n");
    op_exit();
    return 0;
}
