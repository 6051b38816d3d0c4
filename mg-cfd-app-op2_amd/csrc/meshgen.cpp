// Seeded synthetic multigrid meshes (see include/mgcfd_mesh.h).
#include "mgcfd_mesh.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed * 0x9E3779B97F4A7C15ULL + 0x632BE59BD9B4E019ULL) {}
    uint64_t next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        return z ^ (z >> 31);
    }
    double uni() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    double sym() { return 2.0 * uni() - 1.0; }
    uint64_t below(uint64_t n) { return next() % n; }
};

template <typename T>
void shuffle(std::vector<T>& v, Rng& r)
{
    for (size_t i = v.size(); i > 1; i--) std::swap(v[i - 1], v[r.below(i)]);
}

struct Level {
    std::vector<double> xyz;     // [N][3]
    std::vector<int> e2n;        // [E][2]
    std::vector<double> ew;      // [E][3]
    std::vector<int> b2n, bg;    // [B]
    std::vector<double> bw;      // [B][3]
    std::vector<int> n2c;        // [N] (empty on coarsest)
    int n() const { return (int)(xyz.size() / 3); }
    int e() const { return (int)(e2n.size() / 2); }
    int b() const { return (int)b2n.size(); }
};

void add_edge(Level& L, int a, int b, double area)
{
    double d[3], len = 0;
    for (int c = 0; c < 3; c++) { d[c] = L.xyz[3 * (size_t)b + c] - L.xyz[3 * (size_t)a + c]; len += d[c] * d[c]; }
    len = std::sqrt(len);
    L.e2n.push_back(a);
    L.e2n.push_back(b);
    for (int c = 0; c < 3; c++) L.ew.push_back(d[c] / len * area);
}

void add_bnd(Level& L, int node, int group, double nx, double ny, double nz, double area)
{
    L.b2n.push_back(node);
    L.bg.push_back(group);
    L.bw.push_back(nx * area);
    L.bw.push_back(ny * area);
    L.bw.push_back(nz * area);
}

// Random relabelling of nodes, edges (with random orientation) and boundary
// entries of every level; node-->mg_node maps follow both levels' labels.
void permute_all(std::vector<Level>& lv, Rng& r)
{
    std::vector<std::vector<int>> newid(lv.size());
    for (size_t l = 0; l < lv.size(); l++) {
        int n = lv[l].n();
        std::vector<int> p(n);
        std::iota(p.begin(), p.end(), 0);
        shuffle(p, r);                       // p[new] = old
        newid[l].assign(n, 0);
        for (int i = 0; i < n; i++) newid[l][p[i]] = i;
        std::vector<double> x(3 * (size_t)n);
        for (int i = 0; i < n; i++)
            for (int c = 0; c < 3; c++) x[3 * (size_t)i + c] = lv[l].xyz[3 * (size_t)p[i] + c];
        lv[l].xyz.swap(x);
    }
    for (size_t l = 0; l < lv.size(); l++) {
        Level& L = lv[l];
        const auto& id = newid[l];
        int ne = L.e();
        std::vector<int> pe(ne);
        std::iota(pe.begin(), pe.end(), 0);
        shuffle(pe, r);
        std::vector<int> e2n(2 * (size_t)ne);
        std::vector<double> ew(3 * (size_t)ne);
        for (int i = 0; i < ne; i++) {
            int o = pe[i];
            bool flip = r.below(2) != 0;
            int a = id[L.e2n[2 * (size_t)o]], b = id[L.e2n[2 * (size_t)o + 1]];
            e2n[2 * (size_t)i] = flip ? b : a;
            e2n[2 * (size_t)i + 1] = flip ? a : b;
            for (int c = 0; c < 3; c++) ew[3 * (size_t)i + c] = flip ? -L.ew[3 * (size_t)o + c] : L.ew[3 * (size_t)o + c];
        }
        L.e2n.swap(e2n);
        L.ew.swap(ew);
        int nb = L.b();
        std::vector<int> pb(nb);
        std::iota(pb.begin(), pb.end(), 0);
        shuffle(pb, r);
        std::vector<int> b2n(nb), bg(nb);
        std::vector<double> bw(3 * (size_t)nb);
        for (int i = 0; i < nb; i++) {
            b2n[i] = id[L.b2n[pb[i]]];
            bg[i] = L.bg[pb[i]];
            for (int c = 0; c < 3; c++) bw[3 * (size_t)i + c] = L.bw[3 * (size_t)pb[i] + c];
        }
        L.b2n.swap(b2n);
        L.bg.swap(bg);
        L.bw.swap(bw);
        if (!L.n2c.empty()) {
            std::vector<int> m(L.n2c.size());
            for (size_t f = 0; f < L.n2c.size(); f++) m[id[f]] = newid[l + 1][L.n2c[f]];
            L.n2c.swap(m);
        }
    }
}

// ------------------------------------------------------------- box mesh --

Level box_level(int n, int l, double h0, Rng& r)
{
    Level L;
    const double scale = (double)(1 << l);
    const double h = h0 * scale;
    auto id = [n](int i, int j, int k) { return ((size_t)k * n + j) * n + i; };
    L.xyz.resize(3 * (size_t)n * n * n);
    for (int k = 0; k < n; k++)
        for (int j = 0; j < n; j++)
            for (int i = 0; i < n; i++) {
                size_t p = id(i, j, k);
                int I[3] = {i, j, k};
                for (int c = 0; c < 3; c++)
                    L.xyz[3 * p + c] = (I[c] * scale + (scale - 1.0) * 0.5) * h0 + r.sym() * 0.2 * h;
            }
    // dual-cell extent of node p along axis c from its lattice neighbours
    auto ext = [&](int i, int j, int k, int c) {
        int I[3] = {i, j, k};
        int lo[3] = {i, j, k}, hi[3] = {i, j, k};
        lo[c] = std::max(0, I[c] - 1);
        hi[c] = std::min(n - 1, I[c] + 1);
        double xl = L.xyz[3 * id(lo[0], lo[1], lo[2]) + c], xh = L.xyz[3 * id(hi[0], hi[1], hi[2]) + c];
        return 0.5 * (xh - xl);
    };
    for (int k = 0; k < n; k++)
        for (int j = 0; j < n; j++)
            for (int i = 0; i < n; i++) {
                int I[3] = {i, j, k};
                for (int c = 0; c < 3; c++) {
                    if (I[c] + 1 >= n) continue;
                    int J[3] = {i, j, k};
                    J[c]++;
                    int t1 = (c + 1) % 3, t2 = (c + 2) % 3;
                    double area = 0.5 * (ext(i, j, k, t1) * ext(i, j, k, t2) +
                                         ext(J[0], J[1], J[2], t1) * ext(J[0], J[1], J[2], t2));
                    add_edge(L, (int)id(i, j, k), (int)id(J[0], J[1], J[2]), area);
                }
            }
    for (int f = 0; f < 6; f++) {
        int c = f / 2;
        bool hi = f % 2;
        int t1 = (c + 1) % 3, t2 = (c + 2) % 3;
        for (int a = 0; a < n; a++)
            for (int b = 0; b < n; b++) {
                int I[3];
                I[c] = hi ? n - 1 : 0;
                I[t1] = a;
                I[t2] = b;
                double nrm[3] = {0, 0, 0};
                nrm[c] = hi ? -1.0 : 1.0;     // inward
                int group = (c == 2 && !hi) ? 0 : 3;
                double area = ext(I[0], I[1], I[2], t1) * ext(I[0], I[1], I[2], t2);
                add_bnd(L, (int)id(I[0], I[1], I[2]), group, nrm[0], nrm[1], nrm[2], area);
            }
    }
    return L;
}

// ------------------------------------------------------------- kNN mesh --

Level knn_level0(int N, int E, Rng& r, std::string& err)
{
    Level L;
    int ny = std::max(2, (int)std::lround(std::cbrt(N / 2.4)));
    int nz = ny;
    int nx = (N + ny * nz - 1) / (ny * nz);
    const double h = 1.0 / (nz - 1);
    auto lat = [&](int p, int& i, int& j, int& k) { i = p / (ny * nz); j = (p / nz) % ny; k = p % nz; };
    auto pid = [&](int i, int j, int k) -> int {
        if (i < 0 || j < 0 || k < 0 || i >= nx || j >= ny || k >= nz) return -1;
        long long p = ((long long)i * ny + j) * nz + k;
        return p < N ? (int)p : -1;
    };
    L.xyz.resize(3 * (size_t)N);
    for (int p = 0; p < N; p++) {
        int i, j, k;
        lat(p, i, j, k);
        L.xyz[3 * (size_t)p + 0] = i * h + r.sym() * 0.2 * h;
        L.xyz[3 * (size_t)p + 1] = j * h + r.sym() * 0.2 * h;
        L.xyz[3 * (size_t)p + 2] = k * h + r.sym() * 0.2 * h;
    }
    auto d2 = [&](int a, int b) {
        double s = 0;
        for (int c = 0; c < 3; c++) { double t = L.xyz[3 * (size_t)a + c] - L.xyz[3 * (size_t)b + c]; s += t * t; }
        return s;
    };
    std::vector<long long> cand;
    std::vector<std::pair<double, int>> nb;
    for (int kk = 6; kk <= 14; kk += 2) {
        cand.clear();
        for (int p = 0; p < N; p++) {
            int i, j, k;
            lat(p, i, j, k);
            nb.clear();
            for (int di = -1; di <= 1; di++)
                for (int dj = -1; dj <= 1; dj++)
                    for (int dk = -1; dk <= 1; dk++) {
                        if (!di && !dj && !dk) continue;
                        int q = pid(i + di, j + dj, k + dk);
                        if (q >= 0) nb.push_back({d2(p, q), q});
                    }
            std::sort(nb.begin(), nb.end());
            for (int t = 0; t < kk && t < (int)nb.size(); t++) {
                int q = nb[t].second;
                cand.push_back(((long long)std::min(p, q) << 32) | (unsigned)std::max(p, q));
            }
        }
        std::sort(cand.begin(), cand.end());
        cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
        if ((long long)cand.size() >= E) break;
    }
    if ((long long)cand.size() < E) { err = "knn: not enough candidate edges for the requested count"; return L; }
    // trim the longest edges while both ends keep degree > 3
    std::vector<int> deg(N, 0);
    for (long long c : cand) { deg[c >> 32]++; deg[c & 0xffffffff]++; }
    std::vector<std::pair<double, long long>> bylen;
    bylen.reserve(cand.size());
    for (long long c : cand) bylen.push_back({d2((int)(c >> 32), (int)(c & 0xffffffff)), c});
    std::sort(bylen.begin(), bylen.end(), [](auto& x, auto& y) { return x.first > y.first || (x.first == y.first && x.second < y.second); });
    long long excess = (long long)cand.size() - E;
    std::vector<char> keep(bylen.size(), 1);
    for (size_t t = 0; t < bylen.size() && excess > 0; t++) {
        int a = (int)(bylen[t].second >> 32), b = (int)(bylen[t].second & 0xffffffff);
        if (deg[a] > 3 && deg[b] > 3) { keep[t] = 0; deg[a]--; deg[b]--; excess--; }
    }
    if (excess > 0) { err = "knn: cannot trim to the requested edge count"; return L; }
    std::vector<long long> kept;
    for (size_t t = 0; t < bylen.size(); t++)
        if (keep[t]) kept.push_back(bylen[t].second);
    std::sort(kept.begin(), kept.end());
    for (long long c : kept) {
        int a = (int)(c >> 32), b = (int)(c & 0xffffffff);
        double len = std::sqrt(d2(a, b));
        double area = h * h * std::min(1.0, (h / len) * (h / len)) * (1.0 + 0.1 * r.sym());
        add_edge(L, a, b, area);
    }
    for (int p = 0; p < N; p++) {
        int i, j, k;
        lat(p, i, j, k);
        double area = h * h * (1.0 + 0.1 * r.sym());
        if (k == 0) add_bnd(L, p, 0, 0, 0, 1, area);
        if (i == 0) add_bnd(L, p, 3, 1, 0, 0, area);
        if (pid(i + 1, j, k) < 0) add_bnd(L, p, 3, -1, 0, 0, area);
        if (k == nz - 1) add_bnd(L, p, 3, 0, 0, -1, area);
    }
    return L;
}

// Greedy agglomeration of level f into 'target' aggregates; fills f.n2c.
Level agglomerate(Level& f, int target, Rng& r, std::string& err)
{
    const int N = f.n(), E = f.e();
    Level c;
    std::vector<int> ptr(N + 1, 0);
    for (int e = 0; e < E; e++) { ptr[f.e2n[2 * (size_t)e] + 1]++; ptr[f.e2n[2 * (size_t)e + 1] + 1]++; }
    for (int i = 0; i < N; i++) ptr[i + 1] += ptr[i];
    std::vector<int> adj(ptr[N]), fill(ptr.begin(), ptr.end() - 1);
    for (int e = 0; e < E; e++) {
        int a = f.e2n[2 * (size_t)e], b = f.e2n[2 * (size_t)e + 1];
        adj[fill[a]++] = b;
        adj[fill[b]++] = a;
    }
    auto d2 = [&](int a, int b) {
        double s = 0;
        for (int q = 0; q < 3; q++) { double t = f.xyz[3 * (size_t)a + q] - f.xyz[3 * (size_t)b + q]; s += t * t; }
        return s;
    };
    std::vector<int> agg(N, -1), size;
    std::vector<int> order(N);
    std::iota(order.begin(), order.end(), 0);
    shuffle(order, r);
    int count = N;      // aggregates if every unassigned node became a singleton
    for (int u : order) {
        if (count <= target) break;
        if (agg[u] >= 0) continue;
        int best = -1;
        double bd = 0;
        for (int k = ptr[u]; k < ptr[u + 1]; k++) {
            int v = adj[k];
            if (agg[v] >= 0) continue;
            double d = d2(u, v);
            if (best < 0 || d < bd) { best = v; bd = d; }
        }
        if (best < 0) continue;
        agg[u] = agg[best] = (int)size.size();
        size.push_back(2);
        count--;
    }
    for (int u = 0; u < N; u++)
        if (agg[u] < 0) { agg[u] = (int)size.size(); size.push_back(1); }
    for (int limit = 3; count > target && limit <= 64; limit++) {
        shuffle(order, r);
        for (int u : order) {
            if (count <= target) break;
            if (size[agg[u]] != 1) continue;
            int best = -1;
            double bd = 0;
            for (int k = ptr[u]; k < ptr[u + 1]; k++) {
                int v = adj[k];
                int a = agg[v];
                if (a == agg[u] || size[a] >= limit) continue;
                double d = d2(u, v);
                if (best < 0 || d < bd) { best = a; bd = d; }
            }
            if (best < 0) continue;
            size[agg[u]] = 0;
            agg[u] = best;
            size[best]++;
            count--;
        }
    }
    if (count > target) { err = "agglomeration could not reach the target node count"; return c; }
    // compact, randomly ordered coarse ids
    std::vector<int> live;
    for (int a = 0; a < (int)size.size(); a++)
        if (size[a] > 0) live.push_back(a);
    shuffle(live, r);
    std::vector<int> cid(size.size(), -1);
    for (int i = 0; i < (int)live.size(); i++) cid[live[i]] = i;
    const int NC = (int)live.size();
    f.n2c.resize(N);
    for (int u = 0; u < N; u++) f.n2c[u] = cid[agg[u]];
    c.xyz.assign(3 * (size_t)NC, 0.0);
    std::vector<int> cnt(NC, 0);
    for (int u = 0; u < N; u++) {
        int a = f.n2c[u];
        cnt[a]++;
        for (int q = 0; q < 3; q++) c.xyz[3 * (size_t)a + q] += f.xyz[3 * (size_t)u + q];
    }
    for (int a = 0; a < NC; a++)
        for (int q = 0; q < 3; q++) c.xyz[3 * (size_t)a + q] /= cnt[a];
    std::vector<std::pair<long long, double>> ce;
    for (int e = 0; e < E; e++) {
        int A = f.n2c[f.e2n[2 * (size_t)e]], B = f.n2c[f.e2n[2 * (size_t)e + 1]];
        if (A == B) continue;
        double area = 0;
        for (int q = 0; q < 3; q++) area += f.ew[3 * (size_t)e + q] * f.ew[3 * (size_t)e + q];
        ce.push_back({((long long)std::min(A, B) << 32) | (unsigned)std::max(A, B), std::sqrt(area)});
    }
    std::sort(ce.begin(), ce.end());
    for (size_t t = 0; t < ce.size();) {
        size_t u = t;
        double area = 0;
        while (u < ce.size() && ce[u].first == ce[t].first) area += ce[u++].second;
        add_edge(c, (int)(ce[t].first >> 32), (int)(ce[t].first & 0xffffffff), area);
        t = u;
    }
    std::vector<std::pair<long long, int>> cb;
    for (int b = 0; b < f.b(); b++) cb.push_back({((long long)f.n2c[f.b2n[b]] << 8) | (f.bg[b] & 0xff), b});
    std::sort(cb.begin(), cb.end());
    for (size_t t = 0; t < cb.size();) {
        size_t u = t;
        double w[3] = {0, 0, 0};
        while (u < cb.size() && cb[u].first == cb[t].first) {
            for (int q = 0; q < 3; q++) w[q] += f.bw[3 * (size_t)cb[u].second + q];
            u++;
        }
        c.b2n.push_back((int)(cb[t].first >> 8));
        c.bg.push_back((int)(cb[t].first & 0xff));
        for (int q = 0; q < 3; q++) c.bw.push_back(w[q]);
        t = u;
    }
    return c;
}

template <typename T>
T* dup(const std::vector<T>& v)
{
    T* p = (T*)std::malloc(sizeof(T) * std::max<size_t>(1, v.size()));
    if (p && !v.empty()) std::memcpy(p, v.data(), sizeof(T) * v.size());
    return p;
}

int emit(std::vector<Level>& lv, mgcfd_mesh** out)
{
    mgcfd_mesh* m = (mgcfd_mesh*)std::calloc(1, sizeof(mgcfd_mesh));
    if (!m) { g_err = "out of memory"; return -1; }
    m->levels = (int)lv.size();
    for (size_t l = 0; l < lv.size(); l++) {
        mgcfd_mesh_level& o = m->L[l];
        const Level& L = lv[l];
        o.n_nodes = L.n();
        o.n_edges = L.e();
        o.n_bnd = L.b();
        o.coords = dup(L.xyz);
        o.edge_to_node = dup(L.e2n);
        o.edge_weights = dup(L.ew);
        o.bnd_to_node = dup(L.b2n);
        o.bnd_group = dup(L.bg);
        o.bnd_weights = dup(L.bw);
        o.node_to_mg_node = L.n2c.empty() ? nullptr : dup(L.n2c);
    }
    *out = m;
    return 0;
}

}  // namespace

extern "C" {

const char* mgcfd_mesh_last_error(void) { return g_err.c_str(); }

int mgcfd_mesh_box(int n, int levels, uint64_t seed, mgcfd_mesh** out)
{
    g_err.clear();
    if (!out || n < 2 || levels < 1 || levels > MGCFD_MAX_LEVELS) { g_err = "mgcfd_mesh_box: bad arguments"; return -1; }
    if ((n >> (levels - 1)) < 2) { g_err = "mgcfd_mesh_box: too many levels for n"; return -1; }
    Rng r(seed);
    std::vector<Level> lv;
    const double h0 = 1.0 / (n - 1);
    for (int l = 0; l < levels; l++) lv.push_back(box_level(n >> l, l, h0, r));
    for (int l = 0; l + 1 < levels; l++) {
        int nf = n >> l, nc = n >> (l + 1);
        lv[l].n2c.resize((size_t)nf * nf * nf);
        for (int k = 0; k < nf; k++)
            for (int j = 0; j < nf; j++)
                for (int i = 0; i < nf; i++) {
                    int ci = std::min(i / 2, nc - 1), cj = std::min(j / 2, nc - 1), ck = std::min(k / 2, nc - 1);
                    lv[l].n2c[((size_t)k * nf + j) * nf + i] = (ck * nc + cj) * nc + ci;
                }
    }
    permute_all(lv, r);
    return emit(lv, out);
}

int mgcfd_mesh_knn(int n_nodes, int n_edges, int levels, const double* ratio, uint64_t seed, mgcfd_mesh** out)
{
    g_err.clear();
    if (!out || n_nodes < 64 || n_edges < n_nodes || levels < 1 || levels > MGCFD_MAX_LEVELS) {
        g_err = "mgcfd_mesh_knn: bad arguments";
        return -1;
    }
    Rng r(seed);
    std::vector<Level> lv;
    std::string err;
    lv.push_back(knn_level0(n_nodes, n_edges, r, err));
    if (!err.empty()) { g_err = err; return -1; }
    for (int l = 1; l < levels; l++) {
        const double dr[] = {1.0, 0.55, 0.37, 0.27};
        double q = ratio ? ratio[l] : (l < 4 ? dr[l] : 0.27 * std::pow(0.73, l - 3));
        int target = std::max(8, (int)std::lround(n_nodes * q));
        if (target >= lv.back().n()) { g_err = "mgcfd_mesh_knn: level ratios must decrease"; return -1; }
        Level c = agglomerate(lv.back(), target, r, err);
        if (!err.empty()) { g_err = err; return -1; }
        lv.push_back(std::move(c));
    }
    permute_all(lv, r);
    return emit(lv, out);
}

void mgcfd_mesh_free(mgcfd_mesh* m)
{
    if (!m) return;
    for (int l = 0; l < m->levels; l++) {
        mgcfd_mesh_level& o = m->L[l];
        std::free(o.coords);
        std::free(o.edge_to_node);
        std::free(o.edge_weights);
        std::free(o.bnd_to_node);
        std::free(o.bnd_group);
        std::free(o.bnd_weights);
        std::free(o.node_to_mg_node);
    }
    std::free(m);
}

}  // extern "C"
