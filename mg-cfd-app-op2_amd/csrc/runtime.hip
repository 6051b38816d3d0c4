// libmgcfd: OP2-compatible device runtime + op_par_loop entry points for the
// MG-CFD hot path on MI355X (gfx950).  See include/mgcfd_op2.h for the ABI.
//
// Sets, maps and dats live on the device for their whole life.  Node dats are
// stored SoA in an internal node order chosen for locality (reverse
// Cuthill-McKee over the edge maps); edges are ordered by their lower
// internal endpoint and boundary elements by their node.  Host transfers
// (op_fetch_data_char / op_upload_data_char) convert to and from the
// reference's AoS layout and original element order.
//
// Every OP_INC into a node dat runs as a gather over an inverse-map plan
// whose per-target entry lists keep the reference's element order, so the
// summation order equals OP2's sequential backend.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <numeric>
#include <string>
#include <vector>

#include "mgcfd_op2.h"
#include "kernels.h"

namespace mgcfd {

static thread_local std::string g_err;

static int fail(const char* fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return -1;
}

#define HIPCHK(expr)                                                                   \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess) return fail("%s failed: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

static inline int round_up(int x, int m) { return ((x + m - 1) / m) * m; }
static inline unsigned nblk(size_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

// ---------------------------------------------------------------- state --

struct SetImpl {
    int ld = 64;                 // SoA leading dimension (size rounded to 64)
    std::vector<int> perm;       // internal -> original (empty: identity)
    std::vector<int> iperm;      // original -> internal (empty: identity)
    int to_int(int orig) const { return iperm.empty() ? orig : iperm[orig]; }
    int to_orig(int in) const { return perm.empty() ? in : perm[in]; }
};

// Inverse of a map restricted to the targets it touches: for each target
// (internal id, ascending) the source elements in original order.
struct InvPlan {
    bool built = false;
    int n_targets = 0;
    bool dense = false;          // every target present: target == index
    int max_deg = 0;
    int* d_targets = nullptr;    // null when dense
    int* d_ptr = nullptr;        // n_targets + 1
    int2* d_inc2 = nullptr;      // dim-2 maps: {other endpoint, src | col << 31}
    int* d_inc1 = nullptr;       // dim-1 maps: src
};

struct MapImpl {
    std::vector<int> h_orig;     // [from][dim], original ids, 0-based
    int* d_map = nullptr;        // [from internal][dim], internal ids
    InvPlan inv;
};

struct DatImpl {
    bool is_int = false;
    void* d = nullptr;           // SoA [dim][ld]
    std::vector<char> pending;   // AoS host data awaiting upload (before freeze)
    unsigned version = 1;
    double4* d_pack = nullptr;   // packed weights cache (dim-3 edge dats)
    unsigned pack_version = 0;
    double* d_tmp = nullptr;     // per-element temporary (cell volumes)
};

struct MgLevel {
    mgcfd_level_bind b;
};

struct Runtime {
    bool inited = false;
    int device = 0;
    hipStream_t stream = nullptr;
    int base_index = 0;
    bool frozen = false;
    std::deque<op_set_core> sets;
    std::deque<op_map_core> maps;
    std::deque<op_dat_core> dats;
    std::deque<std::string> names;
    Consts k{};
    int flux_variant = MGCFD_FLUX_GATHER_EXACT;
    // reduction scratch
    double* d_part = nullptr;
    int part_cap = 0;
    double* d_scal = nullptr;    // [0] min_dt, [1] sum
    int* d_iscal = nullptr;      // [0] count
    double* h_pin = nullptr;     // pinned host landing zone
    int* h_pini = nullptr;
    // device-resident multigrid
    std::vector<MgLevel> mg;
    int* d_status = nullptr;
    int* d_cycle = nullptr;
    int* d_bad = nullptr;
    double* d_rms = nullptr;
    int rms_cap = 0;
    double* d_mindt = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    hipGraph_t graph = nullptr;
};

static Runtime R;

static SetImpl* S(op_set s) { return static_cast<SetImpl*>(s->impl); }
static MapImpl* M(op_map m) { return static_cast<MapImpl*>(m->impl); }
static DatImpl* D(op_dat d) { return static_cast<DatImpl*>(d->impl); }

static const char* keep_name(const char* n)
{
    R.names.emplace_back(n ? n : "");
    return R.names.back().c_str();
}

static void default_consts()
{
    // Reference defaults const.h:32-33 (Mach 1.2, 0 deg) until op_decl_const.
    const double mach = 1.2;
    R.k.sc = (double)0.2f;
    R.k.ffv[0] = 1.4;
    const double c = sqrt(1.4 * 1.0 / 1.4);
    const double sp = mach * c;
    const double ux = sp * cos(0.0), uy = sp * sin(0.0), uz = 0.0;
    R.k.ffv[1] = 1.4 * ux; R.k.ffv[2] = 1.4 * uy; R.k.ffv[3] = 1.4 * uz;
    R.k.ffv[4] = 1.4 * (0.5 * (sp * sp)) + (1.0 / (1.4 - 1.0));
    const double p = 1.0;
    R.k.ffmx[0] = ux * R.k.ffv[1] + p; R.k.ffmx[1] = ux * R.k.ffv[2]; R.k.ffmx[2] = ux * R.k.ffv[3];
    R.k.ffmy[0] = R.k.ffmx[1]; R.k.ffmy[1] = uy * R.k.ffv[2] + p; R.k.ffmy[2] = uy * R.k.ffv[3];
    R.k.ffmz[0] = R.k.ffmx[2]; R.k.ffmz[1] = R.k.ffmy[2]; R.k.ffmz[2] = uz * R.k.ffv[3] + p;
    const double dep = R.k.ffv[4] + p;
    R.k.ffde[0] = ux * dep; R.k.ffde[1] = uy * dep; R.k.ffde[2] = uz * dep;
}

static int ensure_part(int nblocks)
{
    if (nblocks <= R.part_cap) return 0;
    if (R.d_part) HIPCHK(hipFree(R.d_part));
    R.part_cap = std::max(nblocks, 4096);
    HIPCHK(hipMalloc(&R.d_part, sizeof(double) * R.part_cap));
    return 0;
}

static int check_launch()
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail("kernel launch failed: %s", hipGetErrorString(e));
    return 0;
}

// ---------------------------------------------------------- renumbering --

// Reverse Cuthill-McKee ordering of a node set from all dim-2 maps into it.
static std::vector<int> rcm_order(int n, const std::vector<const MapImpl*>& emaps,
                                  const std::vector<int>& from_sizes)
{
    std::vector<int> deg(n, 0);
    for (size_t m = 0; m < emaps.size(); m++) {
        const auto& h = emaps[m]->h_orig;
        for (int e = 0; e < from_sizes[m]; e++) {
            deg[h[2 * (size_t)e]]++;
            deg[h[2 * (size_t)e + 1]]++;
        }
    }
    std::vector<int> ptr(n + 1, 0);
    for (int i = 0; i < n; i++) ptr[i + 1] = ptr[i] + deg[i];
    std::vector<int> adj(ptr[n]);
    std::vector<int> fill(ptr.begin(), ptr.end() - 1);
    for (size_t m = 0; m < emaps.size(); m++) {
        const auto& h = emaps[m]->h_orig;
        for (int e = 0; e < from_sizes[m]; e++) {
            int a = h[2 * (size_t)e], b = h[2 * (size_t)e + 1];
            adj[fill[a]++] = b;
            adj[fill[b]++] = a;
        }
    }
    std::vector<int> order;
    order.reserve(n);
    std::vector<char> seen(n, 0);
    std::vector<int> level_of(n, -1);
    auto bfs_far = [&](int s) {
        // returns the last node of a BFS from s (a peripheral candidate)
        std::vector<int> q{s};
        std::vector<int> mark;
        level_of[s] = 0;
        mark.push_back(s);
        int last = s;
        for (size_t h = 0; h < q.size(); h++) {
            int u = q[h];
            last = u;
            for (int k = ptr[u]; k < ptr[u + 1]; k++) {
                int v = adj[k];
                if (level_of[v] < 0) { level_of[v] = level_of[u] + 1; q.push_back(v); mark.push_back(v); }
            }
        }
        for (int v : mark) level_of[v] = -1;
        return last;
    };
    std::vector<int> byd(n);
    std::iota(byd.begin(), byd.end(), 0);
    std::stable_sort(byd.begin(), byd.end(), [&](int x, int y) { return deg[x] < deg[y]; });
    std::vector<int> nb;
    for (int s0 : byd) {
        if (seen[s0]) continue;
        int s = bfs_far(bfs_far(s0));
        size_t head = order.size();
        order.push_back(s);
        seen[s] = 1;
        for (size_t h = head; h < order.size(); h++) {
            int u = order[h];
            nb.clear();
            for (int k = ptr[u]; k < ptr[u + 1]; k++)
                if (!seen[adj[k]]) { seen[adj[k]] = 1; nb.push_back(adj[k]); }
            std::stable_sort(nb.begin(), nb.end(), [&](int x, int y) { return deg[x] < deg[y]; });
            order.insert(order.end(), nb.begin(), nb.end());
        }
    }
    std::reverse(order.begin(), order.end());
    return order;   // order[internal] = original
}

static int upload_dat(op_dat d);

static int build_map_device(op_map m)
{
    MapImpl* mi = M(m);
    SetImpl* fs = S(m->from);
    SetImpl* ts = S(m->to);
    const int n = m->from->size, dim = m->dim;
    std::vector<int> h((size_t)n * dim);
    for (int i = 0; i < n; i++) {
        int o = fs->to_orig(i);
        for (int c = 0; c < dim; c++) h[(size_t)i * dim + c] = ts->to_int(mi->h_orig[(size_t)o * dim + c]);
    }
    if (mi->d_map) HIPCHK(hipFree(mi->d_map));
    HIPCHK(hipMalloc(&mi->d_map, sizeof(int) * std::max<size_t>(1, h.size())));
    if (!h.empty()) HIPCHK(hipMemcpy(mi->d_map, h.data(), sizeof(int) * h.size(), hipMemcpyHostToDevice));
    return 0;
}

// Chooses every internal order, then uploads maps and pending dats.
static int freeze()
{
    if (R.frozen) return 0;
    // node-like sets: targets of dim-2 maps
    for (auto& s : R.sets) {
        std::vector<const MapImpl*> ems;
        std::vector<int> fsz;
        for (auto& m : R.maps)
            if (m.to == &s && m.dim == 2 && m.from != &s) { ems.push_back(M(&m)); fsz.push_back(m.from->size); }
        if (ems.empty() || s.size == 0) continue;
        SetImpl* si = S(&s);
        si->perm = rcm_order(s.size, ems, fsz);
        si->iperm.assign(s.size, 0);
        for (int i = 0; i < s.size; i++) si->iperm[si->perm[i]] = i;
    }
    // element sets mapped into a renumbered set: order by internal targets
    for (auto& s : R.sets) {
        SetImpl* si = S(&s);
        if (!si->perm.empty() || s.size == 0) continue;
        op_map key = nullptr;
        for (auto& m : R.maps)
            if (m.from == &s && !S(m.to)->perm.empty() && (key == nullptr || m.dim > key->dim)) key = &m;
        if (!key) continue;
        const MapImpl* km = M(key);
        const SetImpl* ts = S(key->to);
        const int dim = key->dim;
        std::vector<long long> k1(s.size);
        for (int e = 0; e < s.size; e++) {
            long long lo = ts->iperm[km->h_orig[(size_t)e * dim]];
            long long hi = lo;
            for (int c = 1; c < dim; c++) {
                long long v = ts->iperm[km->h_orig[(size_t)e * dim + c]];
                lo = std::min(lo, v);
                hi = std::max(hi, v);
            }
            k1[e] = lo * (long long)(key->to->size + 1) + hi;
        }
        si->perm.resize(s.size);
        std::iota(si->perm.begin(), si->perm.end(), 0);
        std::stable_sort(si->perm.begin(), si->perm.end(), [&](int x, int y) { return k1[x] < k1[y]; });
        si->iperm.assign(s.size, 0);
        for (int i = 0; i < s.size; i++) si->iperm[si->perm[i]] = i;
    }
    R.frozen = true;
    for (auto& m : R.maps)
        if (int rc = build_map_device(&m)) return rc;
    for (auto& d : R.dats)
        if (int rc = upload_dat(&d)) return rc;
    return 0;
}

// Uploads a dat's pending AoS host data (or zeros) into its SoA buffer.
static int upload_dat(op_dat d)
{
    DatImpl* di = D(d);
    SetImpl* si = S(d->set);
    const int n = d->set->size, dim = d->dim, ld = si->ld;
    const size_t esz = di->is_int ? sizeof(int) : sizeof(double);
    if (!di->d) {
        HIPCHK(hipMalloc(&di->d, esz * (size_t)dim * ld));
        HIPCHK(hipMemset(di->d, 0, esz * (size_t)dim * ld));
    }
    if (di->pending.empty()) return 0;
    std::vector<char> soa(esz * (size_t)dim * ld, 0);
    const size_t in_esz = (size_t)d->size;     // declared bytes per component
    for (int i = 0; i < n; i++) {
        const int o = si->to_orig(i);
        for (int c = 0; c < dim; c++) {
            const char* src = di->pending.data() + ((size_t)o * dim + c) * in_esz;
            std::memcpy(soa.data() + ((size_t)c * ld + i) * esz, src, esz);
        }
    }
    HIPCHK(hipMemcpy(di->d, soa.data(), soa.size(), hipMemcpyHostToDevice));
    di->pending.clear();
    di->pending.shrink_to_fit();
    di->version++;
    return 0;
}

// ---------------------------------------------------------------- plans --

static int build_inverse(op_map m)
{
    MapImpl* mi = M(m);
    InvPlan& P = mi->inv;
    if (P.built) return 0;
    if (int rc = freeze()) return rc;
    SetImpl* fs = S(m->from);
    SetImpl* ts = S(m->to);
    const int nf = m->from->size, nt = m->to->size, dim = m->dim;
    std::vector<int> cnt(nt, 0);
    for (size_t q = 0; q < (size_t)nf * dim; q++) cnt[ts->to_int(mi->h_orig[q])]++;
    std::vector<int> tg;
    for (int t = 0; t < nt; t++)
        if (cnt[t]) tg.push_back(t);
    P.dense = ((int)tg.size() == nt);
    P.n_targets = P.dense ? nt : (int)tg.size();
    std::vector<int> slot(nt, -1);
    for (int i = 0; i < (int)tg.size(); i++) slot[tg[i]] = i;
    std::vector<int> ptr(P.n_targets + 1, 0);
    for (int t = 0; t < nt; t++)
        if (cnt[t]) ptr[(P.dense ? t : slot[t]) + 1] = cnt[t];
    for (int i = 0; i < P.n_targets; i++) {
        P.max_deg = std::max(P.max_deg, ptr[i + 1]);
        ptr[i + 1] += ptr[i];
    }
    std::vector<int> fill(ptr.begin(), ptr.end() - 1);
    std::vector<int2> inc2(dim == 2 ? ptr[P.n_targets] : 0);
    std::vector<int> inc1(dim == 2 ? 0 : ptr[P.n_targets]);
    // original element order, column order within an element: OP2 seq order
    for (int e = 0; e < nf; e++) {
        const int ein = fs->to_int(e);
        for (int c = 0; c < dim; c++) {
            const int t = ts->to_int(mi->h_orig[(size_t)e * dim + c]);
            const int s = P.dense ? t : slot[t];
            const int at = fill[s]++;
            if (dim == 2) {
                const int other = ts->to_int(mi->h_orig[(size_t)e * dim + (1 - c)]);
                inc2[at] = make_int2(other, (int)((unsigned)ein | ((unsigned)c << 31)));
            } else {
                inc1[at] = ein;
            }
        }
    }
    if (!P.dense) {
        HIPCHK(hipMalloc(&P.d_targets, sizeof(int) * std::max<size_t>(1, tg.size())));
        if (!tg.empty()) HIPCHK(hipMemcpy(P.d_targets, tg.data(), sizeof(int) * tg.size(), hipMemcpyHostToDevice));
    }
    HIPCHK(hipMalloc(&P.d_ptr, sizeof(int) * ptr.size()));
    HIPCHK(hipMemcpy(P.d_ptr, ptr.data(), sizeof(int) * ptr.size(), hipMemcpyHostToDevice));
    if (dim == 2) {
        HIPCHK(hipMalloc(&P.d_inc2, sizeof(int2) * std::max<size_t>(1, inc2.size())));
        if (!inc2.empty()) HIPCHK(hipMemcpy(P.d_inc2, inc2.data(), sizeof(int2) * inc2.size(), hipMemcpyHostToDevice));
    } else {
        HIPCHK(hipMalloc(&P.d_inc1, sizeof(int) * std::max<size_t>(1, inc1.size())));
        if (!inc1.empty()) HIPCHK(hipMemcpy(P.d_inc1, inc1.data(), sizeof(int) * inc1.size(), hipMemcpyHostToDevice));
    }
    P.built = true;
    return 0;
}

static int packed_weights(op_dat w, const double4** out)
{
    DatImpl* di = D(w);
    const int n = w->set->size;
    if (!di->d_pack) HIPCHK(hipMalloc(&di->d_pack, sizeof(double4) * std::max(1, n)));
    if (di->pack_version != di->version) {
        if (n > 0) {
            k_pack_weights<<<nblk(n), kBlock, 0, R.stream>>>(n, (const double*)di->d, S(w->set)->ld, di->d_pack);
            if (int rc = check_launch()) return rc;
        }
        di->pack_version = di->version;
    }
    *out = di->d_pack;
    return 0;
}

// ------------------------------------------------------------ arg checks --

static double* dptr(op_dat d) { return static_cast<double*>(D(d)->d); }
static int* iptr(op_dat d) { return static_cast<int*>(D(d)->d); }
static int ldof(op_dat d) { return S(d->set)->ld; }
static void touch(op_dat d) { D(d)->version++; }

static int need(bool ok, const char* loop, const char* what)
{
    return ok ? 0 : fail("op_par_loop %s: argument mismatch (%s)", loop, what);
}

static int chk_ready(const char* loop)
{
    if (!R.inited) return fail("op_par_loop %s: op_init has not been called", loop);
    return freeze();
}

static bool is_dbl(op_dat d) { return d && !D(d)->is_int; }

static int direct(const op_arg& a, op_set set, int dim, bool want_int, const char* loop, const char* what)
{
    bool ok = a.argtype == OP_ARG_DAT && a.dat && a.map == OP_ID && a.dat->set == set &&
              a.dat->dim == dim && D(a.dat)->is_int == want_int;
    return need(ok, loop, what);
}

static int indirect(const op_arg& a, op_set set, int idx, int mdim, int dim, const char* loop,
                    const char* what)
{
    bool ok = a.argtype == OP_ARG_DAT && a.dat && a.map && a.map->from == set && a.map->dim == mdim &&
              a.idx == idx && a.dat->set == a.map->to && a.dat->dim == dim && is_dbl(a.dat);
    return need(ok, loop, what);
}

static int gbl(const op_arg& a, const char* type, op_access acc, const char* loop, const char* what)
{
    bool ok = a.argtype == OP_ARG_GBL && a.data && a.acc == acc && a.dim == 1 && a.type &&
              std::strcmp(a.type, type) == 0;
    return need(ok, loop, what);
}

#define TRY(x) do { if (int rc_ = (x)) return rc_; } while (0)

// Host readback of a device scalar, synchronising the stream.
static int read_scalar(const double* d_src, double* out)
{
    HIPCHK(hipMemcpyAsync(R.h_pin, d_src, sizeof(double), hipMemcpyDeviceToHost, R.stream));
    HIPCHK(hipStreamSynchronize(R.stream));
    *out = *R.h_pin;
    return 0;
}

static int read_iscalar(const int* d_src, int* out)
{
    HIPCHK(hipMemcpyAsync(R.h_pini, d_src, sizeof(int), hipMemcpyDeviceToHost, R.stream));
    HIPCHK(hipStreamSynchronize(R.stream));
    *out = *R.h_pini;
    return 0;
}

// ------------------------------------------------------- flux launchers --

static int launch_flux(int variant, op_map map, op_dat var, op_dat w, op_dat flux, const int* status)
{
    const double4* wp = nullptr;
    TRY(packed_weights(w, &wp));
    const int ld = ldof(var);
    if (variant == MGCFD_FLUX_ATOMIC) {
        const int ne = map->from->size;
        if (ne == 0) return 0;
        k_flux_atomic<<<nblk(ne), kBlock, 0, R.stream>>>(ne, (const int2*)M(map)->d_map, wp, dptr(var),
                                                         dptr(flux), ld, R.k, status);
        return check_launch();
    }
    TRY(build_inverse(map));
    const InvPlan& P = M(map)->inv;
    if (P.n_targets == 0) return 0;
    if (variant == MGCFD_FLUX_GATHER_EXACT)
        k_flux_gather<true><<<nblk(P.n_targets), kBlock, 0, R.stream>>>(
            P.n_targets, P.d_targets, P.d_ptr, P.d_inc2, wp, dptr(var), dptr(flux), ld, R.k, status);
    else
        k_flux_gather<false><<<nblk(P.n_targets), kBlock, 0, R.stream>>>(
            P.n_targets, P.d_targets, P.d_ptr, P.d_inc2, wp, dptr(var), dptr(flux), ld, R.k, status);
    return check_launch();
}

static int launch_bnd(op_map b2n, op_dat group, op_dat bw, op_dat var, op_dat flux, const int* status)
{
    TRY(build_inverse(b2n));
    const InvPlan& P = M(b2n)->inv;
    if (P.n_targets == 0) return 0;
    // a dense plan has no target list; materialise identity targets lazily
    if (P.dense && !P.d_targets) {
        InvPlan& Pm = M(b2n)->inv;
        std::vector<int> id(Pm.n_targets);
        std::iota(id.begin(), id.end(), 0);
        HIPCHK(hipMalloc(&Pm.d_targets, sizeof(int) * id.size()));
        HIPCHK(hipMemcpy(Pm.d_targets, id.data(), sizeof(int) * id.size(), hipMemcpyHostToDevice));
    }
    k_bnd_flux<<<nblk(P.n_targets), kBlock, 0, R.stream>>>(P.n_targets, P.d_targets, P.d_ptr, P.d_inc1,
                                                           iptr(group), dptr(bw), ldof(bw), dptr(var),
                                                           dptr(flux), ldof(var), R.k, status);
    return check_launch();
}

}  // namespace mgcfd

using namespace mgcfd;

// ================================================================ C ABI ===

extern "C" {

const char* mgcfd_last_error(void) { return g_err.c_str(); }

int op_init(int argc, char** argv, int diags)
{
    (void)diags;
    if (R.inited) return 0;
    for (int i = 0; i < argc; i++)
        if (argv && argv[i] && std::strncmp(argv[i], "OP_MAPS_BASE_INDEX=", 19) == 0)
            R.base_index = std::atoi(argv[i] + 19);
    const char* dev = std::getenv("MGCFD_DEVICE");
    R.device = dev ? std::atoi(dev) : 0;
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (ndev <= 0) return fail("op_init: no HIP device visible");
    if (R.device >= ndev) return fail("op_init: device %d not present (%d visible)", R.device, ndev);
    HIPCHK(hipSetDevice(R.device));
    HIPCHK(hipStreamCreateWithFlags(&R.stream, hipStreamNonBlocking));
    HIPCHK(hipMalloc(&R.d_scal, 4 * sizeof(double)));
    HIPCHK(hipMalloc(&R.d_iscal, 4 * sizeof(int)));
    HIPCHK(hipHostMalloc(&R.h_pin, 8 * sizeof(double), hipHostMallocDefault));
    HIPCHK(hipHostMalloc(&R.h_pini, 8 * sizeof(int), hipHostMallocDefault));
    HIPCHK(hipMalloc(&R.d_status, sizeof(int)));
    HIPCHK(hipMalloc(&R.d_cycle, sizeof(int)));
    HIPCHK(hipMalloc(&R.d_bad, sizeof(int)));
    HIPCHK(hipMalloc(&R.d_mindt, sizeof(double)));
    TRY(ensure_part(4096));
    default_consts();
    R.inited = true;
    return 0;
}

void op_exit(void)
{
    if (!R.inited) return;
    hipStreamSynchronize(R.stream);
    if (R.graph_exec) hipGraphExecDestroy(R.graph_exec);
    if (R.graph) hipGraphDestroy(R.graph);
    R.graph_exec = nullptr;
    R.graph = nullptr;
    for (auto& d : R.dats) {
        DatImpl* di = D(&d);
        if (di->d) hipFree(di->d);
        if (di->d_pack) hipFree(di->d_pack);
        if (di->d_tmp) hipFree(di->d_tmp);
        delete di;
    }
    for (auto& m : R.maps) {
        MapImpl* mi = M(&m);
        if (mi->d_map) hipFree(mi->d_map);
        InvPlan& P = mi->inv;
        if (P.d_targets) hipFree(P.d_targets);
        if (P.d_ptr) hipFree(P.d_ptr);
        if (P.d_inc2) hipFree(P.d_inc2);
        if (P.d_inc1) hipFree(P.d_inc1);
        delete mi;
    }
    for (auto& s : R.sets) delete S(&s);
    R.dats.clear();
    R.maps.clear();
    R.sets.clear();
    R.names.clear();
    R.mg.clear();
    if (R.d_part) hipFree(R.d_part);
    if (R.d_rms) hipFree(R.d_rms);
    hipFree(R.d_scal);
    hipFree(R.d_iscal);
    hipFree(R.d_status);
    hipFree(R.d_cycle);
    hipFree(R.d_bad);
    hipFree(R.d_mindt);
    hipHostFree(R.h_pin);
    hipHostFree(R.h_pini);
    hipStreamDestroy(R.stream);
    R = Runtime();
}

op_set op_decl_set(int size, const char* name)
{
    if (size < 0) { fail("op_decl_set %s: negative size", name ? name : ""); return nullptr; }
    if (R.frozen) { fail("op_decl_set %s: sets must be declared before the first loop", name ? name : ""); return nullptr; }
    R.sets.push_back(op_set_core{});
    op_set s = &R.sets.back();
    s->index = (int)R.sets.size() - 1;
    s->size = size;
    s->name = keep_name(name);
    SetImpl* si = new SetImpl();
    si->ld = std::max(64, round_up(size, 64));
    s->impl = si;
    return s;
}

op_map op_decl_map(op_set from, op_set to, int dim, const int* imap, const char* name)
{
    if (!from || !to || dim < 1 || (!imap && from->size > 0)) { fail("op_decl_map %s: bad arguments", name ? name : ""); return nullptr; }
    if (R.frozen) { fail("op_decl_map %s: maps must be declared before the first loop", name ? name : ""); return nullptr; }
    MapImpl* mi = new MapImpl();
    mi->h_orig.resize((size_t)from->size * dim);
    for (size_t q = 0; q < mi->h_orig.size(); q++) {
        int v = imap[q] - R.base_index;
        if (v < 0 || v >= to->size) {
            delete mi;
            fail("op_decl_map %s: entry %zu = %d out of range [0,%d) (base index %d)", name ? name : "", q,
                 imap[q], to->size, R.base_index);
            return nullptr;
        }
        mi->h_orig[q] = v;
    }
    if (dim == 2)
        for (int e = 0; e < from->size; e++)
            if (mi->h_orig[2 * (size_t)e] == mi->h_orig[2 * (size_t)e + 1]) {
                delete mi;
                fail("op_decl_map %s: element %d maps both columns to node %d (self-loop)", name ? name : "",
                     e, mi->h_orig[2 * (size_t)e]);
                return nullptr;
            }
    R.maps.push_back(op_map_core{});
    op_map m = &R.maps.back();
    m->index = (int)R.maps.size() - 1;
    m->from = from;
    m->to = to;
    m->dim = dim;
    m->name = keep_name(name);
    m->impl = mi;
    return m;
}

static op_dat decl_dat(op_set set, int dim, const char* type, int size, const char* data, const char* name)
{
    if (!set || dim < 1 || !type) { fail("op_decl_dat %s: bad arguments", name ? name : ""); return nullptr; }
    bool is_int = std::strcmp(type, "int") == 0;
    bool is_dbl = std::strcmp(type, "double") == 0;
    if (!is_int && !is_dbl) { fail("op_decl_dat %s: unsupported type '%s'", name ? name : "", type); return nullptr; }
    if (is_dbl && size != (int)sizeof(double)) { fail("op_decl_dat %s: double element size %d", name ? name : "", size); return nullptr; }
    if (is_int && size < (int)sizeof(int)) { fail("op_decl_dat %s: int element size %d", name ? name : "", size); return nullptr; }
    DatImpl* di = new DatImpl();
    di->is_int = is_int;
    if (data) {
        // 'int' dats declared with 8-byte elements (euler3d.cpp:405) are read
        // through their first 4 bytes, as the kernels see them.
        const size_t in_esz = is_int ? sizeof(int) : sizeof(double);
        di->pending.resize((size_t)set->size * dim * in_esz);
        for (size_t q = 0; q < (size_t)set->size * dim; q++)
            std::memcpy(di->pending.data() + q * in_esz, data + q * (size_t)size, in_esz);
    }
    R.dats.push_back(op_dat_core{});
    op_dat d = &R.dats.back();
    d->index = (int)R.dats.size() - 1;
    d->set = set;
    d->dim = dim;
    d->size = is_int ? (int)sizeof(int) : size;
    d->type = keep_name(type);
    d->name = keep_name(name);
    d->impl = di;
    if (R.frozen && upload_dat(d)) return nullptr;
    return d;
}

op_dat op_decl_dat_char(op_set set, int dim, const char* type, int size, const char* data, const char* name)
{
    if (!data && set && set->size > 0) { fail("op_decl_dat %s: null data", name ? name : ""); return nullptr; }
    return decl_dat(set, dim, type, size, data, name);
}

op_dat op_decl_dat_temp_char(op_set set, int dim, const char* type, int size, const char* name)
{
    return decl_dat(set, dim, type, size, nullptr, name);
}

int op_decl_const_char(int dim, const char* type, int size, const char* dat, const char* name)
{
    (void)size;
    if (!name || !dat) return fail("op_decl_const: null name or data");
    auto cp = [&](double* dst, int n) -> int {
        if (dim != n || std::strcmp(type, "double") != 0) return fail("op_decl_const %s: expected %d doubles", name, n);
        std::memcpy(dst, dat, sizeof(double) * n);
        return 0;
    };
    if (!std::strcmp(name, "smoothing_coefficient")) return cp(&R.k.sc, 1);
    if (!std::strcmp(name, "ff_variable")) return cp(R.k.ffv, 5);
    if (!std::strcmp(name, "ff_flux_contribution_momentum_x")) return cp(R.k.ffmx, 3);
    if (!std::strcmp(name, "ff_flux_contribution_momentum_y")) return cp(R.k.ffmy, 3);
    if (!std::strcmp(name, "ff_flux_contribution_momentum_z")) return cp(R.k.ffmz, 3);
    if (!std::strcmp(name, "ff_flux_contribution_density_energy")) return cp(R.k.ffde, 3);
    if (!std::strcmp(name, "mesh_name")) return 0;     // not read by any kernel
    return fail("op_decl_const: unknown constant '%s'", name);
}

op_arg op_arg_dat(op_dat dat, int idx, op_map map, int dim, const char* type, op_access acc)
{
    op_arg a{};
    a.argtype = OP_ARG_DAT;
    a.dat = dat;
    a.map = map;
    a.idx = map ? idx : -1;
    a.dim = dim;
    a.size = dat ? dat->size : 0;
    a.type = type;
    a.acc = acc;
    return a;
}

op_arg op_arg_gbl_char(char* data, int dim, const char* type, int size, op_access acc)
{
    op_arg a{};
    a.argtype = OP_ARG_GBL;
    a.data = data;
    a.dim = dim;
    a.size = size;
    a.type = type;
    a.acc = acc;
    a.idx = -1;
    return a;
}

int op_get_size(op_set set) { return set ? set->size : -1; }

int op_fetch_data_char(op_dat d, char* dest)
{
    if (!d || !dest) return fail("op_fetch_data: null argument");
    TRY(freeze());
    DatImpl* di = D(d);
    SetImpl* si = S(d->set);
    const int n = d->set->size, dim = d->dim, ld = si->ld;
    const size_t esz = di->is_int ? sizeof(int) : sizeof(double);
    std::vector<char> soa(esz * (size_t)dim * ld);
    HIPCHK(hipStreamSynchronize(R.stream));
    HIPCHK(hipMemcpy(soa.data(), di->d, soa.size(), hipMemcpyDeviceToHost));
    for (int i = 0; i < n; i++) {
        const int o = si->to_orig(i);
        for (int c = 0; c < dim; c++)
            std::memcpy(dest + ((size_t)o * dim + c) * esz, soa.data() + ((size_t)c * ld + i) * esz, esz);
    }
    return 0;
}

int op_upload_data_char(op_dat d, const char* src)
{
    if (!d || !src) return fail("op_upload_data: null argument");
    TRY(freeze());
    DatImpl* di = D(d);
    const size_t esz = di->is_int ? sizeof(int) : sizeof(double);
    di->pending.assign(src, src + (size_t)d->set->size * d->dim * esz);
    HIPCHK(hipStreamSynchronize(R.stream));
    return upload_dat(d);
}

int op_partition(const char* lib, const char* method, op_set set, op_map map, op_dat coords)
{
    (void)lib; (void)method; (void)set; (void)map; (void)coords;
    return freeze();
}

int op_renumber(op_map map)
{
    (void)map;
    return freeze();
}

void op_timers(double* cpu, double* wall)
{
    if (R.inited) hipStreamSynchronize(R.stream);
    const double t = std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
    if (cpu) *cpu = (double)std::clock() / CLOCKS_PER_SEC;
    if (wall) *wall = t;
}

int op_rank(int* rank)
{
    if (rank) *rank = 0;
    return 0;
}

int mgcfd_set_flux_variant(int v)
{
    if (v < 0 || v >= MGCFD_FLUX_NVARIANTS) return fail("mgcfd_set_flux_variant: unknown variant %d", v);
    R.flux_variant = v;
    return 0;
}

int mgcfd_get_flux_variant(void) { return R.flux_variant; }

int mgcfd_sync(void)
{
    if (!R.inited) return fail("mgcfd_sync: not initialised");
    HIPCHK(hipStreamSynchronize(R.stream));
    return 0;
}

int mgcfd_device_info(char* name, int name_len, int* n_cu)
{
    hipDeviceProp_t p;
    HIPCHK(hipGetDeviceProperties(&p, R.device));
    if (name && name_len > 0) {
        std::snprintf(name, name_len, "%s (%s)", p.name, p.gcnArchName);
    }
    if (n_cu) *n_cu = p.multiProcessorCount;
    return 0;
}

// -------------------------------------------------------- op_par_loops --

int op_par_loop_initialize_variables_kernel(const char* name, op_set set, op_arg a0)
{
    const char* L = "initialize_variables_kernel";
    (void)name;
    TRY(chk_ready(L));
    TRY(direct(a0, set, 5, false, L, "arg0 variables[5] OP_WRITE"));
    if (set->size) k_fill5<<<nblk(set->size), kBlock, 0, R.stream>>>(set->size, dptr(a0.dat), ldof(a0.dat), R.k, 1);
    touch(a0.dat);
    return check_launch();
}

int op_par_loop_zero_5d_array_kernel(const char* name, op_set set, op_arg a0)
{
    const char* L = "zero_5d_array_kernel";
    (void)name;
    TRY(chk_ready(L));
    TRY(direct(a0, set, 5, false, L, "arg0 array[5] OP_WRITE"));
    if (set->size) k_fill5<<<nblk(set->size), kBlock, 0, R.stream>>>(set->size, dptr(a0.dat), ldof(a0.dat), R.k, 0);
    touch(a0.dat);
    return check_launch();
}

int op_par_loop_zero_1d_array_kernel(const char* name, op_set set, op_arg a0)
{
    const char* L = "zero_1d_array_kernel";
    (void)name;
    TRY(chk_ready(L));
    TRY(direct(a0, set, 1, false, L, "arg0 array[1] OP_WRITE"));
    if (set->size) k_fill<<<nblk(set->size), kBlock, 0, R.stream>>>((size_t)set->size, dptr(a0.dat), 0.0);
    touch(a0.dat);
    return check_launch();
}

int op_par_loop_calculate_cell_volumes(const char* name, op_set set, op_arg a0, op_arg a1, op_arg a2,
                                       op_arg a3, op_arg a4)
{
    const char* L = "calculate_cell_volumes";
    (void)name;
    TRY(chk_ready(L));
    TRY(indirect(a0, set, 0, 2, 3, L, "arg0 coords via map[0]"));
    TRY(indirect(a1, set, 1, 2, 3, L, "arg1 coords via map[1]"));
    TRY(direct(a2, set, 3, false, L, "arg2 edge weights[3] OP_RW"));
    TRY(indirect(a3, set, 0, 2, 1, L, "arg3 volumes via map[0] OP_INC"));
    TRY(indirect(a4, set, 1, 2, 1, L, "arg4 volumes via map[1] OP_INC"));
    TRY(need(a0.dat == a1.dat && a0.map == a1.map && a3.dat == a4.dat && a3.map == a0.map &&
             a4.map == a0.map && a3.acc == OP_INC && a4.acc == OP_INC, L, "coords/volumes pairing"));
    const int ne = set->size;
    if (ne == 0) return 0;
    DatImpl* wi = D(a2.dat);
    if (!wi->d_tmp) HIPCHK(hipMalloc(&wi->d_tmp, sizeof(double) * ne));
    TRY(build_inverse(a0.map));
    k_cell_volumes_edge<<<nblk(ne), kBlock, 0, R.stream>>>(ne, (const int2*)M(a0.map)->d_map, dptr(a0.dat),
                                                           ldof(a0.dat), dptr(a2.dat), ldof(a2.dat), wi->d_tmp);
    TRY(check_launch());
    const InvPlan& P = M(a0.map)->inv;
    k_cell_volumes_node<<<nblk(P.n_targets), kBlock, 0, R.stream>>>(P.n_targets, P.d_targets, P.d_ptr,
                                                                    P.d_inc2, wi->d_tmp, dptr(a3.dat));
    touch(a2.dat);
    touch(a3.dat);
    return check_launch();
}

int op_par_loop_dampen_ewt(const char* name, op_set set, op_arg a0)
{
    const char* L = "dampen_ewt";
    (void)name;
    TRY(chk_ready(L));
    TRY(direct(a0, set, 3, false, L, "arg0 weights[3] OP_RW"));
    if (set->size) k_scale3<<<nblk(set->size), kBlock, 0, R.stream>>>(set->size, dptr(a0.dat), ldof(a0.dat), 1e-7);
    touch(a0.dat);
    return check_launch();
}

int op_par_loop_copy_double_kernel(const char* name, op_set set, op_arg a0, op_arg a1)
{
    const char* L = "copy_double_kernel";
    (void)name;
    TRY(chk_ready(L));
    TRY(direct(a0, set, 5, false, L, "arg0 variables[5] OP_READ"));
    TRY(direct(a1, set, 5, false, L, "arg1 old_variables[5] OP_WRITE"));
    if (set->size) k_copy5<<<nblk(set->size), kBlock, 0, R.stream>>>(set->size, dptr(a0.dat), dptr(a1.dat), ldof(a0.dat), nullptr);
    touch(a1.dat);
    return check_launch();
}

int op_par_loop_calculate_dt_kernel(const char* name, op_set set, op_arg a0, op_arg a1, op_arg a2)
{
    const char* L = "calculate_dt_kernel";
    (void)name;
    TRY(chk_ready(L));
    TRY(direct(a0, set, 5, false, L, "arg0 variables[5] OP_READ"));
    TRY(direct(a1, set, 1, false, L, "arg1 volumes[1] OP_READ"));
    TRY(direct(a2, set, 1, false, L, "arg2 dt[1] OP_WRITE"));
    if (set->size)
        k_calc_dt<<<nblk(set->size), kBlock, 0, R.stream>>>(set->size, dptr(a0.dat), dptr(a1.dat), dptr(a2.dat),
                                                            nullptr, ldof(a0.dat), nullptr, nullptr);
    touch(a2.dat);
    return check_launch();
}

int op_par_loop_get_min_dt_kernel(const char* name, op_set set, op_arg a0, op_arg a1)
{
    const char* L = "get_min_dt_kernel";
    (void)name;
    TRY(chk_ready(L));
    TRY(direct(a0, set, 1, false, L, "arg0 dt[1] OP_READ"));
    TRY(gbl(a1, "double", OP_MIN, L, "arg1 gbl double OP_MIN"));
    const int n = set->size;
    double init;
    std::memcpy(&init, a1.data, sizeof(double));
    if (n == 0) return 0;
    const int nb = (int)nblk(n);
    TRY(ensure_part(nb));
    k_min_part<<<nb, kBlock, 0, R.stream>>>(n, dptr(a0.dat), R.d_part);
    k_min_final<<<1, kBlock, 0, R.stream>>>(nb, R.d_part, init, R.d_scal, nullptr, nullptr, nullptr);
    TRY(check_launch());
    double m;
    TRY(read_scalar(R.d_scal, &m));
    std::memcpy(a1.data, &m, sizeof(double));
    return 0;
}

int op_par_loop_compute_step_factor_kernel(const char* name, op_set set, op_arg a0, op_arg a1, op_arg a2,
                                           op_arg a3)
{
    const char* L = "compute_step_factor_kernel";
    (void)name;
    TRY(chk_ready(L));
    TRY(direct(a0, set, 5, false, L, "arg0 variables[5] OP_READ"));
    TRY(direct(a1, set, 1, false, L, "arg1 volumes[1] OP_READ"));
    TRY(gbl(a2, "double", OP_READ, L, "arg2 gbl min_dt OP_READ"));
    TRY(direct(a3, set, 1, false, L, "arg3 step_factors[1] OP_WRITE"));
    double m;
    std::memcpy(&m, a2.data, sizeof(double));
    if (set->size)
        k_step_factor<<<nblk(set->size), kBlock, 0, R.stream>>>(set->size, dptr(a1.dat), nullptr, m, dptr(a3.dat), nullptr);
    touch(a3.dat);
    return check_launch();
}

int op_par_loop_compute_flux_edge_kernel(const char* name, op_set set, op_arg a0, op_arg a1, op_arg a2,
                                         op_arg a3, op_arg a4)
{
    const char* L = "compute_flux_edge_kernel";
    (void)name;
    TRY(chk_ready(L));
    TRY(indirect(a0, set, 0, 2, 5, L, "arg0 variables via map[0] OP_READ"));
    TRY(indirect(a1, set, 1, 2, 5, L, "arg1 variables via map[1] OP_READ"));
    TRY(direct(a2, set, 3, false, L, "arg2 edge weights[3] OP_READ"));
    TRY(indirect(a3, set, 0, 2, 5, L, "arg3 fluxes via map[0] OP_INC"));
    TRY(indirect(a4, set, 1, 2, 5, L, "arg4 fluxes via map[1] OP_INC"));
    TRY(need(a0.dat == a1.dat && a3.dat == a4.dat && a0.map == a1.map && a3.map == a0.map && a4.map == a0.map &&
             a3.acc == OP_INC && a4.acc == OP_INC && a0.dat != a3.dat, L, "variables/fluxes pairing"));
    TRY(launch_flux(R.flux_variant, a0.map, a0.dat, a2.dat, a3.dat, nullptr));
    touch(a3.dat);
    return 0;
}

int op_par_loop_compute_bnd_node_flux_kernel(const char* name, op_set set, op_arg a0, op_arg a1, op_arg a2,
                                             op_arg a3)
{
    const char* L = "compute_bnd_node_flux_kernel";
    (void)name;
    TRY(chk_ready(L));
    TRY(direct(a0, set, 1, true, L, "arg0 group[1] int OP_READ"));
    TRY(direct(a1, set, 3, false, L, "arg1 bnd weights[3] OP_READ"));
    TRY(indirect(a2, set, 0, 1, 5, L, "arg2 variables via bnd map OP_READ"));
    TRY(indirect(a3, set, 0, 1, 5, L, "arg3 fluxes via bnd map OP_INC"));
    TRY(need(a2.map == a3.map && a3.acc == OP_INC && a2.dat != a3.dat, L, "variables/fluxes pairing"));
    TRY(launch_bnd(a2.map, a0.dat, a1.dat, a2.dat, a3.dat, nullptr));
    touch(a3.dat);
    return 0;
}

int op_par_loop_time_step_kernel(const char* name, op_set set, op_arg a0, op_arg a1, op_arg a2, op_arg a3,
                                 op_arg a4)
{
    const char* L = "time_step_kernel";
    (void)name;
    TRY(chk_ready(L));
    TRY(gbl(a0, "int", OP_READ, L, "arg0 gbl rkCycle int OP_READ"));
    TRY(direct(a1, set, 1, false, L, "arg1 step_factors[1] OP_READ"));
    TRY(direct(a2, set, 5, false, L, "arg2 fluxes[5] OP_RW"));
    TRY(direct(a3, set, 5, false, L, "arg3 old_variables[5] OP_READ"));
    TRY(direct(a4, set, 5, false, L, "arg4 variables[5] OP_WRITE"));
    int rk;
    std::memcpy(&rk, a0.data, sizeof(int));
    if (set->size)
        k_time_step<<<nblk(set->size), kBlock, 0, R.stream>>>(set->size, rk, dptr(a1.dat), dptr(a2.dat), dptr(a3.dat),
                                                              dptr(a4.dat), ldof(a2.dat), nullptr);
    touch(a2.dat);
    touch(a4.dat);
    return check_launch();
}

int op_par_loop_unstructured_stream_kernel(const char* name, op_set set, op_arg a0, op_arg a1, op_arg a2,
                                           op_arg a3, op_arg a4)
{
    const char* L = "unstructured_stream_kernel";
    (void)name;
    TRY(chk_ready(L));
    TRY(indirect(a0, set, 0, 2, 5, L, "arg0 variables via map[0] OP_READ"));
    TRY(indirect(a1, set, 1, 2, 5, L, "arg1 variables via map[1] OP_READ"));
    TRY(direct(a2, set, 3, false, L, "arg2 edge weights[3] OP_READ"));
    TRY(indirect(a3, set, 0, 2, 5, L, "arg3 dummy fluxes via map[0] OP_INC"));
    TRY(indirect(a4, set, 1, 2, 5, L, "arg4 dummy fluxes via map[1] OP_INC"));
    TRY(need(a0.dat == a1.dat && a3.dat == a4.dat && a0.map == a1.map && a3.map == a0.map && a4.map == a0.map &&
             a0.dat != a3.dat, L, "variables/fluxes pairing"));
    const double4* wp = nullptr;
    TRY(packed_weights(a2.dat, &wp));
    TRY(build_inverse(a0.map));
    const InvPlan& P = M(a0.map)->inv;
    if (P.n_targets)
        k_stream_gather<<<nblk(P.n_targets), kBlock, 0, R.stream>>>(P.n_targets, P.d_targets, P.d_ptr, P.d_inc2, wp,
                                                                    dptr(a0.dat), dptr(a3.dat), ldof(a0.dat));
    touch(a3.dat);
    return check_launch();
}

int op_par_loop_residual_kernel(const char* name, op_set set, op_arg a0, op_arg a1, op_arg a2)
{
    const char* L = "residual_kernel";
    (void)name;
    TRY(chk_ready(L));
    TRY(direct(a0, set, 5, false, L, "arg0 old_variables[5] OP_READ"));
    TRY(direct(a1, set, 5, false, L, "arg1 variables[5] OP_READ"));
    TRY(direct(a2, set, 5, false, L, "arg2 residuals[5] OP_WRITE"));
    if (set->size)
        k_residual<<<nblk(set->size), kBlock, 0, R.stream>>>(set->size, dptr(a0.dat), dptr(a1.dat), dptr(a2.dat),
                                                             ldof(a0.dat), nullptr, nullptr, nullptr);
    touch(a2.dat);
    return check_launch();
}

int op_par_loop_calc_rms_kernel(const char* name, op_set set, op_arg a0, op_arg a1)
{
    const char* L = "calc_rms_kernel";
    (void)name;
    TRY(chk_ready(L));
    TRY(direct(a0, set, 5, false, L, "arg0 residuals[5] OP_READ"));
    TRY(gbl(a1, "double", OP_INC, L, "arg1 gbl double OP_INC"));
    double init;
    std::memcpy(&init, a1.data, sizeof(double));
    const int n = set->size;
    if (n == 0) return 0;
    const int nb = (int)nblk(n);
    TRY(ensure_part(nb));
    k_sumsq_part<<<nb, kBlock, 0, R.stream>>>(n, dptr(a0.dat), ldof(a0.dat), R.d_part);
    k_sum_final<<<1, kBlock, 0, R.stream>>>(nb, R.d_part, init, R.d_scal + 1, n, nullptr, nullptr, nullptr, nullptr);
    TRY(check_launch());
    double s;
    TRY(read_scalar(R.d_scal + 1, &s));
    std::memcpy(a1.data, &s, sizeof(double));
    return 0;
}

static int count_loop(int n, op_dat d, int* out, int which)
{
    k_fill_i<<<1, kBlock, 0, R.stream>>>(1, R.d_iscal, 0);
    if (n) {
        if (which == 0) k_count_bad<<<nblk(n), kBlock, 0, R.stream>>>(n, dptr(d), ldof(d), R.d_iscal);
        else k_count_pos<<<nblk(n), kBlock, 0, R.stream>>>(n, dptr(d), ldof(d), R.d_iscal);
        TRY(check_launch());
    }
    return read_iscalar(R.d_iscal, out);
}

int op_par_loop_count_bad_vals(const char* name, op_set set, op_arg a0, op_arg a1)
{
    const char* L = "count_bad_vals";
    (void)name;
    TRY(chk_ready(L));
    TRY(direct(a0, set, 5, false, L, "arg0 values[5] OP_READ"));
    TRY(gbl(a1, "int", OP_INC, L, "arg1 gbl int OP_INC"));
    int c;
    TRY(count_loop(set->size, a0.dat, &c, 0));
    int init;
    std::memcpy(&init, a1.data, sizeof(int));
    init += c;
    std::memcpy(a1.data, &init, sizeof(int));
    return 0;
}

int op_par_loop_up_pre_kernel(const char* name, op_set set, op_arg a0, op_arg a1)
{
    const char* L = "up_pre_kernel";
    (void)name;
    TRY(chk_ready(L));
    TRY(indirect(a0, set, 0, 1, 5, L, "arg0 coarse variables via mg map OP_WRITE"));
    bool ok = a1.argtype == OP_ARG_DAT && a1.dat && a1.map == a0.map && a1.idx == 0 && a1.dat->dim == 1 &&
              D(a1.dat)->is_int && a1.dat->set == a0.map->to;
    TRY(need(ok, L, "arg1 up_scratch int via mg map OP_WRITE"));
    if (set->size)
        k_up_pre<<<nblk(set->size), kBlock, 0, R.stream>>>(set->size, M(a0.map)->d_map, dptr(a0.dat), ldof(a0.dat),
                                                           iptr(a1.dat), nullptr);
    touch(a0.dat);
    touch(a1.dat);
    return check_launch();
}

int op_par_loop_up_kernel(const char* name, op_set set, op_arg a0, op_arg a1, op_arg a2)
{
    const char* L = "up_kernel";
    (void)name;
    TRY(chk_ready(L));
    TRY(direct(a0, set, 5, false, L, "arg0 fine variables[5] OP_READ"));
    TRY(indirect(a1, set, 0, 1, 5, L, "arg1 coarse variables via mg map OP_INC"));
    bool ok = a2.argtype == OP_ARG_DAT && a2.dat && a2.map == a1.map && a2.idx == 0 && a2.dat->dim == 1 &&
              D(a2.dat)->is_int && a2.dat->set == a1.map->to && a1.acc == OP_INC && a2.acc == OP_INC;
    TRY(need(ok, L, "arg2 up_scratch int via mg map OP_INC"));
    TRY(build_inverse(a1.map));
    const InvPlan& P = M(a1.map)->inv;
    if (P.n_targets)
        k_up_gather<<<nblk(P.n_targets), kBlock, 0, R.stream>>>(P.n_targets, P.d_targets, P.d_ptr, P.d_inc1,
                                                                dptr(a0.dat), ldof(a0.dat), dptr(a1.dat),
                                                                ldof(a1.dat), iptr(a2.dat), 0, nullptr);
    touch(a1.dat);
    touch(a2.dat);
    return check_launch();
}

int op_par_loop_up_post_kernel(const char* name, op_set set, op_arg a0, op_arg a1)
{
    const char* L = "up_post_kernel";
    (void)name;
    TRY(chk_ready(L));
    TRY(direct(a0, set, 5, false, L, "arg0 variables[5] OP_RW"));
    TRY(direct(a1, set, 1, true, L, "arg1 up_scratch int OP_READ"));
    if (set->size)
        k_up_post<<<nblk(set->size), kBlock, 0, R.stream>>>(set->size, dptr(a0.dat), ldof(a0.dat), iptr(a1.dat), nullptr);
    touch(a0.dat);
    return check_launch();
}

int op_par_loop_down_kernel(const char* name, op_set set, op_arg a0, op_arg a1, op_arg a2, op_arg a3, op_arg a4)
{
    const char* L = "down_kernel";
    (void)name;
    TRY(chk_ready(L));
    TRY(direct(a0, set, 5, false, L, "arg0 variables[5] OP_INC"));
    TRY(direct(a1, set, 5, false, L, "arg1 residuals[5] OP_READ"));
    TRY(direct(a2, set, 3, false, L, "arg2 coords[3] OP_READ"));
    TRY(indirect(a3, set, 0, 1, 5, L, "arg3 coarse residuals via mg map"));
    TRY(indirect(a4, set, 0, 1, 3, L, "arg4 coarse coords via mg map"));
    TRY(need(a3.map == a4.map, L, "same mg map"));
    if (set->size)
        k_down<<<nblk(set->size), kBlock, 0, R.stream>>>(set->size, M(a3.map)->d_map, dptr(a0.dat), dptr(a1.dat),
                                                         dptr(a2.dat), ldof(a0.dat), dptr(a3.dat), dptr(a4.dat),
                                                         ldof(a3.dat), nullptr);
    touch(a0.dat);
    return check_launch();
}

int op_par_loop_identify_differences(const char* name, op_set set, op_arg a0, op_arg a1, op_arg a2)
{
    const char* L = "identify_differences";
    (void)name;
    TRY(chk_ready(L));
    TRY(direct(a0, set, 5, false, L, "arg0 test values[5] OP_READ"));
    TRY(direct(a1, set, 5, false, L, "arg1 master values[5] OP_READ"));
    TRY(direct(a2, set, 5, false, L, "arg2 differences[5] OP_WRITE"));
    if (set->size)
        k_identify_differences<<<nblk(set->size), kBlock, 0, R.stream>>>(set->size, dptr(a0.dat), dptr(a1.dat),
                                                                         dptr(a2.dat), ldof(a0.dat), nullptr);
    touch(a2.dat);
    return check_launch();
}

int op_par_loop_count_non_zeros(const char* name, op_set set, op_arg a0, op_arg a1)
{
    const char* L = "count_non_zeros";
    (void)name;
    TRY(chk_ready(L));
    TRY(direct(a0, set, 5, false, L, "arg0 values[5] OP_READ"));
    TRY(gbl(a1, "int", OP_INC, L, "arg1 gbl int OP_INC"));
    int c;
    TRY(count_loop(set->size, a0.dat, &c, 1));
    int init;
    std::memcpy(&init, a1.data, sizeof(int));
    init += c;
    std::memcpy(a1.data, &init, sizeof(int));
    return 0;
}

}  // extern "C"

#include "mgrun.inc"
