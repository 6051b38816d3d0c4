// libmgcfd: OP2-compatible device runtime + op_par_loop entry points for the
// MG-CFD hot path on MI355X (gfx950).  See include/mgcfd_op2.h for the ABI.
//
// Sets, maps and dats live on the device for their whole life.  Node dats are
// stored SoA in an internal node order chosen for locality (recursive graph
// bisection into compact parts of kPart nodes, one part per flux workgroup);
// edges are ordered by their lower internal endpoint and boundary elements
// by their node.  Host transfers
// (op_fetch_data_char / op_upload_data_char) convert to and from the
// reference's AoS layout and original element order.
//
// Every OP_INC into a node dat runs as a gather over an inverse-map plan
// whose per-target entry lists keep the reference's element order, so the
// summation order equals OP2's sequential backend.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <atomic>
#include <thread>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <mutex>
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

static inline unsigned nblk(size_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

// ---------------------------------------------------------------- state --

struct SetImpl {
    std::vector<int> perm;       // internal -> original (empty: identity)
    std::vector<int> iperm;      // original -> internal (empty: identity)
    std::vector<int> seg;        // segment offsets (batched meshes), empty: one segment
    std::vector<int> parts;      // part boundaries of the node order (size parts + 1)
    std::vector<double> geo;     // node coordinates [n][3] (original order) when known: geometric order
    bool frozen = false;         // internal order chosen
    bool priv = false;           // declared by a batch builder, which orders and uploads it itself
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
    int* d_nptr = nullptr;       // dim-1 maps: CSR pointer over every to-set element (d_ptr itself when dense)
    // blocked-gather plan (dim-2 maps onto a partitioned node set)
    int bk_blocks = 0;           // 0 = not available
    int bk_entries = 0;          // jagged-diagonal entries (= incidences)
    size_t bk_lds = 0;           // dynamic LDS bytes per workgroup
    int2* d_bk_bnode = nullptr;  // {first node, count} per block
    int* d_bk_hptr = nullptr;    // blocks + 1
    int* d_bk_halo = nullptr;    // halo node ids per block
    int* d_bk_jbase = nullptr;   // first jagged entry of each block
    int* d_bk_pbase = nullptr;   // start of each block's diagonal offsets in d_bk_jpref
    int* d_bk_jpref = nullptr;   // per block: offset of diagonal k (k = 0..max degree)
    unsigned* d_bk_jslot = nullptr;   // per entry: LDS slot of the other endpoint | b-role << 31
    int* d_bk_eid = nullptr;     // per entry: internal edge id (-1 = padding)
    double* d_bk_w3 = nullptr;   // per entry: weight components w0 | w1 | w2 (SoA)
    int* d_bk_nptr = nullptr;    // CSR pointer over every node of the target set (a node's degree; d_ptr
                                 // counts targets only when some node has no incidence)
    const void* bk_wdat = nullptr;    // weight dat the packed weights came from
    unsigned bk_wver = 0;
    int bk_max_halo = 0;
    // part plan (edge-parallel flux over mesh parts, k_flux_part)
    int pt_parts = 0;            // 0 = not available
    long long pt_nloc = 0;       // local edges over all parts
    int pt_rec_cap = 0;          // max staged nodes (owned + halo) of one part
    int pt_rec_med = 0;          // median staged nodes of a part
    int pt_rs = 6;               // values per FP64 LDS record (6, or 5: see build_partplan)
    std::vector<PartInfo> h_pt_info;   // host copy of the parts
    unsigned pt_gen = 0;         // identifies the part plan built last (process-wide count)
    int pt_max_ne = 0;           // max local edges of one part
    int pt_max_col = 0;
    size_t pt_lds = 0;
    PartInfo* d_pt_info = nullptr;
    int* d_pt_halo = nullptr;
    unsigned* d_pt_code = nullptr;   // slot a | slot b << 10 | own a << 20 | own b << 21 | colour << 22
    int* d_pt_eid = nullptr;         // internal edge id per local edge
    double* d_pt_w = nullptr;        // 3 x pt_nloc weight components
    const void* pt_wdat = nullptr;
    unsigned pt_wver = 0;
    float* d_pt_w32 = nullptr;       // same in FP32 (mixed-precision variants), allocated on first use
    const void* pt_w32dat = nullptr;
    unsigned pt_w32ver = 0;
    size_t pt_lds32 = 0;
    int pt_psize = kPart;            // max nodes per part (autotuned)
    int pt_threads = 0;              // threads per part of the kernel with the fused stage (0: default; autotuned)
    int pt_persist = 0;              // fused-stage kernel: 0 one workgroup per part, else persistent workgroups
                                     // per CU (autotuned)
    int pt_rs_set = 0;               // FP64 record values of the edge-flux kernel: 0 by the rule below, or 5 / 6
    bool bk_built = false, pt_built = false;
    std::vector<int> h_ptr;          // host CSR over every target-set element, kept for re-planning (dim-2 maps)
    std::vector<int2> h_inc2;
};

struct MapImpl {
    bool on_device = false;
    bool composed = false;       // mgcfd_decl_map_compose: may repeat a target, plays no part in ordering
    std::vector<int> h_orig;     // [from][dim], original ids, 0-based
    int* d_map = nullptr;        // [from internal][dim], internal ids
    InvPlan inv;
};

struct DatImpl {
    bool is_int = false;
    void* d = nullptr;           // AoS [size][dim], internal element order
    std::vector<char> pending;   // AoS host data awaiting upload (before freeze)
    unsigned version = 1;
    double4* d_pack = nullptr;   // packed weights cache (dim-3 edge dats)
    unsigned pack_version = 0;
    double* d_tmp = nullptr;     // per-element temporary (cell volumes)
};

struct Runtime {
    bool inited = false;
    int device = 0;
    int n_cu = 256;              // compute units of the device
    hipStream_t lib_stream = nullptr;   // the library's own stream (op_init)
    int base_index = 0;
    bool frozen = false;         // at least one loop has run
    bool dirty = false;          // objects declared since the last freeze
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
};

static Runtime R;

// Serialises the runtime's registries (sets, maps, dats, names) and the
// ordering and upload of newly declared objects: served requests are built
// on worker threads while the request loop drives the running ones.
static std::recursive_mutex g_rt_mu;
using RtLock = std::lock_guard<std::recursive_mutex>;

// The stream the launchers enqueue on: the library's own, unless the calling
// thread has set another (StreamScope: a batch or served request runs its
// work on its context's stream, on whichever host thread builds or drives it).
static thread_local hipStream_t tl_stream = nullptr;
static inline hipStream_t cur_stream() { return tl_stream ? tl_stream : R.lib_stream; }

// Scratch of mgcfd_get_min_dt_noalloc, the entry point for callers that
// cannot allocate device memory: per-workgroup minima in [0, kNoAllocBlocks)
// and the result in [kNoAllocBlocks].  A file-scope device array, so every
// GPU the library runs on holds its own copy until the process ends; host
// code reaches it only through the address hipGetSymbolAddress gives on each
// device (scratch_dev), never through the array's name.
constexpr int kNoAllocBlocks = 65536;           // sets of up to 65536 * 256 = 16,777,216 elements
constexpr int kMaxDevices = 64;
static __device__ __align__(256) double scratch[kNoAllocBlocks + 1];
static double* scratch_dev[kMaxDevices];        // device address of 'scratch' on each device (op_init)

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

// Node order of a set from all dim-2 maps into it: recursive graph bisection
// into parts of exactly kPart nodes (the last part takes the remainder).
// Each split orders the segment by breadth-first distance from a
// pseudo-peripheral node and cuts it at a multiple of kPart, so every part
// is a compact, mostly connected patch of the mesh: the workgroup that owns
// a part finds most neighbours of its nodes inside the part.
// Morton (Z-order) key of each node of [lo, hi): coordinates quantised to
// 21 bits per axis over the range's bounding box, bits interleaved.
// n independent host-only tasks (no HIP calls), at most 16 threads.
template <typename F>
static void host_tasks(int n, F&& f)
{
    if (n <= 1) {
        if (n == 1) f(0);
        return;
    }
    std::atomic<int> next{0};
    std::vector<std::thread> pool;
    for (int t = 0; t < std::min(n, 16); t++)
        pool.emplace_back([&] {
            for (int i; (i = next.fetch_add(1)) < n;) f(i);
        });
    for (auto& th : pool) th.join();
}

// Bit b of a 21-bit value moved to bit 3b.
static inline uint64_t spread3(uint64_t x)
{
    x &= 0x1fffffull;
    x = (x | x << 32) & 0x1f00000000ffffull;
    x = (x | x << 16) & 0x1f0000ff0000ffull;
    x = (x | x << 8) & 0x100f00f00f00f00full;
    x = (x | x << 4) & 0x10c30c30c30c30c3ull;
    x = (x | x << 2) & 0x1249249249249249ull;
    return x;
}

// Z-order of the nodes [lo, hi): 21 bits per axis over the bounding box,
// axis 0 most significant within each bit triple; ties keep node order.
static void morton_order(const std::vector<double>& geo, int lo, int hi, std::vector<int>& order)
{
    double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
    for (int i = lo; i < hi; i++)
        for (int a = 0; a < 3; a++) {
            mn[a] = std::min(mn[a], geo[3 * (size_t)i + a]);
            mx[a] = std::max(mx[a], geo[3 * (size_t)i + a]);
        }
    std::vector<std::pair<uint64_t, int>> key(hi - lo);
    for (int i = lo; i < hi; i++) {
        uint64_t q[3];
        for (int a = 0; a < 3; a++) {
            const double span = mx[a] - mn[a];
            const double f = span > 0 ? (geo[3 * (size_t)i + a] - mn[a]) / span : 0.0;
            q[a] = (uint64_t)std::min(2097151.0, std::max(0.0, f * 2097151.0));
        }
        key[i - lo] = {(spread3(q[0]) << 2) | (spread3(q[1]) << 1) | spread3(q[2]), i};
    }
    std::sort(key.begin(), key.end());
    for (int i = lo; i < hi; i++) order[i] = key[i - lo].second;
}

static std::vector<int> partition_order(int n, const std::vector<const MapImpl*>& emaps,
                                        const std::vector<int>& from_sizes, const std::vector<int>& segs,
                                        std::vector<int>& parts, const std::vector<double>& geo)
{
    std::vector<int> ptr(n + 1, 0);
    for (size_t m = 0; m < emaps.size(); m++) {
        const auto& h = emaps[m]->h_orig;
        for (int e = 0; e < from_sizes[m]; e++) {
            ptr[h[2 * (size_t)e] + 1]++;
            ptr[h[2 * (size_t)e + 1] + 1]++;
        }
    }
    for (int i = 0; i < n; i++) ptr[i + 1] += ptr[i];
    // the adjacency lists serve the graph bisection only; the geometric
    // order needs the degrees alone
    std::vector<int> adj(geo.empty() ? ptr[n] : 0);
    if (geo.empty()) {
        std::vector<int> fill(ptr.begin(), ptr.end() - 1);
        for (size_t m = 0; m < emaps.size(); m++) {
            const auto& h = emaps[m]->h_orig;
            for (int e = 0; e < from_sizes[m]; e++) {
                int a = h[2 * (size_t)e], b = h[2 * (size_t)e + 1];
                adj[fill[a]++] = b;
                adj[fill[b]++] = a;
            }
        }
    }
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::vector<int> seg(n, 0), stamp(n, 0);
    int next_seg = 1, next_stamp = 1;
    std::vector<int> q, comp;
    // breadth-first order of the part of segment 'id' reachable from s,
    // restricted to nodes not yet placed (stamp != placed)
    auto bfs = [&](int s, int id, int mark, std::vector<int>& out) {
        out.clear();
        out.push_back(s);
        stamp[s] = mark;
        for (size_t h = 0; h < out.size(); h++) {
            int u = out[h];
            for (int k = ptr[u]; k < ptr[u + 1]; k++) {
                int v = adj[k];
                if (seg[v] == id && stamp[v] != mark && stamp[v] >= 0) { stamp[v] = mark; out.push_back(v); }
            }
        }
    };
    struct Range { int lo, hi, id; };
    // one independent bisection per segment (batched meshes never share a part)
    std::vector<int> so = segs.empty() ? std::vector<int>{0, n} : segs;
    std::vector<Range> stack;
    if (!geo.empty()) {
        // coordinates known (op_partition with a coordinate dat, batched and
        // served meshes): Z-order per segment, cut into kPart-node parts
        host_tasks((int)so.size() - 1, [&](int k) { morton_order(geo, so[k], so[k + 1], order); });
    } else {
        for (size_t k = so.size() - 1; k-- > 0;) {
            for (int i = so[k]; i < so[k + 1]; i++) seg[i] = next_seg;
            stack.push_back({so[k], so[k + 1], next_seg++});
        }
    }
    std::vector<int> buf;
    while (!stack.empty()) {
        Range r = stack.back();
        stack.pop_back();
        if (r.hi - r.lo <= kPart) continue;
        buf.clear();
        for (int i = r.lo; i < r.hi; i++) {
            int s0 = order[i];
            if (stamp[s0] < 0) continue;            // already placed in this pass
            bfs(s0, r.id, next_stamp++, comp);      // find the far end of the component
            int far = comp.back();
            for (int v : comp) stamp[v] = 0;
            bfs(far, r.id, next_stamp++, comp);
            for (int v : comp) { stamp[v] = -1; buf.push_back(v); }
        }
        for (int v : buf) stamp[v] = 0;
        std::copy(buf.begin(), buf.end(), order.begin() + r.lo);
        int half = (r.hi - r.lo) / 2;
        int mid = r.lo + ((half + kPart - 1) / kPart) * kPart;
        if (mid >= r.hi) mid = r.hi - kPart;
        int idl = next_seg++, idr = next_seg++;
        for (int i = r.lo; i < mid; i++) seg[order[i]] = idl;
        for (int i = mid; i < r.hi; i++) seg[order[i]] = idr;
        stack.push_back({mid, r.hi, idr});
        stack.push_back({r.lo, mid, idl});
    }
    // parts: kPart-node chunks of each segment; inside each part, nodes by
    // decreasing degree (the blocked flux kernel walks incidences in jagged
    // diagonals over a part's threads)
    parts.assign(1, 0);
    for (size_t k = 0; k + 1 < so.size(); k++)
        for (int lo = so[k]; lo < so[k + 1]; lo += kPart) parts.push_back(std::min(so[k + 1], lo + kPart));
    host_tasks((int)so.size() - 1, [&](int k) {
        for (int lo = so[k]; lo < so[k + 1]; lo += kPart) {
            const int hi = std::min(so[k + 1], lo + kPart);
            std::stable_sort(order.begin() + lo, order.begin() + hi, [&](int x, int y) {
                return ptr[x + 1] - ptr[x] > ptr[y + 1] - ptr[y];
            });
        }
    });
    return order;   // order[internal] = original
}

static int upload_dat(op_dat d);
static void mg_release_all();
static void mg_join_admissions();

// Node coordinates [n][3] of a set not yet ordered: its internal order will
// be geometric.
static void set_geometry(op_set s, const double* xyz)
{
    SetImpl* si = static_cast<SetImpl*>(s->impl);
    si->geo.assign(xyz, xyz + 3 * (size_t)s->size);
}

// Host set-up timing (MGCFD_TIMING=1 in the environment): one line per
// partition / plan build on stderr.
struct HostTimer {
    const char* what;
    const char* name;
    std::chrono::steady_clock::time_point t0;
    HostTimer(const char* w, const char* n) : what(w), name(n ? n : ""), t0(std::chrono::steady_clock::now()) {}
    ~HostTimer()
    {
        static const bool on = std::getenv("MGCFD_TIMING") != nullptr;
        if (!on) return;
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        std::fprintf(stderr, "mgcfd timing: %s %s %.1f ms\n", what, name, ms);
    }
};     // multigrid contexts and batches (mgrun.inc)

// Part boundaries of a node set's internal order (kPart-node chunks when the
// set was not partitioned).
static std::vector<int> set_parts(op_set s)
{
    const SetImpl* si = static_cast<const SetImpl*>(s->impl);
    if (!si->parts.empty()) return si->parts;
    std::vector<int> pb{0};
    for (int lo = 0; lo < s->size; lo += kPart) pb.push_back(std::min(s->size, lo + kPart));
    return pb;
}

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

// f(0) .. f(n - 1) on up to 16 host threads, in chunks of 8 (serially below
// 64 items).
template <typename F>
static void parallel_for(int n, F&& f)
{
    const int nth = std::min(16, (int)std::max(1u, std::thread::hardware_concurrency()));
    if (n < 64 || nth <= 1) {
        for (int i = 0; i < n; i++) f(i);
        return;
    }
    std::atomic<int> next{0};
    std::vector<std::thread> pool;
    for (int t = 0; t < nth; t++)
        pool.emplace_back([&] {
            for (;;) {
                const int i0 = next.fetch_add(8);
                if (i0 >= n) break;
                for (int i = i0; i < std::min(n, i0 + 8); i++) f(i);
            }
        });
    for (auto& th : pool) th.join();
}

// n independent host tasks (n small: the levels of a mesh, the sets to
// order), one per thread; returns the first failing task's code and keeps
// its message.  Tasks may launch on the current stream and allocate.
template <typename F>
static int run_tasks(int n, F&& f)
{
    if (n <= 1) return n == 1 ? f(0) : 0;
    std::vector<int> rc(n, 0);
    std::vector<std::string> msg(n);
    std::atomic<int> next{0};
    std::vector<std::thread> pool;
    const int nth = std::min(n, 16);
    const hipStream_t parent = tl_stream;
    for (int t = 0; t < nth; t++)
        pool.emplace_back([&] {
            // tasks may allocate, copy and launch: the caller's device and stream
            hipSetDevice(R.device);
            tl_stream = parent;
            for (int i; (i = next.fetch_add(1)) < n;) {
                rc[i] = f(i);
                if (rc[i]) msg[i] = g_err;       // the worker's own message
            }
        });
    for (auto& th : pool) th.join();
    for (int i = 0; i < n; i++)
        if (rc[i]) {
            g_err = msg[i];
            return rc[i];
        }
    return 0;
}

// Sets declared by this thread are private to it (SetImpl::priv) while set:
// a batch builder declares its objects under the registry lock, then orders
// and uploads them itself without the lock.
static thread_local bool tl_priv_decl = false;

static bool priv_set(op_set s) { return s && s->impl && S(s)->priv; }

// Orders the given sets that are not frozen yet, then uploads the given maps
// and dats that are not on the device.  The lists hold either every object
// of the registry that no builder keeps private (freeze, under the registry
// lock) or the objects one batch builder declared for itself (without the
// lock: no other thread touches them).
static int freeze_objects(const std::vector<op_set>& sets, const std::vector<op_map>& maps,
                          const std::vector<op_dat>& dats)
{
    // node-like sets: targets of dim-2 maps (each set on its own thread)
    struct NodeSet {
        op_set s;
        std::vector<const MapImpl*> ems;
        std::vector<int> fsz;
    };
    std::vector<NodeSet> ns;
    for (op_set sp : sets) {
        op_set_core& s = *sp;
        if (!s.impl || S(&s)->frozen) continue;
        NodeSet x{&s, {}, {}};
        for (op_map mp : maps) {
            op_map_core& m = *mp;
            if (m.impl && m.to == &s && m.dim == 2 && m.from != &s && !M(&m)->composed) {
                x.ems.push_back(M(&m));
                x.fsz.push_back(m.from->size);
            }
        }
        if (x.ems.empty() || s.size == 0) continue;
        ns.push_back(std::move(x));
    }
    if (int rc = run_tasks((int)ns.size(), [&](int k) -> int {
        op_set s = ns[k].s;
        SetImpl* si = S(s);
        {
            HostTimer tm("partition", s->name);
            si->perm = partition_order(s->size, ns[k].ems, ns[k].fsz, si->seg, si->parts, si->geo);
            si->geo.clear();
            si->geo.shrink_to_fit();
        }
        si->iperm.assign(s->size, -1);
        for (int i = 0; i < s->size; i++) si->iperm[si->perm[i]] = i;
        for (int i = 0; i < s->size; i++)
            if (si->iperm[i] < 0) return fail("internal error: node order of set %s is not a permutation", s->name);
        return 0;
    }))
        return rc;
    // element sets mapped into a renumbered set: order by internal targets
    std::vector<std::pair<op_set, op_map>> es;
    for (op_set sp : sets) {
        op_set_core& s = *sp;
        if (!s.impl) continue;
        SetImpl* si = S(&s);
        if (si->frozen || !si->perm.empty() || s.size == 0) continue;
        op_map key = nullptr;
        for (op_map mp : maps) {
            op_map_core& m = *mp;
            if (m.impl && !M(&m)->composed && m.from == &s && m.to->impl && !S(m.to)->perm.empty() &&
                (key == nullptr || m.dim > key->dim))
                key = &m;
        }
        if (key) es.emplace_back(&s, key);
    }
    if (int rc = run_tasks((int)es.size(), [&](int k) -> int {
        op_set_core& s = *es[k].first;
        op_map key = es[k].second;
        SetImpl* si = S(&s);
        HostTimer tm("order", s.name);
        const MapImpl* km = M(key);
        const SetImpl* ts = S(key->to);
        const int dim = key->dim, nt = key->to->size;
        // stable order by (lowest internal endpoint, highest): two counting
        // passes, least significant key first
        std::vector<int> lo(s.size), hi(s.size);
        for (int e = 0; e < s.size; e++) {
            int a = ts->iperm[km->h_orig[(size_t)e * dim]], b = a;
            for (int c = 1; c < dim; c++) {
                const int v = ts->iperm[km->h_orig[(size_t)e * dim + c]];
                a = std::min(a, v);
                b = std::max(b, v);
            }
            lo[e] = a;
            hi[e] = b;
        }
        // counting sort of the elements [e0, e1), whose keys lie in [k0, k1),
        // into out[e0, e1)
        auto counting = [&](const std::vector<int>& key1, const int* in, int* out, int e0, int e1, int k0, int k1) {
            std::vector<int> cnt(k1 - k0 + 1, 0);
            for (int q = e0; q < e1; q++) cnt[key1[in[q]] - k0 + 1]++;
            for (int i = 0; i < k1 - k0; i++) cnt[i + 1] += cnt[i];
            for (int q = e0; q < e1; q++) out[e0 + cnt[key1[in[q]] - k0]++] = in[q];
        };
        std::vector<int> id(s.size), tmp(s.size);
        std::iota(id.begin(), id.end(), 0);
        si->perm.resize(s.size);
        // Elements of a segmented target set (batched meshes) that come
        // segment by segment reference only their segment's nodes, so the
        // order is the segments' own orders one after another: sorted per
        // segment, in parallel.  Otherwise one sort over the whole set.
        const std::vector<int>& tseg = ts->seg;
        std::vector<int> eb;
        if (tseg.size() > 2 && s.size > 0) {
            eb.push_back(0);
            int k = 0;
            bool ok = true;
            for (int e = 0; e < s.size && ok; e++) {
                while (k + 1 < (int)tseg.size() - 1 && lo[e] >= tseg[k + 1]) {
                    k++;
                    eb.push_back(e);
                }
                ok = lo[e] >= tseg[k] && hi[e] < tseg[k + 1];
            }
            while ((int)eb.size() < (int)tseg.size()) eb.push_back(s.size);
            if (!ok) eb.clear();
        }
        if (!eb.empty()) {
            host_tasks((int)eb.size() - 1, [&](int k) {
                counting(hi, id.data(), tmp.data(), eb[k], eb[k + 1], tseg[k], tseg[k + 1]);
                counting(lo, tmp.data(), si->perm.data(), eb[k], eb[k + 1], tseg[k], tseg[k + 1]);
            });
        } else {
            counting(hi, id.data(), tmp.data(), 0, s.size, 0, nt);
            counting(lo, tmp.data(), si->perm.data(), 0, s.size, 0, nt);
        }
        si->iperm.assign(s.size, -1);
        for (int i = 0; i < s.size; i++) si->iperm[si->perm[i]] = i;
        for (int i = 0; i < s.size; i++)
            if (si->iperm[i] < 0) return fail("internal error: element order of set %s is not a permutation", s.name);
        return 0;
    }))
        return rc;
    for (op_set sp : sets)
        if (sp->impl) S(sp)->frozen = true;
    // maps and dats permuted into the internal order and uploaded, each on
    // a host thread of its own (HIP calls are thread-safe; each touches only
    // its own object)
    std::vector<op_map> ml;
    for (op_map m : maps)
        if (m->impl && !M(m)->on_device) ml.push_back(m);
    std::vector<op_dat> dl;
    for (op_dat d : dats)
        if (d->impl && (!D(d)->d || !D(d)->pending.empty())) dl.push_back(d);
    // maps and dats in one pool: the largest of each (a level's edge map and
    // edge weights) upload side by side
    const int nm = (int)ml.size();
    return run_tasks(nm + (int)dl.size(), [&](int k) -> int {
        if (k >= nm) return upload_dat(dl[k - nm]);
        if (int r = build_map_device(ml[k])) return r;
        M(ml[k])->on_device = true;
        return 0;
    });
}

// Chooses the internal order of every set declared since the last call, then
// uploads their maps and pending dats (objects a batch builder keeps private
// are left to it).  Sets, maps and dats may be declared at any time (batched
// and served solves add meshes after the first loop); an order, once
// chosen, never changes.
static int freeze()
{
    RtLock lk(g_rt_mu);
    if (R.frozen && !R.dirty) return 0;
    std::vector<op_set> sets;
    std::vector<op_map> maps;
    std::vector<op_dat> dats;
    for (auto& s : R.sets)
        if (s.impl && !S(&s)->priv) sets.push_back(&s);
    for (auto& m : R.maps)
        if (m.impl && !priv_set(m.from) && !priv_set(m.to)) maps.push_back(&m);
    for (auto& d : R.dats)
        if (d.impl && !priv_set(d.set)) dats.push_back(&d);
    if (int rc = freeze_objects(sets, maps, dats)) return rc;
    R.frozen = true;
    R.dirty = false;
    return 0;
}

// Uploads a dat's pending AoS host data (or zeros) into its device buffer
// (AoS, internal element order).
static int upload_dat(op_dat d)
{
    DatImpl* di = D(d);
    SetImpl* si = S(d->set);
    const int n = d->set->size, dim = d->dim;
    const size_t esz = di->is_int ? sizeof(int) : sizeof(double);
    const size_t bytes = esz * (size_t)dim * std::max(n, 1);
    if (!di->d) {
        HIPCHK(hipMalloc(&di->d, bytes));
        HIPCHK(hipMemset(di->d, 0, bytes));
    }
    if (di->pending.empty()) return 0;
    std::vector<char> aos(bytes, 0);
    auto permute = [&](auto* dst, const auto* src) {
        for (int i = 0; i < n; i++) {
            const size_t o = (size_t)si->to_orig(i) * dim, q = (size_t)i * dim;
            for (int v = 0; v < dim; v++) dst[q + v] = src[o + v];
        }
    };
    if (di->is_int)
        permute(reinterpret_cast<int*>(aos.data()), reinterpret_cast<const int*>(di->pending.data()));
    else
        permute(reinterpret_cast<double*>(aos.data()), reinterpret_cast<const double*>(di->pending.data()));
    HIPCHK(hipMemcpy(di->d, aos.data(), bytes, hipMemcpyHostToDevice));
    di->pending.clear();
    di->pending.shrink_to_fit();
    di->version++;
    return 0;
}

// ---------------------------------------------------------------- plans --

#define TRY(x) do { if (int rc_ = (x)) return rc_; } while (0)

// Blocked-gather plan.  Blocks are the kPart-node parts of the node order
// (halved recursively while a part's halo exceeds kHaloCap: the kernel
// stages one halo node per thread).  Per block: the sorted halo, and the
// incidences of its nodes in jagged-diagonal order -- diagonal k holds the
// k-th incidence (reference edge order) of every node with degree > k;
// nodes are sorted by decreasing degree inside a part, so diagonal k is a
// prefix of the block's threads.
constexpr int kHaloCap = kPart;

// Runs f(i) for i in [0, n) on up to 16 host threads (the plan builders'
// per-part work is independent); f must be thread-safe.

// Blocked-gather plan of one initial part (split while its halo exceeds
// kHaloCap); offsets are local to the part and fixed up on concatenation.
struct BlkOut {
    std::vector<int2> blocks;           // {first node, count}
    std::vector<int> nhalo, halo;       // per block halo count; halos concatenated
    std::vector<int> npref, jpref;      // per block diagonal-offset count; offsets concatenated
    std::vector<int> nent;              // per block entry count
    std::vector<unsigned> jslot;
    std::vector<int> eid;
    std::string err;
};

static void blocked_part(int lo0, int hi0, const std::vector<int>& ptr, const std::vector<int2>& inc2, BlkOut& out)
{
    std::vector<int2> work{make_int2(lo0, hi0)};
    std::vector<int> hb;
    while (!work.empty()) {             // LIFO: blocks come out in node order
        const int lo = work.back().x, hi = work.back().y;
        work.pop_back();
        hb.clear();
        for (int q = ptr[lo]; q < ptr[hi]; q++) {
            const int o = inc2[q].x;
            if (o < lo || o >= hi) hb.push_back(o);
        }
        std::sort(hb.begin(), hb.end());
        hb.erase(std::unique(hb.begin(), hb.end()), hb.end());
        if ((int)hb.size() > kHaloCap && hi - lo > 16) {
            const int mid = lo + (hi - lo) / 2;
            work.push_back(make_int2(mid, hi));
            work.push_back(make_int2(lo, mid));
            continue;
        }
        // degrees are non-increasing inside the block, so diagonal k covers
        // threads [0, cnt_k)
        const int maxdeg = ptr[lo + 1] - ptr[lo];
        for (int t = 1; t < hi - lo; t++)
            if (ptr[lo + t + 1] - ptr[lo + t] > ptr[lo + t] - ptr[lo + t - 1]) {
                out.err = "internal error: node degrees not sorted inside a part";
                return;
            }
        out.blocks.push_back(make_int2(lo, hi - lo));
        const size_t e0 = out.eid.size();
        int off = 0;
        for (int kd = 0; kd < maxdeg; kd++) {
            out.jpref.push_back(off);
            for (int t = 0; t < hi - lo; t++) {
                const int nd = lo + t;
                if (ptr[nd + 1] - ptr[nd] <= kd) break;
                const int2 r = inc2[ptr[nd] + kd];
                const int o = r.x;
                const unsigned slot = (o >= lo && o < hi)
                    ? (unsigned)(o - lo)
                    : (unsigned)(kPart + (std::lower_bound(hb.begin(), hb.end(), o) - hb.begin()));
                out.jslot.push_back(slot | ((unsigned)r.y & 0x80000000u));
                out.eid.push_back(r.y & 0x7fffffff);
                off++;
            }
        }
        out.jpref.push_back(off);
        out.npref.push_back(maxdeg + 1);
        out.nent.push_back((int)(out.eid.size() - e0));
        out.nhalo.push_back((int)hb.size());
        out.halo.insert(out.halo.end(), hb.begin(), hb.end());
    }
}

static int build_blocked(InvPlan& P, op_map m, const std::vector<int>& ptr, const std::vector<int2>& inc2)
{
    HostTimer tm("blocked plan", m->name);
    // the kernel reads a node's degree from ptr at the node's index
    if ((long long)ptr.size() != (long long)m->to->size + 1)
        return fail("internal error: blocked plan of %s needs a CSR over every node", m->name);
    const std::vector<int> pb = set_parts(m->to);
    const int np = (int)pb.size() - 1;
    std::vector<BlkOut> outs(std::max(np, 0));
    parallel_for(np, [&](int k) { blocked_part(pb[k], pb[k + 1], ptr, inc2, outs[k]); });
    std::vector<int2> blocks;
    std::vector<int> hptr{0}, halo, jbase, pbase, jpref, eid;
    std::vector<unsigned> jslot;
    int max_halo = 0;
    for (auto& o : outs) {
        if (!o.err.empty()) return fail("%s", o.err.c_str());
        size_t hp = 0, pp = 0, ep = 0;
        for (size_t b = 0; b < o.blocks.size(); b++) {
            blocks.push_back(o.blocks[b]);
            jbase.push_back((int)(jslot.size() + ep));
            pbase.push_back((int)(jpref.size() + pp));
            hp += o.nhalo[b];
            hptr.push_back((int)(halo.size() + hp));
            pp += o.npref[b];
            ep += o.nent[b];
            max_halo = std::max(max_halo, o.nhalo[b]);
        }
        halo.insert(halo.end(), o.halo.begin(), o.halo.end());
        jpref.insert(jpref.end(), o.jpref.begin(), o.jpref.end());
        jslot.insert(jslot.end(), o.jslot.begin(), o.jslot.end());
        eid.insert(eid.end(), o.eid.begin(), o.eid.end());
        o = BlkOut();
    }
    // invariants the kernel relies on
    for (size_t bi = 0; bi < blocks.size(); bi++) {
        const int nh = hptr[bi + 1] - hptr[bi];
        if (blocks[bi].y > kPart || nh > kHaloCap) return fail("internal error: blocked plan block too large");
        if (bi && blocks[bi].x != blocks[bi - 1].x + blocks[bi - 1].y) return fail("internal error: blocks not contiguous");
    }
    for (unsigned r : jslot) {
        const unsigned sl = r & 0x7fffffffu;
        if (sl >= (unsigned)(kPart + kHaloCap)) return fail("internal error: blocked plan slot out of range");
    }
    auto up = [](auto*& d, const auto& h) -> int {
        using T = typename std::decay<decltype(h[0])>::type;
        HIPCHK(hipMalloc((void**)&d, sizeof(T) * std::max<size_t>(1, h.size())));
        if (!h.empty()) HIPCHK(hipMemcpy((void*)d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice));
        return 0;
    };
    TRY(up(P.d_bk_bnode, blocks));
    TRY(up(P.d_bk_hptr, hptr));
    TRY(up(P.d_bk_halo, halo));
    TRY(up(P.d_bk_jbase, jbase));
    TRY(up(P.d_bk_pbase, pbase));
    TRY(up(P.d_bk_jpref, jpref));
    TRY(up(P.d_bk_jslot, jslot));
    TRY(up(P.d_bk_eid, eid));
    TRY(up(P.d_bk_nptr, ptr));
    HIPCHK(hipMalloc(&P.d_bk_w3, sizeof(double) * 3 * std::max<size_t>(1, eid.size())));
    P.bk_blocks = (int)blocks.size();
    P.bk_entries = (int)eid.size();
    P.bk_max_halo = max_halo;
    P.bk_lds = sizeof(double) * kRec * (size_t)(kPart + max_halo);
    return 0;
}

// Part plan of k_flux_part: the kPart-node parts of the node order (halved
// while owned + halo nodes exceed kPartSlots); per part its sorted halo and
// its local edges (edges with an owned endpoint: interior edges once, cut
// edges on both sides), greedily coloured so that no two edges of a colour
// share an owned node, and sorted by colour.
constexpr int kPartSlots = 704;   // LDS node records per part (10-bit slots)

// Part plan of one initial part (split while owned + halo nodes exceed
// kPartSlots or nodes exceed the part size); h0/e0 local to the part.
struct PartOut {
    std::vector<PartInfo> parts;
    std::vector<int> halo, eid;
    std::vector<unsigned> code;
    std::string err;
};

// LDS bank-aware order of a part's local edges, within each colour (so the
// coloured kernel's colour phases, and with them its sums, are unchanged).
// The kernels give consecutive local edges to consecutive lanes, and the LDS
// serves a wave's lanes in groups: 8 lanes per pass for the 16-byte record
// reads of 48-byte records (whose bank is fixed by the slot modulo 8), 16
// per pass for 8-byte accesses (accumulators at a 40-byte stride and
// 40-byte records: slot modulo 16).  Each position takes, among the next
// candidates of its colour, the edge whose endpoint slots (sa, sb) collide
// least with those already in its 8- and 16-lane groups.  The result is
// left in 'le'; 'out' is scratch space.
template <typename LEv>
static void bank_order(std::vector<LEv>& le, std::vector<LEv>& out)
{
    constexpr size_t kLook = 48;
    out.clear();
    out.reserve(le.size());
    auto sa = [&](size_t i) { return le[i].sa; };
    auto sb = [&](size_t i) { return le[i].sb; };
    std::vector<size_t> rest;
    unsigned m8a = 0, m8b = 0, m16a = 0, m16b = 0;
    for (size_t lo = 0; lo < le.size();) {
        size_t hi = lo;
        while (hi < le.size() && le[hi].col == le[lo].col) hi++;
        rest.clear();
        for (size_t i = lo; i < hi; i++) rest.push_back(i);
        size_t head = 0;
        while (head < rest.size()) {
            const size_t pos = out.size();
            if (pos % 16 == 0) m16a = m16b = 0;
            if (pos % 8 == 0) m8a = m8b = 0;
            size_t best = head;
            int bc = 1 << 30;
            for (size_t k = head; k < rest.size() && k < head + kLook; k++) {
                const size_t i = rest[k];
                const int c = ((m8a >> (sa(i) & 7)) & 1) + ((m8b >> (sb(i) & 7)) & 1) +
                              ((m16a >> (sa(i) & 15)) & 1) + ((m16b >> (sb(i) & 15)) & 1);
                if (c < bc) {
                    bc = c;
                    best = k;
                    if (c == 0) break;
                }
            }
            std::swap(rest[head], rest[best]);
            const size_t i = rest[head++];
            m8a |= 1u << (sa(i) & 7);
            m8b |= 1u << (sb(i) & 7);
            m16a |= 1u << (sa(i) & 15);
            m16b |= 1u << (sb(i) & 15);
            out.push_back(le[i]);
        }
        lo = hi;
    }
    le.swap(out);
}

static void part_part(int lo0, int hi0, int psize, const std::vector<int>& ptr, const std::vector<int2>& inc2,
                      PartOut& out)
{
    std::vector<int2> work{make_int2(lo0, hi0)};
    std::vector<int> hb;
    std::vector<unsigned long long> used(kPart);   // per owned node: colours taken
    // a local edge: endpoints, their LDS slots, edge id, owner flags, colour
    struct LE { int a, b, sa, sb, e; unsigned flags; int col; };
    std::vector<LE> le, tmp;
    std::vector<int> ccount;
    while (!work.empty()) {
        const int lo = work.back().x, hi = work.back().y;
        work.pop_back();
        hb.clear();
        for (int q = ptr[lo]; q < ptr[hi]; q++) {
            const int o = inc2[q].x;
            if (o < lo || o >= hi) hb.push_back(o);
        }
        std::sort(hb.begin(), hb.end());
        hb.erase(std::unique(hb.begin(), hb.end()), hb.end());
        if ((hi - lo) + (int)hb.size() > kPartSlots || (hi - lo) > psize) {
            if (hi - lo <= 8) {
                out.err = "edge-flux part plan: node " + std::to_string(lo) + " has more than " +
                          std::to_string(kPartSlots) + " neighbours";
                return;
            }
            const int mid = lo + (hi - lo) / 2;
            work.push_back(make_int2(mid, hi));
            work.push_back(make_int2(lo, mid));
            continue;
        }
        // owned nodes take slots [0, hi - lo), the sorted halo the rest
        auto slot = [&](int nd) -> int {
            if (nd >= lo && nd < hi) return nd - lo;
            return (hi - lo) + (int)(std::lower_bound(hb.begin(), hb.end(), nd) - hb.begin());
        };
        le.clear();
        for (int nd = lo; nd < hi; nd++)
            for (int q = ptr[nd]; q < ptr[nd + 1]; q++) {
                const int o = inc2[q].x;
                const bool role_b = (unsigned)inc2[q].y >> 31;
                const int e = inc2[q].y & 0x7fffffff;
                const bool other_owned = o >= lo && o < hi;
                if (other_owned && role_b) continue;     // interior edge: taken from its a side
                LE x;
                x.e = e;
                const int so = slot(o);
                if (!role_b) { x.a = nd; x.b = o; x.sa = nd - lo; x.sb = so; x.flags = 1u | (other_owned ? 2u : 0u); }
                else { x.a = o; x.b = nd; x.sa = so; x.sb = nd - lo; x.flags = 2u; }
                le.push_back(x);
            }
        // greedy colouring in edge order
        std::fill(used.begin(), used.begin() + (hi - lo), 0ull);
        std::sort(le.begin(), le.end(), [](const LE& x, const LE& y) { return x.e < y.e; });
        int ncol = 0;
        for (auto& x : le) {
            unsigned long long u = 0;
            if (x.flags & 1u) u |= used[x.a - lo];
            if (x.flags & 2u) u |= used[x.b - lo];
            if (~u == 0ull) {
                out.err = "edge-flux part plan: more than 64 edge colours in a part";
                return;
            }
            const int c = __builtin_ctzll(~u);
            x.col = c;
            if (x.flags & 1u) used[x.a - lo] |= 1ull << c;
            if (x.flags & 2u) used[x.b - lo] |= 1ull << c;
            ncol = std::max(ncol, c + 1);
        }
        // stable counting sort by colour (edge order kept within a colour)
        ccount.assign(ncol + 1, 0);
        for (const auto& x : le) ccount[x.col + 1]++;
        for (int c = 0; c < ncol; c++) ccount[c + 1] += ccount[c];
        tmp.resize(le.size());
        for (const auto& x : le) tmp[ccount[x.col]++] = x;
        le.swap(tmp);
        bank_order(le, tmp);
        PartInfo pi;
        pi.n0 = lo;
        pi.cnt = hi - lo;
        pi.h0 = (int)out.halo.size();
        pi.nh = (int)hb.size();
        pi.e0 = (int)out.code.size();
        pi.ne = (int)le.size();
        pi.ncol = ncol;
        pi.pad = 0;
        for (const auto& x : le) {
            out.code.push_back((unsigned)x.sa | ((unsigned)x.sb << 10) | (x.flags << 20) | ((unsigned)x.col << 22));
            out.eid.push_back(x.e);
        }
        out.halo.insert(out.halo.end(), hb.begin(), hb.end());
        out.parts.push_back(pi);
    }
}

static int build_partplan(InvPlan& P, op_map m, const std::vector<int>& ptr, const std::vector<int2>& inc2)
{
    HostTimer tm("part plan", m->name);
    if ((long long)ptr.size() != (long long)m->to->size + 1)
        return fail("internal error: part plan of %s needs a CSR over every node", m->name);
    const int nt = m->to->size;
    const std::vector<int> pb = set_parts(m->to);
    const int np = (int)pb.size() - 1;
    std::vector<PartOut> outs(std::max(np, 0));
    const int psize = P.pt_psize;
    parallel_for(np, [&](int k) { part_part(pb[k], pb[k + 1], psize, ptr, inc2, outs[k]); });
    std::vector<PartInfo> parts;
    std::vector<int> halo, eid;
    std::vector<unsigned> code;
    int rec_cap = 0, max_ne = 0, max_col = 0;
    for (auto& o : outs) {
        if (!o.err.empty()) return fail("%s", o.err.c_str());
        for (PartInfo pi : o.parts) {
            pi.h0 += (int)halo.size();
            pi.e0 += (int)code.size();
            parts.push_back(pi);
            rec_cap = std::max(rec_cap, pi.cnt + pi.nh);
            max_ne = std::max(max_ne, pi.ne);
            max_col = std::max(max_col, pi.ncol);
        }
        halo.insert(halo.end(), o.halo.begin(), o.halo.end());
        code.insert(code.end(), o.code.begin(), o.code.end());
        eid.insert(eid.end(), o.eid.begin(), o.eid.end());
        o = PartOut();
    }
    // invariants the kernel relies on
    long long covered = 0;
    for (size_t i = 0; i < parts.size(); i++) {
        const PartInfo& pi = parts[i];
        if (pi.cnt > kPart || pi.cnt + pi.nh > kPartSlots) return fail("internal error: part plan part too large");
        if (i && pi.n0 != parts[i - 1].n0 + parts[i - 1].cnt) return fail("internal error: parts not contiguous");
        covered += pi.cnt;
        for (int j = 0; j < pi.ne; j++) {
            const unsigned c = code[pi.e0 + j];
            if ((c & 1023u) >= (unsigned)(pi.cnt + pi.nh) || ((c >> 10) & 1023u) >= (unsigned)(pi.cnt + pi.nh))
                return fail("internal error: part plan slot out of range");
            if (((c & (1u << 20)) && (c & 1023u) >= (unsigned)pi.cnt) ||
                ((c & (1u << 21)) && ((c >> 10) & 1023u) >= (unsigned)pi.cnt))
                return fail("internal error: part plan owner flag on a halo slot");
        }
    }
    if (covered != nt) return fail("internal error: part plan does not cover the node set");
    auto up = [](auto*& d, const auto& h) -> int {
        using T = typename std::decay<decltype(h[0])>::type;
        HIPCHK(hipMalloc((void**)&d, sizeof(T) * std::max<size_t>(1, h.size())));
        if (!h.empty()) HIPCHK(hipMemcpy((void*)d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice));
        return 0;
    };
    TRY(up(P.d_pt_info, parts));
    halo.push_back(0);   // pad: the kernel's clamped halo index loads stay in bounds when nh == 0
    TRY(up(P.d_pt_halo, halo));
    TRY(up(P.d_pt_code, code));
    TRY(up(P.d_pt_eid, eid));
    HIPCHK(hipMalloc(&P.d_pt_w, sizeof(double) * 3 * std::max<size_t>(1, eid.size())));
    P.pt_parts = (int)parts.size();
    P.h_pt_info = parts;
    static std::atomic<unsigned> gen{0};
    P.pt_gen = ++gen;
    P.pt_nloc = (long long)eid.size();
    P.pt_rec_cap = rec_cap;
    {
        std::vector<int> r(parts.size());
        for (size_t i = 0; i < parts.size(); i++) r[i] = parts[i].cnt + parts[i].nh;
        std::nth_element(r.begin(), r.begin() + r.size() / 2, r.end());
        P.pt_rec_med = r.empty() ? 0 : r[r.size() / 2];
    }
    P.pt_max_ne = max_ne;
    P.pt_max_col = max_col;
    // FP64 records: 40 bytes (the wave speed formed per edge) where 48-byte
    // ones would leave fewer than four workgroups per CU, and where the parts
    // run as one round (fewer than the device's 4 workgroups per CU): there a
    // dispatch lasts one workgroup lifetime, which staging no wave speeds
    // shortens; else 48 bytes (the speed formed once per staged node).
    const bool lds_bound = (size_t)rec_cap * kPRec * sizeof(double) + (size_t)kPart * 5 * sizeof(double) > 40960;
    P.pt_rs = (lds_bound || (long long)parts.size() < 4LL * R.n_cu) ? 5 : kPRec;
    if (P.pt_rs_set) P.pt_rs = P.pt_rs_set;
    P.pt_lds = sizeof(double) * ((size_t)rec_cap * P.pt_rs + (size_t)kPart * 5);
    P.pt_lds32 = sizeof(double) * (((size_t)rec_cap * PartRec<float>::kStride * sizeof(float) + 7) / 8 +
                                   (size_t)kPart * 5);   // FP32 records, FP64 accumulators
    return 0;
}

static int build_inverse(op_map m)
{
    MapImpl* mi = M(m);
    InvPlan& P = mi->inv;
    if (P.built) return 0;
    if (int rc = freeze()) return rc;
    HostTimer tm("plans", m->name);
    SetImpl* fs = S(m->from);
    SetImpl* ts = S(m->to);
    const int nf = m->from->size, nt = m->to->size, dim = m->dim;
    // The source elements are split into nch consecutive chunks, each counted
    // and later placed on a host thread of its own: chunk c's incidences of
    // target t go after those of chunks < c, so every target keeps its
    // incidences in element order, as one pass would.  (Counts per chunk and
    // target: at most 128 MB.)
    const long long nq = (long long)nf * dim;
    const int nch = nq < (1 << 18) || nt == 0
                        ? 1
                        : (int)std::max(1LL, std::min(8LL, (128LL << 20) / ((long long)sizeof(int) * nt)));
    std::vector<int> ebound(nch + 1);
    for (int c = 0; c <= nch; c++) ebound[c] = (int)((long long)nf * c / nch);
    std::vector<std::vector<int>> ccnt(nch, std::vector<int>(nt, 0));
    host_tasks(nch, [&](int c) {
        std::vector<int>& k = ccnt[c];
        for (size_t q = (size_t)ebound[c] * dim; q < (size_t)ebound[c + 1] * dim; q++) k[ts->to_int(mi->h_orig[q])]++;
    });
    std::vector<int> cnt(nt, 0);
    for (int c = 0; c < nch; c++)
        for (int t = 0; t < nt; t++) cnt[t] += ccnt[c][t];
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
    std::vector<int2> inc2(dim == 2 ? ptr[P.n_targets] : 0);
    std::vector<int> inc1(dim == 2 ? 0 : ptr[P.n_targets]);
    // each chunk's first position in every target's run
    for (int t = 0; t < nt; t++) {
        if (!cnt[t]) continue;
        int at = ptr[P.dense ? t : slot[t]];
        for (int c = 0; c < nch; c++) {
            const int k = ccnt[c][t];
            ccnt[c][t] = at;
            at += k;
        }
    }
    // original element order, column order within an element: OP2 seq order
    host_tasks(nch, [&](int ch) {
        std::vector<int>& fill = ccnt[ch];
        for (int e = ebound[ch]; e < ebound[ch + 1]; e++) {
            const int ein = fs->to_int(e);
            for (int c = 0; c < dim; c++) {
                const int t = ts->to_int(mi->h_orig[(size_t)e * dim + c]);
                const int at = fill[t]++;
                if (dim == 2) {
                    const int other = ts->to_int(mi->h_orig[(size_t)e * dim + (1 - c)]);
                    inc2[at] = make_int2(other, (int)((unsigned)ein | ((unsigned)c << 31)));
                } else {
                    inc1[at] = ein;
                }
            }
        }
    });
    if (!P.dense) {
        HIPCHK(hipMalloc(&P.d_targets, sizeof(int) * std::max<size_t>(1, tg.size())));
        if (!tg.empty()) HIPCHK(hipMemcpy(P.d_targets, tg.data(), sizeof(int) * tg.size(), hipMemcpyHostToDevice));
    }
    HIPCHK(hipMalloc(&P.d_ptr, sizeof(int) * ptr.size()));
    HIPCHK(hipMemcpy(P.d_ptr, ptr.data(), sizeof(int) * ptr.size(), hipMemcpyHostToDevice));
    if (dim == 2) {
        HIPCHK(hipMalloc(&P.d_inc2, sizeof(int2) * std::max<size_t>(1, inc2.size())));
        if (!inc2.empty()) HIPCHK(hipMemcpy(P.d_inc2, inc2.data(), sizeof(int2) * inc2.size(), hipMemcpyHostToDevice));
        // flux plans are built on first use (ensure_flux_plan) from a CSR
        // over every node of the target set: a node no edge touches has an
        // empty row (the incidences are the same, in the same order)
        if (P.dense) {
            P.h_ptr = std::move(ptr);
        } else {
            P.h_ptr.assign(nt + 1, 0);
            for (int t = 0; t < nt; t++) P.h_ptr[t + 1] = P.h_ptr[t] + cnt[t];
        }
        P.h_inc2 = std::move(inc2);
    } else {
        HIPCHK(hipMalloc(&P.d_inc1, sizeof(int) * std::max<size_t>(1, inc1.size())));
        if (!inc1.empty()) HIPCHK(hipMemcpy(P.d_inc1, inc1.data(), sizeof(int) * inc1.size(), hipMemcpyHostToDevice));
        if (P.dense) {
            P.d_nptr = P.d_ptr;
        } else {       // targets are in increasing element order, so the entries line up
            std::vector<int> nptr(nt + 1, 0);
            for (int t = 0; t < nt; t++) nptr[t + 1] = nptr[t] + cnt[t];
            HIPCHK(hipMalloc(&P.d_nptr, sizeof(int) * nptr.size()));
            HIPCHK(hipMemcpy(P.d_nptr, nptr.data(), sizeof(int) * nptr.size(), hipMemcpyHostToDevice));
        }
    }
    P.built = true;
    return 0;
}

static void free_partplan(InvPlan& P)
{
    for (void* p : {(void*)P.d_pt_info, (void*)P.d_pt_halo, (void*)P.d_pt_code, (void*)P.d_pt_eid, (void*)P.d_pt_w,
                    (void*)P.d_pt_w32})
        if (p) hipFree(p);
    P.d_pt_info = nullptr; P.d_pt_halo = nullptr; P.d_pt_code = nullptr; P.d_pt_eid = nullptr;
    P.d_pt_w = nullptr; P.d_pt_w32 = nullptr;
    P.pt_parts = 0;
    P.h_pt_info.clear();
    P.pt_wdat = P.pt_w32dat = nullptr;
    P.pt_wver = P.pt_w32ver = 0;
}

// Re-plans the edge-parallel flux of an edge map with parts of at most
// 'nodes' nodes (the autotuner's knob).
static int set_part_nodes(op_map m, int nodes)
{
    if (nodes < 16 || nodes > kPart) return fail("set_part_nodes: %d nodes per part outside [16, %d]", nodes, kPart);
    TRY(build_inverse(m));
    InvPlan& P = M(m)->inv;
    if (P.h_ptr.empty()) return fail("set_part_nodes: map %s has no part plan", m->name);
    if (P.pt_psize == nodes && P.pt_built) return 0;
    HIPCHK(hipStreamSynchronize(cur_stream()));
    free_partplan(P);
    P.pt_built = false;
    P.pt_psize = nodes;
    TRY(build_partplan(P, m, P.h_ptr, P.h_inc2));
    P.pt_built = true;
    return 0;
}

static void free_invplan(InvPlan& P)
{
    if (P.d_nptr && P.d_nptr != P.d_ptr) hipFree(P.d_nptr);
    for (void* p : {(void*)P.d_targets, (void*)P.d_ptr, (void*)P.d_inc2, (void*)P.d_inc1, (void*)P.d_bk_bnode,
                    (void*)P.d_bk_hptr, (void*)P.d_bk_halo, (void*)P.d_bk_jbase, (void*)P.d_bk_pbase,
                    (void*)P.d_bk_jpref, (void*)P.d_bk_jslot, (void*)P.d_bk_eid, (void*)P.d_bk_w3,
                    (void*)P.d_bk_nptr})
        if (p) hipFree(p);
    free_partplan(P);
    P = InvPlan();
}

// Releases the device memory and state of one set, map or dat (batched and
// served solves that end); the handle stays valid as an empty, dead object.
static void release_dat(op_dat d)
{
    RtLock lk(g_rt_mu);
    if (!d || !d->impl) return;
    DatImpl* di = D(d);
    for (void* p : {(void*)di->d, (void*)di->d_pack, (void*)di->d_tmp})
        if (p) hipFree(p);
    delete di;
    d->impl = nullptr;
}

static void release_map(op_map m)
{
    RtLock lk(g_rt_mu);
    if (!m || !m->impl) return;
    MapImpl* mi = M(m);
    if (mi->d_map) hipFree(mi->d_map);
    free_invplan(mi->inv);
    delete mi;
    m->impl = nullptr;
}

static void release_set(op_set s)
{
    RtLock lk(g_rt_mu);
    if (!s || !s->impl) return;
    delete S(s);
    s->impl = nullptr;
    s->size = 0;
}

// Builds, on first use, the plan the given edge-flux variant launches on:
// the blocked-gather plan (GATHER_*) or the part plan (PART_*, MIXED_*).
static int ensure_flux_plan(op_map m, int variant)
{
    TRY(build_inverse(m));
    InvPlan& P = M(m)->inv;
    if (P.h_ptr.empty()) return 0;
    const bool gather = variant == MGCFD_FLUX_GATHER_EXACT || variant == MGCFD_FLUX_GATHER_FAST;
    const bool part = variant == MGCFD_FLUX_PART_COLOURED || variant == MGCFD_FLUX_PART_ATOMIC ||
                      variant == MGCFD_FLUX_MIXED_COLOURED || variant == MGCFD_FLUX_MIXED_ATOMIC;
    if (gather && !P.bk_built) {
        TRY(build_blocked(P, m, P.h_ptr, P.h_inc2));
        P.bk_built = true;
    }
    if (part && !P.pt_built) {
        TRY(build_partplan(P, m, P.h_ptr, P.h_inc2));
        P.pt_built = true;
    }
    return 0;
}

static int packed_weights(op_dat w, const double4** out)
{
    DatImpl* di = D(w);
    const int n = w->set->size;
    if (!di->d_pack) HIPCHK(hipMalloc(&di->d_pack, sizeof(double4) * std::max(1, n)));
    if (di->pack_version != di->version) {
        if (n > 0) {
            k_pack_weights<<<nblk(n), kBlock, 0, cur_stream()>>>(n, (const double*)di->d, di->d_pack);
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

// Host readback of a device scalar, synchronising the stream.
static int read_scalar(const double* d_src, double* out)
{
    HIPCHK(hipMemcpyAsync(R.h_pin, d_src, sizeof(double), hipMemcpyDeviceToHost, cur_stream()));
    HIPCHK(hipStreamSynchronize(cur_stream()));
    *out = *R.h_pin;
    return 0;
}

static int read_iscalar(const int* d_src, int* out)
{
    HIPCHK(hipMemcpyAsync(R.h_pini, d_src, sizeof(int), hipMemcpyDeviceToHost, cur_stream()));
    HIPCHK(hipStreamSynchronize(cur_stream()));
    *out = *R.h_pini;
    return 0;
}

// ------------------------------------------------------- flux launchers --

static bool part_variant(int v) { return v == MGCFD_FLUX_PART_COLOURED || v == MGCFD_FLUX_PART_ATOMIC; }
static bool mixed_variant(int v) { return v == MGCFD_FLUX_MIXED_COLOURED || v == MGCFD_FLUX_MIXED_ATOMIC; }

// Re-gathers the plan's per-entry weights when the weight dat has changed.
static int refresh_plan_weights(int variant, InvPlan& P, op_dat w)
{
    if ((variant == MGCFD_FLUX_GATHER_EXACT || variant == MGCFD_FLUX_GATHER_FAST) && P.bk_blocks > 0 &&
        (P.bk_wdat != (const void*)w || P.bk_wver != D(w)->version)) {
        k_pack_incidence_weights3<<<nblk(P.bk_entries), kBlock, 0, cur_stream()>>>(P.bk_entries, P.d_bk_eid, dptr(w),
                                                                              P.d_bk_w3);
        TRY(check_launch());
        P.bk_wdat = w;
        P.bk_wver = D(w)->version;
    }
    if (mixed_variant(variant) && P.pt_parts > 0 &&
        (!P.d_pt_w32 || P.pt_w32dat != (const void*)w || P.pt_w32ver != D(w)->version)) {
        if (!P.d_pt_w32) HIPCHK(hipMalloc(&P.d_pt_w32, sizeof(float) * 3 * std::max<size_t>(1, (size_t)P.pt_nloc)));
        k_pack_part_weights32<<<nblk((size_t)P.pt_nloc), kBlock, 0, cur_stream()>>>(P.pt_nloc, P.d_pt_eid, dptr(w),
                                                                              P.d_pt_w32);
        TRY(check_launch());
        P.pt_w32dat = w;
        P.pt_w32ver = D(w)->version;
    }
    if (part_variant(variant) && P.pt_parts > 0 && (P.pt_wdat != (const void*)w || P.pt_wver != D(w)->version)) {
        k_pack_part_weights<<<nblk((size_t)P.pt_nloc), kBlock, 0, cur_stream()>>>(P.pt_nloc, P.d_pt_eid, dptr(w),
                                                                            P.d_pt_w);
        TRY(check_launch());
        P.pt_wdat = w;
        P.pt_wver = D(w)->version;
    }
    return 0;
}

static int launch_blocked_pf(bool exact, const InvPlan& P, op_dat var, op_dat flux, const int* status, int fresh)
{
    if (exact)
        k_flux_blocked_pf<true, 4><<<P.bk_blocks, kBlock, P.bk_lds, cur_stream()>>>(
            P.d_bk_bnode, P.d_bk_hptr, P.d_bk_halo, P.d_bk_jbase, P.d_bk_pbase, P.d_bk_jpref, P.d_bk_nptr, P.d_bk_jslot,
            P.d_bk_w3, P.bk_entries, dptr(var), dptr(flux), R.k, status, fresh);
    else
        k_flux_blocked_pf<false, 4><<<P.bk_blocks, kBlock, P.bk_lds, cur_stream()>>>(
            P.d_bk_bnode, P.d_bk_hptr, P.d_bk_halo, P.d_bk_jbase, P.d_bk_pbase, P.d_bk_jpref, P.d_bk_nptr, P.d_bk_jslot,
            P.d_bk_w3, P.bk_entries, dptr(var), dptr(flux), R.k, status, fresh);
    return check_launch();
}

// The stage fused into a part kernel (k_flux_part_stage): its input state,
// stage arguments and, for batched solves, the part -> segment tables.
struct PartStage {
    const double* vin = nullptr;
    const StageArgs* a = nullptr;
    const int* pseg = nullptr;
    const int* pfirst = nullptr;
};

template <typename Real, int T, int EPT, int NPT, int RS>
static int launch_part_rs(bool coloured, const InvPlan& P, const Real* w, size_t lds, op_dat var, op_dat flux,
                          const int* status, int fresh, const PartStage* st)
{
    if (st) {
        const int pw = 8 * ((R.n_cu * P.pt_persist + 7) / 8);    // persistent grid: a multiple of the 8 XCDs
        if (P.pt_persist > 0 && pw < P.pt_parts) {
            if (coloured)
                k_flux_part_stage_persistent<Real, T, EPT, NPT, true, RS><<<pw, T, lds, cur_stream()>>>(
                    P.d_pt_info, P.d_pt_halo, P.d_pt_code, w, P.pt_nloc, P.pt_rec_cap, st->vin, fresh, *st->a, R.k,
                    st->pseg, st->pfirst, P.pt_parts);
            else
                k_flux_part_stage_persistent<Real, T, EPT, NPT, false, RS><<<pw, T, lds, cur_stream()>>>(
                    P.d_pt_info, P.d_pt_halo, P.d_pt_code, w, P.pt_nloc, P.pt_rec_cap, st->vin, fresh, *st->a, R.k,
                    st->pseg, st->pfirst, P.pt_parts);
            return check_launch();
        }
        if (coloured)
            k_flux_part_stage<Real, T, EPT, NPT, true, RS><<<P.pt_parts, T, lds, cur_stream()>>>(
                P.d_pt_info, st->pseg, st->pfirst, P.d_pt_halo, P.d_pt_code, w, P.pt_nloc, P.pt_rec_cap, st->vin,
                fresh, *st->a, R.k);
        else
            k_flux_part_stage<Real, T, EPT, NPT, false, RS><<<P.pt_parts, T, lds, cur_stream()>>>(
                P.d_pt_info, st->pseg, st->pfirst, P.d_pt_halo, P.d_pt_code, w, P.pt_nloc, P.pt_rec_cap, st->vin,
                fresh, *st->a, R.k);
        return check_launch();
    }
    if (coloured)
        k_flux_part<Real, T, EPT, NPT, true, RS><<<P.pt_parts, T, lds, cur_stream()>>>(
            P.d_pt_info, P.d_pt_halo, P.d_pt_code, w, P.pt_nloc, P.pt_rec_cap, dptr(var), dptr(flux), R.k, status,
            fresh);
    else
        k_flux_part<Real, T, EPT, NPT, false, RS><<<P.pt_parts, T, lds, cur_stream()>>>(
            P.d_pt_info, P.d_pt_halo, P.d_pt_code, w, P.pt_nloc, P.pt_rec_cap, dptr(var), dptr(flux), R.k, status,
            fresh);
    return check_launch();
}

// Record size: FP32 records have six values; FP64 ones the plan's pt_rs in
// the edge-flux kernel, and five in the kernel with the fused stage (its
// epilogue makes each workgroup longer, and 40-byte records, staging no
// wave speeds, shorten it: a 300K-node cycle 0.440 -> 0.418 ms).
template <typename Real, int T, int EPT, int NPT>
static int launch_part_t(bool coloured, const InvPlan& P, const Real* w, size_t lds, op_dat var, op_dat flux,
                         const int* status, int fresh, const PartStage* st)
{
    if constexpr (sizeof(Real) == sizeof(double)) {
        if (P.pt_rs == 5 || st)
            return launch_part_rs<Real, T, EPT, NPT, 5>(coloured, P, w, lds, var, flux, status, fresh, st);
    }
    return launch_part_rs<Real, T, EPT, NPT, PartRec<Real>::kStride>(coloured, P, w, lds, var, flux, status, fresh,
                                                                    st);
}

// Staged nodes per thread: the smaller register array when the part plan's
// largest part allows it.
template <typename Real, int T, int EPT>
static int launch_part_e(bool coloured, const InvPlan& P, const Real* w, size_t lds, op_dat var, op_dat flux,
                         const int* status, int fresh, const PartStage* st)
{
    constexpr int kSmall = (512 + T - 1) / T, kLarge = (768 + T - 1) / T;
    static_assert(kPartSlots <= kLarge * T, "staged nodes exceed the kernel's registers");
    if (P.pt_rec_cap <= kSmall * T)
        return launch_part_t<Real, T, EPT, kSmall>(coloured, P, w, lds, var, flux, status, fresh, st);
    return launch_part_t<Real, T, EPT, kLarge>(coloured, P, w, lds, var, flux, status, fresh, st);
}

// Edges per thread from the plan's largest part.
template <typename Real, int T>
static int launch_part_n(bool coloured, const InvPlan& P, const Real* w, size_t lds, op_dat var, op_dat flux,
                         const int* status, int fresh, const PartStage* st)
{
    const int ept = (P.pt_max_ne + T - 1) / T;
    if (ept <= 1024 / T) return launch_part_e<Real, T, 1024 / T>(coloured, P, w, lds, var, flux, status, fresh, st);
    if (ept <= 1536 / T) return launch_part_e<Real, T, 1536 / T>(coloured, P, w, lds, var, flux, status, fresh, st);
    if (ept <= 2048 / T) return launch_part_e<Real, T, 2048 / T>(coloured, P, w, lds, var, flux, status, fresh, st);
    if (ept <= 3072 / T) return launch_part_e<Real, T, 3072 / T>(coloured, P, w, lds, var, flux, status, fresh, st);
    return fail("edge-flux part plan: %d local edges in one part exceed the kernel's %d", P.pt_max_ne, 3072);
}

// Real = double: the FP64 part kernel, 256 threads per part.  Real = float:
// mixed precision, whose smaller register footprint lets 512 threads (8
// waves) share one part in the edge-flux kernel, which shortens each
// workgroup's edge phase; the kernel with the fused stage needs ~96 VGPRs
// (its epilogue), which leaves 2 such workgroups per CU, so it runs 256
// threads per part (5 per CU): batched solves 310 -> 323 per second.
// The autotuner may set a plan's fused-stage kernel to 512 threads per part
// (pt_threads), which halves each thread's edges at the cost of registers.
template <typename Real>
static int launch_part(bool coloured, const InvPlan& P, const Real* w, size_t lds, op_dat var, op_dat flux,
                       const int* status, int fresh, const PartStage* st = nullptr)
{
    if (st) {
        if (P.pt_threads == 512) return launch_part_n<Real, 512>(coloured, P, w, lds, var, flux, status, fresh, st);
        return launch_part_n<Real, 256>(coloured, P, w, lds, var, flux, status, fresh, st);
    }
    if constexpr (sizeof(Real) == 4) return launch_part_n<Real, 512>(coloured, P, w, lds, var, flux, status, fresh, st);
    return launch_part_n<Real, 256>(coloured, P, w, lds, var, flux, status, fresh, st);
}

// Plans and packed weights a flux launch needs; allocates and uploads, so it
// runs outside stream capture.
static int prepare_flux(int variant, op_map map, op_dat w)
{
    const double4* wp = nullptr;
    TRY(packed_weights(w, &wp));
    if (variant == MGCFD_FLUX_ATOMIC) return 0;
    TRY(ensure_flux_plan(map, variant));
    InvPlan& P = M(map)->inv;
    TRY(refresh_plan_weights(variant, P, w));
    return 0;
}

// Edge flux of a part variant fused with a Runge-Kutta stage (device
// multigrid driver); plans and weights come from prepare_flux.
static bool fused_stage_variant(int v) { return part_variant(v) || mixed_variant(v); }

static int launch_flux_stage(int variant, op_map map, const PartStage& st, int fresh)
{
    const InvPlan& P = M(map)->inv;
    if (!(part_variant(variant) || mixed_variant(variant)) || P.pt_parts == 0)
        return fail("edge-flux variant %d has no part plan for map %s", variant, map->name);
    if (mixed_variant(variant))
        return launch_part<float>(variant == MGCFD_FLUX_MIXED_COLOURED, P, P.d_pt_w32, P.pt_lds32, nullptr, nullptr,
                                  st.a->status, fresh, &st);
    const size_t lds5 = sizeof(double) * ((size_t)P.pt_rec_cap * 5 + (size_t)kPart * 5);   // 40-byte records
    return launch_part<double>(variant == MGCFD_FLUX_PART_COLOURED, P, P.d_pt_w, lds5, nullptr, nullptr,
                               st.a->status, fresh, &st);
}

// fresh: the edge fluxes start from zero instead of adding to the flux dat
// (the device MG driver's later Runge-Kutta stages, whose flux the previous
// stage left to be overwritten); the HBM-atomic variant always adds.
static int launch_flux(int variant, op_map map, op_dat var, op_dat w, op_dat flux, const int* status,
                       bool fresh = false)
{
    if (variant == MGCFD_FLUX_ATOMIC) {
        const double4* wp = nullptr;
        TRY(packed_weights(w, &wp));
        const int ne = map->from->size;
        if (ne == 0) return 0;
        k_flux_atomic<<<nblk(ne), kBlock, 0, cur_stream()>>>(ne, (const int2*)M(map)->d_map, wp, dptr(var),
                                                         dptr(flux), R.k, status);
        return check_launch();
    }
    TRY(ensure_flux_plan(map, variant));
    InvPlan& P = M(map)->inv;
    if (P.n_targets == 0) return 0;
    TRY(refresh_plan_weights(variant, P, w));
    if (part_variant(variant) || mixed_variant(variant)) {
        if (P.pt_parts == 0) return fail("edge-flux variant %d needs a two-column edge map", variant);
        if (mixed_variant(variant))
            return launch_part<float>(variant == MGCFD_FLUX_MIXED_COLOURED, P, P.d_pt_w32, P.pt_lds32, var, flux,
                                      status, fresh);
        return launch_part<double>(variant == MGCFD_FLUX_PART_COLOURED, P, P.d_pt_w, P.pt_lds, var, flux, status,
                                   fresh);
    }
    if ((variant == MGCFD_FLUX_GATHER_EXACT || variant == MGCFD_FLUX_GATHER_FAST) && P.bk_blocks > 0)
        return launch_blocked_pf(variant == MGCFD_FLUX_GATHER_EXACT, P, var, flux, status, fresh);
    const double4* wp = nullptr;
    TRY(packed_weights(w, &wp));
    if (variant == MGCFD_FLUX_GATHER_FAST)
        k_flux_gather<false><<<nblk(P.n_targets), kBlock, 0, cur_stream()>>>(
            P.n_targets, P.d_targets, P.d_ptr, P.d_inc2, wp, dptr(var), dptr(flux), R.k, status, fresh);
    else
        k_flux_gather<true><<<nblk(P.n_targets), kBlock, 0, cur_stream()>>>(
            P.n_targets, P.d_targets, P.d_ptr, P.d_inc2, wp, dptr(var), dptr(flux), R.k, status, fresh);
    return check_launch();
}

// Identity target list of a dense boundary plan (a boundary map that
// touches every node); the boundary kernel always walks a target list.
static int bnd_targets(op_map b2n)
{
    TRY(build_inverse(b2n));
    InvPlan& P = M(b2n)->inv;
    if (P.n_targets == 0 || P.d_targets) return 0;
    std::vector<int> id(P.n_targets);
    std::iota(id.begin(), id.end(), 0);
    HIPCHK(hipMalloc(&P.d_targets, sizeof(int) * id.size()));
    HIPCHK(hipMemcpy(P.d_targets, id.data(), sizeof(int) * id.size(), hipMemcpyHostToDevice));
    return 0;
}

static int launch_bnd(op_map b2n, op_dat group, op_dat bw, op_dat var, op_dat flux, const int* status)
{
    TRY(bnd_targets(b2n));
    const InvPlan& P = M(b2n)->inv;
    if (P.n_targets == 0) return 0;
    k_bnd_flux<<<nblk(P.n_targets), kBlock, 0, cur_stream()>>>(P.n_targets, P.d_targets, P.d_ptr, P.d_inc1,
                                                           iptr(group), dptr(bw), dptr(var), dptr(flux), R.k,
                                                           status);
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
    if (R.device >= kMaxDevices) return fail("op_init: device %d is beyond the %d this library serves", R.device,
                                             kMaxDevices);
    if (!scratch_dev[R.device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail("op_init: no device address for the no-allocation scratch: %s", hipGetErrorString(e));
        scratch_dev[R.device] = static_cast<double*>(p);
    }
    {
        hipDeviceProp_t prop;
        HIPCHK(hipGetDeviceProperties(&prop, R.device));
        R.n_cu = std::max(1, prop.multiProcessorCount);
    }
    HIPCHK(hipStreamCreateWithFlags(&R.lib_stream, hipStreamNonBlocking));
    HIPCHK(hipMalloc(&R.d_scal, 4 * sizeof(double)));
    HIPCHK(hipMalloc(&R.d_iscal, 4 * sizeof(int)));
    HIPCHK(hipHostMalloc(&R.h_pin, 8 * sizeof(double), hipHostMallocDefault));
    HIPCHK(hipHostMalloc(&R.h_pini, 8 * sizeof(int), hipHostMallocDefault));
    TRY(ensure_part(4096));
    default_consts();
    R.inited = true;
    return 0;
}

void op_exit(void)
{
    if (!R.inited) return;
    mg_join_admissions();      // before the lock, which admissions in flight take
    RtLock lk(g_rt_mu);
    hipStreamSynchronize(cur_stream());
    mg_release_all();
    for (auto& d : R.dats) release_dat(&d);
    for (auto& m : R.maps) release_map(&m);
    for (auto& s : R.sets) delete S(&s);
    R.dats.clear();
    R.maps.clear();
    R.sets.clear();
    R.names.clear();
    if (R.d_part) hipFree(R.d_part);
    hipFree(R.d_scal);
    hipFree(R.d_iscal);
    hipHostFree(R.h_pin);
    hipHostFree(R.h_pini);
    hipStreamDestroy(R.lib_stream);
    R = Runtime();
}

op_set op_decl_set(int size, const char* name)
{
    if (size < 0) { fail("op_decl_set %s: negative size", name ? name : ""); return nullptr; }
    RtLock lk(g_rt_mu);
    R.sets.push_back(op_set_core{});
    op_set s = &R.sets.back();
    s->index = (int)R.sets.size() - 1;
    s->size = size;
    s->name = keep_name(name);
    SetImpl* si = new SetImpl();
    si->priv = tl_priv_decl;
    s->impl = si;
    R.dirty = true;
    return s;
}

static op_map add_map(op_set from, op_set to, int dim, MapImpl* mi, const char* name);

op_map op_decl_map(op_set from, op_set to, int dim, const int* imap, const char* name)
{
    if (!from || !to || dim < 1 || (!imap && from->size > 0)) { fail("op_decl_map %s: bad arguments", name ? name : ""); return nullptr; }
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
    return add_map(from, to, dim, mi, name);
}

// A map from zero-based entries that the caller has range-checked (the
// library's own batched meshes): the vector is taken over, with no copy and
// no OP_MAPS_BASE_INDEX shift (that applies to the user's maps only).
static op_map decl_map_owned(op_set from, op_set to, int dim, std::vector<int>&& v, const char* name)
{
    if (!from || !to || dim < 1 || v.size() != (size_t)from->size * dim) {
        fail("decl_map %s: bad arguments", name ? name : "");
        return nullptr;
    }
    if (dim == 2)
        for (int e = 0; e < from->size; e++)
            if (v[2 * (size_t)e] == v[2 * (size_t)e + 1]) {
                fail("op_decl_map %s: element %d maps both columns to node %d (self-loop)", name ? name : "", e,
                     v[2 * (size_t)e]);
                return nullptr;
            }
    MapImpl* mi = new MapImpl();
    mi->h_orig = std::move(v);
    return add_map(from, to, dim, mi, name);
}

static op_map add_map(op_set from, op_set to, int dim, MapImpl* mi, const char* name)
{
    RtLock lk(g_rt_mu);
    R.maps.push_back(op_map_core{});
    op_map m = &R.maps.back();
    m->index = (int)R.maps.size() - 1;
    m->from = from;
    m->to = to;
    m->dim = dim;
    m->name = keep_name(name);
    m->impl = mi;
    R.dirty = true;
    return m;
}

op_map mgcfd_decl_map_compose(op_map outer, op_map inner, const char* name)
{
    if (!R.inited) { fail("mgcfd_decl_map_compose: op_init has not been called"); return nullptr; }
    if (!outer || !inner || !outer->impl || !inner->impl || inner->dim != 1 || outer->to != inner->from) {
        fail("mgcfd_decl_map_compose %s: needs maps A: X -> Y and B: Y -> Z with B of dimension 1",
             name ? name : "");
        return nullptr;
    }
    const MapImpl* a = M(outer);
    const MapImpl* b = M(inner);
    MapImpl* mi = new MapImpl();
    mi->composed = true;
    mi->h_orig.resize(a->h_orig.size());
    for (size_t q = 0; q < a->h_orig.size(); q++) mi->h_orig[q] = b->h_orig[(size_t)a->h_orig[q]];
    return add_map(outer->from, inner->to, outer->dim, mi, name);
}

static op_dat add_dat(op_set set, int dim, const char* type, int size, const char* name, DatImpl* di);

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
        const size_t nel = (size_t)set->size * dim;
        di->pending.resize(nel * in_esz);
        if ((size_t)size == in_esz)
            std::memcpy(di->pending.data(), data, nel * in_esz);
        else
            for (size_t q = 0; q < nel; q++) std::memcpy(di->pending.data() + q * in_esz, data + q * (size_t)size, in_esz);
    }
    return add_dat(set, dim, type, is_int ? (int)sizeof(int) : size, name, di);
}

// A dat whose AoS bytes (original element order, 4-byte ints or doubles)
// the caller hands over: the library's own batched meshes, whose level
// arrays are built outside the registry lock; no copy.
static op_dat decl_dat_owned(op_set set, int dim, const char* type, std::vector<char>&& bytes, const char* name)
{
    const bool is_int = type && std::strcmp(type, "int") == 0;
    const size_t esz = is_int ? sizeof(int) : sizeof(double);
    if (!set || dim < 1 || !type || (!is_int && std::strcmp(type, "double") != 0) ||
        bytes.size() != esz * dim * (size_t)set->size) {
        fail("decl_dat %s: bad arguments", name ? name : "");
        return nullptr;
    }
    DatImpl* di = new DatImpl();
    di->is_int = is_int;
    di->pending = std::move(bytes);
    return add_dat(set, dim, type, (int)esz, name, di);
}

static op_dat add_dat(op_set set, int dim, const char* type, int size, const char* name, DatImpl* di)
{
    RtLock lk(g_rt_mu);
    R.dats.push_back(op_dat_core{});
    op_dat d = &R.dats.back();
    d->index = (int)R.dats.size() - 1;
    d->set = set;
    d->dim = dim;
    d->size = size;
    d->type = keep_name(type);
    d->name = keep_name(name);
    d->impl = di;
    if (S(set)->frozen) {
        if (upload_dat(d)) return nullptr;
    } else {
        R.dirty = true;
    }
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
    const int n = d->set->size, dim = d->dim;
    const size_t esz = di->is_int ? sizeof(int) : sizeof(double);
    const size_t rec = esz * dim;
    std::vector<char> aos(rec * (size_t)std::max(n, 1));
    HIPCHK(hipStreamSynchronize(cur_stream()));
    HIPCHK(hipMemcpy(aos.data(), di->d, rec * (size_t)n, hipMemcpyDeviceToHost));
    for (int i = 0; i < n; i++) std::memcpy(dest + (size_t)si->to_orig(i) * rec, aos.data() + (size_t)i * rec, rec);
    return 0;
}

int op_upload_data_char(op_dat d, const char* src)
{
    if (!d || !src) return fail("op_upload_data: null argument");
    TRY(freeze());
    DatImpl* di = D(d);
    const size_t esz = di->is_int ? sizeof(int) : sizeof(double);
    di->pending.assign(src, src + (size_t)d->set->size * d->dim * esz);
    HIPCHK(hipStreamSynchronize(cur_stream()));
    return upload_dat(d);
}

// One GPU: nothing to distribute.  A coordinate dat on the set (as
// euler3d.cpp:340-377 passes) selects the geometric (Z-order) internal node
// order for that set instead of the graph bisection; the order itself is
// chosen at the first loop, as for every set.
int op_partition(const char* lib, const char* method, op_set set, op_map map, op_dat coords)
{
    (void)lib; (void)method; (void)map;
    if (!R.inited) return fail("op_partition: op_init has not been called");
    if (set && coords && coords->set == set && coords->dim == 3 && coords->impl && !D(coords)->is_int &&
        set->impl && !S(set)->frozen && !D(coords)->pending.empty())
        set_geometry(set, reinterpret_cast<const double*>(D(coords)->pending.data()));
    return 0;
}

int op_renumber(op_map map)
{
    (void)map;
    return freeze();
}

void op_timers(double* cpu, double* wall)
{
    if (R.inited) hipStreamSynchronize(cur_stream());
    const double t = std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
    if (cpu) *cpu = (double)std::clock() / CLOCKS_PER_SEC;
    if (wall) *wall = t;
}

int op_rank(int* rank)
{
    if (rank) *rank = 0;
    return 0;
}

int mgcfd_plan_stats(op_map map, int* stats)
{
    if (!map || !stats) return fail("mgcfd_plan_stats: null argument");
    TRY(ensure_flux_plan(map, MGCFD_FLUX_GATHER_EXACT));
    const InvPlan& P = M(map)->inv;
    stats[0] = P.bk_blocks;
    stats[1] = kPart + P.bk_max_halo;
    stats[2] = P.bk_entries;
    stats[3] = P.max_deg;
    stats[4] = (int)P.bk_lds;
    return 0;
}

int mgcfd_map_part_config(op_map map, int part_nodes, int part_threads, int record_bytes)
{
    if (!R.inited) return fail("mgcfd_map_part_config: op_init has not been called");
    if (!map || map->dim != 2) return fail("mgcfd_map_part_config: needs an edge map of dimension 2");
    if (part_threads != 0 && part_threads != 256 && part_threads != 512)
        return fail("mgcfd_map_part_config: %d threads per part (0, 256 or 512)", part_threads);
    if (record_bytes != 0 && record_bytes != 40 && record_bytes != 48)
        return fail("mgcfd_map_part_config: %d-byte records (0, 40 or 48)", record_bytes);
    TRY(ensure_flux_plan(map, MGCFD_FLUX_PART_ATOMIC));
    InvPlan& P = M(map)->inv;
    const int rs = record_bytes / 8;
    if (P.pt_rs_set != rs) {
        P.pt_rs_set = rs;
        P.pt_psize = -1;                 // re-plan below even at the same part size
    }
    TRY(set_part_nodes(map, part_nodes));
    P.pt_threads = part_threads;
    return 0;
}

int mgcfd_part_plan_stats(op_map map, int* stats)
{
    if (!map || !stats) return fail("mgcfd_part_plan_stats: null argument");
    TRY(ensure_flux_plan(map, MGCFD_FLUX_PART_ATOMIC));
    const InvPlan& P = M(map)->inv;
    if (P.pt_parts == 0) return fail("mgcfd_part_plan_stats: map %s has no part plan", map->name);
    stats[0] = P.pt_parts;
    stats[1] = P.pt_rec_cap;
    stats[2] = P.pt_rec_med;
    stats[3] = (int)P.pt_nloc;
    stats[4] = P.pt_max_ne;
    stats[5] = (int)P.pt_lds;
    return 0;
}

int mgcfd_stream_copy(long long bytes, int iters, double* ms_per_call)
{
    if (!R.inited) return fail("mgcfd_stream_copy: op_init has not been called");
    if (bytes < 16 || iters < 1 || !ms_per_call) return fail("mgcfd_stream_copy: bad arguments");
    const size_t n = (size_t)bytes / 16;
    double2 *a = nullptr, *b = nullptr;
    HIPCHK(hipMalloc(&a, n * 16));
    if (hipMalloc(&b, n * 16) != hipSuccess) {
        hipFree(a);
        return fail("mgcfd_stream_copy: out of device memory");
    }
    k_fill<<<nblk(2 * n), kBlock, 0, cur_stream()>>>(2 * n, reinterpret_cast<double*>(a), 1.0);
    const unsigned grid = 256 * 32;
    k_stream_copy<<<grid, kBlock, 0, cur_stream()>>>(n, a, b);
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipEventRecord(e0, cur_stream()));
    for (int i = 0; i < iters; i++) k_stream_copy<<<grid, kBlock, 0, cur_stream()>>>(n, a, b);
    HIPCHK(hipEventRecord(e1, cur_stream()));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipFree(a);
    hipFree(b);
    *ms_per_call = (double)ms / iters;
    return check_launch();
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
    HIPCHK(hipStreamSynchronize(cur_stream()));
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
    if (set->size) k_fill5<<<nblk(set->size), kBlock, 0, cur_stream()>>>(set->size, dptr(a0.dat), R.k, 1);
    touch(a0.dat);
    return check_launch();
}

int op_par_loop_zero_5d_array_kernel(const char* name, op_set set, op_arg a0)
{
    const char* L = "zero_5d_array_kernel";
    (void)name;
    TRY(chk_ready(L));
    TRY(direct(a0, set, 5, false, L, "arg0 array[5] OP_WRITE"));
    if (set->size) k_fill5<<<nblk(set->size), kBlock, 0, cur_stream()>>>(set->size, dptr(a0.dat), R.k, 0);
    touch(a0.dat);
    return check_launch();
}

int op_par_loop_zero_1d_array_kernel(const char* name, op_set set, op_arg a0)
{
    const char* L = "zero_1d_array_kernel";
    (void)name;
    TRY(chk_ready(L));
    TRY(direct(a0, set, 1, false, L, "arg0 array[1] OP_WRITE"));
    if (set->size) k_fill<<<nblk(set->size), kBlock, 0, cur_stream()>>>((size_t)set->size, dptr(a0.dat), 0.0);
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
    k_cell_volumes_edge<<<nblk(ne), kBlock, 0, cur_stream()>>>(ne, (const int2*)M(a0.map)->d_map, dptr(a0.dat),
                                                           dptr(a2.dat), wi->d_tmp);
    TRY(check_launch());
    const InvPlan& P = M(a0.map)->inv;
    k_cell_volumes_node<<<nblk(P.n_targets), kBlock, 0, cur_stream()>>>(P.n_targets, P.d_targets, P.d_ptr,
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
    if (set->size) k_scale<<<nblk(3 * (size_t)set->size), kBlock, 0, cur_stream()>>>(3 * (size_t)set->size, dptr(a0.dat), 1e-7);
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
    if (set->size) k_copy<<<nblk(5 * (size_t)set->size), kBlock, 0, cur_stream()>>>(5 * (size_t)set->size, dptr(a0.dat), dptr(a1.dat), nullptr);
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
        k_calc_dt<<<nblk(set->size), kBlock, 0, cur_stream()>>>(set->size, dptr(a0.dat), dptr(a1.dat), dptr(a2.dat),
                                                            nullptr, nullptr, MinPub{});
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
    k_min_part<<<nb, kBlock, 0, cur_stream()>>>(n, dptr(a0.dat), R.d_part);
    k_min_final<<<1, kBlock, 0, cur_stream()>>>(nb, R.d_part, init, R.d_scal, nullptr);
    TRY(check_launch());
    double m;
    TRY(read_scalar(R.d_scal, &m));
    std::memcpy(a1.data, &m, sizeof(double));
    return 0;
}

int mgcfd_get_min_dt_noalloc(const char* name, op_set set, op_arg a0, op_arg a1)
{
    const char* L = "mgcfd_get_min_dt_noalloc";
    (void)name;
    if (!R.inited) return fail("%s: op_init has not been called", L);
    // the first loop after a declaration renumbers and uploads (it allocates)
    if (!R.frozen || R.dirty)
        return fail("%s: sets, maps or dats were declared since the last loop; run any op_par_loop first", L);
    TRY(direct(a0, set, 1, false, L, "arg0 dt[1] OP_READ"));
    TRY(gbl(a1, "double", OP_MIN, L, "arg1 gbl double OP_MIN"));
    const int n = set->size;
    double init;
    std::memcpy(&init, a1.data, sizeof(double));
    if (n == 0) return 0;
    const long long nb = nblk(n);
    if (nb > kNoAllocBlocks)
        return fail("%s: the set has %d elements, more than the %lld its fixed scratch serves", L, n,
                    (long long)kNoAllocBlocks * kBlock);
    double* ptr = scratch_dev[R.device];
    if (!ptr) return fail("%s: no scratch address for device %d", L, R.device);
    k_min_part<<<(unsigned)nb, kBlock, 0, cur_stream()>>>(n, dptr(a0.dat), ptr);
    k_min_final<<<1, kBlock, 0, cur_stream()>>>((int)nb, ptr, init, ptr + kNoAllocBlocks, nullptr);
    TRY(check_launch());
    double m;
    TRY(read_scalar(ptr + kNoAllocBlocks, &m));
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
        k_step_factor<<<nblk(set->size), kBlock, 0, cur_stream()>>>(set->size, dptr(a1.dat), nullptr, m, dptr(a3.dat), nullptr);
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
        k_time_step<<<nblk(set->size), kBlock, 0, cur_stream()>>>(set->size, rk, dptr(a1.dat), dptr(a2.dat), dptr(a3.dat),
                                                              dptr(a4.dat), nullptr);
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
        k_stream_gather<<<nblk(P.n_targets), kBlock, 0, cur_stream()>>>(P.n_targets, P.d_targets, P.d_ptr, P.d_inc2, wp,
                                                                    dptr(a0.dat), dptr(a3.dat));
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
        k_residual<<<nblk(set->size), kBlock, 0, cur_stream()>>>(set->size, dptr(a0.dat), dptr(a1.dat), dptr(a2.dat), nullptr, nullptr, nullptr);
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
    k_sumsq_part<<<nb, kBlock, 0, cur_stream()>>>(n, dptr(a0.dat), R.d_part);
    k_sum_final<<<1, kSumBlock, 0, cur_stream()>>>(nb, R.d_part, init, R.d_scal + 1, n, nullptr, nullptr, nullptr, nullptr);
    TRY(check_launch());
    double s;
    TRY(read_scalar(R.d_scal + 1, &s));
    std::memcpy(a1.data, &s, sizeof(double));
    return 0;
}

static int count_loop(int n, op_dat d, int* out, int which)
{
    k_fill_i<<<1, kBlock, 0, cur_stream()>>>(1, R.d_iscal, 0);
    if (n) {
        if (which == 0) k_count_bad<<<nblk(n), kBlock, 0, cur_stream()>>>(n, dptr(d), R.d_iscal);
        else k_count_pos<<<nblk(n), kBlock, 0, cur_stream()>>>(n, dptr(d), R.d_iscal);
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
        k_up_pre<<<nblk(set->size), kBlock, 0, cur_stream()>>>(set->size, M(a0.map)->d_map, dptr(a0.dat),
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
        k_up_gather<<<nblk(P.n_targets), kBlock, 0, cur_stream()>>>(P.n_targets, P.d_targets, P.d_ptr, P.d_inc1,
                                                                dptr(a0.dat), dptr(a1.dat), iptr(a2.dat), 0, nullptr);
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
        k_up_post<<<nblk(set->size), kBlock, 0, cur_stream()>>>(set->size, dptr(a0.dat), iptr(a1.dat), nullptr);
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
        k_down<<<nblk(set->size), kBlock, 0, cur_stream()>>>(set->size, M(a3.map)->d_map, dptr(a0.dat), dptr(a1.dat),
                                                         dptr(a2.dat), dptr(a3.dat), dptr(a4.dat), nullptr);
    touch(a0.dat);
    return check_launch();
}

int op_par_loop_down_v2_kernel_pre(const char* name, op_set set, op_arg a0, op_arg a1)
{
    const char* L = "down_v2_kernel_pre";
    (void)name;
    TRY(chk_ready(L));
    TRY(direct(a0, set, 1, false, L, "arg0 prolonged weight sums OP_WRITE"));
    TRY(direct(a1, set, 5, false, L, "arg1 prolonged residuals[5] OP_WRITE"));
    const size_t n = (size_t)set->size;
    if (n) {
        k_fill<<<nblk(n), kBlock, 0, cur_stream()>>>(n, dptr(a0.dat), 0.0);
        k_fill<<<nblk(5 * n), kBlock, 0, cur_stream()>>>(5 * n, dptr(a1.dat), 0.0);
    }
    touch(a0.dat);
    touch(a1.dat);
    return check_launch();
}

int op_par_loop_down_v2_kernel(const char* name, op_set set, op_arg a0, op_arg a1, op_arg a2, op_arg a3,
                               op_arg a4, op_arg a5, op_arg a6, op_arg a7, op_arg a8, op_arg a9)
{
    const char* L = "down_v2_kernel";
    (void)name;
    TRY(chk_ready(L));
    TRY(indirect(a0, set, 0, 2, 3, L, "arg0 fine coords via edge-->node[0]"));
    TRY(indirect(a1, set, 1, 2, 3, L, "arg1 fine coords via edge-->node[1]"));
    TRY(indirect(a2, set, 0, 2, 3, L, "arg2 coarse coords via edge-->mg_nodes[0]"));
    TRY(indirect(a3, set, 1, 2, 3, L, "arg3 coarse coords via edge-->mg_nodes[1]"));
    TRY(indirect(a4, set, 0, 2, 5, L, "arg4 coarse residuals via edge-->mg_nodes[0]"));
    TRY(indirect(a5, set, 1, 2, 5, L, "arg5 coarse residuals via edge-->mg_nodes[1]"));
    TRY(indirect(a6, set, 0, 2, 5, L, "arg6 prolonged residuals via edge-->node[0] OP_INC"));
    TRY(indirect(a7, set, 1, 2, 5, L, "arg7 prolonged residuals via edge-->node[1] OP_INC"));
    TRY(indirect(a8, set, 0, 2, 1, L, "arg8 prolonged weight sums via edge-->node[0] OP_INC"));
    TRY(indirect(a9, set, 1, 2, 1, L, "arg9 prolonged weight sums via edge-->node[1] OP_INC"));
    const op_map en = a0.map, em = a2.map;
    TRY(need(a1.map == en && a6.map == en && a7.map == en && a8.map == en && a9.map == en, L,
             "fine arguments through one edge-->node map"));
    TRY(need(a3.map == em && a4.map == em && a5.map == em, L, "coarse arguments through one edge-->mg_nodes map"));
    TRY(need(a0.dat == a1.dat && a2.dat == a3.dat && a4.dat == a5.dat && a6.dat == a7.dat && a8.dat == a9.dat,
             L, "one dat per argument pair"));
    TRY(need(a6.acc == OP_INC && a7.acc == OP_INC && a8.acc == OP_INC && a9.acc == OP_INC, L,
             "prolonged sums OP_INC"));
    TRY(need(a6.dat != a4.dat && a6.dat != a0.dat && a6.dat != a2.dat, L, "prolonged residuals distinct from inputs"));
    TRY(build_inverse(en));
    const InvPlan& P = M(en)->inv;
    if (P.n_targets)
        k_down_v2<<<nblk(P.n_targets), kBlock, 0, cur_stream()>>>(P.n_targets, P.d_targets, P.d_ptr, P.d_inc2,
                                                               M(em)->d_map, dptr(a0.dat), dptr(a2.dat),
                                                               dptr(a4.dat), dptr(a6.dat), dptr(a8.dat));
    touch(a6.dat);
    touch(a8.dat);
    return check_launch();
}

int op_par_loop_down_v2_kernel_post(const char* name, op_set set, op_arg a0, op_arg a1, op_arg a2, op_arg a3)
{
    const char* L = "down_v2_kernel_post";
    (void)name;
    TRY(chk_ready(L));
    TRY(direct(a0, set, 5, false, L, "arg0 prolonged residuals[5] OP_READ"));
    TRY(direct(a1, set, 1, false, L, "arg1 prolonged weight sums OP_READ"));
    TRY(direct(a2, set, 5, false, L, "arg2 residuals[5] OP_READ"));
    TRY(direct(a3, set, 5, false, L, "arg3 variables[5] OP_INC"));
    if (set->size)
        k_down_v2_post<<<nblk(set->size), kBlock, 0, cur_stream()>>>(set->size, dptr(a0.dat), dptr(a1.dat),
                                                                 dptr(a2.dat), dptr(a3.dat));
    touch(a3.dat);
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
        k_identify_differences<<<nblk(set->size), kBlock, 0, cur_stream()>>>(set->size, dptr(a0.dat), dptr(a1.dat),
                                                                         dptr(a2.dat), nullptr);
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

int op_par_loop_by_name(const char* name, op_set set, int nargs, const op_arg* a)
{
    if (!name) return fail("op_par_loop: null loop name");
    struct Entry { const char* n; int k; void* f; };
    static const Entry tab[] = {
        {"initialize_variables_kernel", 1, (void*)op_par_loop_initialize_variables_kernel},
        {"zero_5d_array_kernel", 1, (void*)op_par_loop_zero_5d_array_kernel},
        {"zero_1d_array_kernel", 1, (void*)op_par_loop_zero_1d_array_kernel},
        {"calculate_cell_volumes", 5, (void*)op_par_loop_calculate_cell_volumes},
        {"dampen_ewt", 1, (void*)op_par_loop_dampen_ewt},
        {"copy_double_kernel", 2, (void*)op_par_loop_copy_double_kernel},
        {"calculate_dt_kernel", 3, (void*)op_par_loop_calculate_dt_kernel},
        {"get_min_dt_kernel", 2, (void*)op_par_loop_get_min_dt_kernel},
        {"compute_step_factor_kernel", 4, (void*)op_par_loop_compute_step_factor_kernel},
        {"compute_flux_edge_kernel", 5, (void*)op_par_loop_compute_flux_edge_kernel},
        {"compute_bnd_node_flux_kernel", 4, (void*)op_par_loop_compute_bnd_node_flux_kernel},
        {"time_step_kernel", 5, (void*)op_par_loop_time_step_kernel},
        {"unstructured_stream_kernel", 5, (void*)op_par_loop_unstructured_stream_kernel},
        {"residual_kernel", 3, (void*)op_par_loop_residual_kernel},
        {"calc_rms_kernel", 2, (void*)op_par_loop_calc_rms_kernel},
        {"count_bad_vals", 2, (void*)op_par_loop_count_bad_vals},
        {"up_pre_kernel", 2, (void*)op_par_loop_up_pre_kernel},
        {"up_kernel", 3, (void*)op_par_loop_up_kernel},
        {"up_post_kernel", 2, (void*)op_par_loop_up_post_kernel},
        {"down_kernel", 5, (void*)op_par_loop_down_kernel},
        {"down_v2_kernel_pre", 2, (void*)op_par_loop_down_v2_kernel_pre},
        {"down_v2_kernel", 10, (void*)op_par_loop_down_v2_kernel},
        {"down_v2_kernel_post", 4, (void*)op_par_loop_down_v2_kernel_post},
        {"identify_differences", 3, (void*)op_par_loop_identify_differences},
        {"count_non_zeros", 2, (void*)op_par_loop_count_non_zeros},
    };
    for (const Entry& e : tab) {
        if (std::strcmp(e.n, name) != 0) continue;
        if (nargs != e.k) return fail("op_par_loop %s: expected %d arguments, got %d", name, e.k, nargs);
        using F1 = int (*)(const char*, op_set, op_arg);
        using F2 = int (*)(const char*, op_set, op_arg, op_arg);
        using F3 = int (*)(const char*, op_set, op_arg, op_arg, op_arg);
        using F4 = int (*)(const char*, op_set, op_arg, op_arg, op_arg, op_arg);
        using F5 = int (*)(const char*, op_set, op_arg, op_arg, op_arg, op_arg, op_arg);
        using F10 = int (*)(const char*, op_set, op_arg, op_arg, op_arg, op_arg, op_arg, op_arg, op_arg, op_arg,
                            op_arg, op_arg);
        switch (nargs) {
            case 1: return ((F1)e.f)(name, set, a[0]);
            case 2: return ((F2)e.f)(name, set, a[0], a[1]);
            case 3: return ((F3)e.f)(name, set, a[0], a[1], a[2]);
            case 4: return ((F4)e.f)(name, set, a[0], a[1], a[2], a[3]);
            case 5: return ((F5)e.f)(name, set, a[0], a[1], a[2], a[3], a[4]);
            default: return ((F10)e.f)(name, set, a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7], a[8], a[9]);
        }
    }
    return fail("op_par_loop: no device implementation of loop '%s'", name);
}

}  // extern "C"

#include "mgrun.inc"
