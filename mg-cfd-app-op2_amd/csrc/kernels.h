// gfx950 device kernels for the MG-CFD hot path.
//
// Data layout in HBM (set up by runtime.hip):
//   dats         the reference's own AoS layout x[i * dim + v] (double or
//                int), in the runtime's internal element order; one gathered
//                node is one ~40-byte record, i.e. one or two cache lines
//   edge map     int2 per edge (internal ids)
//   weights      packed double4 {w0, w1, w2, |w|} per edge or per
//                incidence, rebuilt whenever the weight dat changes (|w| is
//                the reference's ewt, flux.h:48-50)
//   inverse map  CSR per target node: the incidences that INC into it, in
//                the reference's element order (OP2 sequential order)
//
// Kernels marked "exact" disable FMA contraction and keep the reference's
// operation order and true divisions, so they are bitwise equal to the
// sequential CPU evaluation.  "fast" kernels use one reciprocal of the
// density, refined hardware reciprocal and square-root estimates, and FMA
// contraction.
#pragma once

#include <hip/hip_runtime.h>
#include <cfloat>
#include <cmath>
#include <cstdint>

namespace mgcfd {

constexpr double kGamma = 1.4;
constexpr int kBlock = 256;
constexpr int kWave = 64;
constexpr int kPart = 256;       // nodes per flux workgroup (one mesh part)
constexpr int kRec = 12;         // doubles per staged node record in LDS

struct Consts {
    double sc;          // smoothing_coefficient
    double ffv[5];      // ff_variable
    double ffmx[3], ffmy[3], ffmz[3], ffde[3];
};

// Flux-tensor and thermodynamic state of one node (inlined_funcs.h:70-129).
struct Th {
    double rho, mx, my, mz, E;
    double ux, uy, uz;
    double speed, c, p;
    double fxx, fxy, fxz, fyy, fyz, fzz;   // momentum flux (symmetric)
    double fdx, fdy, fdz;                  // energy flux
};

// Flux tensor from velocity, momentum and pressure (inlined_funcs.h:70-96).
__device__ __forceinline__ void flux_tensor_exact(Th& t)
{
#pragma clang fp contract(off)
    t.fxx = t.ux * t.mx + t.p;
    t.fxy = t.ux * t.my;
    t.fxz = t.ux * t.mz;
    t.fyy = t.uy * t.my + t.p;
    t.fyz = t.uy * t.mz;
    t.fzz = t.uz * t.mz + t.p;
    const double de_p = t.E + t.p;
    t.fdx = t.ux * de_p;
    t.fdy = t.uy * de_p;
    t.fdz = t.uz * de_p;
}

__device__ __forceinline__ void flux_tensor_fast(Th& t)
{
    t.fxx = t.ux * t.mx + t.p;
    t.fxy = t.ux * t.my;
    t.fxz = t.ux * t.mz;
    t.fyy = t.uy * t.my + t.p;
    t.fyz = t.uy * t.mz;
    t.fzz = t.uz * t.mz + t.p;
    const double de_p = t.E + t.p;
    t.fdx = t.ux * de_p;
    t.fdy = t.uy * de_p;
    t.fdz = t.uz * de_p;
}

template <bool EXACT>
__device__ __forceinline__ void flux_tensor(Th& t)
{
    if constexpr (EXACT) flux_tensor_exact(t);
    else flux_tensor_fast(t);
}

// Reciprocal and square root of the fast kernels (not the correctly rounded
// library expansions).  FP64: the hardware estimates (v_rcp_f64, v_rsq_f64:
// relative error about 2^-24, measured over 2^20 inputs spanning 2^+-20)
// refined by one Newton-Raphson step each -- 1/x: r += r (1 - x r); sqrt z
// from y = 1/sqrt z: g = z y, g += (z - g g) y / 2 -- which leaves about
// 2^-47 relative (the tests hold the fluxes to 1e-12).  The square root is
// taken of z = x + 2^-1000, which changes no x above 2^-947 and keeps x = 0
// finite (2^-500).  FP32: the hardware instructions (1 ulp).
__device__ __forceinline__ double part_rcp(double x)
{
    const double r = __builtin_amdgcn_rcp(x);
    return fma(r, fma(-x, r, 1.0), r);
}
__device__ __forceinline__ double part_sqrt(double x)
{
    const double z = x + 0x1p-1000;
    const double y = __builtin_amdgcn_rsq(z);
    const double g = z * y;
    return fma(fma(-g, g, z), 0.5 * y, g);
}
__device__ __forceinline__ float part_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float part_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

// Reference order: u = m / rho, |u|^2, p, c = sqrt(gamma p / rho).
__device__ __forceinline__ void thermo_exact(const double* U, Th& t)
{
#pragma clang fp contract(off)
    t.rho = U[0]; t.mx = U[1]; t.my = U[2]; t.mz = U[3]; t.E = U[4];
    t.ux = t.mx / t.rho;
    t.uy = t.my / t.rho;
    t.uz = t.mz / t.rho;
    const double ssq = t.ux * t.ux + t.uy * t.uy + t.uz * t.uz;
    t.speed = sqrt(ssq);
    t.p = (kGamma - 1.0) * (t.E - 0.5 * t.rho * ssq);
    t.c = sqrt(kGamma * t.p / t.rho);
    flux_tensor_exact(t);
}

__device__ __forceinline__ void thermo_fast(const double* U, Th& t)
{
    t.rho = U[0]; t.mx = U[1]; t.my = U[2]; t.mz = U[3]; t.E = U[4];
    const double inv = part_rcp(t.rho);
    t.ux = t.mx * inv;
    t.uy = t.my * inv;
    t.uz = t.mz * inv;
    const double ssq = t.ux * t.ux + t.uy * t.uy + t.uz * t.uz;
    t.speed = part_sqrt(ssq);
    t.p = (kGamma - 1.0) * (t.E - 0.5 * t.rho * ssq);
    t.c = part_sqrt(kGamma * t.p * inv);
    flux_tensor_fast(t);
}

template <bool EXACT>
__device__ __forceinline__ void thermo(const double* U, Th& t)
{
    if constexpr (EXACT) thermo_exact(U, t);
    else thermo_fast(U, t);
}

__device__ __forceinline__ void load5(const double* __restrict__ x, int i, double U[5])
{
    const double* p = x + (size_t)i * 5;
#pragma unroll
    for (int v = 0; v < 5; v++) U[v] = p[v];
}

__device__ __forceinline__ void store5(double* __restrict__ x, int i, const double U[5])
{
    double* p = x + (size_t)i * 5;
#pragma unroll
    for (int v = 0; v < 5; v++) p[v] = U[v];
}

// Increment that edge (own, oth) adds to node 'own' (flux.h:139-207).  The
// reference computes factor_a and factor_b with the same operand order taken
// from the endpoint's own side; with fx = -0.5 w (or +0.5 w on the b side,
// whose terms the reference subtracts) one formula serves both roles exactly.
__device__ __forceinline__ void edge_increment_exact(const Th& o, const Th& t, double ewt, double sc,
                                                     double fx, double fy, double fz, double inc[5])
{
#pragma clang fp contract(off)
    const double factor = -ewt * sc * 0.5 * (o.speed + t.speed + o.c + t.c);
    inc[0] = factor * (o.rho - t.rho) + fx * (o.mx + t.mx) + fy * (o.my + t.my) + fz * (o.mz + t.mz);
    inc[4] = factor * (o.E - t.E) + fx * (o.fdx + t.fdx) + fy * (o.fdy + t.fdy) + fz * (o.fdz + t.fdz);
    inc[1] = factor * (o.mx - t.mx) + fx * (o.fxx + t.fxx) + fy * (o.fxy + t.fxy) + fz * (o.fxz + t.fxz);
    inc[2] = factor * (o.my - t.my) + fx * (o.fxy + t.fxy) + fy * (o.fyy + t.fyy) + fz * (o.fyz + t.fyz);
    inc[3] = factor * (o.mz - t.mz) + fx * (o.fxz + t.fxz) + fy * (o.fyz + t.fyz) + fz * (o.fzz + t.fzz);
}

__device__ __forceinline__ void edge_increment_fast(const Th& o, const Th& t, double ewt, double sc,
                                                    double fx, double fy, double fz, double inc[5])
{
    const double factor = -ewt * sc * 0.5 * (o.speed + t.speed + o.c + t.c);
    inc[0] = factor * (o.rho - t.rho) + fx * (o.mx + t.mx) + fy * (o.my + t.my) + fz * (o.mz + t.mz);
    inc[4] = factor * (o.E - t.E) + fx * (o.fdx + t.fdx) + fy * (o.fdy + t.fdy) + fz * (o.fdz + t.fdz);
    inc[1] = factor * (o.mx - t.mx) + fx * (o.fxx + t.fxx) + fy * (o.fxy + t.fxy) + fz * (o.fxz + t.fxz);
    inc[2] = factor * (o.my - t.my) + fx * (o.fxy + t.fxy) + fy * (o.fyy + t.fyy) + fz * (o.fyz + t.fyz);
    inc[3] = factor * (o.mz - t.mz) + fx * (o.fxz + t.fxz) + fy * (o.fyz + t.fyz) + fz * (o.fzz + t.fzz);
}

template <bool EXACT>
__device__ __forceinline__ void edge_increment(const Th& o, const Th& t, double ewt, double sc, double fx,
                                               double fy, double fz, double inc[5])
{
    if constexpr (EXACT) edge_increment_exact(o, t, ewt, sc, fx, fy, fz, inc);
    else edge_increment_fast(o, t, ewt, sc, fx, fy, fz, inc);
}

__device__ __forceinline__ double add_exact(double a, double b)
{
#pragma clang fp contract(off)
    return a + b;
}

// Workgroup index remapped so that each of the 8 XCDs (which receive
// workgroups round-robin) walks one contiguous slice of the iteration space
// and keeps its neighbours in its own L2.
__device__ __forceinline__ int xcd_block()
{
    const int nb = gridDim.x, b = blockIdx.x;
    const int per = nb / 8, rem = nb % 8;
    const int x = b % 8, k = b / 8;
    // XCD x's slice starts at x * per + min(x, rem) (the first rem slices
    // hold one more); no branch, so the kernel's first loads issue together
    return x * per + min(x, rem) + k;
}

__device__ __forceinline__ bool halted(const int* status) { return status && *status != 0; }

// base[i], or 0 when base is null, read by one load with no branch around it
// (a null base reads a zero word).  Kernels that test their context's status
// late (before their first global write, not at entry) read it this way with
// their first loads, so its round trip overlaps theirs instead of preceding
// them.
__device__ int kZeroWord = 0;     // never written
__device__ __forceinline__ int word_or_zero(const int* base, int i) { return *(base ? base + i : &kZeroWord); }
__device__ __forceinline__ int status_of(const int* status) { return word_or_zero(status, 0); }

// ------------------------------------------------------------------------
// Edge flux, blocked gather (default).  Workgroup b owns the nodes
// [bnode.x, bnode.x + bnode.y) of one compact mesh part; inside a part the
// nodes are ordered by decreasing degree.
//  1. Stage: thread t evaluates the thermodynamic state of owned node t and
//     of halo node t (neighbours owned by other parts, at most kPart per
//     block) and stores both as kRec-double records in LDS.
//  2. Walk: thread t adds, in the reference's edge order, the increment of
//     each of its node's edges, reading the other endpoint from LDS.  The
//     k-th incidences of all threads of a block are stored contiguously
//     (jagged-diagonal order: jbase + jpref[k] + t), so the per-incidence
//     loads {LDS slot | b-role << 31} and the weight components w0 | w1 | w2
//     (SoA) are coalesced.  They are double-buffered in registers: chunks
//     of CH incidences, the first in flight with the staging loads, chunk
//     c+1 loading while chunk c is evaluated.  |w| is recomputed with the
//     reference's expression (flux.h:48-50).
// Every node sums its edges in OP2's sequential order; with EXACT (no FMA
// contraction, true divisions) the result is bitwise the reference's.
// Deterministic, no atomics, one launch, one barrier.
// ------------------------------------------------------------------------
template <bool EXACT>
__device__ __forceinline__ void stage_record(double* rec, const double U[5])
{
    Th th;
    thermo<EXACT>(U, th);
    double2* r = reinterpret_cast<double2*>(rec);
    r[0] = make_double2(th.rho, th.mx);
    r[1] = make_double2(th.my, th.mz);
    r[2] = make_double2(th.E, th.ux);
    r[3] = make_double2(th.uy, th.uz);
    r[4] = make_double2(th.p, th.speed);
    r[5] = make_double2(th.c, 0.0);
}

template <bool EXACT>
__device__ __forceinline__ void load_record(const double* rec, Th& t)
{
    const double2* r = reinterpret_cast<const double2*>(rec);
    const double2 a = r[0], b = r[1], c = r[2], d = r[3], e = r[4], f = r[5];
    t.rho = a.x; t.mx = a.y; t.my = b.x; t.mz = b.y; t.E = c.x; t.ux = c.y;
    t.uy = d.x; t.uz = d.y; t.p = e.x; t.speed = e.y; t.c = f.x;
    flux_tensor<EXACT>(t);
}

// CH incidences of one thread: {slot | role} codes and weight components.
template <int CH>
struct IncChunk {
    unsigned r[CH];
    double x0[CH], x1[CH], x2[CH];
};

template <int CH>
__device__ __forceinline__ void load_chunk(IncChunk<CH>& c, int q0, int deg, int jb, const int* __restrict__ jp,
                                           const unsigned* __restrict__ jslot, const double* __restrict__ jw,
                                           int nent)
{
#pragma unroll
    for (int i = 0; i < CH; i++) {
        const int q = q0 + i;
        c.r[i] = 0;
        c.x0[i] = c.x1[i] = c.x2[i] = 0.0;
        if (q < deg) {
            const int at = jb + jp[q];
            c.r[i] = jslot[at];
            c.x0[i] = jw[at];
            c.x1[i] = jw[nent + at];
            c.x2[i] = jw[2 * nent + at];
        }
    }
}

template <bool EXACT, int CH>
__device__ __forceinline__ void walk_chunk(const IncChunk<CH>& c, int q0, int deg, const double* sm, const Th& o,
                                           double sc, double acc[5])
{
#pragma unroll
    for (int i = 0; i < CH; i++) {
        if (q0 + i < deg) {
            const unsigned r = c.r[i];
            Th th;
            load_record<EXACT>(sm + (int)(r & 0x7fffffffu) * kRec, th);
            double ewt;
            {
#pragma clang fp contract(off)
                ewt = sqrt(c.x0[i] * c.x0[i] + c.x1[i] * c.x1[i] + c.x2[i] * c.x2[i]);
            }
            double fx = -0.5 * c.x0[i], fy = -0.5 * c.x1[i], fz = -0.5 * c.x2[i];
            if (r & 0x80000000u) { fx = -fx; fy = -fy; fz = -fz; }   // node is endpoint b
            double d[5];
            edge_increment<EXACT>(o, th, ewt, sc, fx, fy, fz, d);
            if constexpr (EXACT) {
#pragma unroll
                for (int v = 0; v < 5; v++) acc[v] = add_exact(acc[v], d[v]);
            } else {
#pragma unroll
                for (int v = 0; v < 5; v++) acc[v] += d[v];
            }
        }
    }
}

template <bool EXACT, int CH>
__global__ __launch_bounds__(kBlock) void k_flux_blocked_pf(
    const int2* __restrict__ bnode, const int* __restrict__ hptr, const int* __restrict__ halo,
    const int* __restrict__ jbase, const int* __restrict__ pbase, const int* __restrict__ jpref,
    const int* __restrict__ nptr, const unsigned* __restrict__ jslot, const double* __restrict__ jw, int nent,
    const double* __restrict__ var, double* __restrict__ flux, Consts k, const int* __restrict__ status, int fresh)
{
    extern __shared__ double sm[];
    if (halted(status)) return;
    const int b = xcd_block(), t = threadIdx.x;
    const int h0 = hptr[b], nh = hptr[b + 1] - h0;
    const int hn = (t < nh) ? halo[h0 + t] : -1;
    const int2 bn = bnode[b];
    const bool own = t < bn.y;
    const int nd = bn.x + t;
    int deg = 0;
    if (own) deg = nptr[nd + 1] - nptr[nd];
    const int jb = jbase[b] + t;
    const int* jp = jpref + pbase[b];
    IncChunk<CH> A, B;
    load_chunk<CH>(A, 0, deg, jb, jp, jslot, jw, nent);
    double uo[5], uh[5], acc[5];
    if (own) {
        load5(var, nd, uo);
        if (fresh) {
#pragma unroll
            for (int v = 0; v < 5; v++) acc[v] = 0.0;
        } else {
            load5(flux, nd, acc);
        }
    }
    if (hn >= 0) load5(var, hn, uh);
    if (own) stage_record<EXACT>(sm + t * kRec, uo);
    if (hn >= 0) stage_record<EXACT>(sm + (kPart + t) * kRec, uh);
    __syncthreads();
    if (!own) return;
    Th o;
    load_record<EXACT>(sm + t * kRec, o);
    for (int q0 = 0; q0 < deg; q0 += 2 * CH) {
        load_chunk<CH>(B, q0 + CH, deg, jb, jp, jslot, jw, nent);
        walk_chunk<EXACT, CH>(A, q0, deg, sm, o, k.sc, acc);
        if (q0 + CH >= deg) break;
        load_chunk<CH>(A, q0 + 2 * CH, deg, jb, jp, jslot, jw, nent);
        walk_chunk<EXACT, CH>(B, q0 + CH, deg, sm, o, k.sc, acc);
    }
    store5(flux, nd, acc);
}

// Per-incidence weight components (SoA) of k_flux_blocked_pf.
__global__ void k_pack_incidence_weights3(int n, const int* __restrict__ eid, const double* __restrict__ w,
                                          double* __restrict__ out)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int e = eid[k];
    double a = 0.0, b = 0.0, c = 0.0;
    if (e >= 0) { a = w[3 * (size_t)e]; b = w[3 * (size_t)e + 1]; c = w[3 * (size_t)e + 2]; }
    out[k] = a;
    out[n + k] = b;
    out[2 * (size_t)n + k] = c;
}

// ------------------------------------------------------------------------
// Edge flux, edge-parallel over mesh parts with node state and increments
// in LDS.  Workgroup b owns the nodes [n0, n0 + cnt) of one compact part and
// every edge with at least one endpoint among them ("local edges": interior
// edges once, cut edges once on each side).
//  1. Load (all loads issued before the barrier, flat and coalesced): owned
//     variables and halo variables as raw {rho, m, E} records in LDS, the
//     old flux of owned nodes into the LDS accumulators, and this thread's
//     local edges {slot a, slot b, owner flags, colour} with their weights.
//  2. Each thread evaluates its edges once: both endpoint states from their
//     LDS records (one density reciprocal each, fast form), then inc =
//     factor (U_a - U_b) + sum_k f_k (F_a + F_b)_k with f = -w/2; node a
//     gains +inc and node b gains -inc (flux.h:139-207 with factor_b ==
//     factor_a up to rounding).  Raw records keep the LDS footprint at
//     48 B per staged node, so more workgroups share a CU.
//  3. Accumulation into the LDS accumulators: COLOURED walks the edge
//     colours of the part (no two edges of one colour share an owned node),
//     one barrier per colour, so the summation order is fixed by the plan and
//     results are deterministic; otherwise LDS FP64 atomics (ds_add_f64) in
//     arrival order.
//  4. The accumulators are stored as one coalesced block.
// Not bitwise the sequential order (ulp-level drift, as the fast gather).
// ------------------------------------------------------------------------
constexpr int kPRec = 6;    // doubles per LDS node record of k_flux_part {rho, mx, my, mz, E, -}

struct PartInfo {
    int n0, cnt;    // owned nodes [n0, n0 + cnt)
    int h0, nh;     // halo ids halo[h0 .. h0 + nh)
    int e0, ne;     // local edges [e0, e0 + ne)
    int ncol, pad;  // edge colours (COLOURED)
};

template <typename R>
struct PState {
    R rho, mx, my, mz, E, ux, uy, uz, p, s;   // s = |u| + c
};

// Wave speed |u| + c of a node state, computed once per staged node and kept
// in its record (R arithmetic: FP64, or FP32 for mixed precision).
template <typename R>
__device__ __forceinline__ R part_speed(R rho, R mx, R my, R mz, R E)
{
    const R inv = part_rcp(rho);
    const R ux = mx * inv, uy = my * inv, uz = mz * inv;
    const R ssq = ux * ux + uy * uy + uz * uz;
    const R p = R(kGamma - 1.0) * (E - R(0.5) * rho * ssq);
    return part_sqrt(ssq) + part_sqrt(R(kGamma) * p * inv);
}

// An endpoint state from its record {rho, mx, my, mz, E, |u| + c}.
__device__ __forceinline__ PState<double> part_state(const double* r)
{
    const double2* q = reinterpret_cast<const double2*>(r);
    const double2 a = q[0], b = q[1], c = q[2];
    PState<double> t;
    t.rho = a.x; t.mx = a.y; t.my = b.x; t.mz = b.y; t.E = c.x; t.s = c.y;
    const double inv = part_rcp(t.rho);
    t.ux = t.mx * inv; t.uy = t.my * inv; t.uz = t.mz * inv;
    const double ssq = t.ux * t.ux + t.uy * t.uy + t.uz * t.uz;
    t.p = (kGamma - 1.0) * (t.E - 0.5 * t.rho * ssq);
    return t;
}

__device__ __forceinline__ PState<float> part_state32(const float* r)
{
    const float2* q = reinterpret_cast<const float2*>(r);
    const float2 a = q[0], b = q[1], c = q[2];
    PState<float> t;
    t.rho = a.x; t.mx = a.y; t.my = b.x; t.mz = b.y; t.E = c.x; t.s = c.y;
    const float inv = part_rcp(t.rho);
    t.ux = t.mx * inv; t.uy = t.my * inv; t.uz = t.mz * inv;
    const float ssq = t.ux * t.ux + t.uy * t.uy + t.uz * t.uz;
    t.p = (float)(kGamma - 1.0) * (t.E - 0.5f * t.rho * ssq);
    return t;
}

// Increment of node a for edge (a, b) in the fast form (FP64 or FP32
// arithmetic); node b gains -inc.
template <typename R, typename O>
__device__ __forceinline__ void part_increment(const PState<R>& a, const PState<R>& b, R w0, R w1, R w2, R sc,
                                               O inc[5])
{
    const R ewt = part_sqrt(w0 * w0 + w1 * w1 + w2 * w2);
    const R factor = -ewt * sc * R(0.5) * (a.s + b.s);
    const R fx = R(-0.5) * w0, fy = R(-0.5) * w1, fz = R(-0.5) * w2;
    const R dea = a.E + a.p, deb = b.E + b.p;
    // momentum flux tensors F_a + F_b (symmetric) and energy flux
    const R sxx = a.ux * a.mx + a.p + b.ux * b.mx + b.p;
    const R sxy = a.ux * a.my + b.ux * b.my;
    const R sxz = a.ux * a.mz + b.ux * b.mz;
    const R syy = a.uy * a.my + a.p + b.uy * b.my + b.p;
    const R syz = a.uy * a.mz + b.uy * b.mz;
    const R szz = a.uz * a.mz + a.p + b.uz * b.mz + b.p;
    const R sdx = a.ux * dea + b.ux * deb;
    const R sdy = a.uy * dea + b.uy * deb;
    const R sdz = a.uz * dea + b.uz * deb;
    inc[0] = (O)(factor * (a.rho - b.rho) + fx * (a.mx + b.mx) + fy * (a.my + b.my) + fz * (a.mz + b.mz));
    inc[1] = (O)(factor * (a.mx - b.mx) + fx * sxx + fy * sxy + fz * sxz);
    inc[2] = (O)(factor * (a.my - b.my) + fx * sxy + fy * syy + fz * syz);
    inc[3] = (O)(factor * (a.mz - b.mz) + fx * sxz + fy * syz + fz * szz);
    inc[4] = (O)(factor * (a.E - b.E) + fx * sdx + fy * sdy + fz * sdz);
}

// Adds inc to owned endpoint a and -inc to owned endpoint b (FP64 adds; an
// FP32 increment is converted first).
template <typename O>
__device__ __forceinline__ void part_apply(double* acc, unsigned code, const O inc[5])
{
    const int sa = code & 1023u, sb = (code >> 10) & 1023u;
    if (code & (1u << 20)) {
#pragma unroll
        for (int v = 0; v < 5; v++) atomicAdd(acc + sa * 5 + v, (double)inc[v]);
    }
    if (code & (1u << 21)) {
#pragma unroll
        for (int v = 0; v < 5; v++) atomicAdd(acc + sb * 5 + v, -(double)inc[v]);
    }
}

// LDS node records of the part kernels: the raw state {rho, m, E} and the
// wave speed |u| + c (6 values, FP64 or FP32); the rest of each endpoint
// state is derived per edge -- the small record lets more workgroups share a
// CU, which pays more than deriving each state once.
template <typename R>
struct PartRec;
template <>
struct PartRec<double> {
    static constexpr int kStride = 6;
    static __device__ __forceinline__ PState<double> state(const double* r) { return part_state(r); }
};
template <>
struct PartRec<float> {
    static constexpr int kStride = 6;
    static __device__ __forceinline__ PState<float> state(const float* r) { return part_state32(r); }
};

// An endpoint state from a record of RS values: the six-value records above,
// or (FP64, RS = 5) the raw {rho, m, E}, whose wave speed is formed here.
// The FP64 kernel takes 40-byte records when 48-byte ones would leave fewer
// than four workgroups per CU.
template <typename R, int RS>
__device__ __forceinline__ PState<R> part_rec_state(const R* r)
{
    if constexpr (RS == 5) {
        static_assert(sizeof(R) == sizeof(double), "40-byte records are FP64");
        PState<double> t;
        t.rho = r[0]; t.mx = r[1]; t.my = r[2]; t.mz = r[3]; t.E = r[4];
        const double inv = part_rcp(t.rho);
        t.ux = t.mx * inv; t.uy = t.my * inv; t.uz = t.mz * inv;
        const double ssq = t.ux * t.ux + t.uy * t.uy + t.uz * t.uz;
        t.p = (kGamma - 1.0) * (t.E - 0.5 * t.rho * ssq);
        t.s = part_sqrt(ssq) + part_sqrt(kGamma * t.p * inv);
        return t;
    } else {
        return PartRec<R>::state(r);
    }
}

// Marks a loaded register as needed here: the compiler waits for its load at
// this point and cannot sink the load into a later divergent branch (where it
// would be issued alone under its own wait).
template <typename V>
__device__ __forceinline__ void part_ready(const V& v)
{
    asm volatile("" ::"v"(v));
}

// Staging of one part, shared by the part kernels: node slot sl = t + k T
// (k < NPT, NPT * T >= rec_cap) of the owned nodes [0, cnt) and the halo
// nodes [cnt, cnt + nh).  The halo-list loads are unconditional (clamped into
// the part's list, which the plan pads by one entry) so that they issue back
// to back under one wait instead of one wait per divergent branch; every
// slot then names a valid node, so the state loads are unconditional too.
template <int T, int NPT>
__device__ __forceinline__ void part_halo_entries(const PartInfo& P, const int* __restrict__ halo, int t,
                                                  int nd[NPT])
{
    const int hmax = max(P.nh - 1, 0);
#pragma unroll
    for (int k = 0; k < NPT; k++) nd[k] = halo[P.h0 + min(max(t + k * T - P.cnt, 0), hmax)];
}

template <int T, int NPT>
__device__ __forceinline__ void part_load_nodes(const PartInfo& P, const double* __restrict__ var, int t,
                                                int nd[NPT], double U[NPT][5])
{
#pragma unroll
    for (int k = 0; k < NPT; k++) {
        part_ready(nd[k]);
        const int sl = t + k * T;
        nd[k] = sl < P.cnt ? P.n0 + sl : nd[k];
    }
#pragma unroll
    for (int k = 0; k < NPT; k++)
#pragma unroll
        for (int v = 0; v < 5; v++) U[k][v] = var[(size_t)nd[k] * 5 + v];
}

// Records {rho, m, E, |u| + c} of the staged slots.
template <typename R, int T, int NPT, int RS>
__device__ __forceinline__ void part_store_records(const PartInfo& P, int t, const double U[NPT][5], R* rec)
{
    const int nrec = P.cnt + P.nh;
#pragma unroll
    for (int k = 0; k < NPT; k++) {
        const int sl = t + k * T;
        if (sl < nrec) {
            const R x0 = (R)U[k][0], x1 = (R)U[k][1], x2 = (R)U[k][2], x3 = (R)U[k][3], x4 = (R)U[k][4];
            R* r = rec + sl * RS;
            r[0] = x0; r[1] = x1; r[2] = x2; r[3] = x3; r[4] = x4;
            if constexpr (RS > 5) r[5] = part_speed<R>(x0, x1, x2, x3, x4);
        }
    }
}

// The owned nodes' old flux (FL flat entries per thread), or zeros for a
// fresh start, into the FP64 accumulators.
template <int T, int FL>
__device__ __forceinline__ void part_load_flux(const PartInfo& P, const double* __restrict__ flux, int t,
                                               bool fresh, double fl[FL])
{
    const double* fo = flux + (size_t)P.n0 * 5;
#pragma unroll
    for (int i = 0; i < FL; i++) fl[i] = fresh ? 0.0 : fo[min(t + i * T, P.cnt * 5 - 1)];
}

template <int T, int FL>
__device__ __forceinline__ void part_store_flux(const PartInfo& P, int t, const double fl[FL], double* acc)
{
#pragma unroll
    for (int i = 0; i < FL; i++) {
        const int j = t + i * T;
        if (j < P.cnt * 5) acc[j] = fl[i];
    }
}

// Local-edge codes and weights of one part, EPT per thread (code 0: none).
template <typename R, int T, int EPT>
__device__ __forceinline__ void part_load_edges(const PartInfo& P, const unsigned* __restrict__ ecode,
                                                const R* __restrict__ ew, long long nloc, int t, unsigned code[EPT],
                                                R w0[EPT], R w1[EPT], R w2[EPT])
{
#pragma unroll
    for (int i = 0; i < EPT; i++) {
        const int j = t + i * T;
        code[i] = 0;
        w0[i] = w1[i] = w2[i] = R(0);
        if (j < P.ne) {
            const long long e = P.e0 + j;
            code[i] = ecode[e];
            w0[i] = ew[e];
            w1[i] = ew[nloc + e];
            w2[i] = ew[2 * nloc + e];
        }
    }
}

// The part's local edges, each increment added to its owned endpoints with
// LDS atomics.
template <typename R, int EPT, int RS>
__device__ __forceinline__ void part_edges_atomic(const R* rec, double* acc, const unsigned code[EPT],
                                                  const R w0[EPT], const R w1[EPT], const R w2[EPT], R sc)
{
#pragma unroll
    for (int i = 0; i < EPT; i++) {
        if (!code[i]) continue;
        const PState<R> a = part_rec_state<R, RS>(rec + (code[i] & 1023u) * RS);
        const PState<R> b = part_rec_state<R, RS>(rec + ((code[i] >> 10) & 1023u) * RS);
        R inc[5];
        part_increment<R, R>(a, b, w0[i], w1[i], w2[i], sc, inc);
        part_apply(acc, code[i], inc);
    }
}

// Stages one part and accumulates its local edges: returns the FP64
// accumulators (in LDS, after a barrier) holding the edge flux of the part's
// owned nodes, started from the flux dat or, with fresh, from zero.
// LAZY (coloured only): form each edge's increment in its colour's phase
// instead of all of them up front, so no increments are held across the
// phases (fewer registers, for the kernel with the fused stage).
template <typename R, int T, int EPT, int NPT, bool COLOURED, int RS, bool LAZY>
__device__ __forceinline__ double* part_accumulate(const PartInfo& P, const int* __restrict__ halo,
                                                   const unsigned* __restrict__ ecode, const R* __restrict__ ew,
                                                   long long nloc, int rec_cap, const double* __restrict__ var,
                                                   const double* __restrict__ flux, const Consts& k, int fresh,
                                                   double* sm)
{
    constexpr int FL = (5 * kPart + T - 1) / T;
    const int t = threadIdx.x;
    R* rec = reinterpret_cast<R*>(sm);
    double* acc = sm + ((size_t)rec_cap * RS * sizeof(R) + sizeof(double) - 1) / sizeof(double);
    // 1. loads: halo entries, edge codes and weights, then node states and
    // the owned flux; records (with |u| + c) and accumulators into LDS
    int nd[NPT];
    unsigned code[EPT];
    R w0[EPT], w1[EPT], w2[EPT];
    double U[NPT][5], fl[FL];
    part_halo_entries<T, NPT>(P, halo, t, nd);
    part_load_edges<R, T, EPT>(P, ecode, ew, nloc, t, code, w0, w1, w2);
    part_load_nodes<T, NPT>(P, var, t, nd, U);
    part_load_flux<T, FL>(P, flux, t, fresh != 0, fl);
    part_store_records<R, T, NPT, RS>(P, t, U, rec);
    part_store_flux<T, FL>(P, t, fl, acc);
    __syncthreads();
    // 2./3. edges
    const R sc = (R)k.sc;
    if constexpr (COLOURED && LAZY) {
        for (int c = 0; c < P.ncol; c++) {
#pragma unroll
            for (int i = 0; i < EPT; i++)
                if (code[i] && (int)(code[i] >> 22) == c) {
                    const PState<R> a = part_rec_state<R, RS>(rec + (code[i] & 1023u) * RS);
                    const PState<R> b = part_rec_state<R, RS>(rec + ((code[i] >> 10) & 1023u) * RS);
                    R inc[5];
                    part_increment<R, R>(a, b, w0[i], w1[i], w2[i], sc, inc);
                    part_apply(acc, code[i], inc);
                }
            __syncthreads();
        }
    } else if constexpr (COLOURED) {
        R inc[EPT][5];      // held in the arithmetic's precision until the colour walk
#pragma unroll
        for (int i = 0; i < EPT; i++)
            if (code[i]) {
                const PState<R> a = part_rec_state<R, RS>(rec + (code[i] & 1023u) * RS);
                const PState<R> b = part_rec_state<R, RS>(rec + ((code[i] >> 10) & 1023u) * RS);
                part_increment<R, R>(a, b, w0[i], w1[i], w2[i], sc, inc[i]);
            }
        for (int c = 0; c < P.ncol; c++) {
#pragma unroll
            for (int i = 0; i < EPT; i++)
                if (code[i] && (int)(code[i] >> 22) == c) part_apply(acc, code[i], inc[i]);
            __syncthreads();
        }
    } else {
        part_edges_atomic<R, EPT, RS>(rec, acc, code, w0, w1, w2, sc);
        __syncthreads();
    }
    return acc;
}

// LDS: R records [rec_cap * 6] (rounded up to doubles), then FP64
// accumulators [cnt * 5].  R = double: the FP64 kernel; R = float: mixed
// precision (records, weights and increment arithmetic in FP32, each
// increment converted to FP64 and accumulated in FP64 from the FP64 flux).
// One workgroup per part; NPT staged nodes per thread (NPT * T >= rec_cap);
// RS values per record.
template <typename R, int T, int EPT, int NPT, bool COLOURED, int RS>
__global__ __launch_bounds__(T) void k_flux_part(
    const PartInfo* __restrict__ parts, const int* __restrict__ halo, const unsigned* __restrict__ ecode,
    const R* __restrict__ ew, long long nloc, int rec_cap, const double* __restrict__ var,
    double* __restrict__ flux, Consts k, const int* __restrict__ status, int fresh)
{
    extern __shared__ double sm[];
    const int halt = status_of(status);     // tested before the store
    const PartInfo P = parts[xcd_block()];
    const double* acc =
        part_accumulate<R, T, EPT, NPT, COLOURED, RS, false>(P, halo, ecode, ew, nloc, rec_cap, var, flux, k, fresh,
                                                             sm);
    if (halt) return;
    // 4. store
    double* fw = flux + (size_t)P.n0 * 5;
    for (int j = threadIdx.x; j < P.cnt * 5; j += T) fw[j] = acc[j];
}

// Local-edge weights of the mixed-precision part kernel, in FP32.
__global__ void k_pack_part_weights32(long long n, const int* __restrict__ eid, const double* __restrict__ w,
                                      float* __restrict__ out)
{
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const size_t e = (size_t)eid[k];
    out[k] = (float)w[3 * e];
    out[n + k] = (float)w[3 * e + 1];
    out[2 * n + k] = (float)w[3 * e + 2];
}

// Local-edge weights of k_flux_part, gathered from the weight dat into
// three component arrays.
__global__ void k_pack_part_weights(long long n, const int* __restrict__ eid, const double* __restrict__ w,
                                    double* __restrict__ out)
{
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const size_t e = (size_t)eid[k];
    out[k] = w[3 * e];
    out[n + k] = w[3 * e + 1];
    out[2 * n + k] = w[3 * e + 2];
}

// ------------------------------------------------------------------------
// Edge flux, node-centric gather without LDS: one thread per node, every
// incidence re-evaluates the neighbour's state from global memory.
// ------------------------------------------------------------------------
template <bool EXACT>
__global__ __launch_bounds__(kBlock) void k_flux_gather(
    int n_targets, const int* __restrict__ targets, const int* __restrict__ ptr,
    const int2* __restrict__ inc, const double4* __restrict__ wpk,
    const double* __restrict__ var, double* __restrict__ flux, Consts k,
    const int* __restrict__ status, int fresh)
{
    if (halted(status)) return;
    const int tid = xcd_block() * kBlock + threadIdx.x;
    if (tid >= n_targets) return;
    const int n = targets ? targets[tid] : tid;
    double U[5], acc[5];
    load5(var, n, U);
    if (fresh) {
#pragma unroll
        for (int v = 0; v < 5; v++) acc[v] = 0.0;
    } else {
        load5(flux, n, acc);
    }
    Th o;
    thermo<EXACT>(U, o);
    const int beg = ptr[tid], end = ptr[tid + 1];
    for (int q = beg; q < end; q++) {
        const int2 r = inc[q];
        const double4 w = wpk[r.y & 0x7fffffff];
        double X[5];
        load5(var, r.x, X);
        Th t;
        thermo<EXACT>(X, t);
        double fx = -0.5 * w.x, fy = -0.5 * w.y, fz = -0.5 * w.z;
        if (r.y < 0) { fx = -fx; fy = -fy; fz = -fz; }   // node is endpoint b
        double d[5];
        edge_increment<EXACT>(o, t, w.w, k.sc, fx, fy, fz, d);
        if constexpr (EXACT) {
#pragma unroll
            for (int v = 0; v < 5; v++) acc[v] = add_exact(acc[v], d[v]);
        } else {
#pragma unroll
            for (int v = 0; v < 5; v++) acc[v] += d[v];
        }
    }
    store5(flux, n, acc);
}

// Edge flux, edge-parallel with FP64 hardware atomics (global_atomic_add_f64).
__global__ __launch_bounds__(kBlock) void k_flux_atomic(
    int n_edges, const int2* __restrict__ e2n, const double4* __restrict__ wpk,
    const double* __restrict__ var, double* __restrict__ flux, Consts k, const int* __restrict__ status)
{
    if (halted(status)) return;
    const int e = xcd_block() * kBlock + threadIdx.x;
    if (e >= n_edges) return;
    const int2 ab = e2n[e];
    const double4 w = wpk[e];
    double Ua[5], Ub[5];
    load5(var, ab.x, Ua);
    load5(var, ab.y, Ub);
    Th a, b;
    thermo_fast(Ua, a);
    thermo_fast(Ub, b);
    const double fx = -0.5 * w.x, fy = -0.5 * w.y, fz = -0.5 * w.z;
    double da[5], db[5];
    edge_increment_fast(a, b, w.w, k.sc, fx, fy, fz, da);
    edge_increment_fast(b, a, w.w, k.sc, -fx, -fy, -fz, db);
#pragma unroll
    for (int v = 0; v < 5; v++) {
        unsafeAtomicAdd(&flux[5 * (size_t)ab.x + v], da[v]);
        unsafeAtomicAdd(&flux[5 * (size_t)ab.y + v], db[v]);
    }
}

// unstructured_stream.h:7-57 as a gather: same data movement as the flux
// kernel with (almost) no arithmetic.
__global__ __launch_bounds__(kBlock) void k_stream_gather(
    int n_targets, const int* __restrict__ targets, const int* __restrict__ ptr,
    const int2* __restrict__ inc, const double4* __restrict__ wpk,
    const double* __restrict__ var, double* __restrict__ flux)
{
#pragma clang fp contract(off)
    const int tid = xcd_block() * kBlock + threadIdx.x;
    if (tid >= n_targets) return;
    const int n = targets ? targets[tid] : tid;
    double acc[5];
    load5(flux, n, acc);
    for (int q = ptr[tid]; q < ptr[tid + 1]; q++) {
        const int2 r = inc[q];
        double u[5];
        load5(var, r.x, u);
        if (r.y >= 0) {     // node is a: fluxes_a += b-values (+ weights)
            const double4 w = wpk[r.y];
            acc[0] += u[0] + w.x;
            acc[1] += u[1] + w.z;
            acc[2] += u[2];
            acc[3] += u[3];
            acc[4] += u[4] + w.y;
        } else {            // node is b: fluxes_b += a-values
#pragma unroll
            for (int v = 0; v < 5; v++) acc[v] += u[v];
        }
    }
    store5(flux, n, acc);
}

// Plain device-to-device copy (16-byte elements, grid-stride): the
// calibration kernel for bandwidth and for the HBM traffic counters.
__global__ __launch_bounds__(kBlock) void k_stream_copy(size_t n, const double2* __restrict__ a,
                                                        double2* __restrict__ b)
{
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) b[i] = a[i];
}

// Packed edge weights {w0, w1, w2, |w|} (flux.h:48-50).
__global__ void k_pack_weights(int n, const double* __restrict__ w, double4* __restrict__ out)
{
#pragma clang fp contract(off)
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const double a = w[3 * (size_t)e], b = w[3 * (size_t)e + 1], c = w[3 * (size_t)e + 2];
    out[e] = make_double4(a, b, c, sqrt(a * a + b * b + c * c));
}

// ------------------------------------------------------------------------
// Boundary / far-field flux (flux.h:14-39): one thread per boundary node,
// entries in boundary-element order.
// ------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_bnd_flux(
    int n_targets, const int* __restrict__ targets, const int* __restrict__ ptr,
    const int* __restrict__ ent, const int* __restrict__ group, const double* __restrict__ bw,
    const double* __restrict__ var, double* __restrict__ flux, Consts k, const int* __restrict__ status)
{
#pragma clang fp contract(off)
    if (halted(status)) return;
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= n_targets) return;
    const int n = targets[tid];
    double U[5], acc[5];
    load5(var, n, U);
    load5(flux, n, acc);
    Th b;
    thermo_exact(U, b);
    for (int q = ptr[tid]; q < ptr[tid + 1]; q++) {
        const int e = ent[q];
        const int g = group[e];
        const double w0 = bw[3 * (size_t)e], w1 = bw[3 * (size_t)e + 1], w2 = bw[3 * (size_t)e + 2];
        if (g <= 2) {
            // flux_boundary.elem_func:50-54
            acc[0] = acc[0] + 0.0;
            acc[1] = acc[1] + w0 * b.p;
            acc[2] = acc[2] + w1 * b.p;
            acc[3] = acc[3] + w2 * b.p;
            acc[4] = acc[4] + 0.0;
        } else if (g <= 7) {
            // flux_wall.elem_func:103-130
            const double fx = 0.5 * w0, fy = 0.5 * w1, fz = 0.5 * w2;
            acc[0] += fx * (k.ffv[1] + b.mx) + fy * (k.ffv[2] + b.my) + fz * (k.ffv[3] + b.mz);
            acc[4] += fx * (k.ffde[0] + b.fdx) + fy * (k.ffde[1] + b.fdy) + fz * (k.ffde[2] + b.fdz);
            acc[1] += fx * (k.ffmx[0] + b.fxx) + fy * (k.ffmx[1] + b.fxy) + fz * (k.ffmx[2] + b.fxz);
            acc[2] += fx * (k.ffmy[0] + b.fxy) + fy * (k.ffmy[1] + b.fyy) + fz * (k.ffmy[2] + b.fyz);
            acc[3] += fx * (k.ffmz[0] + b.fxz) + fy * (k.ffmz[1] + b.fyz) + fz * (k.ffmz[2] + b.fzz);
        }
    }
    store5(flux, n, acc);
}

// ------------------------------------------------------------------------
// Element-wise kernels
// ------------------------------------------------------------------------
__global__ void k_fill5(int n, double* __restrict__ x, Consts k, int use_ff)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
#pragma unroll
    for (int v = 0; v < 5; v++) x[5 * (size_t)i + v] = use_ff ? k.ffv[v] : 0.0;
}

__global__ void k_fill(size_t n, double* __restrict__ x, double val)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = val;
}

__global__ void k_fill_u64(size_t n, unsigned long long* __restrict__ x, unsigned long long val)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = val;
}

__global__ void k_fill_i(size_t n, int* __restrict__ x, int val)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = val;
}

__global__ void k_scale(size_t n, double* __restrict__ w, double s)
{
#pragma clang fp contract(off)
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) w[i] *= s;
}

__global__ void k_copy(size_t n, const double* __restrict__ src, double* __restrict__ dst,
                       const int* __restrict__ status)
{
    if (halted(status)) return;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

// Block-wide reductions over a 256-thread workgroup (4 waves of 64).
template <int NT = kBlock>
__device__ __forceinline__ double block_min(double x)
{
    __shared__ double sm[NT / kWave];
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) {
        const double y = __shfl_down(x, o, kWave);
        x = (y < x) ? y : x;
    }
    const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
    if (lane == 0) sm[w] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        x = sm[0];
        for (int i = 1; i < NT / kWave; i++) x = (sm[i] < x) ? sm[i] : x;
    }
    return x;
}

// block_min with the result in every thread.
template <int NT = kBlock>
__device__ __forceinline__ double block_min_all(double x)
{
    __shared__ double bc;
    x = block_min<NT>(x);
    if (threadIdx.x == 0) bc = x;
    __syncthreads();
    return bc;
}

template <int NT = kBlock>
__device__ __forceinline__ double block_sum(double x)
{
#pragma clang fp contract(off)
    __shared__ double sm[NT / kWave];
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) x += __shfl_down(x, o, kWave);
    const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
    if (lane == 0) sm[w] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        x = sm[0];
        for (int i = 1; i < NT / kWave; i++) x += sm[i];
    }
    return x;
}

template <int NT = kBlock>
__device__ __forceinline__ int block_sum_i(int x)
{
    __shared__ int smi[NT / kWave];
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) x += __shfl_down(x, o, kWave);
    const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
    if (lane == 0) smi[w] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        x = smi[0];
        for (int i = 1; i < NT / kWave; i++) x += smi[i];
    }
    return x;
}

// Publication of a level's minimum dt by the kernels that compute dt
// (get_min_dt_kernel, time_stepping_kernels.h:34-41, fused into them): every
// workgroup folds its minimum into one of kMinSlots slots of its mesh with a
// 64-bit atomic min on an order-preserving integer image of the double, so
// the first Runge-Kutta stage reduces kMinSlots values instead of one per
// workgroup of the level.  The atomics need no fence (the stage is the next
// kernel on the stream), and spreading the workgroups over slots keeps them
// off one address.  The second stage resets the slots to the empty value
// for the next smoothing step.  min is exact: the order of the atomics does
// not change the result.
constexpr int kMinSlots = 64;
constexpr unsigned long long kMinEmpty = 0xFFEFFFFFFFFFFFFFull;    // min_key(DBL_MAX)

// Order-preserving image of a double (not NaN): unsigned order of the images
// is the order of the values; -0.0 and +0.0 share the image of +0.0.
__device__ __forceinline__ unsigned long long min_key(double d)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(d == 0.0 ? 0.0 : d);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

__device__ __forceinline__ double min_unkey(unsigned long long k)
{
    return __longlong_as_double((long long)((k >> 63) ? (k & 0x7FFFFFFFFFFFFFFFull) : ~k));
}

struct MinPub {
    const int* sblk;            // first block of each segment (null: one segment of gridDim.x blocks)
    unsigned long long* slots;  // kMinSlots per segment
};

// m: the workgroup's minimum (valid in thread 0).
__device__ __forceinline__ void min_publish(const MinPub& p, int seg, double m)
{
    if (threadIdx.x == 0 && p.slots) {
        const int lo = p.sblk ? p.sblk[seg] : 0;
        atomicMin(p.slots + (size_t)seg * kMinSlots + ((int)blockIdx.x - lo) % kMinSlots, min_key(m));
    }
}

// Candidates for get_min_dt_kernel: like the reference's 'dt < min_dt' test
// starting from DBL_MAX, NaN and values >= DBL_MAX never win.
__device__ __forceinline__ double min_candidate(double d) { return (d < DBL_MAX) ? d : DBL_MAX; }

// calculate_dt_kernel (time_stepping_kernels.h:13-32).  With 'old' non-null
// it also performs copy_double_kernel (copy_double_kernel.h:6-13) and with
// 'part' non-null it emits the per-workgroup min for get_min_dt_kernel.
// dt of node i with state U (time_stepping_kernels.h:13-32), optionally
// saving old = U.
__device__ __forceinline__ double calc_dt_state(int i, const double U[5], const double* __restrict__ vol,
                                                double* __restrict__ dt, double* __restrict__ old)
{
#pragma clang fp contract(off)
    if (old) store5(old, i, U);
    const double rho = U[0], E = U[4];
    const double ux = U[1] / rho, uy = U[2] / rho, uz = U[3] / rho;
    const double ssq = ux * ux + uy * uy + uz * uz;
    const double p = (kGamma - 1.0) * (E - 0.5 * rho * ssq);
    const double c = sqrt(kGamma * p / rho);
    const double v = 0.5 * (cbrt(vol[i]) / (sqrt(ssq) + c));
    dt[i] = v;
    return v;
}

// dt of node i (time_stepping_kernels.h:13-32), optionally saving old = var.
__device__ __forceinline__ double calc_dt_node(int i, const double* __restrict__ var,
                                               const double* __restrict__ vol, double* __restrict__ dt,
                                               double* __restrict__ old)
{
    double U[5];
    load5(var, i, U);
    return calc_dt_state(i, U, vol, dt, old);
}

// The next smoothing step's calculate_dt + copy (time_stepping_kernels.h:13-32,
// copy_double_kernel.h:6-13) on the level a restriction or prolongation has
// just produced, done by that kernel's thread for its node: dt, old = var and
// the per-workgroup minimum, over plain blocks or a segment block table
// ({segment, first, end} per workgroup, batched solves).
struct DtArgs {
    const int4* blk;
    const double* vol;
    double* dt;
    double* old;
    MinPub pub;
};

__global__ __launch_bounds__(kBlock) void k_calc_dt(
    int n, const double* __restrict__ var, const double* __restrict__ vol,
    double* __restrict__ dt, double* __restrict__ old, const int* __restrict__ status, MinPub pub)
{
#pragma clang fp contract(off)
    if (halted(status)) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double d = DBL_MAX;
    if (i < n) d = min_candidate(calc_dt_node(i, var, vol, dt, old));
    if (pub.slots) min_publish(pub, 0, block_min(d));
}

// get_min_dt_kernel as a standalone loop: per-workgroup min of dt.
__global__ __launch_bounds__(kBlock) void k_min_part(int n, const double* __restrict__ x,
                                                      double* __restrict__ part)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const double d = (i < n) ? min_candidate(x[i]) : DBL_MAX;
    const double m = block_min(d);
    if (threadIdx.x == 0) part[blockIdx.x] = m;
}

// Final min over the partials; *out = min(init, partials).  With 'status'
// set, a negative minimum raises abort code 1 (euler3d.cpp:480-484).
__global__ __launch_bounds__(kBlock) void k_min_final(int np, const double* __restrict__ part,
                                                       double init, double* __restrict__ out,
                                                       int* __restrict__ status)
{
    if (status && *status != 0) return;
    double m = DBL_MAX;
    for (int i = threadIdx.x; i < np; i += kBlock) m = (part[i] < m) ? part[i] : m;
    m = block_min(m);
    if (threadIdx.x == 0) {
        m = (m < init) ? m : init;
        *out = m;
        if (status && m < 0.0) *status = 1;
    }
}

// compute_step_factor_kernel (time_stepping_kernels.h:43-64): sf = min_dt / V.
__global__ void k_step_factor(int n, const double* __restrict__ vol, const double* __restrict__ dmin,
                              double hmin, double* __restrict__ sf, const int* __restrict__ status)
{
    if (halted(status)) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double m = dmin ? *dmin : hmin;
    sf[i] = m / vol[i];
}

// time_step_kernel (time_stepping_kernels.h:66-86).
__global__ void k_time_step(int n, int rk, const double* __restrict__ sf, double* __restrict__ flux,
                            const double* __restrict__ old, double* __restrict__ var,
                            const int* __restrict__ status)
{
#pragma clang fp contract(off)
    if (halted(status)) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double factor = sf[i] / (double)(3 + 1 - rk);
    double f[5], o[5], u[5];
    load5(flux, i, f);
    load5(old, i, o);
#pragma unroll
    for (int v = 0; v < 5; v++) u[v] = o[v] + factor * f[v];
    store5(var, i, u);
    const double z[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    store5(flux, i, z);
}

// One Runge-Kutta stage of a multigrid smoothing step after its edge flux,
// fused into one pass over the nodes: the node's boundary and far-field
// increments (flux.h:14-39, k_bnd_flux's arithmetic in boundary-element
// order), on the first stage the minimum dt (each workgroup reduces the
// per-workgroup minima of k_calc_dt over its mesh, as k_min_final does,
// euler3d.cpp:480-484) and the step factor (time_stepping_kernels.h:43-64),
// the stage update (time_stepping_kernels.h:66-86) and on the last stage the
// residual with its per-workgroup sum of squares and NaN/Inf count
// (validation.h:27-44, 102-115).  Each node sees the same operations in the
// same order as with the separate loops.
struct StageArgs {
    int n, rk;
    const int4* blk;           // batched solves: {segment, first, end} per workgroup; null: plain blocks over n
    const int* sblk;           // batched solves: first workgroup of each segment
    unsigned long long* mslots;   // minimum-dt slots (kMinSlots per segment): read by the first stage,
    int nmslots;                  // reset by the second
    int* mstatus;              // abort code 1 on a negative minimum: one status, or per segment
    int* zero;                 // first stage: a counter the last stage accumulates into, cleared
    const double* vol;
    double* sf;
    double* flux;
    const double* old;
    double* var;
    const int* bptr;           // boundary elements of node i: bnw[bptr[i] .. bptr[i + 1]) (null: none)
    const double4* bnw;        // per boundary element in that order: {w0, w1, w2, group}
    const Consts* kseg;        // batched solves: far-field constants per segment
    double* res;               // last stage: residual, or null
    double* part;              // last stage, finest level: per-workgroup sum of squares
    int* bad;                  // ... NaN/Inf count: one counter (atomic) ...
    int* badpart;              // ... or per workgroup
    int zero_flux;             // reset the flux to zero (else the next edge flux starts fresh)
    const int* status;
};

// First stage: the minimum dt of the workgroup's mesh, reduced from the
// mesh's slots that the kernel computing dt filled (min_publish); the mesh's
// first workgroup raises abort code 1 on a negative value
// (euler3d.cpp:476-484).  Second stage: the slots are reset for the next
// smoothing step.
template <int NT>
__device__ __forceinline__ double stage_min_dt(const StageArgs& a, int seg, bool first)
{
    // the slot minimum over the images (unsigned order), with uniform loads:
    // the reduction stays in scalar registers, off the vector registers the
    // edge phase needs
    const unsigned* sl = reinterpret_cast<const unsigned*>(a.mslots + (size_t)seg * kMinSlots);
    unsigned khi = (unsigned)(kMinEmpty >> 32), klo = (unsigned)kMinEmpty;
#pragma unroll 8
    for (int j = 0; j < kMinSlots; j++) {
        const unsigned yhi = __builtin_amdgcn_readfirstlane(sl[2 * j + 1]);
        const unsigned ylo = __builtin_amdgcn_readfirstlane(sl[2 * j]);
        const bool lt = yhi < khi || (yhi == khi && ylo < klo);
        khi = lt ? yhi : khi;
        klo = lt ? ylo : klo;
    }
    const double m = min_unkey(((unsigned long long)khi << 32) | klo);
    const double mdt = (m < DBL_MAX) ? m : DBL_MAX;
    if (first && threadIdx.x == 0) {
        if (mdt < 0.0 && a.mstatus[seg] == 0) a.mstatus[seg] = 1;
        if (a.zero) *a.zero = 0;
    }
    return mdt;
}

template <int NT>
__device__ __forceinline__ void stage_reset_min(const StageArgs& a)
{
    if (a.rk == 1 && blockIdx.x == 0)
        for (int i = threadIdx.x; i < a.nmslots; i += NT) a.mslots[i] = kMinEmpty;
}

// One node of a Runge-Kutta stage, in the reference's order: f holds node
// i's edge flux; the node's boundary and far-field increments are added
// (flux.h:14-39, its state read from vin), the step factor is formed on the
// first stage (time_stepping_kernels.h:43-64), the update written to a.var
// (:66-86), the flux reset when asked, and on the last stage the residual
// with its square sum and NaN/Inf count (validation.h:27-44,102-115).
// o: node i's old state; vs: its volume on the first stage, else its step
// factor; [q0, q1): its boundary elements (all loaded by the caller, which
// may issue the loads early).
__device__ __forceinline__ void stage_node(const StageArgs& a, const Consts& k0, int seg, int i,
                                           const double* __restrict__ vin, double f[5], double mdt,
                                           const double o[5], double vs, int q0, int q1, double& s, int& nb)
{
#pragma clang fp contract(off)
    double u[5];
    if (q0 < q1) {
        const Consts& k = a.kseg ? a.kseg[seg] : k0;
        double U[5];
        load5(vin, i, U);
        Th b;
        thermo_exact(U, b);
        for (int q = q0; q < q1; q++) {
            const double4 bq = a.bnw[q];
            const int g = (int)bq.w;
            const double w0 = bq.x, w1 = bq.y, w2 = bq.z;
            if (g <= 2) {
                f[0] = f[0] + 0.0;
                f[1] = f[1] + w0 * b.p;
                f[2] = f[2] + w1 * b.p;
                f[3] = f[3] + w2 * b.p;
                f[4] = f[4] + 0.0;
            } else if (g <= 7) {
                const double fx = 0.5 * w0, fy = 0.5 * w1, fz = 0.5 * w2;
                f[0] += fx * (k.ffv[1] + b.mx) + fy * (k.ffv[2] + b.my) + fz * (k.ffv[3] + b.mz);
                f[4] += fx * (k.ffde[0] + b.fdx) + fy * (k.ffde[1] + b.fdy) + fz * (k.ffde[2] + b.fdz);
                f[1] += fx * (k.ffmx[0] + b.fxx) + fy * (k.ffmx[1] + b.fxy) + fz * (k.ffmx[2] + b.fxz);
                f[2] += fx * (k.ffmy[0] + b.fxy) + fy * (k.ffmy[1] + b.fyy) + fz * (k.ffmy[2] + b.fyz);
                f[3] += fx * (k.ffmz[0] + b.fxz) + fy * (k.ffmz[1] + b.fyz) + fz * (k.ffmz[2] + b.fzz);
            }
        }
    }
    double sfi;
    if (a.rk == 0) {
        sfi = mdt / vs;
        a.sf[i] = sfi;
    } else {
        sfi = vs;
    }
    const double factor = sfi / (double)(3 + 1 - a.rk);
#pragma unroll
    for (int v = 0; v < 5; v++) u[v] = o[v] + factor * f[v];
    store5(a.var, i, u);
    if (a.zero_flux) {
        const double z[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        store5(a.flux, i, z);
    }
    if (a.res) {
        double r[5];
#pragma unroll
        for (int v = 0; v < 5; v++) {
            r[v] = u[v] - o[v];
            s += r[v] * r[v];
            nb += (isnan(u[v]) || isinf(u[v])) ? 1 : 0;
        }
        store5(a.res, i, r);
    }
}

// The last stage's per-workgroup square sum and NaN/Inf count (slot w).
template <int NT>
__device__ __forceinline__ void stage_sums(const StageArgs& a, int w, double s, int nb)
{
    if (a.part) {
        const double t = block_sum<NT>(s);
        if (threadIdx.x == 0) a.part[w] = t;
    }
    if (a.bad || a.badpart) {
        const int t = block_sum_i<NT>(nb);
        if (threadIdx.x == 0) {
            if (a.badpart) a.badpart[w] = t;
            else if (t) atomicAdd(a.bad, t);
        }
    }
}

// The stage's boundary table: the weights and group of each boundary
// element, in the node -> boundary-element CSR order of 'bent', so a node
// reads its elements as one contiguous run.
__global__ void k_pack_bnd(int nq, const int* __restrict__ bent, const double* __restrict__ bw,
                           const int* __restrict__ bgroup, double4* __restrict__ out)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    const size_t e = (size_t)bent[q];
    out[q] = make_double4(bw[3 * e], bw[3 * e + 1], bw[3 * e + 2], (double)bgroup[e]);
}

__global__ __launch_bounds__(kBlock) void k_stage(StageArgs a, Consts k0)
{
    if (halted(a.status)) return;
    stage_reset_min<kBlock>(a);
    int i, lim, seg = 0;
    bool first;                // first workgroup of its mesh
    if (a.blk) {
        const int4 B = a.blk[blockIdx.x];
        seg = B.x;
        i = B.y + threadIdx.x;
        lim = B.z;
        first = (int)blockIdx.x == a.sblk[seg];
    } else {
        i = blockIdx.x * blockDim.x + threadIdx.x;
        lim = a.n;
        first = blockIdx.x == 0;
    }
    double mdt = 0.0;
    if (a.rk == 0) {
        mdt = stage_min_dt<kBlock>(a, seg, first);
        if (mdt < 0.0 && a.status) return;   // a halting context aborts before any update
    }
    double s = 0.0;
    int nb = 0;
    if (i < lim) {
        double f[5], o[5];
        load5(a.flux, i, f);
        load5(a.old, i, o);
        const int q0 = a.bptr ? a.bptr[i] : 0, q1 = a.bptr ? a.bptr[i + 1] : 0;
        stage_node(a, k0, seg, i, a.var, f, mdt, o, a.rk == 0 ? a.vol[i] : a.sf[i], q0, q1, s, nb);
    }
    stage_sums<kBlock>(a, blockIdx.x, s, nb);
}

// The FP64 / mixed part kernel fused with its Runge-Kutta stage (device
// multigrid driver): a part's accumulators hold the complete edge flux of its
// owned nodes, so the stage update of those nodes runs in the same
// workgroup.  Other parts read the stage's input state as their halo while
// this one writes its update, so the input (vin) and the update (a.var) are
// different buffers.  The edge flux starts from the flux dat on the first
// stage (fresh == 0) and from zero after it; it is not written back.
// pseg / pfirst: segment of each part and first part of each segment
// (batched solves), null for one mesh.  Returns false when the context has
// halted or aborts (negative minimum dt).
template <typename R, int T, int EPT, int NPT, bool COLOURED, int RS>
__device__ __forceinline__ bool part_stage_one(
    int p, const PartInfo* __restrict__ parts, const int* __restrict__ halo, const unsigned* __restrict__ ecode,
    const R* __restrict__ ew, long long nloc, int rec_cap, const double* __restrict__ vin, int fresh,
    const StageArgs& a, const Consts& k0, const int* __restrict__ pseg, const int* __restrict__ pfirst, double* sm)
{
    // a halted context writes nothing: the status is tested after the edge
    // phase (which writes only LDS), so its load overlaps the staging loads
    const int halt = status_of(a.status);
    const PartInfo P = parts[p];
    const int seg = word_or_zero(pseg, p);
    static_assert(T >= kPart, "one owned node per thread");
    // the owned node's old state and volume / step factor, in flight through
    // the edge phase (clamped to a valid node for threads without one)
    const int t = threadIdx.x;
    const int i = P.n0 + min(t, max(P.cnt - 1, 0));
    double o[5];
    load5(a.old, i, o);
    const double vs = a.rk == 0 ? a.vol[i] : a.sf[i];
    const int q0 = a.bptr ? a.bptr[i] : 0, q1 = a.bptr ? a.bptr[i + 1] : 0;
    double* acc = part_accumulate<R, T, EPT, NPT, COLOURED, RS, true>(P, halo, ecode, ew, nloc, rec_cap, vin,
                                                                      a.flux, k0, fresh, sm);
    if (halt) return false;
    // the minimum dt after the edge phase (which writes only LDS), before any
    // update: fewer live registers through the edge phase
    double mdt = 0.0;
    if (a.rk == 0) {
        mdt = stage_min_dt<T>(a, seg, pseg ? p == pfirst[seg] : p == 0);
        if (mdt < 0.0 && a.status) return false;
    }
    double s = 0.0;
    int nb = 0;
    if (t < P.cnt) {
        double f[5];
#pragma unroll
        for (int v = 0; v < 5; v++) f[v] = acc[t * 5 + v];
        stage_node(a, k0, seg, i, vin, f, mdt, o, vs, q0, q1, s, nb);
    }
    stage_sums<T>(a, p, s, nb);
    return true;
}

// One workgroup per part.
template <typename R, int T, int EPT, int NPT, bool COLOURED, int RS>
__global__ __launch_bounds__(T) void k_flux_part_stage(
    const PartInfo* __restrict__ parts, const int* __restrict__ pseg, const int* __restrict__ pfirst,
    const int* __restrict__ halo, const unsigned* __restrict__ ecode, const R* __restrict__ ew, long long nloc,
    int rec_cap, const double* __restrict__ vin, int fresh, StageArgs a, Consts k0)
{
    extern __shared__ double sm[];
    if (part_stage_one<R, T, EPT, NPT, COLOURED, RS>(xcd_block(), parts, halo, ecode, ew, nloc, rec_cap, vin, fresh,
                                                     a, k0, pseg, pfirst, sm))
        stage_reset_min<T>(a);
}

// Persistent workgroups (a grid of a multiple of 8, below nparts): the
// workgroups of each XCD walk that XCD's contiguous slice of the parts with
// a stride, so the slice's halo stays in the XCD's L2 as it does with one
// workgroup per part.  A kernel of its own: the loop costs the one-part
// kernel ~30 VGPRs, a wave per SIMD.
template <typename R, int T, int EPT, int NPT, bool COLOURED, int RS>
__global__ __launch_bounds__(T) void k_flux_part_stage_persistent(
    const PartInfo* __restrict__ parts, const int* __restrict__ halo, const unsigned* __restrict__ ecode,
    const R* __restrict__ ew, long long nloc, int rec_cap, const double* __restrict__ vin, int fresh,
    StageArgs a, Consts k0, const int* __restrict__ pseg, const int* __restrict__ pfirst, int nparts)
{
    extern __shared__ double sm[];
    if (halted(a.status)) return;
    stage_reset_min<T>(a);
    const int x = blockIdx.x % 8, k = blockIdx.x / 8, g = gridDim.x / 8;
    const int per = nparts / 8, rem = nparts % 8;
    const int lo = (x < rem) ? x * (per + 1) : rem * (per + 1) + (x - rem) * per;
    const int hi = lo + per + (x < rem ? 1 : 0);
    for (int p = lo + k; p < hi; p += g) {
        if (!part_stage_one<R, T, EPT, NPT, COLOURED, RS>(p, parts, halo, ecode, ew, nloc, rec_cap, vin, fresh, a,
                                                          k0, pseg, pfirst, sm))
            return;
        __syncthreads();        // the next part reuses the LDS
    }
}

// residual_kernel (validation.h:27-35); with 'part' set also the per-node
// sum of squares of calc_rms_kernel (validation.h:37-44) and with 'bad' the
// NaN/Inf count of count_bad_vals (validation.h:102-115).
__global__ __launch_bounds__(kBlock) void k_residual(
    int n, const double* __restrict__ old, const double* __restrict__ var, double* __restrict__ res,
    double* __restrict__ part, int* __restrict__ bad, const int* __restrict__ status)
{
#pragma clang fp contract(off)
    if (halted(status)) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double s = 0.0;
    int nb = 0;
    if (i < n) {
        double o[5], u[5], r[5];
        load5(old, i, o);
        load5(var, i, u);
#pragma unroll
        for (int v = 0; v < 5; v++) {
            r[v] = u[v] - o[v];
            s += r[v] * r[v];
            nb += (isnan(u[v]) || isinf(u[v])) ? 1 : 0;
        }
        if (res) store5(res, i, r);
    }
    if (part) {
        const double t = block_sum(s);
        if (threadIdx.x == 0) part[blockIdx.x] = t;
    }
    if (bad) {
        const int t = block_sum_i(nb);
        if (threadIdx.x == 0 && t) atomicAdd(bad, t);
    }
}

// calc_rms_kernel as a standalone loop (sum of squares per workgroup).
__global__ __launch_bounds__(kBlock) void k_sumsq_part(int n, const double* __restrict__ r,
                                                        double* __restrict__ part)
{
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double s = 0.0;
    if (i < n) {
        double x[5];
        load5(r, i, x);
#pragma unroll
        for (int v = 0; v < 5; v++) s += x[v] * x[v];
    }
    const double t = block_sum(s);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// Final reductions of per-workgroup partials run as one workgroup per sum.
// Each thread keeps four partial sums over entries NT apart, so four loads
// are in flight per thread instead of one (a level-0 sum of an 8M-node mesh
// has 31K partials); the order is fixed, so the sum is deterministic.
constexpr int kSumBlock = 1024;

template <int NT>
__device__ __forceinline__ double strided_sum(const double* __restrict__ x, const int* __restrict__ cnt, int lo,
                                              int hi, int& nb)
{
#pragma clang fp contract(off)
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    int c = 0;
    int i = lo + (int)threadIdx.x;
    for (; i + 3 * NT < hi; i += 4 * NT) {
        s0 += x[i];
        s1 += x[i + NT];
        s2 += x[i + 2 * NT];
        s3 += x[i + 3 * NT];
        if (cnt) c += cnt[i] + cnt[i + NT] + cnt[i + 2 * NT] + cnt[i + 3 * NT];
    }
    for (; i < hi; i += NT) {
        s0 += x[i];
        if (cnt) c += cnt[i];
    }
    nb = c;
    return (s0 + s1) + (s2 + s3);
}

// Sum of partials; *out = init + sum.  With 'hist' the L0 RMS
// sqrt(sum / n) is stored at hist[*cycle] (euler3d.cpp:532-537) and a
// non-zero bad-value count raises abort code 2 (euler3d.cpp:539-548).
__global__ __launch_bounds__(kSumBlock) void k_sum_final(int np, const double* __restrict__ part,
                                                          double init, double* __restrict__ out,
                                                          int nnodes, double* __restrict__ hist,
                                                          const int* __restrict__ cycle,
                                                          const int* __restrict__ bad,
                                                          int* __restrict__ status)
{
#pragma clang fp contract(off)
    if (status && *status != 0) return;
    int unused = 0;
    double s = strided_sum<kSumBlock>(part, nullptr, 0, np, unused);
    s = block_sum<kSumBlock>(s);
    if (threadIdx.x == 0) {
        s = init + s;
        if (out) *out = s;
        if (hist) hist[*cycle] = sqrt(s / (double)nnodes);
        if (status && bad && *bad > 0) *status = 2;
    }
}

__global__ __launch_bounds__(kBlock) void k_count_bad(int n, const double* __restrict__ x,
                                                       int* __restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    int c = 0;
    if (i < n) {
        double u[5];
        load5(x, i, u);
#pragma unroll
        for (int v = 0; v < 5; v++) c += (isnan(u[v]) || isinf(u[v])) ? 1 : 0;
    }
    c = block_sum_i(c);
    if (threadIdx.x == 0 && c) atomicAdd(out, c);
}

// identify_differences (validation.h:46-89) fused with count_non_zeros
// (validation.h:91-100) when 'count' is set.
__global__ __launch_bounds__(kBlock) void k_identify_differences(
    int n, const double* __restrict__ test, const double* __restrict__ master,
    double* __restrict__ diff, int* __restrict__ count)
{
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    int c = 0;
    if (i < n) {
        double tv[5], mv[5], r[5];
        load5(test, i, tv);
        load5(master, i, mv);
#pragma unroll
        for (int v = 0; v < 5; v++) {
            double acc = mv[v] * 10.0e-8;
            if (acc < 0.0) acc *= -1.0;
            if (acc < 3.0e-19) acc = 3.0e-19;
            double d = tv[v] - mv[v];
            if (d < 0.0) d *= -1.0;
            r[v] = (d > acc) ? d : 0.0;
            c += (r[v] > 0.0) ? 1 : 0;
        }
        if (diff) store5(diff, i, r);
    }
    if (count) {
        c = block_sum_i(c);
        if (threadIdx.x == 0 && c) atomicAdd(count, c);
    }
}

__global__ __launch_bounds__(kBlock) void k_count_pos(int n, const double* __restrict__ x,
                                                       int* __restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    int c = 0;
    if (i < n) {
        double u[5];
        load5(x, i, u);
#pragma unroll
        for (int v = 0; v < 5; v++) c += (u[v] > 0.0) ? 1 : 0;
    }
    c = block_sum_i(c);
    if (threadIdx.x == 0 && c) atomicAdd(out, c);
}

// ------------------------------------------------------------------------
// Multigrid transfer (mg.h:29-88)
// ------------------------------------------------------------------------

// up_pre_kernel: zero the coarse targets of every fine node (writes of the
// same value from several children are benign).
__global__ void k_up_pre(int nf, const int* __restrict__ f2c, double* __restrict__ vc,
                         int* __restrict__ scratch, const int* __restrict__ status)
{
    if (halted(status)) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nf) return;
    const int c = f2c[i];
    const double z[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    store5(vc, c, z);
    scratch[c] = 0;
}

// up_kernel as a gather over the children of each coarse node, children in
// fine-node order; with 'fused' the up_pre zeroing and the up_post mean are
// folded in (one pass, same arithmetic).
__global__ __launch_bounds__(kBlock) void k_up_gather(
    int n_targets, const int* __restrict__ targets, const int* __restrict__ ptr,
    const int* __restrict__ child, const double* __restrict__ vf, double* __restrict__ vc,
    int* __restrict__ scratch, int fused, const int* __restrict__ status)
{
#pragma clang fp contract(off)
    if (halted(status)) return;
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= n_targets) return;
    const int c = targets ? targets[tid] : tid;
    const int beg = ptr[tid], end = ptr[tid + 1];
    double acc[5];
    int cnt;
    if (fused) {
        if (beg == end) return;           // no children: stale value kept (mg.h:58)
#pragma unroll
        for (int v = 0; v < 5; v++) acc[v] = 0.0;
        cnt = 0;
    } else {
        load5(vc, c, acc);
        cnt = scratch[c];
    }
    for (int q = beg; q < end; q++) {
        double x[5];
        load5(vf, child[q], x);
#pragma unroll
        for (int v = 0; v < 5; v++) acc[v] += x[v];
        cnt += 1;
    }
    if (fused) {
        const double avg = cnt == 0 ? 1.0 : 1.0 / (double)cnt;
#pragma unroll
        for (int v = 0; v < 5; v++) acc[v] *= avg;
    }
    store5(vc, c, acc);
    scratch[c] = cnt;
}

// up_post_kernel: mean over children (count 0 keeps the value).
__global__ void k_up_post(int nc, double* __restrict__ vc, const int* __restrict__ scratch,
                          const int* __restrict__ status)
{
#pragma clang fp contract(off)
    if (halted(status)) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nc) return;
    const int s = scratch[i];
    const double avg = s == 0 ? 1.0 : 1.0 / (double)s;
    double x[5];
    load5(vc, i, x);
#pragma unroll
    for (int v = 0; v < 5; v++) x[v] *= avg;
    store5(vc, i, x);
}

// down_kernel (prolongation).
// up_pre + up + up_post (mg.h:29-64) as one gather per coarse node of a
// restriction whose every coarse node is a target (dense plan), followed by
// the coarse level's calculate_dt.
__global__ __launch_bounds__(kBlock) void k_up_gather_dt(
    int nc, const int* __restrict__ ptr, const int* __restrict__ child, const double* __restrict__ vf,
    double* __restrict__ vc, int* __restrict__ scratch, const int* __restrict__ status, DtArgs d)
{
#pragma clang fp contract(off)
    if (halted(status)) return;
    int c, lim;
    if (d.blk) {
        const int4 B = d.blk[blockIdx.x];
        c = B.y + threadIdx.x;
        lim = B.z;
    } else {
        c = blockIdx.x * blockDim.x + threadIdx.x;
        lim = nc;
    }
    double m = DBL_MAX;
    if (c < lim) {
        const int beg = ptr[c], end = ptr[c + 1];
        double acc[5];
        if (beg == end) {
            load5(vc, c, acc);             // no children: stale value kept (mg.h:58)
        } else {
#pragma unroll
            for (int v = 0; v < 5; v++) acc[v] = 0.0;
            int cnt = 0;
            for (int q = beg; q < end; q++) {
                double x[5];
                load5(vf, child[q], x);
#pragma unroll
                for (int v = 0; v < 5; v++) acc[v] += x[v];
                cnt += 1;
            }
            const double avg = 1.0 / (double)cnt;
#pragma unroll
            for (int v = 0; v < 5; v++) acc[v] *= avg;
            store5(vc, c, acc);
            scratch[c] = cnt;
        }
        m = min_candidate(calc_dt_state(c, acc, d.vol, d.dt, d.old));
    }
    min_publish(d.pub, d.blk ? d.blk[blockIdx.x].x : 0, block_min(m));
}

__global__ void k_down(int nf, const int* __restrict__ f2c, double* __restrict__ vf,
                       const double* __restrict__ rf, const double* __restrict__ xf,
                       const double* __restrict__ rc, const double* __restrict__ xc,
                       const int* __restrict__ status)
{
#pragma clang fp contract(off)
    if (halted(status)) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nf) return;
    const int c = f2c[i];
    const double dx = fabs(xf[3 * (size_t)i] - xc[3 * (size_t)c]);
    const double dy = fabs(xf[3 * (size_t)i + 1] - xc[3 * (size_t)c + 1]);
    const double dz = fabs(xf[3 * (size_t)i + 2] - xc[3 * (size_t)c + 2]);
    const double dm = sqrt(dx * dx + dy * dy + dz * dz);
    double u[5], r[5], ra[5];
    load5(vf, i, u);
    load5(rf, i, r);
    load5(rc, c, ra);
    u[0] -= dm * (ra[0] - r[0]);
    u[1] -= dx * (ra[1] - r[1]);
    u[2] -= dy * (ra[2] - r[2]);
    u[3] -= dz * (ra[3] - r[3]);
    u[4] -= dm * (ra[4] - r[4]);
    store5(vf, i, u);
}

// Inverse distance 1 / |p - q| as mg.h:134,147,170,183 form it (exact:
// no contraction, IEEE square root and division).
__device__ __forceinline__ double down_v2_idist(const double* __restrict__ p, const double* __restrict__ q)
{
#pragma clang fp contract(off)
    const double dx = p[0] - q[0], dy = p[1] - q[1], dz = p[2] - q[2];
    return 1.0 / sqrt(dx * dx + dy * dy + dz * dz);
}

// down_v2_kernel (mg.h:101-191) as a gather: one thread per fine node walks
// its incidences of the edge map in element order (the OP2 sequential order)
// and applies, for its own side of each edge, what the reference applies
// through its OP_INC pointers:
//   - the node coincides with its coarse parent: the sums are set to the
//     parent's residual and 1 (mg.h:124-132, 160-167), on every such edge;
//   - side a: + w(a, parent a) r(parent a) + w(parent b, a) r(parent b);
//   - side b: + w(b, parent b) r(parent b) + w(parent a, b) r(parent b) --
//     the reference's b side takes b's parent residual in both terms.
// The sums start from the dats' values (OP_INC), so the result is bitwise
// the sequential one.  e2mg is the edge-->mg_nodes map, internal ids.
__global__ __launch_bounds__(kBlock) void k_down_v2(
    int n_targets, const int* __restrict__ targets, const int* __restrict__ ptr, const int2* __restrict__ inc,
    const int* __restrict__ e2mg, const double* __restrict__ xf, const double* __restrict__ xc,
    const double* __restrict__ rc, double* __restrict__ rp, double* __restrict__ wsum)
{
#pragma clang fp contract(off)
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= n_targets) return;
    const int n = targets ? targets[tid] : tid;
    const double x2[3] = {xf[3 * (size_t)n], xf[3 * (size_t)n + 1], xf[3 * (size_t)n + 2]};
    double acc[5], w = wsum[n];
    load5(rp, n, acc);
    for (int q = ptr[tid]; q < ptr[tid + 1]; q++) {
        const unsigned code = (unsigned)inc[q].y;
        const size_t e = code & 0x7fffffffu;
        const int side = (int)(code >> 31);
        const int own = e2mg[2 * e + side], oth = e2mg[2 * e + (1 - side)];
        const double* x1 = xc + 3 * (size_t)own;
        double r1[5];
        load5(rc, own, r1);
        if (x2[0] - x1[0] == 0.0 && x2[1] - x1[1] == 0.0 && x2[2] - x1[2] == 0.0) {
#pragma unroll
            for (int v = 0; v < 5; v++) acc[v] = r1[v];
            w = 1.0;
            continue;
        }
        double d = down_v2_idist(x2, x1);
#pragma unroll
        for (int v = 0; v < 5; v++) acc[v] += d * r1[v];
        w += d;
        d = down_v2_idist(xc + 3 * (size_t)oth, x2);
        if (side == 0) load5(rc, oth, r1);
#pragma unroll
        for (int v = 0; v < 5; v++) acc[v] += d * r1[v];
        w += d;
    }
    store5(rp, n, acc);
    wsum[n] = w;
}

// down_v2_kernel_post (mg.h:193-204): var += res - rp / wsum.
__global__ void k_down_v2_post(int n, const double* __restrict__ rp, const double* __restrict__ wsum,
                               const double* __restrict__ rf, double* __restrict__ vf)
{
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double u[5], r[5], p[5];
    load5(vf, i, u);
    load5(rf, i, r);
    load5(rp, i, p);
    const double w = wsum[i];
#pragma unroll
    for (int v = 0; v < 5; v++) u[v] += r[v] - (p[v] / w);
    store5(vf, i, u);
}

// down_kernel (mg.h:66-88) followed by the fine level's calculate_dt.
__global__ __launch_bounds__(kBlock) void k_down_dt(
    int nf, const int* __restrict__ f2c, double* __restrict__ vf, const double* __restrict__ rf,
    const double* __restrict__ xf, const double* __restrict__ rc, const double* __restrict__ xc,
    const int* __restrict__ status, DtArgs d)
{
#pragma clang fp contract(off)
    if (halted(status)) return;
    int i, lim;
    if (d.blk) {
        const int4 B = d.blk[blockIdx.x];
        i = B.y + threadIdx.x;
        lim = B.z;
    } else {
        i = blockIdx.x * blockDim.x + threadIdx.x;
        lim = nf;
    }
    double m = DBL_MAX;
    if (i < lim) {
        const int c = f2c[i];
        const double dx = fabs(xf[3 * (size_t)i] - xc[3 * (size_t)c]);
        const double dy = fabs(xf[3 * (size_t)i + 1] - xc[3 * (size_t)c + 1]);
        const double dz = fabs(xf[3 * (size_t)i + 2] - xc[3 * (size_t)c + 2]);
        const double dm = sqrt(dx * dx + dy * dy + dz * dz);
        double u[5], r[5], ra[5];
        load5(vf, i, u);
        load5(rf, i, r);
        load5(rc, c, ra);
        u[0] -= dm * (ra[0] - r[0]);
        u[1] -= dx * (ra[1] - r[1]);
        u[2] -= dy * (ra[2] - r[2]);
        u[3] -= dz * (ra[3] - r[3]);
        u[4] -= dm * (ra[4] - r[4]);
        store5(vf, i, u);
        m = min_candidate(calc_dt_state(i, u, d.vol, d.dt, d.old));
    }
    min_publish(d.pub, d.blk ? d.blk[blockIdx.x].x : 0, block_min(m));
}

// ------------------------------------------------------------------------
// Setup (misc.h:40-84)
// ------------------------------------------------------------------------

// calculate_cell_volumes, edge part: tetra volume per edge and the weight
// rewrite w <- (d/|d|)|w| / |d|.
__global__ void k_cell_volumes_edge(int ne, const int2* __restrict__ e2n, const double* __restrict__ x,
                                    double* __restrict__ w, double* __restrict__ tv)
{
#pragma clang fp contract(off)
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    const int2 ab = e2n[e];
    double d[3];
    double dist = 0.0;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        d[i] = x[3 * (size_t)ab.y + i] - x[3 * (size_t)ab.x + i];
        dist += d[i] * d[i];
    }
    dist = sqrt(dist);
    double* we = w + 3 * (size_t)e;
    double area = 0.0;
#pragma unroll
    for (int i = 0; i < 3; i++) area += we[i] * we[i];
    area = sqrt(area);
    tv[e] = (1.0 / 3.0) * 0.5 * dist * area;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        double r = (d[i] / dist) * area;
        r /= dist;
        we[i] = r;
    }
}

// calculate_cell_volumes, node part: volumes INC in edge order.
__global__ void k_cell_volumes_node(int n_targets, const int* __restrict__ targets,
                                    const int* __restrict__ ptr, const int2* __restrict__ inc,
                                    const double* __restrict__ tv, double* __restrict__ vol)
{
#pragma clang fp contract(off)
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= n_targets) return;
    const int n = targets ? targets[tid] : tid;
    double acc = vol[n];
    for (int q = ptr[tid]; q < ptr[tid + 1]; q++) acc += tv[inc[q].y & 0x7fffffff];
    vol[n] = acc;
}

// Multigrid cycle counter, bumped once per cycle on the device.
// ------------------------------------------------------------------------
// Batched solves: the per-node reductions over a segment block table
// {segment, lo, hi} whose blocks never straddle two meshes, then one final
// reduction block per mesh.
// ------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_calc_dt_seg(const int4* __restrict__ blk, const double* __restrict__ var,
                                                        const double* __restrict__ vol, double* __restrict__ dt,
                                                        double* __restrict__ old, const int* __restrict__ halt,
                                                        MinPub pub)
{
    if (halted(halt)) return;
    const int4 B = blk[blockIdx.x];
    const int i = B.y + threadIdx.x;
    double d = DBL_MAX;
    if (i < B.z) d = min_candidate(calc_dt_node(i, var, vol, dt, old));
    min_publish(pub, B.x, block_min(d));
}

__global__ __launch_bounds__(kSumBlock) void k_sum_final_seg(const int* __restrict__ sblk,
                                                             const double* __restrict__ part,
                                                             const int* __restrict__ badpart,
                                                             const int* __restrict__ nnodes, double* __restrict__ hist,
                                                             int cap, const int* __restrict__ cycle,
                                                             int* __restrict__ sstatus, const int* __restrict__ halt)
{
#pragma clang fp contract(off)
    if (halted(halt)) return;
    const int s = blockIdx.x;
    int nb = 0;
    double x = strided_sum<kSumBlock>(part, badpart, sblk[s], sblk[s + 1], nb);
    x = block_sum<kSumBlock>(x);
    nb = block_sum_i<kSumBlock>(nb);
    if (threadIdx.x == 0) {
        const int c = *cycle;
        if (c < cap) hist[(size_t)s * cap + c] = sqrt(x / (double)nnodes[s]);
        if (nb > 0 && sstatus[s] == 0) sstatus[s] = 2;
    }
}

// Cycle counter: counts cycles that ran (a halted context stops counting).
__global__ void k_bump(int* c, const int* __restrict__ status)
{
    if (threadIdx.x == 0 && !halted(status)) *c += 1;
}

// Convergence stop of served solves, run after the cycle counter moved on:
// segment s whose L0 RMS of the cycle that just ran has dropped to 'drop'
// times its first cycle's gets status 3, which halts every later kernel of a
// one-segment context (the state stays that of the converged cycle).  In a
// context of several segments nothing halts: the segment's cycle count goes
// to segcyc[s], k_snap_converged copies its state this cycle, and the next
// call marks it copied (snapflag).
__global__ void k_converge(int nseg, int* __restrict__ sstatus, const double* __restrict__ rms, int cap,
                           const int* __restrict__ cycle, double drop, int* __restrict__ snapflag,
                           int* __restrict__ segcyc)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nseg) return;
    if (sstatus[s] == 3) {
        if (snapflag) snapflag[s] = 1;
        return;
    }
    if (sstatus[s] != 0) return;
    const int c = *cycle - 1;       // the cycle that just ran
    if (c >= 1 && c < cap && rms[(size_t)s * cap + c] <= drop * rms[(size_t)s * cap]) {
        sstatus[s] = 3;
        if (segcyc) segcyc[s] = c + 1;
    }
}

// Level-0 state of the segments that converged in the cycle that just ran
// (status 3, not yet copied), one workgroup per segment block {segment, lo,
// hi}: the result of a served request that keeps running with the other
// members of its group.
__global__ __launch_bounds__(kBlock) void k_snap_converged(int nblocks, const int4* __restrict__ blk,
                                                           const int* __restrict__ sstatus,
                                                           const int* __restrict__ snapflag,
                                                           const double* __restrict__ var, double* __restrict__ snap)
{
    const int b = blockIdx.x;
    if (b >= nblocks) return;
    const int4 q = blk[b];
    if (sstatus[q.x] != 3 || snapflag[q.x]) return;
    for (int i = 5 * q.y + threadIdx.x; i < 5 * q.z; i += kBlock) snap[i] = var[i];
}

}  // namespace mgcfd
