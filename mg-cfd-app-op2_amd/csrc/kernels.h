// gfx950 device kernels for the MG-CFD hot path.
//
// Data layout in HBM (set up by runtime.hip):
//   node dats   SoA, component-major: x[v * ld + i], ld = size rounded up to 64
//   edge map    int2 per edge (internal ids), edges ordered by lower endpoint
//   weights     packed double4 {w0, w1, w2, |w|} per edge, rebuilt whenever
//               the weight dat changes (|w| is the reference's ewt, flux.h:48)
//   inverse map CSR per target node: the incidences that INC into it, in the
//               reference's element order (OP2 sequential order), each an
//               int2 {other endpoint, element | column << 31}
//
// Kernels marked "exact" disable FMA contraction and keep the reference's
// operation order and true divisions, so they are bitwise equal to the
// sequential CPU evaluation.  "fast" kernels use one reciprocal of the density
// and FMA contraction.
#pragma once

#include <hip/hip_runtime.h>
#include <cfloat>
#include <cmath>
#include <cstdint>

namespace mgcfd {

constexpr double kGamma = 1.4;
constexpr int kBlock = 256;
constexpr int kWave = 64;

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

// Reference order: u = m / rho, |u|^2, p, c = sqrt(gamma p / rho).
__device__ __forceinline__ void thermo_exact(double rho, double mx, double my, double mz,
                                             double E, Th& t)
{
#pragma clang fp contract(off)
    t.rho = rho; t.mx = mx; t.my = my; t.mz = mz; t.E = E;
    t.ux = mx / rho;
    t.uy = my / rho;
    t.uz = mz / rho;
    double ssq = t.ux * t.ux + t.uy * t.uy + t.uz * t.uz;
    t.speed = sqrt(ssq);
    t.p = (kGamma - 1.0) * (E - 0.5 * rho * ssq);
    t.c = sqrt(kGamma * t.p / rho);
    t.fxx = t.ux * mx + t.p;
    t.fxy = t.ux * my;
    t.fxz = t.ux * mz;
    t.fyy = t.uy * my + t.p;
    t.fyz = t.uy * mz;
    t.fzz = t.uz * mz + t.p;
    double de_p = E + t.p;
    t.fdx = t.ux * de_p;
    t.fdy = t.uy * de_p;
    t.fdz = t.uz * de_p;
}

__device__ __forceinline__ void thermo_fast(double rho, double mx, double my, double mz,
                                            double E, Th& t)
{
    t.rho = rho; t.mx = mx; t.my = my; t.mz = mz; t.E = E;
    double inv = 1.0 / rho;
    t.ux = mx * inv;
    t.uy = my * inv;
    t.uz = mz * inv;
    double ssq = t.ux * t.ux + t.uy * t.uy + t.uz * t.uz;
    t.speed = sqrt(ssq);
    t.p = (kGamma - 1.0) * (E - 0.5 * rho * ssq);
    t.c = sqrt(kGamma * t.p * inv);
    t.fxx = t.ux * mx + t.p;
    t.fxy = t.ux * my;
    t.fxz = t.ux * mz;
    t.fyy = t.uy * my + t.p;
    t.fyz = t.uy * mz;
    t.fzz = t.uz * mz + t.p;
    double de_p = E + t.p;
    t.fdx = t.ux * de_p;
    t.fdy = t.uy * de_p;
    t.fdz = t.uz * de_p;
}

template <bool EXACT>
__device__ __forceinline__ void thermo(double rho, double mx, double my, double mz, double E, Th& t)
{
    if constexpr (EXACT) thermo_exact(rho, mx, my, mz, E, t);
    else thermo_fast(rho, mx, my, mz, E, t);
}

// Increment that edge (own, oth) adds to node 'own' (flux.h:139-207).  The
// reference computes factor_a and factor_b with the same operand order taken
// from the endpoint's own side; with fx = -0.5 w (or +0.5 w on the b side,
// whose terms the reference subtracts) one formula serves both roles exactly.
__device__ __forceinline__ void edge_increment_exact(const Th& o, const Th& t, double ewt, double sc,
                                                     double fx, double fy, double fz, double inc[5])
{
#pragma clang fp contract(off)
    double factor = -ewt * sc * 0.5 * (o.speed + t.speed + o.c + t.c);
    inc[0] = factor * (o.rho - t.rho) + fx * (o.mx + t.mx) + fy * (o.my + t.my) + fz * (o.mz + t.mz);
    inc[4] = factor * (o.E - t.E) + fx * (o.fdx + t.fdx) + fy * (o.fdy + t.fdy) + fz * (o.fdz + t.fdz);
    inc[1] = factor * (o.mx - t.mx) + fx * (o.fxx + t.fxx) + fy * (o.fxy + t.fxy) + fz * (o.fxz + t.fxz);
    inc[2] = factor * (o.my - t.my) + fx * (o.fxy + t.fxy) + fy * (o.fyy + t.fyy) + fz * (o.fyz + t.fyz);
    inc[3] = factor * (o.mz - t.mz) + fx * (o.fxz + t.fxz) + fy * (o.fyz + t.fyz) + fz * (o.fzz + t.fzz);
}

__device__ __forceinline__ void edge_increment_fast(const Th& o, const Th& t, double ewt, double sc,
                                                    double fx, double fy, double fz, double inc[5])
{
    double factor = -ewt * sc * 0.5 * (o.speed + t.speed + o.c + t.c);
    inc[0] = factor * (o.rho - t.rho) + fx * (o.mx + t.mx) + fy * (o.my + t.my) + fz * (o.mz + t.mz);
    inc[4] = factor * (o.E - t.E) + fx * (o.fdx + t.fdx) + fy * (o.fdy + t.fdy) + fz * (o.fdz + t.fdz);
    inc[1] = factor * (o.mx - t.mx) + fx * (o.fxx + t.fxx) + fy * (o.fxy + t.fxy) + fz * (o.fxz + t.fxz);
    inc[2] = factor * (o.my - t.my) + fx * (o.fxy + t.fxy) + fy * (o.fyy + t.fyy) + fz * (o.fyz + t.fyz);
    inc[3] = factor * (o.mz - t.mz) + fx * (o.fxz + t.fxz) + fy * (o.fyz + t.fyz) + fz * (o.fzz + t.fzz);
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
    return (x < rem) ? x * (per + 1) + k : rem * (per + 1) + (x - rem) * per + k;
}

__device__ __forceinline__ bool halted(const int* status) { return status && *status != 0; }

// ------------------------------------------------------------------------
// Edge flux, node-centric gather.  One thread per target node; the thread
// evaluates every incident edge in the reference's edge order and adds the
// increments into a register accumulator seeded with the node's flux, which
// reproduces OP2's sequential INC order.  No atomics, one launch.
// ------------------------------------------------------------------------
template <bool EXACT>
__global__ __launch_bounds__(kBlock) void k_flux_gather(
    int n_targets, const int* __restrict__ targets, const int* __restrict__ ptr,
    const int2* __restrict__ inc, const double4* __restrict__ wpk,
    const double* __restrict__ var, double* __restrict__ flux, int ld, Consts k,
    const int* __restrict__ status)
{
    if (halted(status)) return;
    const int tid = xcd_block() * kBlock + threadIdx.x;
    if (tid >= n_targets) return;
    const int n = targets ? targets[tid] : tid;
    Th o;
    thermo<EXACT>(var[n], var[ld + n], var[2 * ld + n], var[3 * ld + n], var[4 * ld + n], o);
    double acc[5];
#pragma unroll
    for (int v = 0; v < 5; v++) acc[v] = flux[v * ld + n];
    const int beg = ptr[tid], end = ptr[tid + 1];
    for (int q = beg; q < end; q++) {
        const int2 r = inc[q];
        const int j = r.x;
        const int e = r.y & 0x7fffffff;
        const double4 w = wpk[e];
        Th t;
        thermo<EXACT>(var[j], var[ld + j], var[2 * ld + j], var[3 * ld + j], var[4 * ld + j], t);
        double fx = -0.5 * w.x, fy = -0.5 * w.y, fz = -0.5 * w.z;
        if (r.y < 0) { fx = -fx; fy = -fy; fz = -fz; }   // node is endpoint b
        double d[5];
        if constexpr (EXACT) {
            edge_increment_exact(o, t, w.w, k.sc, fx, fy, fz, d);
#pragma unroll
            for (int v = 0; v < 5; v++) acc[v] = add_exact(acc[v], d[v]);
        } else {
            edge_increment_fast(o, t, w.w, k.sc, fx, fy, fz, d);
#pragma unroll
            for (int v = 0; v < 5; v++) acc[v] += d[v];
        }
    }
#pragma unroll
    for (int v = 0; v < 5; v++) flux[v * ld + n] = acc[v];
}

// Edge flux, edge-parallel with FP64 hardware atomics (global_atomic_add_f64).
__global__ __launch_bounds__(kBlock) void k_flux_atomic(
    int n_edges, const int2* __restrict__ e2n, const double4* __restrict__ wpk,
    const double* __restrict__ var, double* __restrict__ flux, int ld, Consts k,
    const int* __restrict__ status)
{
    if (halted(status)) return;
    const int e = xcd_block() * kBlock + threadIdx.x;
    if (e >= n_edges) return;
    const int2 ab = e2n[e];
    const double4 w = wpk[e];
    Th a, b;
    thermo_fast(var[ab.x], var[ld + ab.x], var[2 * ld + ab.x], var[3 * ld + ab.x], var[4 * ld + ab.x], a);
    thermo_fast(var[ab.y], var[ld + ab.y], var[2 * ld + ab.y], var[3 * ld + ab.y], var[4 * ld + ab.y], b);
    const double fx = -0.5 * w.x, fy = -0.5 * w.y, fz = -0.5 * w.z;
    double da[5], db[5];
    edge_increment_fast(a, b, w.w, k.sc, fx, fy, fz, da);
    edge_increment_fast(b, a, w.w, k.sc, -fx, -fy, -fz, db);
#pragma unroll
    for (int v = 0; v < 5; v++) {
        unsafeAtomicAdd(&flux[v * ld + ab.x], da[v]);
        unsafeAtomicAdd(&flux[v * ld + ab.y], db[v]);
    }
}

// unstructured_stream.h:7-57 as a gather: same data movement as the flux
// kernel with (almost) no arithmetic.
__global__ __launch_bounds__(kBlock) void k_stream_gather(
    int n_targets, const int* __restrict__ targets, const int* __restrict__ ptr,
    const int2* __restrict__ inc, const double4* __restrict__ wpk,
    const double* __restrict__ var, double* __restrict__ flux, int ld)
{
#pragma clang fp contract(off)
    const int tid = xcd_block() * kBlock + threadIdx.x;
    if (tid >= n_targets) return;
    const int n = targets ? targets[tid] : tid;
    double acc[5];
#pragma unroll
    for (int v = 0; v < 5; v++) acc[v] = flux[v * ld + n];
    for (int q = ptr[tid]; q < ptr[tid + 1]; q++) {
        const int2 r = inc[q];
        const int j = r.x;
        const double u0 = var[j], u1 = var[ld + j], u2 = var[2 * ld + j], u3 = var[3 * ld + j],
                     u4 = var[4 * ld + j];
        if (r.y >= 0) {     // node is a: fluxes_a += b-values (+ weights)
            const double4 w = wpk[r.y];
            acc[0] += u0 + w.x;
            acc[1] += u1 + w.z;
            acc[2] += u2;
            acc[3] += u3;
            acc[4] += u4 + w.y;
        } else {            // node is b: fluxes_b += a-values
            acc[0] += u0;
            acc[1] += u1;
            acc[2] += u2;
            acc[3] += u3;
            acc[4] += u4;
        }
    }
#pragma unroll
    for (int v = 0; v < 5; v++) flux[v * ld + n] = acc[v];
}

// Packed weights {w0, w1, w2, |w|}: |w| = sqrt(w0^2 + w1^2 + w2^2) (flux.h:48-50).
__global__ void k_pack_weights(int n, const double* __restrict__ w, int ld, double4* __restrict__ out)
{
#pragma clang fp contract(off)
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const double a = w[e], b = w[ld + e], c = w[2 * ld + e];
    out[e] = make_double4(a, b, c, sqrt(a * a + b * b + c * c));
}

// ------------------------------------------------------------------------
// Boundary / far-field flux (flux.h:14-39): one thread per boundary node,
// entries in boundary-element order.
// ------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_bnd_flux(
    int n_targets, const int* __restrict__ targets, const int* __restrict__ ptr,
    const int* __restrict__ ent, const int* __restrict__ group, const double* __restrict__ bw,
    int ldb, const double* __restrict__ var, double* __restrict__ flux, int ld, Consts k,
    const int* __restrict__ status)
{
#pragma clang fp contract(off)
    if (halted(status)) return;
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= n_targets) return;
    const int n = targets[tid];
    Th b;
    thermo_exact(var[n], var[ld + n], var[2 * ld + n], var[3 * ld + n], var[4 * ld + n], b);
    double acc[5];
#pragma unroll
    for (int v = 0; v < 5; v++) acc[v] = flux[v * ld + n];
    for (int q = ptr[tid]; q < ptr[tid + 1]; q++) {
        const int e = ent[q];
        const int g = group[e];
        const double w0 = bw[e], w1 = bw[ldb + e], w2 = bw[2 * ldb + e];
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
#pragma unroll
    for (int v = 0; v < 5; v++) flux[v * ld + n] = acc[v];
}

// ------------------------------------------------------------------------
// Node-wise kernels
// ------------------------------------------------------------------------
__global__ void k_fill5(int n, double* __restrict__ x, int ld, Consts k, int use_ff)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
#pragma unroll
    for (int v = 0; v < 5; v++) x[v * ld + i] = use_ff ? k.ffv[v] : 0.0;
}

__global__ void k_fill(size_t n, double* __restrict__ x, double val)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = val;
}

__global__ void k_fill_i(size_t n, int* __restrict__ x, int val)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = val;
}

__global__ void k_scale3(int n, double* __restrict__ w, int ld, double s)
{
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    w[i] *= s;
    w[ld + i] *= s;
    w[2 * ld + i] *= s;
}

__global__ void k_copy5(int n, const double* __restrict__ src, double* __restrict__ dst, int ld,
                        const int* __restrict__ status)
{
    if (halted(status)) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
#pragma unroll
    for (int v = 0; v < 5; v++) dst[v * ld + i] = src[v * ld + i];
}

// Block-wide reductions over a 256-thread workgroup (4 waves of 64).
__device__ __forceinline__ double block_min(double x)
{
    __shared__ double sm[kBlock / kWave];
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) {
        double y = __shfl_down(x, o, kWave);
        x = (y < x) ? y : x;
    }
    const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
    if (lane == 0) sm[w] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        x = sm[0];
        for (int i = 1; i < kBlock / kWave; i++) x = (sm[i] < x) ? sm[i] : x;
    }
    return x;
}

__device__ __forceinline__ double block_sum(double x)
{
#pragma clang fp contract(off)
    __shared__ double sm[kBlock / kWave];
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) x += __shfl_down(x, o, kWave);
    const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
    if (lane == 0) sm[w] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        x = sm[0];
        for (int i = 1; i < kBlock / kWave; i++) x += sm[i];
    }
    return x;
}

__device__ __forceinline__ int block_sum_i(int x)
{
    __shared__ int smi[kBlock / kWave];
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) x += __shfl_down(x, o, kWave);
    const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
    if (lane == 0) smi[w] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        x = smi[0];
        for (int i = 1; i < kBlock / kWave; i++) x += smi[i];
    }
    return x;
}

// calculate_dt_kernel (time_stepping_kernels.h:13-32).  With 'old' non-null
// it also performs copy_double_kernel (copy_double_kernel.h:6-13) and with
// 'part' non-null it emits the per-workgroup min for get_min_dt_kernel
// (NaN never wins, as with the reference's 'dt < min_dt' test).
__global__ __launch_bounds__(kBlock) void k_calc_dt(
    int n, const double* __restrict__ var, const double* __restrict__ vol,
    double* __restrict__ dt, double* __restrict__ old, int ld, double* __restrict__ part,
    const int* __restrict__ status)
{
#pragma clang fp contract(off)
    if (halted(status)) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double d = DBL_MAX;
    if (i < n) {
        const double rho = var[i], mx = var[ld + i], my = var[2 * ld + i], mz = var[3 * ld + i],
                     E = var[4 * ld + i];
        if (old) {
            old[i] = rho; old[ld + i] = mx; old[2 * ld + i] = my; old[3 * ld + i] = mz;
            old[4 * ld + i] = E;
        }
        const double ux = mx / rho, uy = my / rho, uz = mz / rho;
        const double ssq = ux * ux + uy * uy + uz * uz;
        const double p = (kGamma - 1.0) * (E - 0.5 * rho * ssq);
        const double c = sqrt(kGamma * p / rho);
        d = 0.5 * (cbrt(vol[i]) / (sqrt(ssq) + c));
        dt[i] = d;
    }
    if (part) {
        const double m = block_min(d);
        if (threadIdx.x == 0) part[blockIdx.x] = m;
    }
}

// get_min_dt_kernel as a standalone loop: per-workgroup min of dt.
__global__ __launch_bounds__(kBlock) void k_min_part(int n, const double* __restrict__ x,
                                                      double* __restrict__ part)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const double d = (i < n) ? x[i] : DBL_MAX;
    const double m = block_min(d);
    if (threadIdx.x == 0) part[blockIdx.x] = m;
}

// Final min over the partials; slot receives min(init, partials).  With
// 'status' set, a negative minimum raises abort code 1 (euler3d.cpp:480-484).
__global__ __launch_bounds__(kBlock) void k_min_final(int np, const double* __restrict__ part,
                                                       double init, double* __restrict__ out,
                                                       int* __restrict__ status,
                                                       double* __restrict__ hist, const int* step)
{
    if (status && *status != 0) return;
    double m = DBL_MAX;
    for (int i = threadIdx.x; i < np; i += kBlock) m = (part[i] < m) ? part[i] : m;
    m = block_min(m);
    if (threadIdx.x == 0) {
        m = (m < init) ? m : init;
        *out = m;
        if (hist && step) hist[*step] = m;
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
                            const double* __restrict__ old, double* __restrict__ var, int ld,
                            const int* __restrict__ status)
{
#pragma clang fp contract(off)
    if (halted(status)) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double factor = sf[i] / (double)(3 + 1 - rk);
#pragma unroll
    for (int v = 0; v < 5; v++) {
        var[v * ld + i] = old[v * ld + i] + factor * flux[v * ld + i];
        flux[v * ld + i] = 0.0;
    }
}

// residual_kernel (validation.h:27-35); with 'part' set also the per-node
// sum of squares of calc_rms_kernel (validation.h:37-44) and with 'bad' the
// NaN/Inf count of count_bad_vals (validation.h:102-115).
__global__ __launch_bounds__(kBlock) void k_residual(
    int n, const double* __restrict__ old, const double* __restrict__ var,
    double* __restrict__ res, int ld, double* __restrict__ part, int* __restrict__ bad,
    const int* __restrict__ status)
{
#pragma clang fp contract(off)
    if (halted(status)) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double s = 0.0;
    int nb = 0;
    if (i < n) {
#pragma unroll
        for (int v = 0; v < 5; v++) {
            const double x = var[v * ld + i];
            const double r = x - old[v * ld + i];
            if (res) res[v * ld + i] = r;
            s += r * r;
            nb += (isnan(x) || isinf(x)) ? 1 : 0;
        }
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
__global__ __launch_bounds__(kBlock) void k_sumsq_part(int n, const double* __restrict__ r, int ld,
                                                        double* __restrict__ part)
{
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double s = 0.0;
    if (i < n) {
#pragma unroll
        for (int v = 0; v < 5; v++) s += r[v * ld + i] * r[v * ld + i];
    }
    const double t = block_sum(s);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// Sum of partials; *out = init + sum.  With 'hist' the L0 RMS
// sqrt(sum / n) is stored at hist[*cycle] (euler3d.cpp:532-537) and a
// non-zero bad-value count raises abort code 2 (euler3d.cpp:539-548).
__global__ __launch_bounds__(kBlock) void k_sum_final(int np, const double* __restrict__ part,
                                                       double init, double* __restrict__ out,
                                                       int nnodes, double* __restrict__ hist,
                                                       const int* __restrict__ cycle,
                                                       const int* __restrict__ bad,
                                                       int* __restrict__ status)
{
#pragma clang fp contract(off)
    if (status && *status != 0) return;
    double s = 0.0;
    for (int i = threadIdx.x; i < np; i += kBlock) s += part[i];
    s = block_sum(s);
    if (threadIdx.x == 0) {
        s = init + s;
        if (out) *out = s;
        if (hist) hist[*cycle] = sqrt(s / (double)nnodes);
        if (status && bad && *bad > 0) *status = 2;
    }
}

__global__ __launch_bounds__(kBlock) void k_count_bad(int n, const double* __restrict__ x, int ld,
                                                       int* __restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    int c = 0;
    if (i < n) {
#pragma unroll
        for (int v = 0; v < 5; v++) c += (isnan(x[v * ld + i]) || isinf(x[v * ld + i])) ? 1 : 0;
    }
    c = block_sum_i(c);
    if (threadIdx.x == 0 && c) atomicAdd(out, c);
}

// identify_differences (validation.h:46-89) fused with count_non_zeros
// (validation.h:91-100) when 'count' is set.
__global__ __launch_bounds__(kBlock) void k_identify_differences(
    int n, const double* __restrict__ test, const double* __restrict__ master,
    double* __restrict__ diff, int ld, int* __restrict__ count)
{
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    int c = 0;
    if (i < n) {
#pragma unroll
        for (int v = 0; v < 5; v++) {
            double acc = master[v * ld + i] * 10.0e-8;
            if (acc < 0.0) acc *= -1.0;
            if (acc < 3.0e-19) acc = 3.0e-19;
            double d = test[v * ld + i] - master[v * ld + i];
            if (d < 0.0) d *= -1.0;
            const double r = (d > acc) ? d : 0.0;
            if (diff) diff[v * ld + i] = r;
            c += (r > 0.0) ? 1 : 0;
        }
    }
    if (count) {
        c = block_sum_i(c);
        if (threadIdx.x == 0 && c) atomicAdd(count, c);
    }
}

__global__ __launch_bounds__(kBlock) void k_count_pos(int n, const double* __restrict__ x, int ld,
                                                       int* __restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    int c = 0;
    if (i < n) {
#pragma unroll
        for (int v = 0; v < 5; v++) c += (x[v * ld + i] > 0.0) ? 1 : 0;
    }
    c = block_sum_i(c);
    if (threadIdx.x == 0 && c) atomicAdd(out, c);
}

// ------------------------------------------------------------------------
// Multigrid transfer (mg.h:29-88)
// ------------------------------------------------------------------------

// up_pre_kernel: zero the coarse targets of every fine node (writes of the
// same value from several children are benign).
__global__ void k_up_pre(int nf, const int* __restrict__ f2c, double* __restrict__ vc, int ldc,
                         int* __restrict__ scratch, const int* __restrict__ status)
{
    if (halted(status)) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nf) return;
    const int c = f2c[i];
#pragma unroll
    for (int v = 0; v < 5; v++) vc[v * ldc + c] = 0.0;
    scratch[c] = 0;
}

// up_kernel as a gather over the children of each coarse node, children in
// fine-node order; with 'fused' the up_pre zeroing and the up_post mean are
// folded in (one pass, same arithmetic).
__global__ __launch_bounds__(kBlock) void k_up_gather(
    int n_targets, const int* __restrict__ targets, const int* __restrict__ ptr,
    const int* __restrict__ child, const double* __restrict__ vf, int ldf,
    double* __restrict__ vc, int ldc, int* __restrict__ scratch, int fused,
    const int* __restrict__ status)
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
#pragma unroll
        for (int v = 0; v < 5; v++) acc[v] = vc[v * ldc + c];
        cnt = scratch[c];
    }
    for (int q = beg; q < end; q++) {
        const int f = child[q];
#pragma unroll
        for (int v = 0; v < 5; v++) acc[v] += vf[v * ldf + f];
        cnt += 1;
    }
    if (fused) {
        const double avg = cnt == 0 ? 1.0 : 1.0 / (double)cnt;
#pragma unroll
        for (int v = 0; v < 5; v++) acc[v] *= avg;
    }
#pragma unroll
    for (int v = 0; v < 5; v++) vc[v * ldc + c] = acc[v];
    scratch[c] = cnt;
}

// up_post_kernel: mean over children (count 0 keeps the value).
__global__ void k_up_post(int nc, double* __restrict__ vc, int ldc, const int* __restrict__ scratch,
                          const int* __restrict__ status)
{
#pragma clang fp contract(off)
    if (halted(status)) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nc) return;
    const int s = scratch[i];
    const double avg = s == 0 ? 1.0 : 1.0 / (double)s;
#pragma unroll
    for (int v = 0; v < 5; v++) vc[v * ldc + i] *= avg;
}

// down_kernel (prolongation).
__global__ void k_down(int nf, const int* __restrict__ f2c, double* __restrict__ vf,
                       const double* __restrict__ rf, const double* __restrict__ xf, int ldf,
                       const double* __restrict__ rc, const double* __restrict__ xc, int ldc,
                       const int* __restrict__ status)
{
#pragma clang fp contract(off)
    if (halted(status)) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nf) return;
    const int c = f2c[i];
    const double dx = fabs(xf[i] - xc[c]);
    const double dy = fabs(xf[ldf + i] - xc[ldc + c]);
    const double dz = fabs(xf[2 * ldf + i] - xc[2 * ldc + c]);
    const double dm = sqrt(dx * dx + dy * dy + dz * dz);
    vf[i] -= dm * (rc[c] - rf[i]);
    vf[ldf + i] -= dx * (rc[ldc + c] - rf[ldf + i]);
    vf[2 * ldf + i] -= dy * (rc[2 * ldc + c] - rf[2 * ldf + i]);
    vf[3 * ldf + i] -= dz * (rc[3 * ldc + c] - rf[3 * ldf + i]);
    vf[4 * ldf + i] -= dm * (rc[4 * ldc + c] - rf[4 * ldf + i]);
}

// ------------------------------------------------------------------------
// Setup (misc.h:40-84)
// ------------------------------------------------------------------------

// calculate_cell_volumes, edge part: tetra volume per edge and the weight
// rewrite w <- (d/|d|)|w| / |d|.
__global__ void k_cell_volumes_edge(int ne, const int2* __restrict__ e2n, const double* __restrict__ x,
                                    int ldn, double* __restrict__ w, int lde, double* __restrict__ tv)
{
#pragma clang fp contract(off)
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    const int2 ab = e2n[e];
    double d[3];
    double dist = 0.0;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        d[i] = x[i * ldn + ab.y] - x[i * ldn + ab.x];
        dist += d[i] * d[i];
    }
    dist = sqrt(dist);
    double area = 0.0;
#pragma unroll
    for (int i = 0; i < 3; i++) area += w[i * lde + e] * w[i * lde + e];
    area = sqrt(area);
    tv[e] = (1.0 / 3.0) * 0.5 * dist * area;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        double r = (d[i] / dist) * area;
        r /= dist;
        w[i * lde + e] = r;
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
__global__ void k_bump(int* c) { if (threadIdx.x == 0) *c += 1; }

}  // namespace mgcfd
