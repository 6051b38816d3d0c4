// Reference-kernel harness -- TEST INFRASTRUCTURE ONLY.
//
// Compiles the reference's own per-element kernel headers where they lie
// (include path = the reference checkout, see oracle/Makefile) and drives
// them with OP2 sequential loop semantics: each op_par_loop visits its set in
// element order and every OP_INC lands in that order (the seq backend's
// behaviour for the loop sites in euler3d.cpp:412-641).  The loop drivers
// here are this project's code; the kernel bodies are the reference's.
//
// The exported ref_* functions take the same arguments as the oracle's orc_*
// functions so the tests can compare the two bit for bit.  The globals the
// reference's euler3d.cpp defines (euler3d.cpp:47-57) are defined here and
// filled from the caller's constants before each call.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <sstream>
#include <vector>

#include "const.h"
#include "structures.h"
#include "inlined_funcs.h"
#include "config.h"
#include "global.h"

double smoothing_coefficient = double(0.2f);
double ff_variable[NVAR];
double ff_flux_contribution_momentum_x[NDIM];
double ff_flux_contribution_momentum_y[NDIM];
double ff_flux_contribution_momentum_z[NDIM];
double ff_flux_contribution_density_energy[NDIM];
int mesh_name;
int levels;
int current_level;
config conf;

#include "flux.h"
#include "mg.h"
#include "time_stepping_kernels.h"
#include "validation.h"
#include "unstructured_stream.h"
#include "misc.h"
#include "copy_double_kernel.h"

extern "C" {
#include "mgcfd_oracle.h"
}

static void load_consts(const orc_consts* c)
{
    smoothing_coefficient = c->smoothing_coefficient;
    std::memcpy(ff_variable, c->ff_variable, sizeof ff_variable);
    std::memcpy(ff_flux_contribution_momentum_x, c->ff_flux_contribution_momentum_x, 3 * sizeof(double));
    std::memcpy(ff_flux_contribution_momentum_y, c->ff_flux_contribution_momentum_y, 3 * sizeof(double));
    std::memcpy(ff_flux_contribution_momentum_z, c->ff_flux_contribution_momentum_z, 3 * sizeof(double));
    std::memcpy(ff_flux_contribution_density_energy, c->ff_flux_contribution_density_energy, 3 * sizeof(double));
}

extern "C" {

// Freestream block of euler3d.cpp:156-189 with runtime Mach / angle, built on
// the reference's compute_flux_contribution.
void ref_freestream(double mach, double deg_aoa, orc_consts* c)
{
    std::memset(c, 0, sizeof(*c));
    c->smoothing_coefficient = double(0.2f);
    const double angle_of_attack = double(PI / 180.0) * double(deg_aoa);
    double fv[NVAR];
    fv[VAR_DENSITY] = double(1.4);
    double ff_pressure = double(1.0);
    double ff_speed_of_sound = sqrt(GAMMA * ff_pressure / fv[VAR_DENSITY]);
    double ff_speed = double(mach) * ff_speed_of_sound;
    double3 ff_velocity;
    ff_velocity.x = ff_speed * double(cos((double)angle_of_attack));
    ff_velocity.y = ff_speed * double(sin((double)angle_of_attack));
    ff_velocity.z = 0.0;
    fv[VAR_MOMENTUM + 0] = fv[VAR_DENSITY] * ff_velocity.x;
    fv[VAR_MOMENTUM + 1] = fv[VAR_DENSITY] * ff_velocity.y;
    fv[VAR_MOMENTUM + 2] = fv[VAR_DENSITY] * ff_velocity.z;
    fv[VAR_DENSITY_ENERGY] = fv[VAR_DENSITY] * (double(0.5) * (ff_speed * ff_speed))
                           + (ff_pressure / double(GAMMA - 1.0));
    double3 ff_momentum;
    ff_momentum.x = fv[VAR_MOMENTUM + 0];
    ff_momentum.y = fv[VAR_MOMENTUM + 1];
    ff_momentum.z = fv[VAR_MOMENTUM + 2];
    compute_flux_contribution(fv[VAR_DENSITY], ff_momentum, fv[VAR_DENSITY_ENERGY],
                              ff_pressure, ff_velocity,
                              c->ff_flux_contribution_momentum_x,
                              c->ff_flux_contribution_momentum_y,
                              c->ff_flux_contribution_momentum_z,
                              c->ff_flux_contribution_density_energy);
    std::memcpy(c->ff_variable, fv, sizeof fv);
}

void ref_initialize_variables(int n, const orc_consts* c, double* var)
{
    load_consts(c);
    for (int i = 0; i < n; i++) initialize_variables_kernel(var + 5 * (size_t)i);
}

void ref_calculate_cell_volumes(int n_edges, const int* e2n, const double* coords,
                                double* ewt, double* vol)
{
    for (size_t e = 0; e < (size_t)n_edges; e++) {
        int a = e2n[2 * e], b = e2n[2 * e + 1];
        calculate_cell_volumes(coords + 3 * (size_t)a, coords + 3 * (size_t)b, ewt + 3 * e,
                               vol + a, vol + b);
    }
}

void ref_dampen_ewt(int n, double* w)
{
    for (size_t i = 0; i < (size_t)n; i++) dampen_ewt(w + 3 * i);
}

void ref_calculate_dt(int n, const double* var, const double* vol, double* dt)
{
    for (size_t i = 0; i < (size_t)n; i++) calculate_dt_kernel(var + 5 * i, vol + i, dt + i);
}

double ref_get_min_dt(int n, const double* dt)
{
    double m = std::numeric_limits<double>::max();
    for (size_t i = 0; i < (size_t)n; i++) get_min_dt_kernel(dt + i, &m);
    return m;
}

void ref_compute_step_factor(int n, const double* var, const double* vol, double min_dt, double* sf)
{
    for (size_t i = 0; i < (size_t)n; i++)
        compute_step_factor_kernel(var + 5 * i, vol + i, &min_dt, sf + i);
}

void ref_compute_flux_edge(int n_edges, const int* e2n, const double* var, const double* ewt,
                           double* flux, const orc_consts* c)
{
    load_consts(c);
    for (size_t e = 0; e < (size_t)n_edges; e++) {
        size_t a = (size_t)e2n[2 * e], b = (size_t)e2n[2 * e + 1];
        compute_flux_edge_kernel(var + 5 * a, var + 5 * b, ewt + 3 * e, flux + 5 * a, flux + 5 * b);
    }
}

// Edge colouring for the multi-threaded CPU baseline: greedy, in edge order,
// no two edges of one colour share a node (the scheme OP2's OpenMP backend
// applies to its blocks).  order receives the edges grouped by colour,
// cptr[c]..cptr[c+1] the range of colour c; returns the colour count (at
// most max_colours, -1 if more are needed).
int ref_colour_edges(int n_edges, int n_nodes, const int* e2n, int max_colours, int* order, int* cptr)
{
    std::vector<unsigned long long> used((size_t)n_nodes, 0ull);
    std::vector<int> col((size_t)n_edges);
    int nc = 0;
    for (int e = 0; e < n_edges; e++) {
        const int a = e2n[2 * (size_t)e], b = e2n[2 * (size_t)e + 1];
        const unsigned long long u = used[a] | used[b];
        if (~u == 0ull) return -1;
        const int c = __builtin_ctzll(~u);
        if (c >= max_colours) return -1;
        col[e] = c;
        used[a] |= 1ull << c;
        used[b] |= 1ull << c;
        if (c + 1 > nc) nc = c + 1;
    }
    std::vector<int> cnt((size_t)nc + 1, 0);
    for (int e = 0; e < n_edges; e++) cnt[col[e] + 1]++;
    for (int c = 0; c < nc; c++) cnt[c + 1] += cnt[c];
    for (int c = 0; c <= nc; c++) cptr[c] = cnt[c];
    for (int e = 0; e < n_edges; e++) order[cnt[col[e]]++] = e;
    return nc;
}

// The reference's flux kernel over coloured edges, one OpenMP parallel loop
// per colour: every node's increments still come one edge at a time, with no
// two threads touching one node.
void ref_compute_flux_edge_omp(int n_colours, const int* cptr, const int* order, const int* e2n,
                               const double* var, const double* ewt, double* flux, const orc_consts* c,
                               int threads)
{
    load_consts(c);
    for (int k = 0; k < n_colours; k++) {
#pragma omp parallel for num_threads(threads) schedule(static)
        for (int q = cptr[k]; q < cptr[k + 1]; q++) {
            const size_t e = (size_t)order[q];
            const size_t a = (size_t)e2n[2 * e], b = (size_t)e2n[2 * e + 1];
            compute_flux_edge_kernel(var + 5 * a, var + 5 * b, ewt + 3 * e, flux + 5 * a, flux + 5 * b);
        }
    }
}

void ref_compute_bnd_node_flux(int n_bnd, const int* group, const int* b2n, const double* bw,
                               const double* var, double* flux, const orc_consts* c)
{
    load_consts(c);
    for (size_t i = 0; i < (size_t)n_bnd; i++) {
        size_t n = (size_t)b2n[i];
        compute_bnd_node_flux_kernel(group + i, bw + 3 * i, var + 5 * n, flux + 5 * n);
    }
}

void ref_time_step(int n, int rk, const double* sf, double* flux, const double* old, double* var)
{
    for (size_t i = 0; i < (size_t)n; i++)
        time_step_kernel(&rk, sf + i, flux + 5 * i, old + 5 * i, var + 5 * i);
}

void ref_unstructured_stream(int n_edges, const int* e2n, const double* var, const double* ewt,
                             double* flux)
{
    for (size_t e = 0; e < (size_t)n_edges; e++) {
        size_t a = (size_t)e2n[2 * e], b = (size_t)e2n[2 * e + 1];
        unstructured_stream_kernel(var + 5 * a, var + 5 * b, ewt + 3 * e, flux + 5 * a, flux + 5 * b);
    }
}

void ref_copy_double(int n, const double* var, double* old)
{
    for (size_t i = 0; i < (size_t)n; i++) copy_double_kernel(var + 5 * i, old + 5 * i);
}

void ref_residual(int n, const double* old, const double* var, double* res)
{
    for (size_t i = 0; i < (size_t)n; i++) residual_kernel(old + 5 * i, var + 5 * i, res + 5 * i);
}

double ref_calc_rms_sum(int n, const double* res)
{
    double rms = 0.0;
    for (size_t i = 0; i < (size_t)n; i++) calc_rms_kernel(res + 5 * i, &rms);
    return rms;
}

int ref_count_bad_vals(int n, const double* var)
{
    int count = 0;
    for (size_t i = 0; i < (size_t)n; i++) count_bad_vals(var + 5 * i, &count);
    return count;
}

void ref_up_pre(int n_fine, const int* f2c, double* var_c, int* scratch_c)
{
    for (size_t i = 0; i < (size_t)n_fine; i++) {
        size_t c = (size_t)f2c[i];
        up_pre_kernel(var_c + 5 * c, scratch_c + c);
    }
}

void ref_up(int n_fine, const int* f2c, const double* var_f, double* var_c, int* scratch_c)
{
    for (size_t i = 0; i < (size_t)n_fine; i++) {
        size_t c = (size_t)f2c[i];
        up_kernel(var_f + 5 * i, var_c + 5 * c, scratch_c + c);
    }
}

void ref_up_post(int n_coarse, double* var_c, const int* scratch_c)
{
    for (size_t i = 0; i < (size_t)n_coarse; i++) up_post_kernel(var_c + 5 * i, scratch_c + i);
}

void ref_down(int n_fine, const int* f2c, double* var_f, const double* res_f, const double* xyz_f,
              const double* res_c, const double* xyz_c)
{
    for (size_t i = 0; i < (size_t)n_fine; i++) {
        size_t c = (size_t)f2c[i];
        down_kernel(var_f + 5 * i, res_f + 5 * i, xyz_f + 3 * i, res_c + 5 * c, xyz_c + 3 * c);
    }
}

void ref_down_v2_pre(int n, double* wsum, double* rp)
{
    for (size_t i = 0; i < (size_t)n; i++) down_v2_kernel_pre(wsum + i, rp + 5 * i);
}

// OP2 sequential semantics: the OP_INC arguments point into the dats.
void ref_down_v2(int n_edges, const int* e2n, const int* e2mg, const double* xyz_f, const double* xyz_c,
                 const double* res_c, double* rp, double* wsum)
{
    for (size_t e = 0; e < (size_t)n_edges; e++) {
        size_t a = (size_t)e2n[2 * e], b = (size_t)e2n[2 * e + 1];
        size_t ca = (size_t)e2mg[2 * e], cb = (size_t)e2mg[2 * e + 1];
        down_v2_kernel(xyz_f + 3 * a, xyz_f + 3 * b, xyz_c + 3 * ca, xyz_c + 3 * cb, res_c + 5 * ca,
                       res_c + 5 * cb, rp + 5 * a, rp + 5 * b, wsum + a, wsum + b);
    }
}

void ref_down_v2_post(int n, const double* rp, const double* wsum, const double* res_f, double* var_f)
{
    for (size_t i = 0; i < (size_t)n; i++) down_v2_kernel_post(rp + 5 * i, wsum + i, res_f + 5 * i, var_f + 5 * i);
}

void ref_identify_differences(int n, const double* test, const double* master, double* diff)
{
    for (size_t i = 0; i < (size_t)n; i++)
        identify_differences(test + 5 * i, master + 5 * i, diff + 5 * i);
}

int ref_count_non_zeros(int n, const double* diff)
{
    int count = 0;
    for (size_t i = 0; i < (size_t)n; i++) count_non_zeros(diff + 5 * i, &count);
    return count;
}

// Setup + multigrid loop of euler3d.cpp:412-641 driven through the
// reference kernels (same loop order as the driver, OP2 seq semantics).
void ref_setup(orc_level* L, int nlev, const orc_consts* c)
{
    load_consts(c);
    for (int l = 0; l < nlev; l++) {
        ref_initialize_variables(L[l].n_nodes, c, L[l].variables);
        for (size_t i = 0; i < (size_t)L[l].n_nodes; i++) zero_5d_array_kernel(L[l].fluxes + 5 * i);
        for (size_t i = 0; i < (size_t)L[l].n_nodes; i++) zero_1d_array_kernel(L[l].volumes + i);
        ref_calculate_cell_volumes(L[l].n_edges, L[l].edge_to_node, L[l].coords, L[l].edge_weights,
                                   L[l].volumes);
    }
    for (int l = 0; l < nlev; l++) {
        ref_dampen_ewt(L[l].n_edges, L[l].edge_weights);
        ref_dampen_ewt(L[l].n_bnd, L[l].bnd_weights);
    }
}

int ref_mg_run(orc_level* L, int nlev, int cycles, const orc_consts* c, double* rms_hist,
               double* min_dt_hist, int* steps_done)
{
    load_consts(c);
    int level = 0, mg_dir = MG_UP, i = 0, steps = 0;
    while (i < cycles) {
        orc_level* l = &L[level];
        int n = l->n_nodes;
        ref_copy_double(n, l->variables, l->old_variables);
        ref_calculate_dt(n, l->variables, l->volumes, l->step_factors);
        double min_dt = ref_get_min_dt(n, l->step_factors);
        if (min_dt_hist) min_dt_hist[steps] = min_dt;
        steps++;
        if (min_dt < 0.0f) { if (steps_done) *steps_done = steps; return 1; }
        ref_compute_step_factor(n, l->variables, l->volumes, min_dt, l->step_factors);
        for (int rk = 0; rk < RK; rk++) {
            ref_compute_flux_edge(l->n_edges, l->edge_to_node, l->variables, l->edge_weights,
                                  l->fluxes, c);
            ref_compute_bnd_node_flux(l->n_bnd, l->bnd_group, l->bnd_to_node, l->bnd_weights,
                                      l->variables, l->fluxes, c);
            ref_time_step(n, rk, l->step_factors, l->fluxes, l->old_variables, l->variables);
        }
        ref_residual(n, l->old_variables, l->variables, l->residuals);
        if (level == 0) {
            double rms = ref_calc_rms_sum(n, l->residuals);
            rms = sqrt(rms / double(n));
            if (rms_hist) rms_hist[i] = rms;
            if (ref_count_bad_vals(n, l->variables) > 0) { if (steps_done) *steps_done = steps; return 2; }
        }
        if (nlev <= 1) {
            i++;
        } else if (mg_dir == MG_UP) {
            level++;
            orc_level* f = &L[level - 1];
            orc_level* cl = &L[level];
            ref_up_pre(f->n_nodes, f->node_to_mg_node, cl->variables, cl->up_scratch);
            ref_up(f->n_nodes, f->node_to_mg_node, f->variables, cl->variables, cl->up_scratch);
            ref_up_post(cl->n_nodes, cl->variables, cl->up_scratch);
            if (level == nlev - 1) mg_dir = MG_DOWN;
        } else {
            level--;
            orc_level* f = &L[level];
            orc_level* cl = &L[level + 1];
            if (f->edge_to_mg_nodes) {
                ref_down_v2_pre(f->n_nodes, f->prolonged_wsum, f->prolonged);
                ref_down_v2(f->n_edges, f->edge_to_node, f->edge_to_mg_nodes, f->coords, cl->coords,
                            cl->residuals, f->prolonged, f->prolonged_wsum);
                ref_down_v2_post(f->n_nodes, f->prolonged, f->prolonged_wsum, f->residuals, f->variables);
            } else {
                ref_down(f->n_nodes, f->node_to_mg_node, f->variables, f->residuals, f->coords,
                         cl->residuals, cl->coords);
            }
            if (level == 0) { mg_dir = MG_UP; i++; }
        }
    }
    if (steps_done) *steps_done = steps;
    return 0;
}

// ---- multi-threaded CPU baseline of the whole solve ----
//
// The same multigrid loop with every op_par_loop spread over host threads the
// way OP2's OpenMP backend spreads it: direct loops split over the set; the
// edge flux one parallel loop per colour of edge blocks (no two blocks of a
// colour share a node), each block in edge order; the indirect loops onto a
// node (boundary flux, restriction) one thread per target node, each walking
// its elements in element order;
// the min / sum reductions per thread, then combined.  The kernel bodies are
// the reference's.  Timing baseline only: the block colours change the order
// of each node's flux increments (ulp-level), and the rms sum order differs.
struct RefOmpLevel {
    std::vector<int> cptr, order;                 // block colours: blocks of colour k are order[cptr[k] ..
                                                  // cptr[k + 1]), block b = edges [b * kEB, (b + 1) * kEB)
    std::vector<int> bptr, bent;                  // boundary elements per node (CSR, element order)
    std::vector<int> bnodes;                      // nodes with boundary elements
    std::vector<int> fptr, fine;                  // fine children per coarse node (into this level's
                                                  // coarser neighbour), element order
};

struct RefOmpPlan {
    std::vector<RefOmpLevel> L;
};

// OP2's OpenMP backend executes an indirect loop as blocks of consecutive set
// elements, coloured so that no two blocks of one colour increment the same
// node; blocks of a colour run in parallel, each block in element order.
static const int kEB = 256;

// Greedy block colouring of an edge set; returns the colour count, -1 if
// more than 256 are needed.
static int colour_edge_blocks(int n_edges, int n_nodes, const int* e2n, std::vector<int>& order,
                              std::vector<int>& cptr)
{
    const int nb = (n_edges + kEB - 1) / kEB;
    constexpr int W = 4;                      // 4 x 64 colour bits per node
    std::vector<unsigned long long> used((size_t)n_nodes * W, 0ull);
    std::vector<int> col((size_t)nb);
    int nc = 0;
    for (int b = 0; b < nb; b++) {
        unsigned long long u[W] = {0, 0, 0, 0};
        const int e1 = std::min(n_edges, (b + 1) * kEB);
        for (int e = b * kEB; e < e1; e++)
            for (int k = 0; k < 2; k++) {
                const size_t v = (size_t)e2n[2 * (size_t)e + k];
                for (int w = 0; w < W; w++) u[w] |= used[v * W + w];
            }
        int c = -1;
        for (int w = 0; w < W && c < 0; w++)
            if (~u[w]) c = 64 * w + __builtin_ctzll(~u[w]);
        if (c < 0) return -1;
        col[b] = c;
        for (int e = b * kEB; e < e1; e++)
            for (int k = 0; k < 2; k++) {
                const size_t v = (size_t)e2n[2 * (size_t)e + k];
                used[v * W + c / 64] |= 1ull << (c % 64);
            }
        nc = std::max(nc, c + 1);
    }
    cptr.assign((size_t)nc + 1, 0);
    for (int b = 0; b < nb; b++) cptr[col[b] + 1]++;
    for (int c = 0; c < nc; c++) cptr[c + 1] += cptr[c];
    order.resize((size_t)nb);
    std::vector<int> pos(cptr.begin(), cptr.end() - 1);
    for (int b = 0; b < nb; b++) order[(size_t)pos[col[b]]++] = b;
    return nc;
}

static void csr_by_target(int n_src, const int* tgt, int n_tgt, std::vector<int>& ptr, std::vector<int>& ent)
{
    ptr.assign((size_t)n_tgt + 1, 0);
    for (int i = 0; i < n_src; i++) ptr[(size_t)tgt[i] + 1]++;
    for (int t = 0; t < n_tgt; t++) ptr[t + 1] += ptr[t];
    ent.resize((size_t)n_src);
    std::vector<int> pos(ptr.begin(), ptr.end() - 1);
    for (int i = 0; i < n_src; i++) ent[(size_t)pos[tgt[i]]++] = i;
}

void* ref_omp_plan_create(const orc_level* L, int nlev)
{
    RefOmpPlan* P = new RefOmpPlan();
    P->L.resize(nlev);
    for (int l = 0; l < nlev; l++) {
        RefOmpLevel& q = P->L[l];
        if (colour_edge_blocks(L[l].n_edges, L[l].n_nodes, L[l].edge_to_node, q.order, q.cptr) < 0) {
            delete P;
            return nullptr;
        }
        csr_by_target(L[l].n_bnd, L[l].bnd_to_node, L[l].n_nodes, q.bptr, q.bent);
        for (int i = 0; i < L[l].n_nodes; i++)
            if (q.bptr[i + 1] > q.bptr[i]) q.bnodes.push_back(i);
        if (l + 1 < nlev) csr_by_target(L[l].n_nodes, L[l].node_to_mg_node, L[l + 1].n_nodes, q.fptr, q.fine);
    }
    return P;
}

void ref_omp_plan_free(void* plan) { delete static_cast<RefOmpPlan*>(plan); }

static void omp_smooth_level(const RefOmpLevel& q, orc_level* l, int threads, double* min_dt_out)
{
    const int n = l->n_nodes;
#pragma omp parallel for num_threads(threads) schedule(static)
    for (int i = 0; i < n; i++) copy_double_kernel(l->variables + 5 * (size_t)i, l->old_variables + 5 * (size_t)i);
#pragma omp parallel for num_threads(threads) schedule(static)
    for (int i = 0; i < n; i++)
        calculate_dt_kernel(l->variables + 5 * (size_t)i, l->volumes + i, l->step_factors + i);
    double m = std::numeric_limits<double>::max();
#pragma omp parallel num_threads(threads)
    {
        double mt = std::numeric_limits<double>::max();
#pragma omp for schedule(static)
        for (int i = 0; i < n; i++) get_min_dt_kernel(l->step_factors + i, &mt);
#pragma omp critical
        get_min_dt_kernel(&mt, &m);
    }
    *min_dt_out = m;
}

static void omp_stages(const RefOmpLevel& q, orc_level* l, double min_dt, int threads)
{
    const int n = l->n_nodes;
#pragma omp parallel for num_threads(threads) schedule(static)
    for (int i = 0; i < n; i++)
        compute_step_factor_kernel(l->variables + 5 * (size_t)i, l->volumes + i, &min_dt, l->step_factors + i);
    const int nc = (int)q.cptr.size() - 1;
    const int nbn = (int)q.bnodes.size();
    for (int rk = 0; rk < RK; rk++) {
        for (int k = 0; k < nc; k++) {
#pragma omp parallel for num_threads(threads) schedule(dynamic)
            for (int j = q.cptr[k]; j < q.cptr[k + 1]; j++) {
                const int bk = q.order[j];
                const int e1 = std::min(l->n_edges, (bk + 1) * kEB);
                for (int ei = bk * kEB; ei < e1; ei++) {
                    const size_t e = (size_t)ei;
                    const size_t a = (size_t)l->edge_to_node[2 * e], b = (size_t)l->edge_to_node[2 * e + 1];
                    compute_flux_edge_kernel(l->variables + 5 * a, l->variables + 5 * b, l->edge_weights + 3 * e,
                                             l->fluxes + 5 * a, l->fluxes + 5 * b);
                }
            }
        }
#pragma omp parallel for num_threads(threads) schedule(static)
        for (int t = 0; t < nbn; t++) {
            const size_t v = (size_t)q.bnodes[t];
            for (int j = q.bptr[v]; j < q.bptr[v + 1]; j++) {
                const size_t bi = (size_t)q.bent[j];
                compute_bnd_node_flux_kernel(l->bnd_group + bi, l->bnd_weights + 3 * bi, l->variables + 5 * v,
                                             l->fluxes + 5 * v);
            }
        }
#pragma omp parallel for num_threads(threads) schedule(static)
        for (int i = 0; i < n; i++)
            time_step_kernel(&rk, l->step_factors + i, l->fluxes + 5 * (size_t)i, l->old_variables + 5 * (size_t)i,
                             l->variables + 5 * (size_t)i);
    }
#pragma omp parallel for num_threads(threads) schedule(static)
    for (int i = 0; i < n; i++)
        residual_kernel(l->old_variables + 5 * (size_t)i, l->variables + 5 * (size_t)i, l->residuals + 5 * (size_t)i);
}

int ref_mg_run_omp(void* plan, orc_level* L, int nlev, int cycles, const orc_consts* c, double* rms_hist,
                   int threads)
{
    const RefOmpPlan& P = *static_cast<const RefOmpPlan*>(plan);
    load_consts(c);
    int level = 0, mg_dir = MG_UP, i = 0;
    while (i < cycles) {
        orc_level* l = &L[level];
        const RefOmpLevel& q = P.L[level];
        const int n = l->n_nodes;
        double min_dt;
        omp_smooth_level(q, l, threads, &min_dt);
        if (min_dt < 0.0f) return 1;
        omp_stages(q, l, min_dt, threads);
        if (level == 0) {
            double rms = 0.0;
            int bad = 0;
#pragma omp parallel num_threads(threads)
            {
                double rt = 0.0;
                int bt = 0;
#pragma omp for schedule(static)
                for (int j = 0; j < n; j++) {
                    calc_rms_kernel(l->residuals + 5 * (size_t)j, &rt);
                    count_bad_vals(l->variables + 5 * (size_t)j, &bt);
                }
#pragma omp critical
                {
                    rms += rt;
                    bad += bt;
                }
            }
            if (rms_hist) rms_hist[i] = sqrt(rms / double(n));
            if (bad > 0) return 2;
        }
        if (nlev <= 1) {
            i++;
        } else if (mg_dir == MG_UP) {
            level++;
            orc_level* f = &L[level - 1];
            orc_level* cl = &L[level];
            const RefOmpLevel& fq = P.L[level - 1];
            const int ncn = cl->n_nodes;
            // up_pre + up + up_post (mg.h:29-64) per coarse node over its children in fine order
#pragma omp parallel for num_threads(threads) schedule(static)
            for (int cc = 0; cc < ncn; cc++) {
                const size_t cv = (size_t)cc;
                if (fq.fptr[cv + 1] == fq.fptr[cv]) continue;
                up_pre_kernel(cl->variables + 5 * cv, cl->up_scratch + cv);
                for (int j = fq.fptr[cv]; j < fq.fptr[cv + 1]; j++)
                    up_kernel(f->variables + 5 * (size_t)fq.fine[j], cl->variables + 5 * cv, cl->up_scratch + cv);
            }
#pragma omp parallel for num_threads(threads) schedule(static)
            for (int cc = 0; cc < ncn; cc++) up_post_kernel(cl->variables + 5 * (size_t)cc, cl->up_scratch + cc);
            if (level == nlev - 1) mg_dir = MG_DOWN;
        } else {
            level--;
            orc_level* f = &L[level];
            orc_level* cl = &L[level + 1];
            const int nf = f->n_nodes;
#pragma omp parallel for num_threads(threads) schedule(static)
            for (int j = 0; j < nf; j++) {
                const size_t fi = (size_t)j, cc = (size_t)f->node_to_mg_node[j];
                down_kernel(f->variables + 5 * fi, f->residuals + 5 * fi, f->coords + 3 * fi, cl->residuals + 5 * cc,
                            cl->coords + 3 * cc);
            }
            if (level == 0) { mg_dir = MG_UP; i++; }
        }
    }
    return 0;
}

}  // extern "C"
