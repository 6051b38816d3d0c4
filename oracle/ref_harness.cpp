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

}  // extern "C"
