/*
 * CPU oracle for the MG-CFD multigrid Euler hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * kernels (flux.h, flux_boundary.elem_func, flux_wall.elem_func,
 * inlined_funcs.h, time_stepping_kernels.h, mg.h, misc.h,
 * copy_double_kernel.h, validation.h, unstructured_stream.h) and of the
 * driver's multigrid loop (euler3d.cpp:412-641), executed with OP2's
 * sequential semantics: every op_par_loop visits its set in element order and
 * every OP_INC lands in that order.  Only tests/, __graft_entry__.smoke() and
 * bench.py's CPU-baseline leg may load it, and only as the checker.
 *
 * All arrays are the reference's host layout: AoS [count][dim], 0-based maps.
 *
 * Pinning (see tests/test_oracle.py):
 *   - every kernel below is compared bit-for-bit against the reference's own
 *     kernel headers compiled in place (oracle/_ref, built by oracle/Makefile);
 *   - the MG driver composition is compared bit-for-bit against the same
 *     loop order driven through the reference kernels;
 *   - closed forms: freestream (c_inf = 1), uniform-state dissipation = 0,
 *     w = 0 no-op, a<->b swap symmetry, conservation, restrict = mean,
 *     prolong formula, cell volume of a regular box, -v tolerance function.
 */
#ifndef MGCFD_ORACLE_H
#define MGCFD_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_NVAR 5
#define ORC_NDIM 3

/* Freestream and constants (euler3d.cpp:47-52,156-189, const.h:27-33). */
typedef struct {
    double smoothing_coefficient;          /* double(0.2f) */
    double ff_variable[5];
    double ff_flux_contribution_momentum_x[3];
    double ff_flux_contribution_momentum_y[3];
    double ff_flux_contribution_momentum_z[3];
    double ff_flux_contribution_density_energy[3];
} orc_consts;

void orc_freestream(double mach, double aoa_deg, orc_consts* c);

/* One multigrid level in the reference's host layout. */
typedef struct {
    int n_nodes, n_edges, n_bnd;
    const int* edge_to_node;      /* [E][2] */
    const int* bnd_to_node;       /* [B]    */
    const int* bnd_group;         /* [B]    */
    const int* node_to_mg_node;   /* [N] -> level+1, NULL on the coarsest */
    const double* coords;         /* [N][3] */
    double* edge_weights;         /* [E][3] rewritten by setup */
    double* bnd_weights;          /* [B][3] rewritten by setup */
    double* variables;            /* [N][5] */
    double* old_variables;        /* [N][5] */
    double* residuals;            /* [N][5] */
    double* fluxes;               /* [N][5] */
    double* step_factors;         /* [N]    */
    double* volumes;              /* [N]    */
    int*    up_scratch;           /* [N]    */
    /* Inverse-distance prolongation (mg.h:90-204) onto this level when
     * edge_to_mg_nodes is set ([E][2] -> level+1); NULL: down_kernel. */
    const int* edge_to_mg_nodes;  /* [E][2] */
    double* prolonged;            /* [N][5] */
    double* prolonged_wsum;       /* [N]    */
} orc_level;

/* ---- per-loop restatements (OP2 seq order) ---- */
void orc_initialize_variables(int n, const orc_consts* c, double* var);
void orc_zero_5d(int n, double* a);
void orc_zero_1d(int n, double* a);
void orc_calculate_cell_volumes(int n_edges, const int* e2n, const double* coords,
                                double* ewt, double* vol);
void orc_dampen_ewt(int n, double* w);
void orc_copy_double(int n, const double* var, double* old);
void orc_calculate_dt(int n, const double* var, const double* vol, double* dt);
double orc_get_min_dt(int n, const double* dt);
void orc_compute_step_factor(int n, const double* var, const double* vol,
                             double min_dt, double* sf);
void orc_compute_flux_edge(int n_edges, const int* e2n, const double* var,
                           const double* ewt, double* flux, const orc_consts* c);
void orc_compute_bnd_node_flux(int n_bnd, const int* group, const int* b2n,
                               const double* bw, const double* var, double* flux,
                               const orc_consts* c);
void orc_time_step(int n, int rk, const double* sf, double* flux,
                   const double* old, double* var);
void orc_unstructured_stream(int n_edges, const int* e2n, const double* var,
                             const double* ewt, double* flux);
void orc_residual(int n, const double* old, const double* var, double* res);
double orc_calc_rms_sum(int n, const double* res);
int orc_count_bad_vals(int n, const double* var);
void orc_up_pre(int n_fine, const int* f2c, double* var_c, int* scratch_c);
void orc_up(int n_fine, const int* f2c, const double* var_f, double* var_c, int* scratch_c);
void orc_up_post(int n_coarse, double* var_c, const int* scratch_c);
void orc_down(int n_fine, const int* f2c, double* var_f, const double* res_f,
              const double* xyz_f, const double* res_c, const double* xyz_c);
void orc_down_v2_pre(int n, double* wsum, double* rp);
void orc_down_v2(int n_edges, const int* e2n, const int* e2mg, const double* xyz_f,
                 const double* xyz_c, const double* res_c, double* rp, double* wsum);
void orc_down_v2_post(int n, const double* rp, const double* wsum, const double* res_f, double* var_f);
void orc_identify_differences(int n, const double* test, const double* master, double* diff);
int orc_count_non_zeros(int n, const double* diff);

/* Setup sequence euler3d.cpp:412-441 (non-legacy mode) for all levels. */
void orc_setup(orc_level* L, int levels, const orc_consts* c);

/*
 * Multigrid loop euler3d.cpp:452-641.  Returns 0 on success, 1 for the
 * negative min_dt abort (:480-484), 2 for the bad-value abort (:544-548).
 * rms_hist (may be NULL) receives the L0 RMS residual per cycle
 * (euler3d.cpp:532-537), min_dt_hist (may be NULL) the min_dt of every
 * smoothing step in order.  *steps_done receives the smoothing steps run.
 */
int orc_mg_run(orc_level* L, int levels, int cycles, const orc_consts* c,
               double* rms_hist, double* min_dt_hist, int* steps_done);

#ifdef __cplusplus
}
#endif
#endif
