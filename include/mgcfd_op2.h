/*
 * mgcfd_op2.h -- C ABI of the MI355X (gfx950) MG-CFD hot path.
 *
 * The reference application (euler3d.cpp) reaches its kernels only through
 * the OP2 API: sets, maps and dats are declared, then every loop is an
 * op_par_loop that OP2's code generator turns into a call of
 *     op_par_loop_<kernel>(const char* name, op_set set, op_arg arg0, ...)
 * This header is that binding surface, implemented by libmgcfd.so with
 * hand-written HIP kernels and device-resident dats:
 *   - the OP2 core calls euler3d.cpp makes (decl / arg / fetch / timers),
 *   - one op_par_loop_<kernel> entry per reference loop site,
 *   - a few mgcfd_* extensions (flux-variant selection, device-resident
 *     multigrid run, kernel timing, last error).
 * All entry points are plain C: pointers, sizes and small structs by value.
 * Every op_par_loop_* returns 0 on success and a negative code on failure
 * (argument mismatch, HIP error); mgcfd_last_error() describes the failure.
 * The C++ wrapper (op_seq.h) turns a failure into OP2's behaviour: message
 * and exit(1).
 *
 * Threading: one host thread drives one library instance; calls are
 * stream-ordered on the library's own HIP stream and a call returns only
 * when any host-visible result (global reductions) is valid.
 */
#ifndef MGCFD_OP2_H
#define MGCFD_OP2_H

#include <stdint.h>

#include "mgcfd_mesh.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- OP2 access codes (the values OP2 uses in op_arg_dat / op_arg_gbl) ---- */
typedef enum { OP_READ = 0, OP_WRITE = 1, OP_RW = 2, OP_INC = 3, OP_MIN = 4, OP_MAX = 5 } op_access;

typedef struct op_set_core {
    int         index;       /* declaration order */
    int         size;        /* element count */
    const char* name;
    void*       impl;        /* library-private state (renumbering, plans) */
} op_set_core;
typedef op_set_core* op_set;

typedef struct op_map_core {
    int         index;
    op_set      from;
    op_set      to;
    int         dim;
    const char* name;
    void*       impl;
} op_map_core;
typedef op_map_core* op_map;

/* 'name' is public and mutable: euler3d.cpp:556-570,736-771 renames dats
 * before writing them out. */
typedef struct op_dat_core {
    int         index;
    op_set      set;
    int         dim;
    int         size;        /* bytes per element as declared */
    const char* type;        /* "double" or "int" */
    const char* name;
    void*       impl;
} op_dat_core;
typedef op_dat_core* op_dat;

#define OP_ARG_DAT 0
#define OP_ARG_GBL 1

typedef struct op_arg {
    int       argtype;       /* OP_ARG_DAT or OP_ARG_GBL */
    op_dat    dat;
    op_map    map;           /* NULL == OP_ID (direct access) */
    int       idx;           /* map column, -1 when direct */
    int       dim;
    int       size;          /* bytes per element */
    char*     data;          /* host pointer of an op_arg_gbl */
    const char* type;
    op_access acc;
} op_arg;

#define OP_ID ((op_map)0)

/* ---- OP2 core (euler3d.cpp call sites given per function) ---- */

/* op_init(argc, argv, diags): euler3d.cpp:119-127.  Picks HIP device 0
 * (or MGCFD_DEVICE) and parses "OP_MAPS_BASE_INDEX=<n>" from argv
 * (euler3d.cpp:104-117).  Returns 0 or a negative code. */
int  op_init(int argc, char** argv, int diags);
/* op_exit(): euler3d.cpp:482,546,824.  Frees every set, map and dat. */
void op_exit(void);

/* Host-array declarations used by the HDF5 loaders (op_decl_*_hdf5,
 * euler3d.cpp:243-336).  Data is copied in; maps use the base index given
 * to op_init (default 0). */
op_set op_decl_set(int size, const char* name);
op_map op_decl_map(op_set from, op_set to, int dim, const int* imap, const char* name);
op_dat op_decl_dat_char(op_set set, int dim, const char* type, int size,
                        const char* data, const char* name);
/* op_decl_dat_temp_char: euler3d.cpp:379-409,688 (zero-initialised). */
op_dat op_decl_dat_temp_char(op_set set, int dim, const char* type, int size,
                             const char* name);
/* op_decl_const: euler3d.cpp:232-238, by name as OP2's generated code passes
 * it: smoothing_coefficient, ff_variable, ff_flux_contribution_momentum_x/y/z,
 * ff_flux_contribution_density_energy, mesh_name. */
int  op_decl_const_char(int dim, const char* type, int size, const char* dat,
                        const char* name);

op_arg op_arg_dat(op_dat dat, int idx, op_map map, int dim, const char* type, op_access acc);
op_arg op_arg_gbl_char(char* data, int dim, const char* type, int size, op_access acc);

int  op_get_size(op_set set);                  /* euler3d.cpp:537,700,728-729 */
/* Device -> host copy in the reference's AoS layout and original element
 * order (the data behind op_fetch_data_hdf5_file, euler3d.cpp:564,740-770). */
int  op_fetch_data_char(op_dat dat, char* dest);
/* Host -> device overwrite of a dat (AoS, original order). */
int  op_upload_data_char(op_dat dat, const char* src);
/* op_partition / op_renumber (euler3d.cpp:340-377): one GPU has nothing to
 * distribute.  Every node set gets a locality order at the first loop --
 * compact parts of 256 nodes, one per flux workgroup -- and edge and
 * boundary sets follow their nodes.  The order is geometric (Z-order of the
 * coordinates) for a set passed to op_partition with a dim-3 coordinate dat
 * before its first loop, else a recursive graph bisection over its edge
 * maps.  Results are unaffected; op_fetch_data always returns the original
 * order. */
int  op_partition(const char* lib, const char* method, op_set set, op_map map, op_dat coords);
int  op_renumber(op_map map);
void op_timers(double* cpu, double* wall);     /* euler3d.cpp:450,645 */
int  op_rank(int* rank);

/* ---- op_par_loop entry points: one per reference loop site ----
 * Argument lists and access modes are exactly those of the reference call
 * sites; a mismatch returns a negative code. */

/* misc.h:10-16, site euler3d.cpp:414-415 */
int op_par_loop_initialize_variables_kernel(const char* name, op_set set, op_arg a0);
/* misc.h:18-24, sites euler3d.cpp:416-420 */
int op_par_loop_zero_5d_array_kernel(const char* name, op_set set, op_arg a0);
/* misc.h:34-38, site euler3d.cpp:424-425 */
int op_par_loop_zero_1d_array_kernel(const char* name, op_set set, op_arg a0);
/* misc.h:40-76, site euler3d.cpp:426-431 */
int op_par_loop_calculate_cell_volumes(const char* name, op_set set, op_arg a0, op_arg a1,
                                       op_arg a2, op_arg a3, op_arg a4);
/* misc.h:78-84, sites euler3d.cpp:437-440 */
int op_par_loop_dampen_ewt(const char* name, op_set set, op_arg a0);
/* copy_double_kernel.h:6-13, site euler3d.cpp:467-469 */
int op_par_loop_copy_double_kernel(const char* name, op_set set, op_arg a0, op_arg a1);
/* time_stepping_kernels.h:13-32, site euler3d.cpp:472-475 */
int op_par_loop_calculate_dt_kernel(const char* name, op_set set, op_arg a0, op_arg a1, op_arg a2);
/* time_stepping_kernels.h:34-41, site euler3d.cpp:477-479 (gbl OP_MIN) */
int op_par_loop_get_min_dt_kernel(const char* name, op_set set, op_arg a0, op_arg a1);
/* get_min_dt_kernel without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, checks and result as op_par_loop_get_min_dt_kernel, but
 * the per-workgroup minima live in a fixed device array of the library
 * (65536 workgroups: sets of at most 16,777,216 elements; larger sets are
 * refused with an error) instead of a buffer the library grows with
 * hipMalloc.  It allocates nothing, so it refuses to run while sets, maps
 * or dats declared since the last loop are pending (the next loop's
 * renumbering allocates).  Every call on one GPU shares that array: calls
 * that overlap on one GPU race. */
int mgcfd_get_min_dt_noalloc(const char* name, op_set set, op_arg a0, op_arg a1);
/* time_stepping_kernels.h:43-64, site euler3d.cpp:485-489 */
int op_par_loop_compute_step_factor_kernel(const char* name, op_set set, op_arg a0, op_arg a1,
                                           op_arg a2, op_arg a3);
/* flux.h:41-208, site euler3d.cpp:498-503 -- the hot loop */
int op_par_loop_compute_flux_edge_kernel(const char* name, op_set set, op_arg a0, op_arg a1,
                                         op_arg a2, op_arg a3, op_arg a4);
/* flux.h:14-39, site euler3d.cpp:505-509 */
int op_par_loop_compute_bnd_node_flux_kernel(const char* name, op_set set, op_arg a0, op_arg a1,
                                             op_arg a2, op_arg a3);
/* time_stepping_kernels.h:66-86, site euler3d.cpp:511-516 */
int op_par_loop_time_step_kernel(const char* name, op_set set, op_arg a0, op_arg a1, op_arg a2,
                                 op_arg a3, op_arg a4);
/* unstructured_stream.h:7-57, site euler3d.cpp:519-524 */
int op_par_loop_unstructured_stream_kernel(const char* name, op_set set, op_arg a0, op_arg a1,
                                           op_arg a2, op_arg a3, op_arg a4);
/* validation.h:27-35, site euler3d.cpp:528-531 */
int op_par_loop_residual_kernel(const char* name, op_set set, op_arg a0, op_arg a1, op_arg a2);
/* validation.h:37-44, site euler3d.cpp:534-536 (gbl OP_INC) */
int op_par_loop_calc_rms_kernel(const char* name, op_set set, op_arg a0, op_arg a1);
/* validation.h:102-115, sites euler3d.cpp:540-542,665-667 (gbl OP_INC) */
int op_par_loop_count_bad_vals(const char* name, op_set set, op_arg a0, op_arg a1);
/* mg.h:29-39, site euler3d.cpp:581-583 */
int op_par_loop_up_pre_kernel(const char* name, op_set set, op_arg a0, op_arg a1);
/* mg.h:41-52, site euler3d.cpp:585-588 */
int op_par_loop_up_kernel(const char* name, op_set set, op_arg a0, op_arg a1, op_arg a2);
/* mg.h:54-64, site euler3d.cpp:590-592 */
int op_par_loop_up_post_kernel(const char* name, op_set set, op_arg a0, op_arg a1);
/* mg.h:66-88, site euler3d.cpp:626-631 */
int op_par_loop_down_kernel(const char* name, op_set set, op_arg a0, op_arg a1, op_arg a2,
                            op_arg a3, op_arg a4);
/* Inverse-distance prolongation, mg.h:90-204, sites euler3d.cpp:606-623.
 * The reference driver runs these only when its edge-->mg_nodes map is
 * declared, which it never is (euler3d.cpp:286-292); mgcfd_decl_map_compose
 * builds that map.  down_v2_kernel takes the ten arguments of
 * euler3d.cpp:609-619 (fine coords and prolonged sums through edge-->node,
 * coarse coords and residuals through edge-->mg_nodes) and is bitwise OP2's
 * sequential backend, including its OP_INC-by-pointer overwrite for fine
 * nodes that coincide with their coarse parent (mg.h:124-132,160-167). */
int op_par_loop_down_v2_kernel_pre(const char* name, op_set set, op_arg a0, op_arg a1);
int op_par_loop_down_v2_kernel(const char* name, op_set set, op_arg a0, op_arg a1, op_arg a2, op_arg a3,
                               op_arg a4, op_arg a5, op_arg a6, op_arg a7, op_arg a8, op_arg a9);
int op_par_loop_down_v2_kernel_post(const char* name, op_set set, op_arg a0, op_arg a1, op_arg a2,
                                    op_arg a3);
/* validation.h:46-89, site euler3d.cpp:690-693 */
int op_par_loop_identify_differences(const char* name, op_set set, op_arg a0, op_arg a1, op_arg a2);
/* validation.h:91-100, site euler3d.cpp:696-698 (gbl OP_INC) */
int op_par_loop_count_non_zeros(const char* name, op_set set, op_arg a0, op_arg a1);

/* Dispatch by loop name (the second argument of every op_par_loop call in
 * euler3d.cpp) to the op_par_loop_<name> entry above; nargs must match.
 * Unknown names return a negative code: there is no host fallback. */
int op_par_loop_by_name(const char* name, op_set set, int nargs, const op_arg* args);

/* ---- mgcfd_* extensions (no OP2 counterpart) ---- */

/* The composition C[e][k] = B[A[e][k]] of a map A: X -> Y and a dimension-1
 * map B: Y -> Z, as a new map X -> Z of A's dimension: the edge-->mg_nodes
 * map that euler3d.cpp:286-292 describes (edge-->node, then node-->mg_node)
 * and leaves undeclared.  Its columns may name one node twice.  Returns
 * NULL on error (mgcfd_last_error). */
op_map mgcfd_decl_map_compose(op_map outer, op_map inner, const char* name);

/* Edge-flux implementation used by op_par_loop_compute_flux_edge_kernel:
 *   MGCFD_FLUX_GATHER_EXACT  blocked gather: one workgroup per mesh part
 *                            stages the part's and its halo's thermodynamic
 *                            state in LDS, each node sums its incidences in
 *                            edge order with the reference's operation order
 *                            and no FMA contraction: bitwise equal to OP2's
 *                            sequential backend (default);
 *   MGCFD_FLUX_GATHER_FAST   same, one reciprocal per density and FMA
 *                            contraction (deterministic, ulp-level drift);
 *   MGCFD_FLUX_ATOMIC        edge-parallel with FP64 hardware atomics
 *                            (non-deterministic summation order);
 *   MGCFD_FLUX_NODAL_EXACT   node-centric gather without LDS staging, every
 *                            incidence re-evaluates the neighbour's state
 *                            (bitwise equal to the sequential backend);
 *   MGCFD_FLUX_PART_COLOURED edge-parallel over mesh parts: node state and
 *                            increments in LDS, each edge evaluated once,
 *                            accumulation colour by colour (deterministic,
 *                            ulp-level drift from the sequential order);
 *   MGCFD_FLUX_PART_ATOMIC   same with LDS FP64 atomics instead of colours
 *                            (non-deterministic summation order);
 *   MGCFD_FLUX_MIXED_COLOURED / MGCFD_FLUX_MIXED_ATOMIC
 *                            the part kernels with the flux arithmetic in
 *                            FP32 (node records, weights, increments) and
 *                            FP64 accumulation; every other loop stays FP64
 *                            (step factor, time step, restriction sums). */
#define MGCFD_FLUX_GATHER_EXACT   0
#define MGCFD_FLUX_GATHER_FAST    1
#define MGCFD_FLUX_ATOMIC         2
#define MGCFD_FLUX_NODAL_EXACT    3
#define MGCFD_FLUX_PART_COLOURED  4
#define MGCFD_FLUX_PART_ATOMIC    5
#define MGCFD_FLUX_MIXED_COLOURED 6
#define MGCFD_FLUX_MIXED_ATOMIC   7
#define MGCFD_FLUX_NVARIANTS      8
int mgcfd_set_flux_variant(int variant);
int mgcfd_get_flux_variant(void);

/* Blocked-gather plan of an edge->node map (variants GATHER_*): stats[0]
 * workgroups, [1] staged node slots per workgroup, [2] incidences (2 per
 * edge), [3] max node degree, [4] dynamic LDS bytes per workgroup. */
int mgcfd_plan_stats(op_map map, int* stats);

/* Part plan of an edge->node map (variants PART_* / MIXED_*), built if
 * needed: stats[0] parts (one workgroup each), [1] largest and [2] median
 * staged node records of a part (owned + halo), [3] local edges over all
 * parts (cut edges count once per side), [4] largest local edges of a part,
 * [5] dynamic LDS bytes per workgroup of the FP64 kernel. */
int mgcfd_part_plan_stats(op_map map, int* stats);

/* Part plan configuration of an edge->node map (variants PART_* / MIXED_*):
 * part_nodes in [16, 256] owned nodes per part; part_threads 0 (default),
 * 256 or 512 for the kernel with the fused Runge-Kutta stage; record_bytes
 * 0 (by the plan's rule), 40 or 48 for the FP64 edge-flux kernel's LDS node
 * records.  Re-plans the map. */
int mgcfd_map_part_config(op_map map, int part_nodes, int part_threads, int record_bytes);
/* Last error message of this thread ("" if none). */
const char* mgcfd_last_error(void);

/* Blocks until all queued device work is complete. */
int mgcfd_sync(void);

/* Device name and compute-unit count of the current device. */
int mgcfd_device_info(char* name, int name_len, int* n_cu);

/*
 * Device-resident multigrid (the loop of euler3d.cpp:452-641 without host
 * round trips).  The caller binds, per level, the sets/maps/dats that play
 * the reference roles, then runs cycles.  min_dt, the RMS residual and the
 * abort conditions stay on the device; the host reads them once per run.
 * Levels are bound in order 0..levels-1.
 */
typedef struct mgcfd_level_bind {
    op_set nodes, edges, bnd_nodes;
    op_map edge_to_node, bnd_to_node, node_to_mg_node;   /* last is NULL on the coarsest */
    op_dat variables, old_variables, residuals, fluxes, step_factors, volumes;
    op_dat coords, edge_weights, bnd_weights, bnd_groups, up_scratch;
} mgcfd_level_bind;

int mgcfd_mg_bind(int levels, const mgcfd_level_bind* binds);
/* Runs 'cycles' multigrid cycles.  use_graph != 0 captures one cycle in a
 * hipGraph and replays it.  rms_hist (host, may be NULL) receives the L0 RMS
 * per cycle.  *status: 0 ok, 1 negative min_dt (euler3d.cpp:480-484),
 * 2 NaN/Inf at L0 (euler3d.cpp:544-548); on an abort the state stops
 * changing at the failing smoothing step.  Returns 0 or a negative code. */
int mgcfd_mg_run(int cycles, int use_graph, double* rms_hist, int* status);

/* Times 'iters' back-to-back launches of the edge-flux loop with HIP events
 * on the library stream; *ms_per_call receives the mean.  The launches are
 * captured into one graph and replayed once between the events, so the
 * host's per-launch work does not leave the device idle between them.  The
 * copies (n_copies >= 1 maps with their variables, weights and fluxes)
 * rotate so that the working set can exceed the last-level cache. */
int mgcfd_time_flux(int variant, int n_copies, const op_map* edge_maps, const op_dat* vars,
                    const op_dat* weights, const op_dat* fluxes, int iters, double* ms_per_call);

/* The same launches with an event between consecutive calls: ms_each[iters]
 * receives each call's time (for the distribution; the events between calls
 * add a little to each). */
int mgcfd_time_flux_calls(int variant, int n_copies, const op_map* edge_maps, const op_dat* vars,
                          const op_dat* weights, const op_dat* fluxes, int iters, double* ms_each);

/* The same timing for the unstructured-stream loop (unstructured_stream.h:7-57,
 * which the reference runs beside the flux under -b, euler3d.cpp:518-523):
 * the flux loop's gathers and increments with almost no arithmetic, the
 * access pattern's own ceiling.  'fluxes' receive the stream's dummy sums. */
int mgcfd_time_stream(int n_copies, const op_map* edge_maps, const op_dat* vars, const op_dat* weights,
                      const op_dat* fluxes, int iters, double* ms_per_call);

/*
 * Batched solves (no counterpart in the reference, which runs one mesh per
 * process): n independent multigrid meshes, all with the same number of
 * levels, concatenated level by level behind offset tables and advanced
 * together -- one launch per loop for the whole batch, one captured graph per
 * cycle.  Each mesh has its own freestream (mach[i], aoa_deg[i]), min dt, RMS
 * history and status; a mesh's result equals its solve alone up to the
 * summation order of the chosen flux variant (bitwise for GATHER_EXACT).
 * Batches run on their own HIP stream.
 */
typedef struct mgcfd_batch_s* mgcfd_batch;
/* Declares the concatenated sets/maps/dats and runs the setup loops
 * (euler3d.cpp:412-441) for every mesh. */
int mgcfd_batch_create(int n, const mgcfd_mesh* meshes, const double* mach, const double* aoa_deg,
                       int flux_variant, mgcfd_batch* out);
/* Autotuner of the part variants' kernels (the edge flux with its fused
 * Runge-Kutta stage, which every cycle runs): per level, the nodes per part
 * (256, 192 or 128), threads per part (256 or 512) and launch form (one
 * workgroup per part, or 2 or 4 persistent workgroups per CU walking their
 * XCD's parts) that give the fastest captured multigrid cycle of the whole
 * batch ('iters' timed cycles per candidate, levels tuned in turn).  The
 * batch's state is left as it was.  part_nodes / part_threads /
 * persistent_per_cu [levels] (nullable) receive the choice; variants
 * without a part plan have nothing to tune (0s).  mgcfd_batch_autotune is
 * the same tuning, reporting level 0's part size. */
int mgcfd_batch_autotune(mgcfd_batch b, int iters, int* part_nodes);
int mgcfd_batch_autotune_levels(mgcfd_batch b, int iters, int* part_nodes, int* part_threads,
                                int* persistent_per_cu);
/* Sets one level's part configuration directly (a tuning found earlier):
 * part_nodes in [16, 256], part_threads 256 or 512, persistent_per_cu 0
 * (one workgroup per part) or 1..8.  Re-plans the level. */
int mgcfd_batch_set_part_config(mgcfd_batch b, int level, int part_nodes, int part_threads, int persistent_per_cu);
/* Runs 'cycles' multigrid cycles from the setup state; rms receives
 * [n][cycles] L0 RMS residuals and status [n] (0 ok, 1 negative min dt,
 * 2 NaN/Inf; either may be NULL). */
int mgcfd_batch_run(mgcfd_batch b, int cycles, int use_graph, double* rms, int* status);
/* Copies one mesh's dat ("variables", "fluxes", "step_factors", "volumes")
 * at one level, in the mesh's own node order. */
int mgcfd_batch_fetch(mgcfd_batch b, int mesh, int level, const char* what, double* out);
/* Frees the batch's device memory. */
int mgcfd_batch_destroy(mgcfd_batch b);

/*
 * Served solves (no counterpart in the reference): a request loop with
 * continuous batching.  Each request brings its own mesh (copied at submit),
 * freestream, flux variant (FP64 or mixed precision), cycle cap and RMS-drop
 * target (stop at the first cycle with rms[c] <= rms_drop * rms[0]; 0 = run
 * to the cap).  Requests are admitted in arrival order while the estimated
 * device memory of the admitted ones stays within hbm_budget_bytes and at
 * most max_active contexts run at once.  The request at the head of the
 * queue is admitted together with the compatible requests queued behind it
 * (same flux variant, level count, cycle cap and RMS-drop target; see
 * mgcfd_server_set_grouping): the group runs as one device-resident context
 * whose mesh segments are its members, on its own stream, so one launch per
 * loop serves them all; its cycles are issued as plain launches in chunks of
 * cycles_per_chunk.  Groups that arrive while others run overlap with them
 * on the device.  A request's result is the same alone and in any mix
 * (bitwise with the deterministic variants: exact gathers and coloured part
 * kernels; within rounding with the atomic ones): a member that converges
 * has its state copied at that cycle, and one that aborts (negative dt, NaN)
 * is served again alone.
 * The caller drives the loop with mgcfd_server_progress (non-blocking).
 */
typedef struct mgcfd_server_s* mgcfd_server;
int mgcfd_server_create(double hbm_budget_bytes, int max_active, int cycles_per_chunk, mgcfd_server* out);
/* Group limits of the server: at most max_group requests (default 8) and
 * max_group_nodes level-0 nodes (default 4,000,000; a lone request above it
 * still runs) per group.  max_group = 1 runs every request in a context of
 * its own. */
int mgcfd_server_set_grouping(mgcfd_server s, int max_group, long long max_group_nodes);
int mgcfd_server_submit(mgcfd_server s, const mgcfd_mesh* mesh, double mach, double aoa_deg, int flux_variant,
                        int max_cycles, double rms_drop, int* request_id);
/* Finishes, continues and admits requests; *n_active / *n_queued (nullable)
 * receive the counts afterwards. */
int mgcfd_server_progress(mgcfd_server s, int* n_active, int* n_queued);   /* n_active: requests */
/* *state: 0 queued, 1 running, 2 done (then status -- 0 ran to the cap,
 * 1 negative min dt, 2 NaN/Inf, 3 converged --, cycles run, latency from
 * submit to completion in seconds, level-0 variables [N][5] and the RMS
 * history [cycles] are filled; any output pointer may be NULL).  A request
 * that could not be set up returns an error. */
int mgcfd_server_result(mgcfd_server s, int request_id, int* state, int* status, int* cycles, double* latency_s,
                        double* variables, double* rms);
/* Device timeline of a finished request: when its stream reached its first
 * cycle and when its last cycle completed, in ms since the server was
 * created (HIP events on the request's stream).  Two requests overlapped on
 * the device when their intervals intersect. */
int mgcfd_server_timing(mgcfd_server s, int request_id, double* device_start_ms, double* device_end_ms);
/* The device-memory estimate that admission charges for a mesh.
 * mgcfd_server_submit refuses a request whose estimate exceeds the whole
 * budget. */
int mgcfd_server_estimate_bytes(const mgcfd_mesh* mesh, double* bytes);
int mgcfd_server_destroy(mgcfd_server s);

/* Times 'iters' launches of a plain device-to-device copy of 'bytes' bytes
 * (read 'bytes', write 'bytes'); *ms_per_call receives the mean.  The
 * calibration for achieved bandwidth and for the HBM traffic counters. */
int mgcfd_stream_copy(long long bytes, int iters, double* ms_per_call);

#ifdef __cplusplus
}
#endif
#endif
