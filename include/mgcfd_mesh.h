/*
 * mgcfd_mesh.h -- seeded synthetic multigrid meshes in the reference's
 * per-level layout (the datasets euler3d.cpp:243-336 loads from HDF5):
 *   node_coordinates [N][3] f64, edge-->node [E][2] i32, edge_weights [E][3],
 *   bnd_node-->node [B] i32, bnd_node-->group [B] i32, bnd_node_weights [B][3],
 *   node-->mg_node [N] i32 (into level+1, absent on the coarsest level).
 * Maps are 0-based.  The reference ships no generator (its meshes are an
 * external download, README.md:80-103); these generators produce the
 * workloads the benchmark configurations describe.
 *
 * Boundary groups follow flux.h:29-37: group 0 is the pressure-only surface
 * branch (here the z = 0 "wall" face), group 3 the far-field branch.
 */
#ifndef MGCFD_MESH_H
#define MGCFD_MESH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mgcfd_mesh_level {
    int n_nodes, n_edges, n_bnd;
    double* coords;          /* [N][3] */
    int*    edge_to_node;    /* [E][2] */
    double* edge_weights;    /* [E][3] */
    int*    bnd_to_node;     /* [B]    */
    int*    bnd_group;       /* [B]    */
    double* bnd_weights;     /* [B][3] */
    int*    node_to_mg_node; /* [N] or NULL on the coarsest level */
} mgcfd_mesh_level;

#define MGCFD_MAX_LEVELS 8

typedef struct mgcfd_mesh {
    int levels;
    mgcfd_mesh_level L[MGCFD_MAX_LEVELS];
} mgcfd_mesh;

/* Structured n^3 box on [0,1]^3, coordinates jittered by +-20% of the local
 * spacing, node / edge / boundary order and edge orientation randomly
 * permuted, axis edges only (3 n^2 (n-1) edges).  Coarser levels halve n per
 * axis (n/2^l) with the fine-to-coarse map (i/2, j/2, k/2).  Boundary: one
 * entry per node and box face (6 n^2 in total); z = 0 face group 0, the other
 * faces group 3; weights are inward face normals times the dual face area. */
int mgcfd_mesh_box(int n, int levels, uint64_t seed, mgcfd_mesh** out);

/* Jittered-point nearest-neighbour graph: n_nodes points on a jittered
 * lattice (+-20% of spacing), each joined to its 6 nearest candidates, then
 * trimmed (longest first, never below degree 3) to n_edges edges.  Boundary:
 * z = 0 face group 0 (~n^(2/3) "wall" nodes), the x = 0, x = max and z = max
 * faces group 3.  Coarser levels come from greedy pairwise agglomeration
 * (shortest edge first, then merging singletons) to
 * round(n_nodes * ratio[l]) nodes; coarse coordinates are aggregate
 * centroids, coarse edges the unique aggregate pairs of fine edges with
 * summed face areas, coarse boundary entries the summed fine entries per
 * (aggregate, group).  ratio may be NULL (1, 0.55, 0.37, 0.27, ...). */
int mgcfd_mesh_knn(int n_nodes, int n_edges, int levels, const double* ratio, uint64_t seed,
                   mgcfd_mesh** out);

void mgcfd_mesh_free(mgcfd_mesh* m);

/* Last error of the mesh functions ("" if none). */
const char* mgcfd_mesh_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
