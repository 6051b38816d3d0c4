/*
 * op_hdf5.h -- OP2's HDF5 declarations (euler3d.cpp:243-336,557-570,736-771)
 * implemented by libmgcfd_hdf5.so on top of libmgcfd.
 *
 * On-disk layout (this project's choice; OP2 defines its own internally and
 * it is not part of the reference): one 2-D dataset per set/map/dat,
 * [count][dim] row-major, H5T_NATIVE_INT for maps and "int" dats,
 * H5T_NATIVE_DOUBLE for "double" dats.  Map entries are read with the base
 * index given to op_init ("OP_MAPS_BASE_INDEX=<n>", euler3d.cpp:104-117).
 * Every call prints a message and exits with status 1 on failure, as OP2
 * does.
 */
#ifndef MGCFD_OP_HDF5_H
#define MGCFD_OP_HDF5_H

#include "mgcfd_op2.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Set whose size is the first dimension of 'dataset' (euler3d.cpp:248-267). */
op_set op_decl_set_hdf5_infer_size(const char* file, const char* name, const char* dataset);
/* Map [from->size][dim] read from 'dataset' (euler3d.cpp:270-283). */
op_map op_decl_map_hdf5(op_set from, op_set to, int dim, const char* file, const char* dataset);
/* Dat [set->size][dim] of 'type' read from 'dataset' (euler3d.cpp:276,297-311,327). */
op_dat op_decl_dat_hdf5(op_set set, int dim, const char* type, const char* file, const char* dataset);
/* Writes dat (original element order, AoS) as dataset dat->name into 'file',
 * creating the file if needed and replacing a dataset of the same name
 * (euler3d.cpp:564,740-770). */
void op_fetch_data_hdf5_file(op_dat dat, const char* file);

/* Plain helpers for tools and tests (0 = ok, negative = error). */
int mgcfd_h5_shape(const char* file, const char* dataset, long long* rows, long long* cols);
int mgcfd_h5_read(const char* file, const char* dataset, const char* type, void* out, long long rows,
                  long long cols);
int mgcfd_h5_write(const char* file, const char* dataset, const char* type, const void* data, long long rows,
                   long long cols, int truncate_file);
const char* mgcfd_h5_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
