// libmgcfd_hdf5: OP2's HDF5 declarations over libmgcfd (see include/op_hdf5.h).
#include "op_hdf5.h"

#include <hdf5.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

int fail(const std::string& m)
{
    g_err = m;
    return -1;
}

[[noreturn]] void die(const std::string& m)
{
    std::fprintf(stderr, "%s\n", m.c_str());
    std::fflush(stderr);
    op_exit();
    std::exit(1);
}

struct H5Quiet {
    H5E_auto2_t f;
    void* d;
    H5Quiet() { H5Eget_auto2(H5E_DEFAULT, &f, &d); H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr); }
    ~H5Quiet() { H5Eset_auto2(H5E_DEFAULT, f, d); }
};

hid_t native(const char* type)
{
    if (!type) return -1;
    if (!std::strcmp(type, "double")) return H5T_NATIVE_DOUBLE;
    if (!std::strcmp(type, "int")) return H5T_NATIVE_INT;
    return -1;
}

}  // namespace

extern "C" {

const char* mgcfd_h5_last_error(void) { return g_err.c_str(); }

int mgcfd_h5_shape(const char* file, const char* dataset, long long* rows, long long* cols)
{
    H5Quiet q;
    hid_t f = H5Fopen(file, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0) return fail(std::string("cannot open HDF5 file ") + file);
    hid_t d = H5Dopen2(f, dataset, H5P_DEFAULT);
    if (d < 0) { H5Fclose(f); return fail(std::string("no dataset '") + dataset + "' in " + file); }
    hid_t sp = H5Dget_space(d);
    int nd = H5Sget_simple_extent_ndims(sp);
    hsize_t dims[2] = {0, 1};
    int rc = 0;
    if (nd < 1 || nd > 2) rc = fail(std::string("dataset '") + dataset + "' is not 1-D or 2-D");
    else H5Sget_simple_extent_dims(sp, dims, nullptr);
    H5Sclose(sp);
    H5Dclose(d);
    H5Fclose(f);
    if (rc) return rc;
    if (rows) *rows = (long long)dims[0];
    if (cols) *cols = nd == 2 ? (long long)dims[1] : 1;
    return 0;
}

int mgcfd_h5_read(const char* file, const char* dataset, const char* type, void* out, long long rows,
                  long long cols)
{
    long long r = 0, c = 0;
    if (mgcfd_h5_shape(file, dataset, &r, &c)) return -1;
    if (r != rows || c != cols)
        return fail(std::string("dataset '") + dataset + "' in " + file + " has shape " + std::to_string(r) + "x" +
                    std::to_string(c) + ", expected " + std::to_string(rows) + "x" + std::to_string(cols));
    hid_t t = native(type);
    if (t < 0) return fail(std::string("unsupported type ") + (type ? type : "(null)"));
    H5Quiet q;
    hid_t f = H5Fopen(file, H5F_ACC_RDONLY, H5P_DEFAULT);
    hid_t d = H5Dopen2(f, dataset, H5P_DEFAULT);
    herr_t e = H5Dread(d, t, H5S_ALL, H5S_ALL, H5P_DEFAULT, out);
    H5Dclose(d);
    H5Fclose(f);
    if (e < 0) return fail(std::string("read of '") + dataset + "' failed");
    return 0;
}

int mgcfd_h5_write(const char* file, const char* dataset, const char* type, const void* data, long long rows,
                   long long cols, int truncate_file)
{
    hid_t t = native(type);
    if (t < 0) return fail(std::string("unsupported type ") + (type ? type : "(null)"));
    H5Quiet q;
    hid_t f = -1;
    if (!truncate_file) f = H5Fopen(file, H5F_ACC_RDWR, H5P_DEFAULT);
    if (f < 0) f = H5Fcreate(file, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (f < 0) return fail(std::string("cannot create HDF5 file ") + file);
    if (H5Lexists(f, dataset, H5P_DEFAULT) > 0) H5Ldelete(f, dataset, H5P_DEFAULT);
    hsize_t dims[2] = {(hsize_t)rows, (hsize_t)cols};
    hid_t sp = H5Screate_simple(2, dims, nullptr);
    hid_t d = H5Dcreate2(f, dataset, t, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    herr_t e = (d < 0) ? -1 : H5Dwrite(d, t, H5S_ALL, H5S_ALL, H5P_DEFAULT, data);
    if (d >= 0) H5Dclose(d);
    H5Sclose(sp);
    H5Fclose(f);
    if (e < 0) return fail(std::string("write of '") + dataset + "' to " + file + " failed");
    return 0;
}

op_set op_decl_set_hdf5_infer_size(const char* file, const char* name, const char* dataset)
{
    long long rows = 0, cols = 0;
    if (mgcfd_h5_shape(file, dataset, &rows, &cols)) die("op_decl_set_hdf5_infer_size: " + g_err);
    op_set s = op_decl_set((int)rows, name);
    if (!s) die(std::string("op_decl_set_hdf5_infer_size: ") + mgcfd_last_error());
    return s;
}

op_map op_decl_map_hdf5(op_set from, op_set to, int dim, const char* file, const char* dataset)
{
    std::vector<int> buf((size_t)from->size * dim);
    if (mgcfd_h5_read(file, dataset, "int", buf.data(), from->size, dim)) die("op_decl_map_hdf5: " + g_err);
    op_map m = op_decl_map(from, to, dim, buf.data(), dataset);
    if (!m) die(std::string("op_decl_map_hdf5: ") + mgcfd_last_error());
    return m;
}

op_dat op_decl_dat_hdf5(op_set set, int dim, const char* type, const char* file, const char* dataset)
{
    const bool is_int = type && !std::strcmp(type, "int");
    const size_t esz = is_int ? sizeof(int) : sizeof(double);
    std::vector<char> buf((size_t)set->size * dim * esz);
    if (mgcfd_h5_read(file, dataset, type, buf.data(), set->size, dim)) die("op_decl_dat_hdf5: " + g_err);
    op_dat d = op_decl_dat_char(set, dim, type, (int)esz, buf.data(), dataset);
    if (!d) die(std::string("op_decl_dat_hdf5: ") + mgcfd_last_error());
    return d;
}

void op_fetch_data_hdf5_file(op_dat dat, const char* file)
{
    const bool is_int = dat->type && !std::strcmp(dat->type, "int");
    const size_t esz = is_int ? sizeof(int) : sizeof(double);
    std::vector<char> buf((size_t)dat->set->size * dat->dim * esz);
    if (op_fetch_data_char(dat, buf.data())) die(std::string("op_fetch_data_hdf5_file: ") + mgcfd_last_error());
    if (mgcfd_h5_write(file, dat->name, dat->type, buf.data(), dat->set->size, dat->dim, 0))
        die("op_fetch_data_hdf5_file: " + g_err);
}

}  // extern "C"
