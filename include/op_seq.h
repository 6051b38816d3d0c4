/*
 * op_seq.h -- OP2 user API (C++) over libmgcfd's C ABI (mgcfd_op2.h).
 *
 * A driver written against OP2 (the reference's euler3d.cpp includes
 * "op_seq.h" and "op_hdf5.h", euler3d.cpp:33-35) compiles against these two
 * headers instead and runs its loops on the MI355X:
 *   - op_par_loop(kernel, "name", set, args...) dispatches on the loop name
 *     (the kernel function itself is not called: every loop of the
 *     reference has a device implementation, and an unknown loop is an
 *     error, not a host fallback);
 *   - op_decl_const(dim, type, var) passes the variable's name, as OP2's
 *     code generator does;
 *   - op_arg_gbl / op_decl_dat are the typed forms of the C calls;
 *   - any failure prints the library's message and exits with status 1,
 *     OP2's own behaviour on a failed call.
 * Drivers that do not include the reference's kernel headers name loops with
 * OP_KERNEL(name) tags.
 */
#ifndef MGCFD_OP_SEQ_H
#define MGCFD_OP_SEQ_H

#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "mgcfd_op2.h"

inline void op_check_(int rc, const char* what)
{
    if (rc != 0) {
        std::fprintf(stderr, "%s: %s\n", what, mgcfd_last_error());
        std::fflush(stderr);
        op_exit();
        std::exit(1);
    }
}

template <typename T> struct op_type_size { static constexpr int value = sizeof(T); };

template <typename T>
inline op_arg op_arg_gbl(T* data, int dim, const char* type, op_access acc)
{
    return op_arg_gbl_char(reinterpret_cast<char*>(data), dim, type, (int)sizeof(T), acc);
}

template <typename T>
inline op_dat op_decl_dat(op_set set, int dim, const char* type, T* data, const char* name)
{
    op_dat d = op_decl_dat_char(set, dim, type, (int)sizeof(T), reinterpret_cast<const char*>(data), name);
    if (!d) op_check_(-1, "op_decl_dat");
    return d;
}

inline void op_decl_const_named_(const char* expr, int dim, const char* type, const void* data)
{
    std::string n;
    for (const char* p = expr; *p; p++)
        if (*p != '&' && *p != ' ' && *p != '(' && *p != ')') n += *p;
    const int size = std::strcmp(type, "int") == 0 ? (int)sizeof(int) : (int)sizeof(double);
    op_check_(op_decl_const_char(dim, type, size, static_cast<const char*>(data), n.c_str()), "op_decl_const");
}

/* OP2 generated code passes the name explicitly: */
template <typename T>
inline void op_decl_const2(const char* name, int dim, const char* type, T* data)
{
    op_check_(op_decl_const_char(dim, type, (int)sizeof(T), reinterpret_cast<const char*>(data), name),
              "op_decl_const2");
}
#define op_decl_const(dim, type, dat) op_decl_const_named_(#dat, dim, type, dat)

template <typename K, typename... A>
inline void op_par_loop(K /*kernel*/, const char* name, op_set set, A... args)
{
    const op_arg a[] = {args...};
    op_check_(op_par_loop_by_name(name, set, (int)sizeof...(A), a), name);
}

struct op_kernel_tag {
    const char* name;
};
#define OP_KERNEL(n) constexpr op_kernel_tag n{#n}

inline void op_printf(const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    std::vprintf(fmt, ap);
    va_end(ap);
    std::fflush(stdout);
}

#endif
