/*
 * CPU oracle for the MG-CFD hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * Plain-C restatement of the reference kernels and multigrid driver, with
 * OP2 sequential loop semantics (element order, INC in element order).  Every
 * expression keeps the reference's operand order; build with
 * -ffp-contract=off so no fused multiply-adds change the rounding.
 * Each function cites the reference file:line it restates.
 */
#include "mgcfd_oracle.h"

#include <float.h>
#include <math.h>
#include <stddef.h>
#include <string.h>

#define GAMMA 1.4
#define RK 3
#define PI_CONST 3.1415926535897931

/* ---------------- thermodynamic helpers: inlined_funcs.h:70-129 ---------- */

typedef struct { double x, y, z; } v3;

/* compute_velocity (non-IDIVIDE), inlined_funcs.h:120-125 */
static inline v3 th_velocity(double rho, v3 m)
{
    v3 u;
    u.x = m.x / rho;
    u.y = m.y / rho;
    u.z = m.z / rho;
    return u;
}

/* compute_speed_sqd, inlined_funcs.h:98-101 */
static inline double th_speed_sqd(v3 u)
{
    return u.x * u.x + u.y * u.y + u.z * u.z;
}

/* compute_pressure, inlined_funcs.h:103-106 */
static inline double th_pressure(double rho, double rhoE, double ssq)
{
    return (GAMMA - 1.0) * (rhoE - 0.5 * rho * ssq);
}

/* compute_speed_of_sound (non-IDIVIDE), inlined_funcs.h:126-129 */
static inline double th_sound(double rho, double p)
{
    return sqrt(GAMMA * p / rho);
}

/* compute_flux_contribution, inlined_funcs.h:70-96 */
typedef struct { double mx[3], my[3], mz[3], de[3]; } fcontrib;

static inline void th_flux_contribution(double rhoE, double p, v3 m, v3 u, fcontrib* f)
{
    f->mx[0] = u.x * m.x + p;
    f->mx[1] = u.x * m.y;
    f->mx[2] = u.x * m.z;

    f->my[0] = f->mx[1];
    f->my[1] = u.y * m.y + p;
    f->my[2] = u.y * m.z;

    f->mz[0] = f->mx[2];
    f->mz[1] = f->my[2];
    f->mz[2] = u.z * m.z + p;

    double de_p = rhoE + p;
    f->de[0] = u.x * de_p;
    f->de[1] = u.y * de_p;
    f->de[2] = u.z * de_p;
}

/* ---------------- freestream: euler3d.cpp:47,156-189 -------------------- */

void orc_freestream(double mach, double deg_angle_of_attack, orc_consts* c)
{
    memset(c, 0, sizeof(*c));
    c->smoothing_coefficient = (double)0.2f;             /* euler3d.cpp:47 */

    const double aoa = (double)(PI_CONST / 180.0) * (double)deg_angle_of_attack;
    c->ff_variable[0] = 1.4;
    double ff_pressure = 1.0;
    double ff_c = sqrt(GAMMA * ff_pressure / c->ff_variable[0]);
    double ff_speed = mach * ff_c;

    v3 ff_u;
    ff_u.x = ff_speed * cos(aoa);
    ff_u.y = ff_speed * sin(aoa);
    ff_u.z = 0.0;

    c->ff_variable[1] = c->ff_variable[0] * ff_u.x;
    c->ff_variable[2] = c->ff_variable[0] * ff_u.y;
    c->ff_variable[3] = c->ff_variable[0] * ff_u.z;
    c->ff_variable[4] = c->ff_variable[0] * (0.5 * (ff_speed * ff_speed))
                      + (ff_pressure / (GAMMA - 1.0));

    v3 ff_m = { c->ff_variable[1], c->ff_variable[2], c->ff_variable[3] };
    fcontrib f;
    th_flux_contribution(c->ff_variable[4], ff_pressure, ff_m, ff_u, &f);
    memcpy(c->ff_flux_contribution_momentum_x, f.mx, sizeof f.mx);
    memcpy(c->ff_flux_contribution_momentum_y, f.my, sizeof f.my);
    memcpy(c->ff_flux_contribution_momentum_z, f.mz, sizeof f.mz);
    memcpy(c->ff_flux_contribution_density_energy, f.de, sizeof f.de);
}

/* ---------------- setup kernels: misc.h:10-84 --------------------------- */

void orc_initialize_variables(int n, const orc_consts* c, double* var)
{
    for (int i = 0; i < n; i++)
        for (int j = 0; j < ORC_NVAR; j++)
            var[5 * (size_t)i + j] = c->ff_variable[j];
}

void orc_zero_5d(int n, double* a)
{
    for (size_t i = 0; i < 5 * (size_t)n; i++) a[i] = 0.0;
}

void orc_zero_1d(int n, double* a)
{
    for (int i = 0; i < n; i++) a[i] = 0.0;
}

/* calculate_cell_volumes, misc.h:40-76; loop site euler3d.cpp:426-431 */
void orc_calculate_cell_volumes(int n_edges, const int* e2n, const double* coords,
                                double* ewt_all, double* vol)
{
    for (int e = 0; e < n_edges; e++) {
        const double* c1 = coords + 3 * (size_t)e2n[2 * (size_t)e];
        const double* c2 = coords + 3 * (size_t)e2n[2 * (size_t)e + 1];
        double* ewt = ewt_all + 3 * (size_t)e;
        double d[3];
        double dist = 0.0;
        for (int i = 0; i < 3; i++) {
            d[i] = c2[i] - c1[i];
            dist += d[i] * d[i];
        }
        dist = sqrt(dist);

        double area = 0.0;
        for (int i = 0; i < 3; i++) area += ewt[i] * ewt[i];
        area = sqrt(area);

        double tetra_volume = (1.0 / 3.0) * 0.5 * dist * area;
        vol[e2n[2 * (size_t)e]] += tetra_volume;
        vol[e2n[2 * (size_t)e + 1]] += tetra_volume;

        for (int i = 0; i < 3; i++) ewt[i] = (d[i] / dist) * area;
        for (int i = 0; i < 3; i++) ewt[i] /= dist;
    }
}

/* dampen_ewt, misc.h:78-84 */
void orc_dampen_ewt(int n, double* w)
{
    for (size_t i = 0; i < (size_t)n; i++) {
        w[3 * i + 0] *= 1e-7;
        w[3 * i + 1] *= 1e-7;
        w[3 * i + 2] *= 1e-7;
    }
}

/* copy_double_kernel, copy_double_kernel.h:6-13 */
void orc_copy_double(int n, const double* var, double* old)
{
    memcpy(old, var, sizeof(double) * 5 * (size_t)n);
}

/* ---------------- time stepping: time_stepping_kernels.h:13-86 ---------- */

/* calculate_dt_kernel, :13-32 */
void orc_calculate_dt(int n, const double* var, const double* vol, double* dt)
{
    for (size_t i = 0; i < (size_t)n; i++) {
        const double* v = var + 5 * i;
        double rho = v[0];
        v3 m = { v[1], v[2], v[3] };
        double rhoE = v[4];
        v3 u = th_velocity(rho, m);
        double ssq = th_speed_sqd(u);
        double p = th_pressure(rho, rhoE, ssq);
        double c = th_sound(rho, p);
        dt[i] = 0.5 * (cbrt(vol[i]) / (sqrt(ssq) + c));
    }
}

/* get_min_dt_kernel, :34-41, reduced from DBL_MAX (euler3d.cpp:476-479) */
double orc_get_min_dt(int n, const double* dt)
{
    double m = DBL_MAX;
    for (int i = 0; i < n; i++)
        if (dt[i] < m) m = dt[i];
    return m;
}

/* compute_step_factor_kernel, :43-64 (thermo values there are unused) */
void orc_compute_step_factor(int n, const double* var, const double* vol,
                             double min_dt, double* sf)
{
    (void)var;
    for (int i = 0; i < n; i++) sf[i] = min_dt / vol[i];
}

/* time_step_kernel, :66-86 */
void orc_time_step(int n, int rk, const double* sf, double* flux,
                   const double* old, double* var)
{
    for (size_t i = 0; i < (size_t)n; i++) {
        double factor = sf[i] / (double)(RK + 1 - rk);
        for (int v = 0; v < 5; v++) var[5 * i + v] = old[5 * i + v] + factor * flux[5 * i + v];
        for (int v = 0; v < 5; v++) flux[5 * i + v] = 0.0;
    }
}

/* ---------------- edge flux: flux.h:41-208 ------------------------------ */

typedef struct {
    double rho, rhoE, p, speed, speed_sqd, c;
    v3 m, u;
    fcontrib f;
} endpoint;

static inline void load_endpoint(const double* v, endpoint* s)
{
    s->rho = v[0];
    s->m.x = v[1];
    s->m.y = v[2];
    s->m.z = v[3];
    s->rhoE = v[4];
    s->u = th_velocity(s->rho, s->m);
    s->speed_sqd = th_speed_sqd(s->u);
    s->speed = sqrt(s->speed_sqd);
    s->p = th_pressure(s->rho, s->rhoE, s->speed_sqd);
    s->c = th_sound(s->rho, s->p);
    th_flux_contribution(s->rhoE, s->p, s->m, s->u, &s->f);
}

static inline void flux_edge_one(const double* va, const double* vb, const double* w,
                                 double* fa, double* fb, double sc)
{
    double ewt = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
    endpoint b, a;
    load_endpoint(vb, &b);                /* flux.h:52-92  */
    load_endpoint(va, &a);                /* flux.h:96-136 */

    double factor_a = -ewt * sc * 0.5 * (a.speed + sqrt(b.speed_sqd) + a.c + b.c);
    double factor_b = -ewt * sc * 0.5 * (b.speed + sqrt(a.speed_sqd) + b.c + a.c);
    double fx = -0.5 * w[0], fy = -0.5 * w[1], fz = -0.5 * w[2];

    /* flux.h:149-177, statement order 0,4,1,2,3 */
    fa[0] += factor_a * (a.rho - b.rho)
           + fx * (a.m.x + b.m.x) + fy * (a.m.y + b.m.y) + fz * (a.m.z + b.m.z);
    fa[4] += factor_a * (a.rhoE - b.rhoE)
           + fx * (a.f.de[0] + b.f.de[0]) + fy * (a.f.de[1] + b.f.de[1]) + fz * (a.f.de[2] + b.f.de[2]);
    fa[1] += factor_a * (a.m.x - b.m.x)
           + fx * (a.f.mx[0] + b.f.mx[0]) + fy * (a.f.mx[1] + b.f.mx[1]) + fz * (a.f.mx[2] + b.f.mx[2]);
    fa[2] += factor_a * (a.m.y - b.m.y)
           + fx * (a.f.my[0] + b.f.my[0]) + fy * (a.f.my[1] + b.f.my[1]) + fz * (a.f.my[2] + b.f.my[2]);
    fa[3] += factor_a * (a.m.z - b.m.z)
           + fx * (a.f.mz[0] + b.f.mz[0]) + fy * (a.f.mz[1] + b.f.mz[1]) + fz * (a.f.mz[2] + b.f.mz[2]);

    /* flux.h:179-207 */
    fb[0] += factor_b * (b.rho - a.rho)
           - fx * (a.m.x + b.m.x) - fy * (a.m.y + b.m.y) - fz * (a.m.z + b.m.z);
    fb[4] += factor_b * (b.rhoE - a.rhoE)
           - fx * (a.f.de[0] + b.f.de[0]) - fy * (a.f.de[1] + b.f.de[1]) - fz * (a.f.de[2] + b.f.de[2]);
    fb[1] += factor_b * (b.m.x - a.m.x)
           - fx * (a.f.mx[0] + b.f.mx[0]) - fy * (a.f.mx[1] + b.f.mx[1]) - fz * (a.f.mx[2] + b.f.mx[2]);
    fb[2] += factor_b * (b.m.y - a.m.y)
           - fx * (a.f.my[0] + b.f.my[0]) - fy * (a.f.my[1] + b.f.my[1]) - fz * (a.f.my[2] + b.f.my[2]);
    fb[3] += factor_b * (b.m.z - a.m.z)
           - fx * (a.f.mz[0] + b.f.mz[0]) - fy * (a.f.mz[1] + b.f.mz[1]) - fz * (a.f.mz[2] + b.f.mz[2]);
}

/* op_par_loop site euler3d.cpp:498-503 */
void orc_compute_flux_edge(int n_edges, const int* e2n, const double* var,
                           const double* ewt, double* flux, const orc_consts* c)
{
    for (size_t e = 0; e < (size_t)n_edges; e++) {
        size_t a = (size_t)e2n[2 * e], b = (size_t)e2n[2 * e + 1];
        flux_edge_one(var + 5 * a, var + 5 * b, ewt + 3 * e, flux + 5 * a, flux + 5 * b,
                      c->smoothing_coefficient);
    }
}

/* ---------------- boundary flux: flux.h:14-39 --------------------------- */

/* op_par_loop site euler3d.cpp:505-509 */
void orc_compute_bnd_node_flux(int n_bnd, const int* group, const int* b2n,
                               const double* bw, const double* var, double* flux,
                               const orc_consts* c)
{
    for (size_t i = 0; i < (size_t)n_bnd; i++) {
        int g = group[i];
        const double* w = bw + 3 * i;
        const double* v = var + 5 * (size_t)b2n[i];
        double* f = flux + 5 * (size_t)b2n[i];
        if (g <= 2) {
            /* flux_boundary.elem_func:9-54 (pressure-only surface) */
            endpoint b;
            load_endpoint(v, &b);
            f[0] += 0;
            f[1] += w[0] * b.p;
            f[2] += w[1] * b.p;
            f[3] += w[2] * b.p;
            f[4] += 0;
        } else if (g == 3 || (g >= 4 && g <= 7)) {
            /* flux_wall.elem_func:8-76 (far field) */
            endpoint b;
            load_endpoint(v, &b);
            double fx = 0.5 * w[0], fy = 0.5 * w[1], fz = 0.5 * w[2];
            const double* ffv = c->ff_variable;
            f[0] += fx * (ffv[1] + b.m.x) + fy * (ffv[2] + b.m.y) + fz * (ffv[3] + b.m.z);
            f[4] += fx * (c->ff_flux_contribution_density_energy[0] + b.f.de[0])
                  + fy * (c->ff_flux_contribution_density_energy[1] + b.f.de[1])
                  + fz * (c->ff_flux_contribution_density_energy[2] + b.f.de[2]);
            f[1] += fx * (c->ff_flux_contribution_momentum_x[0] + b.f.mx[0])
                  + fy * (c->ff_flux_contribution_momentum_x[1] + b.f.mx[1])
                  + fz * (c->ff_flux_contribution_momentum_x[2] + b.f.mx[2]);
            f[2] += fx * (c->ff_flux_contribution_momentum_y[0] + b.f.my[0])
                  + fy * (c->ff_flux_contribution_momentum_y[1] + b.f.my[1])
                  + fz * (c->ff_flux_contribution_momentum_y[2] + b.f.my[2]);
            f[3] += fx * (c->ff_flux_contribution_momentum_z[0] + b.f.mz[0])
                  + fy * (c->ff_flux_contribution_momentum_z[1] + b.f.mz[1])
                  + fz * (c->ff_flux_contribution_momentum_z[2] + b.f.mz[2]);
        }
    }
}

/* ---------------- unstructured_stream.h:7-57 ---------------------------- */

void orc_unstructured_stream(int n_edges, const int* e2n, const double* var,
                             const double* ewt, double* flux)
{
    for (size_t e = 0; e < (size_t)n_edges; e++) {
        size_t a = (size_t)e2n[2 * e], b = (size_t)e2n[2 * e + 1];
        const double* va = var + 5 * a;
        const double* vb = var + 5 * b;
        const double* w = ewt + 3 * e;
        double* fa = flux + 5 * a;
        double* fb = flux + 5 * b;
        double ex = w[0], ey = w[1], ez = w[2];
        double p_a = va[0], mxa = va[1], mya = va[2], mza = va[3], pe_a = va[4];
        double p_b = vb[0], mxb = vb[1], myb = vb[2], mzb = vb[3], pe_b = vb[4];
        fa[0] += p_b + ex;
        fa[1] += mxb + ez;
        fa[2] += myb;
        fa[3] += mzb;
        fa[4] += pe_b + ey;
        fb[0] += p_a;
        fb[1] += mxa;
        fb[2] += mya;
        fb[3] += mza;
        fb[4] += pe_a;
    }
}

/* ---------------- residual / rms / bad values: validation.h:27-115 ------- */

void orc_residual(int n, const double* old, const double* var, double* res)
{
    for (size_t i = 0; i < 5 * (size_t)n; i++) res[i] = var[i] - old[i];
}

/* calc_rms_kernel :37-44 accumulated in node order (OP2 seq gbl INC) */
double orc_calc_rms_sum(int n, const double* res)
{
    double rms = 0.0;
    for (size_t i = 0; i < (size_t)n; i++)
        for (int v = 0; v < 5; v++) rms += res[5 * i + v] * res[5 * i + v];
    return rms;
}

int orc_count_bad_vals(int n, const double* var)
{
    int count = 0;
    for (size_t i = 0; i < 5 * (size_t)n; i++)
        if (isnan(var[i]) || isinf(var[i])) count += 1;
    return count;
}

/* identify_differences :46-89 */
void orc_identify_differences(int n, const double* test, const double* master, double* diff)
{
    const double rel = 10.0e-8;
    for (size_t i = 0; i < 5 * (size_t)n; i++) {
        double acc = master[i] * rel;
        if (acc < 0.0) acc *= -1.0;
        if (acc < 3.0e-19) acc = 3.0e-19;
        double d = test[i] - master[i];
        if (d < 0.0) d *= -1.0;
        diff[i] = (d > acc) ? d : 0.0;
    }
}

/* count_non_zeros :91-100 */
int orc_count_non_zeros(int n, const double* diff)
{
    int count = 0;
    for (size_t i = 0; i < 5 * (size_t)n; i++)
        if (diff[i] > 0.0) count++;
    return count;
}

/* ---------------- multigrid transfer: mg.h:29-88 ------------------------ */

/* up_pre_kernel :29-39 over fine nodes, coarse dats written through the map */
void orc_up_pre(int n_fine, const int* f2c, double* var_c, int* scratch_c)
{
    for (int i = 0; i < n_fine; i++) {
        size_t c = (size_t)f2c[i];
        for (int v = 0; v < 5; v++) var_c[5 * c + v] = 0.0;
        scratch_c[c] = 0;
    }
}

/* up_kernel :41-52 */
void orc_up(int n_fine, const int* f2c, const double* var_f, double* var_c, int* scratch_c)
{
    for (size_t i = 0; i < (size_t)n_fine; i++) {
        size_t c = (size_t)f2c[i];
        for (int v = 0; v < 5; v++) var_c[5 * c + v] += var_f[5 * i + v];
        scratch_c[c] += 1;
    }
}

/* up_post_kernel :54-64 */
void orc_up_post(int n_coarse, double* var_c, const int* scratch_c)
{
    for (size_t i = 0; i < (size_t)n_coarse; i++) {
        double avg = scratch_c[i] == 0 ? 1.0 : 1.0 / (double)scratch_c[i];
        for (int v = 0; v < 5; v++) var_c[5 * i + v] *= avg;
    }
}

/* down_kernel :66-88 */
void orc_down(int n_fine, const int* f2c, double* var_f, const double* res_f,
              const double* xyz_f, const double* res_c, const double* xyz_c)
{
    for (size_t i = 0; i < (size_t)n_fine; i++) {
        size_t c = (size_t)f2c[i];
        const double* x = xyz_f + 3 * i;
        const double* xa = xyz_c + 3 * c;
        const double* r = res_f + 5 * i;
        const double* ra = res_c + 5 * c;
        double* v = var_f + 5 * i;
        double dx = fabs(x[0] - xa[0]);
        double dy = fabs(x[1] - xa[1]);
        double dz = fabs(x[2] - xa[2]);
        double dm = sqrt(dx * dx + dy * dy + dz * dz);
        v[0] -= dm * (ra[0] - r[0]);
        v[1] -= dx * (ra[1] - r[1]);
        v[2] -= dy * (ra[2] - r[2]);
        v[3] -= dz * (ra[3] - r[3]);
        v[4] -= dm * (ra[4] - r[4]);
    }
}

/* Inverse-distance prolongation (mg.h:90-204, sites euler3d.cpp:603-624).
 * The reference driver never declares the edge-->mg_nodes map it needs
 * (euler3d.cpp:286-292 leaves it NULL), so these loops only run where a
 * caller supplies it: e2mg[e][k] = node_to_mg_node[e2n[e][k]].
 * down_v2_kernel_pre :90-100 over fine nodes */
void orc_down_v2_pre(int n, double* wsum, double* rp)
{
    for (size_t i = 0; i < (size_t)n; i++) {
        wsum[i] = 0.0;
        for (int v = 0; v < 5; v++) rp[5 * i + v] = 0.0;
    }
}

static inline double inv_dist(const double* p, const double* q)
{
    double dx = p[0] - q[0], dy = p[1] - q[1], dz = p[2] - q[2];
    return 1.0 / sqrt(dx * dx + dy * dy + dz * dz);
}

/* down_v2_kernel :101-191 over edges in element order.  Its OP_INC
 * arguments are pointers into the dats (OP2 sequential backend), so a fine
 * node that coincides with its coarse parent has its sums overwritten
 * (:124-132, :160-167) -- by the same values on every edge.  Side a adds
 * its parent's residual and the residual of b's parent (:133-152); side b
 * adds its own parent's residual twice, the second time weighted by the
 * distance to a's parent (:178-189: residuals1b where a's would be
 * expected). */
void orc_down_v2(int n_edges, const int* e2n, const int* e2mg, const double* xyz_f,
                 const double* xyz_c, const double* res_c, double* rp, double* wsum)
{
    for (size_t e = 0; e < (size_t)n_edges; e++) {
        size_t a = (size_t)e2n[2 * e], b = (size_t)e2n[2 * e + 1];
        size_t ca = (size_t)e2mg[2 * e], cb = (size_t)e2mg[2 * e + 1];
        const double *x2a = xyz_f + 3 * a, *x2b = xyz_f + 3 * b;
        const double *x1a = xyz_c + 3 * ca, *x1b = xyz_c + 3 * cb;
        const double *r1a = res_c + 5 * ca, *r1b = res_c + 5 * cb;
        double *pa = rp + 5 * a, *pb = rp + 5 * b;
        if (x2a[0] - x1a[0] == 0.0 && x2a[1] - x1a[1] == 0.0 && x2a[2] - x1a[2] == 0.0) {
            for (int v = 0; v < 5; v++) pa[v] = r1a[v];
            wsum[a] = 1.0;
        } else {
            double d = inv_dist(x2a, x1a);
            for (int v = 0; v < 5; v++) pa[v] += d * r1a[v];
            wsum[a] += d;
            d = inv_dist(x1b, x2a);
            for (int v = 0; v < 5; v++) pa[v] += d * r1b[v];
            wsum[a] += d;
        }
        if (x2b[0] - x1b[0] == 0.0 && x2b[1] - x1b[1] == 0.0 && x2b[2] - x1b[2] == 0.0) {
            for (int v = 0; v < 5; v++) pb[v] = r1b[v];
            wsum[b] = 1.0;
        } else {
            double d = inv_dist(x2b, x1b);
            for (int v = 0; v < 5; v++) pb[v] += d * r1b[v];
            wsum[b] += d;
            d = inv_dist(x1a, x2b);
            for (int v = 0; v < 5; v++) pb[v] += d * r1b[v];
            wsum[b] += d;
        }
    }
}

/* down_v2_kernel_post :193-204 over fine nodes */
void orc_down_v2_post(int n, const double* rp, const double* wsum, const double* res_f, double* var_f)
{
    for (size_t i = 0; i < (size_t)n; i++)
        for (int v = 0; v < 5; v++) var_f[5 * i + v] += res_f[5 * i + v] - (rp[5 * i + v] / wsum[i]);
}

/* ---------------- driver: euler3d.cpp:412-641 --------------------------- */

void orc_setup(orc_level* L, int levels, const orc_consts* c)
{
    for (int l = 0; l < levels; l++) {                       /* :413-433 */
        orc_initialize_variables(L[l].n_nodes, c, L[l].variables);
        orc_zero_5d(L[l].n_nodes, L[l].fluxes);
        orc_zero_1d(L[l].n_nodes, L[l].volumes);
        orc_calculate_cell_volumes(L[l].n_edges, L[l].edge_to_node, L[l].coords,
                                   L[l].edge_weights, L[l].volumes);
    }
    for (int l = 0; l < levels; l++) {                       /* :436-441 */
        orc_dampen_ewt(L[l].n_edges, L[l].edge_weights);
        orc_dampen_ewt(L[l].n_bnd, L[l].bnd_weights);
    }
}

static void smooth(orc_level* l, const orc_consts* c, double* min_dt_out, int* fail)
{
    int n = l->n_nodes;
    orc_copy_double(n, l->variables, l->old_variables);                   /* :467 */
    orc_calculate_dt(n, l->variables, l->volumes, l->step_factors);       /* :472 */
    double min_dt = orc_get_min_dt(n, l->step_factors);                  /* :476-479 */
    *min_dt_out = min_dt;
    if (min_dt < 0.0f) { *fail = 1; return; }                            /* :480-484 */
    orc_compute_step_factor(n, l->variables, l->volumes, min_dt, l->step_factors);
    for (int rk = 0; rk < RK; rk++) {                                    /* :492-526 */
        orc_compute_flux_edge(l->n_edges, l->edge_to_node, l->variables, l->edge_weights,
                              l->fluxes, c);
        orc_compute_bnd_node_flux(l->n_bnd, l->bnd_group, l->bnd_to_node, l->bnd_weights,
                                  l->variables, l->fluxes, c);
        orc_time_step(n, rk, l->step_factors, l->fluxes, l->old_variables, l->variables);
    }
    orc_residual(n, l->old_variables, l->variables, l->residuals);        /* :528 */
}

int orc_mg_run(orc_level* L, int levels, int cycles, const orc_consts* c,
               double* rms_hist, double* min_dt_hist, int* steps_done)
{
    int level = 0, up = 1, i = 0, steps = 0, fail = 0;
    while (i < cycles) {
        double min_dt;
        smooth(&L[level], c, &min_dt, &fail);
        if (min_dt_hist) min_dt_hist[steps] = min_dt;
        steps++;
        if (fail) { if (steps_done) *steps_done = steps; return 1; }
        if (level == 0) {                                                 /* :532-550 */
            double rms = orc_calc_rms_sum(L[0].n_nodes, L[0].residuals);
            rms = sqrt(rms / (double)L[0].n_nodes);
            if (rms_hist) rms_hist[i] = rms;
            if (orc_count_bad_vals(L[0].n_nodes, L[0].variables) > 0) {
                if (steps_done) *steps_done = steps;
                return 2;
            }
        }
        if (levels <= 1) {                                                /* :573-575 */
            i++;
        } else if (up) {                                                  /* :577-598 */
            level++;
            orc_level* f = &L[level - 1];
            orc_level* cl = &L[level];
            orc_up_pre(f->n_nodes, f->node_to_mg_node, cl->variables, cl->up_scratch);
            orc_up(f->n_nodes, f->node_to_mg_node, f->variables, cl->variables, cl->up_scratch);
            orc_up_post(cl->n_nodes, cl->variables, cl->up_scratch);
            if (level == levels - 1) up = 0;
        } else {                                                          /* :599-639 */
            level--;
            orc_level* f = &L[level];
            orc_level* cl = &L[level + 1];
            if (f->edge_to_mg_nodes) {                                    /* :603-623 */
                orc_down_v2_pre(f->n_nodes, f->prolonged_wsum, f->prolonged);
                orc_down_v2(f->n_edges, f->edge_to_node, f->edge_to_mg_nodes, f->coords, cl->coords,
                            cl->residuals, f->prolonged, f->prolonged_wsum);
                orc_down_v2_post(f->n_nodes, f->prolonged, f->prolonged_wsum, f->residuals,
                                 f->variables);
            } else {                                                      /* :625-631 */
                orc_down(f->n_nodes, f->node_to_mg_node, f->variables, f->residuals, f->coords,
                         cl->residuals, cl->coords);
            }
            if (level == 0) { up = 1; i++; }
        }
    }
    if (steps_done) *steps_done = steps;
    return 0;
}
