/*
 * ORACLE — test infrastructure only.
 *
 * A CPU restatement of the reference's 3-D staggered-grid acoustic FDTD
 * (Nano560/acoustic-disguising, Julia), written from the reference source as
 * text.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it, and only as the checker or the timed CPU baseline; the product
 * library never links or calls it.
 *
 * Every function cites the reference file:line it restates.  Julia's
 * evaluation order and Float32/Float64 promotion rules are kept so that the
 * Float64 instantiation is the reference algorithm run with number type
 * Float64 and the "r32" instantiation reproduces the reference's own
 * Float32-array / Float64-scalar arithmetic.
 *
 * Pins (what checks each part) are listed in oracle/README.md; functions
 * marked "parity unpinned at the last ulp" rely on libm where Julia uses its
 * own elementary functions or TwicePrecision ranges.
 *
 * Build: see oracle/Makefile (gcc -O2 -ffp-contract=off, no fast-math).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* simulation.jl:25-44 — same field order as the Julia isbits struct. */
typedef struct {
    double tmax, xmax, ymax, zmax;
    int64_t nx, ny, nz;
    double c0, r0, cf, dx, dy, dz, fact_m0, fact_m1, dt;
    int64_t nt;
    double z0;
} ref_domain;

/* GF map for the in-loop extrapolation (extrapolation.jl:36-87). */
typedef struct {
    int32_t kind;      /* 0 none, 1 PVFromPin, 2 PVFromPV, 3 PFromPin */
    int32_t writeback; /* 1 intended write-through, 0 reference no-op (copy) */
    int64_t nt_gf, nsrc, nrec;  /* dims (2nd, 3rd) of the GF tensors */
    const double* g[4];
    const int64_t* iRec;
    const int64_t* iSrc;
    int64_t n_irec, n_isrc;
} ref_gfmap;

#define IX(i, j, k) (((i) - 1) + nx * (((j) - 1) + ny * ((k) - 1)))

/* ---------------------------------------------------------------- setup */

/* simulation.jl:46-79 (Domain constructor). */
void ref_domain_make(ref_domain* d, double tmax, double xmax, double ymax, double zmax,
                     int64_t nx, int64_t ny, int64_t nz, double c0, double r0, double cf) {
    d->tmax = tmax; d->xmax = xmax; d->ymax = ymax; d->zmax = zmax;
    d->nx = nx; d->ny = ny; d->nz = nz;
    d->c0 = c0; d->r0 = r0; d->cf = cf;
    d->dx = 2 * xmax / (double)(nx - 1);
    d->dy = 2 * ymax / (double)(ny - 1);
    d->dz = 2 * zmax / (double)(nz - 1);
    double m = d->dx < d->dy ? d->dx : d->dy;
    m = m < d->dz ? m : d->dz;
    d->dt = cf * m / c0 / sqrt(3.0);
    d->nt = (int64_t)ceil(tmax / d->dt);
    d->fact_m0 = d->dt * r0 * (c0 * c0);
    d->fact_m1 = d->dt / r0;
    d->z0 = r0 * c0;
}

/* simulation.jl:89-100: x(d) = range(-xmax, xmax, length=n) and
 * xv(d) = x(d) .- dx/2.  Evaluated as the correctly rounded value of the
 * exact expression xmax*(2i-1-n)/(n-1) [- dx/2] (binary128 intermediates are
 * exact up to the single final division, which then rounds correctly to
 * double).  Julia's TwicePrecision range agrees except possibly in the last
 * ulp: parity unpinned at the last ulp. */
double ref_coord(double xmax, int64_t n, int64_t i, double dx, int32_t staggered) {
    const int64_t m = 2 * i - 1 - n;
    if (!staggered) {
        __float128 num = (__float128)xmax * (__float128)m;
        return (double)(num / (__float128)(n - 1));
    }
    __float128 num = (__float128)2.0 * (__float128)xmax * (__float128)m -
                     (__float128)dx * (__float128)(n - 1);
    return (double)(num / ((__float128)2.0 * (__float128)(n - 1)));
}

/* simulation.jl:86: t(d) = range(0, tmax, length=nt); same treatment. */
double ref_time(double tmax, int64_t nt, int64_t i) {
    if (nt == 1) return 0.0;
    __float128 num = (__float128)tmax * (__float128)(i - 1);
    return (double)(num / (__float128)(nt - 1));
}

/* cpmlcoeffs.jl:65-126 (calc_Kab_CPML with K_max_pml = nothing). */
static void calc_kab(int64_t halo, double dt, double npower, double d0, double alpha_max,
                     int halfgrid, double* a_l, double* a_r, double* b_l, double* b_r) {
    int64_t n = halo + (halfgrid ? 1 : 0);
    double shift = halfgrid ? 0.5 : 0.0;
    if (n <= 0) return;
    double* dist = (double*)malloc(sizeof(double) * (size_t)n);
    const double start = 0 - shift;
    const double stop = ((double)n - shift) - 1;
    const int64_t lendiv = (n - 1) > 1 ? (n - 1) : 1;
    for (int64_t i = 0; i < n; ++i) {           /* LinRange lerpi */
        double t = (double)i / (double)lendiv;
        dist[i] = (1 - t) * start + t * stop;
    }
    if (halfgrid) dist[0] = 0;
    for (int64_t i = 0; i < n; ++i) {
        double ndl = dist[n - 1 - i], ndr = dist[i];
        if (halo != 0) { ndl = ndl / (double)halo; ndr = ndr / (double)halo; }
        /* npower = 2.0: Julia's x^2.0 is the correctly rounded x*x */
        double dl = d0 * (ndl * ndl), dr = d0 * (ndr * ndr);
        double al = alpha_max * (1.0 - ndl), ar = alpha_max * (1.0 - ndr);
        double bl = exp((-(dl / 1.0 + al)) * dt), br = exp((-(dr / 1.0 + ar)) * dt);
        b_l[i] = bl; b_r[i] = br;
        a_l[i] = dl * (bl - 1.0) / (1.0 * (dl + 1.0 * al));
        a_r[i] = dr * (br - 1.0) / (1.0 * (dr + 1.0 * ar));
    }
    free(dist);
}

/* cpmlcoeffs.jl:36-63 (compute_CPML_coefficients!), called from
 * simulation.jl:223 with thickness = dx*npml.  exp from libm: parity
 * unpinned at the last ulp (Julia has its own exp). */
void ref_cpml_coefficients(double vel_max, double dt, int64_t halo, double rcoef,
                           double thickness, double f0, double* a_l, double* a_r,
                           double* a_hl, double* a_hr, double* b_l, double* b_r,
                           double* b_hl, double* b_hr) {
    const double alpha_max = M_PI * f0;
    const double npower = 2.0;
    double d0 = -(npower + 1) * vel_max * log(rcoef) / (2.0 * thickness);
    if (halo == 0) d0 = 0.0;
    calc_kab(halo, dt, npower, d0, alpha_max, 0, a_l, a_r, b_l, b_r);
    calc_kab(halo, dt, npower, d0, alpha_max, 1, a_hl, a_hr, b_hl, b_hr);
}

/* findlast(xk .< xp) over a coordinate vector (simulation.jl:266,318-319);
 * returns 0 for `nothing`. */
static int64_t findlast_lt(const double* xk, int64_t n, double xp) {
    int64_t r = 0;
    for (int64_t i = 1; i <= n; ++i)
        if (xk[i - 1] < xp) r = i;
    return r;
}

/* simulation.jl:254-341 (trilinearInterpolation + bilinearInterpolation).
 * idx: 4 channels x 3 (1-based), w: 4 channels x 8 (column-major 2x2x2).
 * Returns -1 where the reference would throw (a point below every node). */
int ref_trilinear(const ref_domain* d, const double* pt, int64_t* idx, double* w) {
    const int64_t nx = d->nx, ny = d->ny, nz = d->nz;
    double* xs = (double*)malloc(sizeof(double) * (size_t)(2 * (nx + ny + nz)));
    double *x = xs, *xv = xs + nx, *y = xv + nx, *yv = y + ny, *z = yv + ny, *zv = z + nz;
    for (int64_t i = 1; i <= nx; ++i) { x[i - 1] = ref_coord(d->xmax, nx, i, d->dx, 0); xv[i - 1] = ref_coord(d->xmax, nx, i, d->dx, 1); }
    for (int64_t i = 1; i <= ny; ++i) { y[i - 1] = ref_coord(d->ymax, ny, i, d->dy, 0); yv[i - 1] = ref_coord(d->ymax, ny, i, d->dy, 1); }
    for (int64_t i = 1; i <= nz; ++i) { z[i - 1] = ref_coord(d->zmax, nz, i, d->dz, 0); zv[i - 1] = ref_coord(d->zmax, nz, i, d->dz, 1); }
    const double xp = pt[0], yp = pt[1], zp = pt[2];
    const double ds = d->dx * d->dy;
    const double* cx[4] = {x, xv, x, x};
    const double* cy[4] = {y, y, yv, y};
    const double* cz[4] = {z, z, z, zv};
    int rc = 0;
    for (int ch = 0; ch < 4; ++ch) {
        /* vz reuses p's x-y stencil (simulation.jl:281) */
        const int xy = (ch == 3) ? 0 : ch;
        int64_t ix = findlast_lt(cx[xy], nx, xp);
        int64_t iy = findlast_lt(cy[xy], ny, yp);
        int64_t iz = findlast_lt(cz[ch], nz, zp);
        if (ix == 0 || iy == 0 || iz == 0) { rc = -1; }
        double wxy[4] = {0, 0, 0, 0};
        if (ix >= 1 && iy >= 1 && ix < nx && iy < ny) {
            double x1 = cx[xy][ix - 1], x2 = cx[xy][ix];
            double y1 = cy[xy][iy - 1], y2 = cy[xy][iy];
            double dx0 = xp - x1, dx1 = x2 - xp, dy0 = yp - y1, dy1 = y2 - yp;
            wxy[0] = dx1 * dy1 / ds;
            wxy[1] = dx0 * dy1 / ds;
            wxy[2] = dx1 * dy0 / ds;
            wxy[3] = dx0 * dy0 / ds;
        }
        double* wc = w + 8 * ch;
        for (int q = 0; q < 8; ++q) wc[q] = 0.0;
        if (iz >= 1 && iz < nz) {
            double z1 = cz[ch][iz - 1], z2 = cz[ch][iz];
            for (int q = 0; q < 4; ++q) {
                wc[q] = wxy[q] * (z2 - zp) / d->dz;
                wc[q + 4] = wxy[q] * (zp - z1) / d->dz;
            }
        }
        idx[3 * ch + 0] = ix; idx[3 * ch + 1] = iy; idx[3 * ch + 2] = iz;
    }
    free(xs);
    return rc;
}

/* geometry.jl:86-143 (_outside_predicate).  shape: 0 none, 1 sphere(r),
 * 2 cube(edge), 3 cross(length), 4 ellipsoid(a,b,c; not in the reference:
 * (x/a)^2+(y/b)^2+(z/c)^2 > 1 evaluated as x*x/(a*a)+...). */
static int outside(int shape, const double* prm, double x, double y, double z) {
    if (shape == 0) return 1;
    if (shape == 1) { double r = prm[0]; return x * x + y * y + z * z > r * r; }
    if (shape == 2) { double r = prm[0]; return fabs(x) > r / 2 || fabs(y) > r / 2 || fabs(z) > r / 2; }
    if (shape == 3) {
        double l = prm[0], s = l / 2, sl = 2.0 * l / 3;
#define INC(X, Y, Z) (fabs(X) < s && fabs(Y) < s && fabs(Z) < s)
        if (INC(x, y, z)) {
            if (INC(x + sl, y + sl, z) || INC(x - sl, y - sl, z) || INC(x + sl, y - sl, z) || INC(x - sl, y + sl, z)) return 1;
            if (INC(x + sl, y, z + sl) || INC(x - sl, y, z - sl) || INC(x + sl, y, z - sl) || INC(x - sl, y, z + sl)) return 1;
            if (INC(x, y + sl, z + sl) || INC(x, y - sl, z - sl) || INC(x, y + sl, z - sl) || INC(x, y - sl, z + sl)) return 1;
            return 0;
        }
        return 1;
#undef INC
    }
    if (shape == 4) {
        double a = prm[0], b = prm[1], c = prm[2];
        return x * x / (a * a) + y * y / (b * b) + z * z / (c * c) > 1.0;
    }
    return 1;
}

/* geometry.jl:159-176 (build_scatterer) and :250-280 (build_update_mask).
 * mask4: (nx,ny,nz,4) uint8, channel 1 fluid, 2..4 open faces. */
int ref_build_update_mask(const ref_domain* d, int shape, const double* prm, uint8_t* mask4) {
    const int64_t nx = d->nx, ny = d->ny, nz = d->nz, n3 = nx * ny * nz;
    uint8_t* s = (uint8_t*)malloc((size_t)n3);
    double* xs = (double*)malloc(sizeof(double) * (size_t)(nx + ny + nz));
    for (int64_t i = 1; i <= nx; ++i) xs[i - 1] = ref_coord(d->xmax, nx, i, d->dx, 0);
    for (int64_t i = 1; i <= ny; ++i) xs[nx + i - 1] = ref_coord(d->ymax, ny, i, d->dy, 0);
    for (int64_t i = 1; i <= nz; ++i) xs[nx + ny + i - 1] = ref_coord(d->zmax, nz, i, d->dz, 0);
    for (int64_t k = 1; k <= nz; ++k)
        for (int64_t j = 1; j <= ny; ++j)
            for (int64_t i = 1; i <= nx; ++i)
                s[IX(i, j, k)] = shape == 0 ? 0 : !outside(shape, prm, xs[i - 1], xs[nx + j - 1], xs[nx + ny + k - 1]);
    memset(mask4, 0, (size_t)(4 * n3));
    for (int64_t k = 1; k <= nz; ++k)
        for (int64_t j = 1; j <= ny; ++j)
            for (int64_t i = 1; i <= nx; ++i) {
                const int64_t c = IX(i, j, k);
                mask4[c] = !s[c];
                if (i >= 2) mask4[n3 + c] = !s[c - 1] && !s[c];
                if (j >= 2) mask4[2 * n3 + c] = !s[c - nx] && !s[c];
                if (k >= 2) mask4[3 * n3 + c] = !s[c - nx * ny] && !s[c];
            }
    free(s); free(xs);
    return 0;
}

/* illumination.jl:19-46 (fibonacci_sphere).  Output (n,3) column-major.
 * cos/sin from libm: parity unpinned at the last ulp. */
void ref_fibonacci_sphere(int64_t n, double radius, double* points, double* normals) {
    const double golden = M_PI * (3.0 - sqrt(5.0));
    for (int64_t i = 1; i <= n; ++i) {
        double z = 1.0 - (2.0 * (double)(i - 1) + 1.0) / (double)n;
        double q = 1.0 - z * z;
        double r_xy = sqrt(q > 0.0 ? q : 0.0);
        double th = golden * (double)(i - 1);
        double nx = r_xy * cos(th), ny = r_xy * sin(th), nz = z;
        normals[i - 1] = nx; normals[n + i - 1] = ny; normals[2 * n + i - 1] = nz;
        points[i - 1] = radius * nx; points[n + i - 1] = radius * ny; points[2 * n + i - 1] = radius * nz;
    }
}

/* illumination.jl:167-184 (impulsive_wavelet).  Returns the timeshift;
 * sum is sequential (Julia's SIMD pairwise sum may differ in the last ulp). */
double ref_impulsive_wavelet(const ref_domain* d, double f_3db_hz, double* tf) {
    const double sig = sqrt(log(2.0)) / (2 * M_PI * f_3db_hz);
    double x = 10 * sig / d->dt;
    int64_t pl = (int64_t)nearbyint(x);  /* round(Int, x): ties to even */
    if (pl > d->nt) pl = d->nt;
    const double center = (double)pl / 2, width = (double)pl / 10;
    for (int64_t i = 0; i < d->nt; ++i) tf[i] = 0.0;
    for (int64_t k = 1; k <= pl; ++k) {
        double u = (double)k - center;
        tf[k - 1] = exp(-(u * u) / (2 * (width * width)));
    }
    double s = 0.0;
    for (int64_t i = 0; i < d->nt; ++i) s += tf[i];
    const double den = s * d->dt;
    for (int64_t i = 0; i < d->nt; ++i) tf[i] = tf[i] / den;
    return center * d->dt;
}

/* illumination.jl:199-207 (ricker_wavelet) on t(d) (simulation.jl:86). */
void ref_ricker_wavelet(const ref_domain* d, double fc, int32_t injection, double* tf) {
    for (int64_t i = 1; i <= d->nt; ++i) {
        double t = ref_time(d->tmax, d->nt, i);
        double wn = M_PI * fc * (t - 1.0 / fc);
        double w2 = wn * wn;
        tf[i - 1] = (1.0 - 2.0 * w2) * exp(-w2);
    }
    if (injection) {
        double acc = 0.0;
        const double sc = d->dt / d->dx;
        for (int64_t i = 0; i < d->nt; ++i) { acc += tf[i]; tf[i] = acc * sc; }
    }
}

/* simulation.jl:468-518 (initialField, :planeWave branch), in double;
 * the caller rounds to its storage type.  field (nx,ny,nz,4). */
void ref_initial_field(const ref_domain* d, double fc, double distance, double* field) {
    const int64_t nx = d->nx, ny = d->ny, nz = d->nz, n3 = nx * ny * nz;
    const double k = fc / d->c0;
    memset(field, 0, sizeof(double) * (size_t)(4 * n3));
    for (int64_t i = 1; i <= nx; ++i) {
        double cx = ref_coord(d->xmax, nx, i, d->dx, 0);
        double cxv = ref_coord(d->xmax, nx, i, d->dx, 1);
        double w1 = M_PI * k * (cx + distance);
        double w2 = M_PI * k * (cxv + distance - d->c0 * d->dt / 2);
        double p = (1 - 2 * (w1 * w1)) * exp(-(w1 * w1));
        double vx = (1 - 2 * (w2 * w2)) * exp(-(w2 * w2)) / d->z0;
        for (int64_t kk = 1; kk <= nz; ++kk)
            for (int64_t j = 1; j <= ny; ++j) {
                field[IX(i, j, kk)] = p;
                field[n3 + IX(i, j, kk)] = vx;
            }
    }
}

/* diagnostics/_check_inject_kh/injection_coefficients.jl:467-472 (C_for_N). */
void ref_c_for_n(int64_t N, double r_inner, double dt, double r0, double dx, double z0,
                 double* out3) {
    const double dS = 4 * M_PI * (r_inner * r_inner) / (double)N;
    const double C_f = dS * dt / (r0 * (dx * dx * dx));
    out3[0] = C_f;
    out3[1] = C_f * (z0 * z0);
    out3[2] = dS;
}

/* CPML boundary slab lookup (derivatives.jl:66-79, getBoundaryIndex). */
static inline void boundary_index(int64_t i, int64_t ndim, int64_t npml, int64_t* l, int* side) {
    if (i <= npml) { *side = 1; *l = i; return; }
    *l = i - (ndim - npml);
    *side = *l > 0 ? 2 : 0;
}

/* ------------------------------------------------------ analytic KH kernels */

/* kh_kernels.jl:19-28 (A&S 7.1.26 erf, Gaussian f, f', F). */
static double kh_erf(double x) {
    double s = (x > 0) - (x < 0), ax = fabs(x);
    double t = 1.0 / (1.0 + 0.3275911 * ax);
    double y = 1.0 - (((((1.061405429 * t - 1.453152027) * t) + 1.421413741) * t -
                       0.284496736) * t + 0.254829592) * t * exp(-ax * ax);
    return s * y;
}
static double kh_f(double tau, double sg) {
    return (1 / (sqrt(2 * M_PI) * sg)) * exp(-(tau * tau) / (2 * (sg * sg)));
}
static double kh_df(double tau, double sg) { return -(tau / (sg * sg)) * kh_f(tau, sg); }
static double kh_F(double tau, double sg) { return 0.5 * (1 + kh_erf(tau / (sg * sqrt(2.0)))); }

double ref_kh_erf(double x) { return kh_erf(x); }

/* kh_kernels.jl:106-157 (eval_one_kernel), output K (nt, n_outer, n_inner)
 * in double.  which: 0 p_p, 1 p_v, 2 v_p, 3 v_v. */
void ref_eval_one_kernel(int which, int64_t n_outer, const double* outer_pts,
                         const double* outer_nrm, int64_t n_inner, const double* inner_pts,
                         const double* inner_nrm, int64_t nt, const double* t_grid, double sg,
                         double c0, double rho0, double* K) {
    for (int64_t ir = 0; ir < n_inner; ++ir)
        for (int64_t is = 0; is < n_outer; ++is) {
            double dx = inner_pts[ir] - outer_pts[is];
            double dy = inner_pts[n_inner + ir] - outer_pts[n_outer + is];
            double dz = inner_pts[2 * n_inner + ir] - outer_pts[2 * n_outer + is];
            double r = sqrt(dx * dx + dy * dy + dz * dz);
            double ihx = dx / r, ihy = dy / r, ihz = dz / r;
            double cr = inner_nrm[ir] * ihx + inner_nrm[n_inner + ir] * ihy + inner_nrm[2 * n_inner + ir] * ihz;
            double cs = outer_nrm[is] * ihx + outer_nrm[n_outer + is] * ihy + outer_nrm[2 * n_outer + is] * ihz;
            double retard = r / c0;
            double i1 = 1 / (4 * M_PI * r), i2 = 1 / (4 * M_PI * (r * r)), i3 = 1 / (4 * M_PI * (r * r * r));
            for (int64_t it = 0; it < nt; ++it) {
                double tau = t_grid[it] - retard;
                double v;
                if (which == 0) v = kh_df(tau, sg) * i1;
                else if (which == 1) v = cs * (kh_f(tau, sg) * i2 + kh_df(tau, sg) * i1 / c0);
                else if (which == 2) v = (cr / rho0) * (kh_f(tau, sg) * i2 + kh_df(tau, sg) * i1 / c0);
                else v = (cr * cs / rho0) * (2 * kh_F(tau, sg) * i3 + 2 * kh_f(tau, sg) * i2 / c0 +
                                             kh_df(tau, sg) * i1 / (c0 * c0));
                K[it + nt * (is + n_outer * ir)] = v;
            }
        }
}

/* kh_kernels.jl:178-203 (eval_kernel_pedestal), (n_outer, n_inner). */
void ref_eval_kernel_pedestal(int which, int64_t n_outer, const double* outer_pts,
                              const double* outer_nrm, int64_t n_inner, const double* inner_pts,
                              const double* inner_nrm, double rho0, double* Kdc) {
    for (int64_t ir = 0; ir < n_inner; ++ir)
        for (int64_t is = 0; is < n_outer; ++is) {
            if (which != 3) { Kdc[is + n_outer * ir] = 0.0; continue; }
            double dx = inner_pts[ir] - outer_pts[is];
            double dy = inner_pts[n_inner + ir] - outer_pts[n_outer + is];
            double dz = inner_pts[2 * n_inner + ir] - outer_pts[2 * n_outer + is];
            double r = sqrt(dx * dx + dy * dy + dz * dz);
            double ihx = dx / r, ihy = dy / r, ihz = dz / r;
            double cr = inner_nrm[ir] * ihx + inner_nrm[n_inner + ir] * ihy + inner_nrm[2 * n_inner + ir] * ihz;
            double cs = outer_nrm[is] * ihx + outer_nrm[n_outer + is] * ihy + outer_nrm[2 * n_outer + is] * ihz;
            double i3 = 1 / (4 * M_PI * (r * r * r));
            Kdc[is + n_outer * ir] = (cr * cs / rho0) * 2 * i3;
        }
}

/* kh_kernels.jl:256-296 (direct_kh_accumulate!).  K (ntK, n_outer, n_inner),
 * in (nt, n_outer), out (nt, n_inner).  Kdc / cumsum_in may be NULL. */
int ref_direct_kh_accumulate(double* out, int64_t nt, int64_t n_inner, const double* K,
                             int64_t ntK, int64_t n_outer, const double* in, double alpha,
                             const double* Kdc, const double* cumsum_in) {
    if (ntK > nt) return -1;
    for (int64_t s = 0; s < n_inner; ++s)
        for (int64_t r = 0; r < n_outer; ++r)
            for (int64_t t = 1; t <= nt; ++t) {
                double acc = 0.0;
                int64_t m = t < ntK ? t : ntK;
                for (int64_t tt = 1; tt <= m; ++tt)
                    acc += K[(tt - 1) + ntK * (r + n_outer * s)] * in[(t - tt) + nt * r];
                out[(t - 1) + nt * s] += alpha * acc;
            }
    if (Kdc && cumsum_in)
        for (int64_t s = 0; s < n_inner; ++s)
            for (int64_t r = 0; r < n_outer; ++r) {
                double k = Kdc[r + n_outer * s];
                if (k == 0) continue;
                double ak = alpha * k;
                for (int64_t t = 0; t < nt; ++t) out[t + nt * s] += ak * cumsum_in[t + nt * r];
            }
    return 0;
}

/* ------------------------------------------------------------- time loop */

#define T double
#define SUF _f64
#include "fdtd_ref_body.h"
#undef T
#undef SUF

#define T float
#define SUF _r32
#include "fdtd_ref_body.h"
#undef T
#undef SUF
