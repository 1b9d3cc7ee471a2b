/*
 * ORACLE — test infrastructure only (included twice by fdtd_ref.c).
 *
 * One FDTD time step of the reference, restated in plain C, templated on the
 * storage type T.  Arithmetic follows Julia's promotion rules for the
 * reference's Float32 arrays and Float64 domain scalars:
 *   - array-array ops (neighbour differences, b*psi, transceiver ops) in T,
 *   - anything touching a Float64 scalar (dx, fq, fv, a*dv) in double,
 *   - stores round to T.
 * With T = double this is exactly the Float64 restatement of the algorithm.
 *
 * Index conventions are the reference's: 1-based, column-major, field
 * (nx,ny,nz,4) = P,Vx,Vy,Vz.  Never linked into the product.
 */

#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)
#define FN(name) CAT(name, SUF)

/* forward.jl:67-68 + updates.jl:105-142 (update_p!) and :288-318
 * (update_p_mask!): rigid-wall derivative at i==nx / j==ny / k==nz. */
static inline void FN(p_plain)(const ref_domain* d, T* P, const T* Vx, const T* Vy,
                               const T* Vz, int64_t i, int64_t j, int64_t k,
                               const uint8_t* m1) {
    const int64_t nx = d->nx, ny = d->ny, nz = d->nz;
    const int64_t c = IX(i, j, k);
    double dvx, dvy, dvz;
    if (i == nx) dvx = (double)((T)(-2) * Vx[c]) / d->dx;
    else         dvx = (double)(T)(Vx[c + 1] - Vx[c]) / d->dx;
    if (j == ny) dvy = (double)((T)(-2) * Vy[c]) / d->dy;
    else         dvy = (double)(T)(Vy[c + nx] - Vy[c]) / d->dy;
    if (k == nz) dvz = (double)((T)(-2) * Vz[c]) / d->dz;
    else         dvz = (double)(T)(Vz[c + nx * ny] - Vz[c]) / d->dz;
    double coeff = d->fact_m0;
    if (m1) coeff = m1[c] ? d->fact_m0 : 0.0;
    P[c] = (T)((double)P[c] - coeff * ((dvx + dvy) + dvz));
}

/* One CPML memory update: psi = b*psi + a*dv; dv += psi
 * (updates.jl:58-59 and the analogous lines of every update kernel). */
static inline double FN(psi_step)(T* m, T bcoef, T acoef, double dv) {
    T bm = bcoef * (*m);                       /* Float32 * Float32 -> Float32 */
    *m = (T)((double)bm + (double)acoef * dv); /* + Float64, store rounds to T */
    return dv + (double)(*m);
}

/* updates.jl:21-103 (update_p_cpml!, guard npml>1) and :320-378
 * (update_p_cpml_mask!, no guard). */
static inline void FN(p_cpml)(const ref_domain* d, T* P, const T* Vx, const T* Vy,
                              const T* Vz, int64_t i, int64_t j, int64_t k, int64_t p,
                              T* mvx, T* mvy, T* mvz, const T* a, const T* b,
                              const uint8_t* m1) {
    const int64_t nx = d->nx, ny = d->ny, nz = d->nz;
    const int64_t c = IX(i, j, k);
    double dvx = (double)(T)(Vx[c + 1] - Vx[c]) / d->dx;
    double dvy = (double)(T)(Vy[c + nx] - Vy[c]) / d->dy;
    double dvz = (double)(T)(Vz[c + nx * ny] - Vz[c]) / d->dz;
    const int guard = m1 ? 1 : (p > 1);
    int64_t l; int side;
    if (guard) {
        boundary_index(i, nx, p, &l, &side);
        if (side > 0) {
            T* m = &mvx[(l - 1) + p * ((j - 1) + ny * ((k - 1) + nz * (side - 1)))];
            dvx = FN(psi_step)(m, b[(l - 1) + p * (side - 1)], a[(l - 1) + p * (side - 1)], dvx);
        }
        boundary_index(j, ny, p, &l, &side);
        if (side > 0) {
            T* m = &mvy[(i - 1) + nx * ((l - 1) + p * ((k - 1) + nz * (side - 1)))];
            dvy = FN(psi_step)(m, b[(l - 1) + p * (side - 1)], a[(l - 1) + p * (side - 1)], dvy);
        }
        boundary_index(k, nz, p, &l, &side);
        if (side > 0) {
            T* m = &mvz[(i - 1) + nx * ((j - 1) + ny * ((l - 1) + p * (side - 1)))];
            dvz = FN(psi_step)(m, b[(l - 1) + p * (side - 1)], a[(l - 1) + p * (side - 1)], dvz);
        }
    }
    double coeff = d->fact_m0;
    if (m1) coeff = m1[c] ? d->fact_m0 : 0.0;
    P[c] = (T)((double)P[c] - coeff * ((dvx + dvy) + dvz));
}

/* updates.jl:192-275 (update_v{x,y,z}!) and :380-501 (update_v*_mask!).
 * axis 0/1/2 = x/y/z; psi has extent p+1 along the axis; guard p+1 > 1. */
static inline void FN(v_one)(const ref_domain* d, const T* P, T* V, int axis, int64_t i,
                             int64_t j, int64_t k, int64_t p, T* mp, const T* ah,
                             const T* bh, const uint8_t* mc) {
    const int64_t nx = d->nx, ny = d->ny, nz = d->nz;
    const int64_t c = IX(i, j, k);
    const int64_t q = p + 1;
    double dp;
    int64_t idx, n, stride;
    double h;
    if (axis == 0)      { idx = i; n = nx; stride = 1;       h = d->dx; }
    else if (axis == 1) { idx = j; n = ny; stride = nx;      h = d->dy; }
    else                { idx = k; n = nz; stride = nx * ny; h = d->dz; }
    dp = (double)(T)(P[c] - P[c - stride]) / h;
    if (q > 1) {
        int64_t l; int side;
        boundary_index(idx, n, q, &l, &side);
        if (side > 0) {
            T* m;
            if (axis == 0)      m = &mp[(l - 1) + q * ((j - 1) + ny * ((k - 1) + nz * (side - 1)))];
            else if (axis == 1) m = &mp[(i - 1) + nx * ((l - 1) + q * ((k - 1) + nz * (side - 1)))];
            else                m = &mp[(i - 1) + nx * ((j - 1) + ny * ((l - 1) + q * (side - 1)))];
            dp = FN(psi_step)(m, bh[(l - 1) + q * (side - 1)], ah[(l - 1) + q * (side - 1)], dp);
        }
    }
    double coeff = d->fact_m1;
    if (mc) coeff = mc[c] ? d->fact_m1 : 0.0;
    V[c] = (T)((double)V[c] - coeff * dp);
}

/* derivatives.jl:30-60 (interp_trilinear!), one transceiver, sequential
 * corner order (1,1,1),(2,1,1),(1,2,1),... ; record before inject. */
static inline void FN(interp_one)(const ref_domain* d, T* F, const int64_t* idx3,
                                  const T* w8, T* rec, const T* src) {
    const int64_t nx = d->nx, ny = d->ny, nz = d->nz;
    const int64_t I = idx3[0], J = idx3[1], K = idx3[2];
    if (!(I >= 1 && I < nx && J >= 1 && J < ny && K >= 1 && K < nz)) return;
    const T s = *src;
    for (int cc = 0; cc < 2; ++cc)
        for (int bb = 0; bb < 2; ++bb)
            for (int aa = 0; aa < 2; ++aa) {
                const T w = w8[aa + 2 * bb + 4 * cc];
                T* f = &F[IX(I + aa, J + bb, K + cc)];
                *rec = *rec + w * (*f);
                *f = *f + w * s;
            }
}

/* extrapolation.jl:101-249 (extrapolate! and its three kernels), evaluated
 * for one step it < nt.  The GF tensors are double; accumulation in double.
 * gf->writeback = 1 writes the sources at it+1 (the documented intent,
 * docs/impulsive_gf_hologram_path.md:154-171); 0 reproduces the reference's
 * copy-then-discard behaviour (extrapolation.jl:215-216,226-232,246-247),
 * i.e. nothing is written. */
static int FN(ref_extrapolate_step)(const ref_domain* d, const ref_gfmap* gf, int64_t ntx,
                                    const T* rec, const T* vn_rec, T* src, T* vn_src,
                                    int64_t it) {
    const int64_t nt = d->nt;
    const int64_t ng = gf->nt_gf;
    const int64_t Tn = it < ng ? it : ng;     /* T = min(it, nt_gf) */
    const int64_t toff = it - Tn + 1;         /* field index for tt = 1 */
    const T* p_rec = rec;                     /* channel 0 */
    for (int64_t r = 0; r < gf->n_irec; ++r)
        if (gf->iRec[r] < 1 || gf->iRec[r] > ntx) return -3;
    for (int64_t s = 0; s < gf->n_isrc; ++s)
        if (gf->iSrc[s] < 1 || gf->iSrc[s] > ntx) return -3;
    if (gf->kind == 1 || gf->kind == 3) {
        /* pin = (p_rec - vn_rec*z0)/2 over the used history (extrapolation.jl:224,241) */
        const int64_t nr = gf->n_irec;
        double* pin = (double*)malloc(sizeof(double) * (size_t)(Tn * (nr ? nr : 1)));
        if (!pin) return -2;
        for (int64_t r = 0; r < nr; ++r) {
            const int64_t col = gf->iRec[r] - 1;
            for (int64_t tt = 1; tt <= Tn; ++tt) {
                const int64_t t = toff + tt - 1;
                pin[(tt - 1) + Tn * r] =
                    ((double)p_rec[(t - 1) + nt * col] - (double)vn_rec[(t - 1) + nt * col] * d->z0) / 2.0;
            }
        }
        if (gf->kind == 1) {
            /* extrapolate_pv_from_pin! (:139-170); pp,vp are (nt_gf, nsrc, nrec) */
            const int64_t ns = gf->nsrc;
            if (ns != gf->n_isrc || gf->nrec != nr) { free(pin); return -3; }
            for (int64_t s = 0; s < ns; ++s) {
                double psum = 0.0, vsum = 0.0;
                for (int64_t r = 0; r < nr; ++r)
                    for (int64_t tt = 1; tt <= Tn; ++tt) {
                        const int64_t gi = (Tn - tt) + ng * (s + ns * r);
                        const double pv = pin[(tt - 1) + Tn * r];
                        psum += gf->g[0][gi] * pv;
                        vsum += gf->g[1][gi] * pv;
                    }
                if (gf->writeback) {
                    const int64_t col = gf->iSrc[s] - 1;
                    src[it + nt * col] = (T)vsum;     /* p_src[it+1,s] <- vp-derived */
                    vn_src[it + nt * col] = (T)psum;  /* v_src[it+1,s] <- pp-derived */
                }
            }
        } else {
            /* extrapolate_p_from_pin (:172-192); pos,neg are (nt_gf, nrec, nsrc), so
             * their second dimension (held in gf->nsrc, named for the PVFromPin
             * layout) is #iRec and the third (gf->nrec) is #iSrc */
            const int64_t ns = gf->n_isrc;
            if (gf->nsrc != nr || gf->nrec != ns) { free(pin); return -3; }
            for (int64_t s = 0; s < ns; ++s) {
                double pos = 0.0, neg = 0.0;
                for (int64_t r = 0; r < nr; ++r)
                    for (int64_t tt = 1; tt <= Tn; ++tt) {
                        const int64_t gi = (Tn - tt) + ng * (r + nr * s);
                        const double pv = pin[(tt - 1) + Tn * r];
                        pos += gf->g[0][gi] * pv;
                        neg += gf->g[1][gi] * pv;
                    }
                if (gf->writeback) {
                    const int64_t col = gf->iSrc[s] - 1;
                    src[it + nt * col] = (T)(pos - neg);
                    vn_src[it + nt * col] = (T)((pos + neg) / d->z0);
                }
            }
        }
        free(pin);
        return 0;
    }
    if (gf->kind == 2) {
        /* extrapolate_pv_from_pv (:101-137); pp,pv,vp,vv are (nt_gf, nm, nrec) with
         * nm = #iRec (contraction) and nrec = #iSrc (outputs) */
        const int64_t nm = gf->n_irec, no = gf->n_isrc;
        if (gf->nsrc != nm || gf->nrec != no) return -3;
        const double *pp = gf->g[0], *pv = gf->g[1], *vp = gf->g[2], *vv = gf->g[3];
        for (int64_t r = 0; r < no; ++r) {
            double psum = 0.0, vsum = 0.0;
            for (int64_t m = 0; m < nm; ++m) {
                const int64_t col = gf->iRec[m] - 1;
                for (int64_t tt = 1; tt <= Tn; ++tt) {
                    const int64_t gi = (Tn - tt) + ng * (m + nm * r);
                    const int64_t t = toff + tt - 1;
                    const double p_t = (double)p_rec[(t - 1) + nt * col];
                    const double vn_t = (double)vn_rec[(t - 1) + nt * col];
                    psum += pv[gi] * p_t + pp[gi] * vn_t;
                    vsum += vv[gi] * p_t + vp[gi] * vn_t;
                }
            }
            if (gf->writeback) {
                const int64_t col = gf->iSrc[r] - 1;
                src[it + nt * col] = (T)vsum;
                vn_src[it + nt * col] = (T)psum;
            }
        }
        return 0;
    }
    return -4;
}

/* forward.jl:13-139 (forward_onestep!). Returns 0 on success. */
int FN(ref_forward_onestep)(const ref_domain* d, T* field, int64_t p, T* mv_x, T* mv_y,
                            T* mv_z, T* mp_x, T* mp_y, T* mp_z, const T* a, const T* b,
                            const T* ah, const T* bh, const uint8_t* mask, int64_t ntx,
                            const int64_t* idx, const T* w, T* src, T* rec, const T* dir,
                            T* vn_src, T* vn_rec, int64_t it, const ref_gfmap* gf) {
    const int64_t nx = d->nx, ny = d->ny, nz = d->nz, nt = d->nt;
    const int64_t n3 = nx * ny * nz;
    T* P = field;
    T* Vx = field + n3;
    T* Vy = field + 2 * n3;
    T* Vz = field + 3 * n3;
    const uint8_t* m1 = mask ? mask : NULL;
    const uint8_t* m2 = mask ? mask + n3 : NULL;
    const uint8_t* m3 = mask ? mask + 2 * n3 : NULL;
    const uint8_t* m4 = mask ? mask + 3 * n3 : NULL;
    if (it < 1 || it > nt) return -1;

    /* rigid planes (forward.jl:66-70): no CPML, no mask */
    for (int64_t k = 1; k <= nz - 1; ++k)
        for (int64_t i = 1; i <= nx - 1; ++i) FN(p_plain)(d, P, Vx, Vy, Vz, i, ny, k, NULL);
    for (int64_t j = 1; j <= ny - 1; ++j)
        for (int64_t i = 1; i <= nx - 1; ++i) FN(p_plain)(d, P, Vx, Vy, Vz, i, j, nz, NULL);

    /* interior pressure (forward.jl:76-89) */
    #pragma omp parallel for schedule(static)
    for (int64_t k = 1; k <= nz - 1; ++k)
        for (int64_t j = 1; j <= ny - 1; ++j)
            for (int64_t i = 1; i <= nx - 1; ++i) {
                if (p == 0) FN(p_plain)(d, P, Vx, Vy, Vz, i, j, k, m1);
                else FN(p_cpml)(d, P, Vx, Vy, Vz, i, j, k, p, mv_x, mv_y, mv_z, a, b, m1);
            }

    /* pressure transceivers (forward.jl:92), channel 0 */
    for (int64_t is = 0; is < ntx; ++is)
        FN(interp_one)(d, P, idx + 3 * is, w + 8 * is, &rec[(it - 1) + nt * is],
                       &src[(it - 1) + nt * is]);

    /* velocities (forward.jl:99-107) */
    #pragma omp parallel for schedule(static)
    for (int64_t k = 1; k <= nz; ++k)
        for (int64_t j = 1; j <= ny; ++j)
            for (int64_t i = 2; i <= nx; ++i)
                FN(v_one)(d, P, Vx, 0, i, j, k, p, mp_x, ah, bh, m2);
    #pragma omp parallel for schedule(static)
    for (int64_t k = 1; k <= nz; ++k)
        for (int64_t j = 2; j <= ny; ++j)
            for (int64_t i = 1; i <= nx; ++i)
                FN(v_one)(d, P, Vy, 1, i, j, k, p, mp_y, ah, bh, m3);
    #pragma omp parallel for schedule(static)
    for (int64_t k = 2; k <= nz; ++k)
        for (int64_t j = 1; j <= ny; ++j)
            for (int64_t i = 1; i <= nx; ++i)
                FN(v_one)(d, P, Vz, 2, i, j, k, p, mp_z, ah, bh, m4);

    /* velocity transceivers (forward.jl:111-115), channels 1..3 in order */
    if (ntx > 0) {
        T* Vc[3] = {Vx, Vy, Vz};
        for (int ch = 1; ch <= 3; ++ch)
            for (int64_t is = 0; is < ntx; ++is)
                FN(interp_one)(d, Vc[ch - 1], idx + 3 * (is + ntx * ch), w + 8 * (is + ntx * ch),
                               &rec[(it - 1) + nt * (is + ntx * ch)],
                               &src[(it - 1) + nt * (is + ntx * ch)]);
    }

    /* v2vn! (forward.jl:120; derivatives.jl:85-90) */
    for (int64_t is = 0; is < ntx; ++is) {
        const int64_t o = (it - 1) + nt * is;
        vn_rec[o] = rec[o + nt * ntx * 1] * dir[1 + 4 * is] +
                    rec[o + nt * ntx * 2] * dir[2 + 4 * is] +
                    rec[o + nt * ntx * 3] * dir[3 + 4 * is];
    }

    /* extrapolate! (forward.jl:126-128) */
    if (gf && gf->kind != 0 && it < nt) {
        int rc = FN(ref_extrapolate_step)(d, gf, ntx, rec, vn_rec, src, vn_src, it);
        if (rc) return rc;
    }

    /* vn2v! into step it+1 (forward.jl:130-132; derivatives.jl:92-97) */
    if (it < nt) {
        for (int64_t is = 0; is < ntx; ++is) {
            const int64_t o = it + nt * is; /* 0-based row it == 1-based it+1 */
            for (int ch = 1; ch <= 3; ++ch)
                src[o + nt * ntx * ch] = src[o + nt * ntx * ch] + vn_src[o] * dir[ch + 4 * is];
        }
    }

    /* rigid mirrors (forward.jl:134-137) */
    for (int64_t k = 1; k <= nz; ++k)
        for (int64_t i = 1; i <= nx; ++i) Vy[IX(i, 1, k)] = -Vy[IX(i, 2, k)];
    for (int64_t j = 1; j <= ny; ++j)
        for (int64_t i = 1; i <= nx; ++i) Vz[IX(i, j, 1)] = -Vz[IX(i, j, 2)];
    return 0;
}

/* simulation.jl:542-557 (run_fdtd!) without on_step. */
int FN(ref_run_fdtd)(const ref_domain* d, T* field, int64_t p, T* mv_x, T* mv_y, T* mv_z,
                     T* mp_x, T* mp_y, T* mp_z, const T* a, const T* b, const T* ah,
                     const T* bh, const uint8_t* mask, int64_t ntx, const int64_t* idx,
                     const T* w, T* src, T* rec, const T* dir, T* vn_src, T* vn_rec,
                     int64_t start_step, int64_t end_step, const ref_gfmap* gf) {
    for (int64_t it = start_step; it <= end_step; ++it) {
        int rc = FN(ref_forward_onestep)(d, field, p, mv_x, mv_y, mv_z, mp_x, mp_y, mp_z, a, b,
                                         ah, bh, mask, ntx, idx, w, src, rec, dir, vn_src,
                                         vn_rec, it, gf);
        if (rc) return rc;
    }
    return 0;
}

#undef FN
#undef CAT
#undef CAT2
