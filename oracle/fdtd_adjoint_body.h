/*
 * CPU ORACLE, adjoint part (test infrastructure only), instantiated for float
 * and double by fdtd_oracle.c after fdtd_oracle_body.h.
 *
 * Reverse-mode derivative of one forward step of fdtd_oracle_body.h, written
 * in SCATTER form: every loop walks the cells of the FORWARD operator's
 * output, repeats that operator's guards and taps, and pushes the output's
 * adjoint back onto the inputs the forward statement read.  That makes each
 * routine the transpose of the forward routine it sits next to by
 * construction (in exact arithmetic); the GPU kernels gather instead, so the
 * two are independent restatements.  Parity pins (tests/test_oracle_adjoint.py):
 * the dot-product identity <L x, y> = <x, L^T y> for one step on random
 * states, media and damping (every edge guard is exercised); the identity
 * <F w, r> = <w, F^T r> for the whole wavelet-to-trace map; central finite
 * differences of the misfit in the stiffness table.
 *
 * The upstream repository's only exact adjoint is the vendored array-path
 * one (vendor/ringfwi/elastic3d.py:348-463); this file follows its order of
 * operations (recording, velocity update, stiffness gradient, stress update;
 * elastic3d.py:386-461) for the label-indexed, deferred-damped step.
 *
 * Forward step n, with v^n and s^n its start state (fdtd_oracle_body.h):
 *   A  s <- s * d + dt * C[m] M G v          (stress, moved strains)
 *   B  s1..3[src] += w[n] * w_k              (source)
 *   C  y_n = R s  (pressure record)          or, after D, y_n = R_v v
 *   D  v <- (v + dt_h * D s / rho_face) * d  (velocity)
 * Reverse step n with adjoints (lv, ls) of (v, s):
 *   E^T (velocity record) ; D^T ; C^T (pressure record) ; B^T (wavelet
 *   gradient) ; A^T with the stiffness gradient taken before ls is damped.
 * With a CPML the state also holds the memory variables psi, and the
 * reverse step carries their adjoints (S->cpml_psi points at them): D^T and
 * A^T each add the transposes of their band cells' corrections (pml_adj,
 * pml_scatter), and the stiffness gradient reads the strains of step n from
 * S->g, since they depend on the forward memory variables.
 */
#define CAT_(a, b) a##b
#define CAT(a, b) CAT_(a, b)

/* CPML (no upstream counterpart).  A band cell's forward correction of one
 * derivative dd (fdtd_oracle_body.h, pml_corr) is
 *   psi' = b psi + a dd,   corr = (1/kappa - 1) dd + psi',
 * with corr added to the strain (acceleration) the derivative belongs to.
 * Given the adjoint ebar of that strain (acceleration) and the adjoint pbar
 * of psi' (carried back from the later step), with P = ebar + pbar:
 *   adjoint of dd  = (1/kappa - 1) ebar + a P   (returned),
 *   adjoint of psi = b P                        (stored over pbar).
 * S->cpml_psi holds the ADJOINT memory variables in the reverse step. */
static REAL CAT(pml_adj, SUF)(const orc_shot *S, int a, int k, int st, int iz, int iy, int ix,
                              REAL ebar)
{
    const int n = a == 0 ? S->nz : (a == 1 ? S->ny : S->nx);
    const int ia = a == 0 ? iz : (a == 1 ? iy : ix);
    const REAL *pr = (const REAL *)S->cpml_prof[a] + (int64_t)st * 3 * n;
    const int64_t f = (int64_t)k * pml_vol(S, a) + pml_flat(S, a, iz, iy, ix);
    REAL *pb = (REAL *)S->cpml_psi[a];
    const REAL P = ebar + pb[f];
    pb[f] = pr[ia] * P;
    if (orc_adj_plant == 4) return (pr[2 * n + ia] - (REAL)1) * ebar;
    return (pr[2 * n + ia] - (REAL)1) * ebar + pr[n + ia] * P;
}

/* Transpose of one band derivative, pushed into field f: the forward
 * derivative at cell i along an axis of stride st and length n (ia the
 * cell's index on it) is sum_j c_j (f(i+j) - f(i-j-1)) for a backward
 * difference (fwd = 0, kept iff j+1 <= ia < n-j) or sum_j c_j (f(i+j+1) -
 * f(i-j)) for a forward one (fwd = 1, kept iff j <= ia < n-j-1), times
 * scale. */
static void CAT(pml_scatter, SUF)(REAL *f, int64_t i, int ia, int n, int64_t st, int fwd,
                                  REAL w, const REAL *c, int nc, REAL scale)
{
    for (int j = 0; j < nc; ++j) {
        const REAL cw = c[j] * scale * w;
        if (!fwd) {
            if (ia >= j + 1 && ia < n - j) {
                f[i + j * st] += cw;
                f[i - (j + 1) * st] -= cw;
            }
        } else if (ia >= j && ia < n - j - 1) {
            f[i + (j + 1) * st] += cw;
            f[i - j * st] -= cw;
        }
    }
}

/* D^T: t = d * lv; e_a = dt_h * t_a / rho_face_a; ls += D^T e; lv <- t. */
static void CAT(adj_velocity, SUF)(const orc_shot *S, REAL *lv[3], REAL *ls[6])
{
    const int nz = S->nz, ny = S->ny, nx = S->nx, nc = S->nc;
    const int64_t sy = nx, sz = (int64_t)nx * ny;
    REAL *s1 = ls[0], *s2 = ls[1], *s3 = ls[2], *s4 = ls[3], *s5 = ls[4], *s6 = ls[5];
    const REAL *rho = (const REAL *)S->rho_t;
    const REAL *damp = (const REAL *)S->damp;
    const int32_t *mats = S->mats;
    REAL c[4];
    for (int j = 0; j < 4; ++j) c[j] = (REAL)S->coeffs[j];
    const REAL dt_h = (REAL)S->dt_h;
    const REAL half = (REAL)0.5;
    FP_ENTER
    for (int iz = 0; iz < nz; ++iz)
        for (int iy = 0; iy < ny; ++iy)
            for (int ix = 0; ix < nx; ++ix) {
                const int64_t i = ((int64_t)iz * ny + iy) * nx + ix;
                const REAL r0 = rho[mats[i]];
                REAL rvx = ix < nx - 1 ? half * (r0 + rho[mats[i + 1]]) : r0;
                REAL rvy = iy < ny - 1 ? half * (r0 + rho[mats[i + sy]]) : r0;
                const REAL rvz = iz < nz - 1 ? half * (r0 + rho[mats[i + sz]]) : r0;
                if (orc_adj_plant == 1) rvx = rvy = r0;
                const REAL dmp = damp ? damp[i] : (REAL)1;
                const REAL tx = lv[0][i] * dmp, ty = lv[1][i] * dmp, tz = lv[2][i] * dmp;
                lv[0][i] = tx;
                lv[1][i] = ty;
                lv[2][i] = tz;
                const REAL ex = dt_h * tx / rvx, ey = dt_h * ty / rvy, ez = dt_h * tz / rvz;
                for (int j = 0; j < nc; ++j) {
                    const REAL cj = c[j];
                    if (ix >= j && ix < nx - j - 1) {
                        s1[i + (j + 1)] += cj * ex;
                        s1[i - j] -= cj * ex;
                    }
                    if (ix >= j + 1 && ix < nx - j) {
                        s6[i + j] += cj * ey;
                        s6[i - (j + 1)] -= cj * ey;
                        s5[i + j] += cj * ez;
                        s5[i - (j + 1)] -= cj * ez;
                    }
                    if (iy >= j && iy < ny - j - 1) {
                        s2[i + (j + 1) * sy] += cj * ey;
                        s2[i - j * sy] -= cj * ey;
                    }
                    if (iy >= j + 1 && iy < ny - j) {
                        s6[i + j * sy] += cj * ex;
                        s6[i - (j + 1) * sy] -= cj * ex;
                        s4[i + j * sy] += cj * ez;
                        s4[i - (j + 1) * sy] -= cj * ez;
                    }
                    if (iz >= j && iz < nz - j - 1) {
                        s3[i + (j + 1) * sz] += cj * ez;
                        s3[i - j * sz] -= cj * ez;
                    }
                    if (iz >= j + 1 && iz < nz - j) {
                        s5[i + j * sz] += cj * ex;
                        s5[i - (j + 1) * sz] -= cj * ex;
                        s4[i + j * sz] += cj * ey;
                        s4[i - (j + 1) * sz] -= cj * ey;
                    }
                }
                /* CPML band cell: the transposes of its acceleration
                 * corrections (pml_accel_corr: z band D_z s5 -> ax, D_z s4 ->
                 * ay, D_z s3 -> az; y band D_y s6, D_y s2, D_y s4; x band D_x
                 * s1, D_x s6, D_x s5; memory variables 3..5) */
                if (pml_cell(S, iz, iy, ix)) {
                    const int L = S->cpml_width;
                    const REAL one = (REAL)1;
                    if (pml_in(iz, nz, L)) {
                        CAT(pml_scatter, SUF)(s5, i, iz, nz, sz, 0,
                                              CAT(pml_adj, SUF)(S, 0, 3, 0, iz, iy, ix, ex), c, nc, one);
                        CAT(pml_scatter, SUF)(s4, i, iz, nz, sz, 0,
                                              CAT(pml_adj, SUF)(S, 0, 4, 0, iz, iy, ix, ey), c, nc, one);
                        CAT(pml_scatter, SUF)(s3, i, iz, nz, sz, 1,
                                              CAT(pml_adj, SUF)(S, 0, 5, 1, iz, iy, ix, ez), c, nc, one);
                    }
                    if (pml_in(iy, ny, L)) {
                        CAT(pml_scatter, SUF)(s6, i, iy, ny, sy, 0,
                                              CAT(pml_adj, SUF)(S, 1, 3, 0, iz, iy, ix, ex), c, nc, one);
                        CAT(pml_scatter, SUF)(s2, i, iy, ny, sy, 1,
                                              CAT(pml_adj, SUF)(S, 1, 4, 1, iz, iy, ix, ey), c, nc, one);
                        CAT(pml_scatter, SUF)(s4, i, iy, ny, sy, 0,
                                              CAT(pml_adj, SUF)(S, 1, 5, 0, iz, iy, ix, ez), c, nc, one);
                    }
                    if (pml_in(ix, nx, L)) {
                        CAT(pml_scatter, SUF)(s1, i, ix, nx, 1, 1,
                                              CAT(pml_adj, SUF)(S, 2, 3, 1, iz, iy, ix, ex), c, nc, one);
                        CAT(pml_scatter, SUF)(s6, i, ix, nx, 1, 0,
                                              CAT(pml_adj, SUF)(S, 2, 4, 0, iz, iy, ix, ey), c, nc, one);
                        CAT(pml_scatter, SUF)(s5, i, ix, nx, 1, 0,
                                              CAT(pml_adj, SUF)(S, 2, 5, 0, iz, iy, ix, ez), c, nc, one);
                    }
                }
            }
    FP_EXIT
}

/* C^T / E^T: the transpose of the receiver sums.  Pressure: y = sum_k w_k *
 * (-(s1+s2+s3)/3)[k] gives ls_I[k] -= r * w_k / 3; velocity mode a gives
 * lv_a[k] += r * w_k. */
static void CAT(adj_record, SUF)(const orc_shot *S, REAL *lv[3], REAL *ls[6], const REAL *res)
{
    const REAL *w = (const REAL *)S->rec_w;
    FP_ENTER
    for (int r = 0; r < S->n_rec; ++r)
        for (int k = S->rec_ptr[r]; k < S->rec_ptr[r + 1]; ++k) {
            const int64_t q = S->rec_idx[k];
            if (S->record_mode == 0) {
                const REAL a = res[r] * w[k] / (REAL)3;
                ls[0][q] -= a;
                ls[1][q] -= a;
                ls[2][q] -= a;
            } else {
                lv[S->record_mode - 1][q] += res[r] * w[k];
            }
        }
    FP_EXIT
}

/* A^T, with S->v = v^n and S->g scratch: g = G v^n, q = M g (the forward
 * routines); gC[m][6I+J] += dt * ls_I * q_{VP[I],J} for all 36 entries
 * (also those no material uses: the derivative of the forward map in that
 * entry); u_{VP[I],J} = dt * C_IJ[m] * ls_I pushed back through the moves
 * (gbar) and the strain stencils (into lv); finally ls <- ls * d. */
static void CAT(adj_stress, SUF)(const orc_shot *S, REAL *lv[3], REAL *ls[6], double *gC,
                                 REAL *gbar)
{
    const int nz = S->nz, ny = S->ny, nx = S->nx, nc = S->nc;
    const int64_t sy = nx, sz = (int64_t)nx * ny, ncell = sz * nz;
    const REAL *g = (const REAL *)S->g;
    const REAL *Ct = (const REAL *)S->Ct;
    const REAL *damp = (const REAL *)S->damp;
    const REAL dt = (REAL)S->dt;
    const uint64_t act = S->act_mask;
    const int dims[3] = {nz, ny, nx};
    move_t mv[4][6];
    move_table(mv, nz, ny, nx);
    /* strains of v^n for the gradient; with a CPML they depend on the
     * forward memory variables too, and the caller passes them in S->g */
    if (gC && S->cpml_width == 0) CAT(strains, SUF)(S);
    memset(gbar, 0, sizeof(REAL) * 6 * ncell);
    REAL c[4];
    for (int j = 0; j < 4; ++j) c[j] = (REAL)S->coeffs[j];
    const REAL inv_h = (REAL)S->inv_h;
    FP_ENTER
    for (int iz = 0; iz < nz; ++iz)
        for (int iy = 0; iy < ny; ++iy)
            for (int ix = 0; ix < nx; ++ix) {
                const int64_t i = ((int64_t)iz * ny + iy) * nx + ix;
                const int idx3[3] = {iz, iy, ix};
                const int32_t m = S->mats[i];
                const REAL *Cm = Ct + (int64_t)m * 36;
                double *gm = gC ? gC + (int64_t)m * 36 : NULL;
                for (int I = 0; I < 6; ++I) {
                    const int p = VP[I];
                    const REAL lI = ls[I][i];
                    for (int J = 0; J < 6; ++J) {
                        const move_t *t = &mv[p][J];
                        if (gm) {
                            const REAL qv = CAT(moved, SUF)(g + J * ncell, t, idx3, dims, i);
                            gm[6 * I + J] += (double)dt * (double)lI * (double)qv;
                        }
                        if (!(act >> (6 * I + J) & 1u)) continue;
                        const REAL u = dt * Cm[6 * I + J] * lI;
                        REAL *gJ = gbar + J * ncell;
                        if (t->ntap == 1) {
                            gJ[i] += u;
                            continue;
                        }
                        int ok = 1;
                        for (int a = 0; a < 3; ++a) {
                            if (t->guard[a] > 0 && !(idx3[a] <= dims[a] - 2)) ok = 0;
                            if (t->guard[a] < 0 && !(idx3[a] >= 1)) ok = 0;
                        }
                        if (!ok) continue;
                        REAL f = t->ntap == 4 ? (REAL)0.25 : (REAL)0.5;
                        if (orc_adj_plant == 2) f = (REAL)0.5;
                        for (int k = 0; k < t->ntap; ++k) gJ[i + t->off[k]] += f * u;
                    }
                }
            }
    /* strain stencils transposed: g_J = inv_h * sum_j c_j (difference) */
    REAL *vx = lv[0], *vy = lv[1], *vz = lv[2];
    const REAL *b1 = gbar, *b2 = gbar + ncell, *b3 = gbar + 2 * ncell, *b4 = gbar + 3 * ncell,
               *b5 = gbar + 4 * ncell, *b6 = gbar + 5 * ncell;
    /* CPML band cells: the transposes of their strain corrections
     * (pml_strain_corr: z band Db_z vz -> g3, Df_z vy -> g4, Df_z vx -> g5;
     * y band Db_y vy -> g2, Df_y vz -> g4, Df_y vx -> g6; x band Db_x vx ->
     * g1, Df_x vz -> g5, Df_x vy -> g6; memory variables 0..2, staggers 0, 1,
     * 1), scaled by 1/h like the strains */
    if (S->cpml_width > 0) {
        const int L = S->cpml_width;
        for (int iz = 0; iz < nz; ++iz)
            for (int iy = 0; iy < ny; ++iy)
                for (int ix = 0; ix < nx; ++ix) {
                    if (!pml_cell(S, iz, iy, ix)) continue;
                    const int64_t i = ((int64_t)iz * ny + iy) * nx + ix;
                    if (pml_in(iz, nz, L)) {
                        CAT(pml_scatter, SUF)(vz, i, iz, nz, sz, 0,
                                              CAT(pml_adj, SUF)(S, 0, 0, 0, iz, iy, ix, b3[i]), c, nc, inv_h);
                        CAT(pml_scatter, SUF)(vy, i, iz, nz, sz, 1,
                                              CAT(pml_adj, SUF)(S, 0, 1, 1, iz, iy, ix, b4[i]), c, nc, inv_h);
                        CAT(pml_scatter, SUF)(vx, i, iz, nz, sz, 1,
                                              CAT(pml_adj, SUF)(S, 0, 2, 1, iz, iy, ix, b5[i]), c, nc, inv_h);
                    }
                    if (pml_in(iy, ny, L)) {
                        CAT(pml_scatter, SUF)(vy, i, iy, ny, sy, 0,
                                              CAT(pml_adj, SUF)(S, 1, 0, 0, iz, iy, ix, b2[i]), c, nc, inv_h);
                        CAT(pml_scatter, SUF)(vz, i, iy, ny, sy, 1,
                                              CAT(pml_adj, SUF)(S, 1, 1, 1, iz, iy, ix, b4[i]), c, nc, inv_h);
                        CAT(pml_scatter, SUF)(vx, i, iy, ny, sy, 1,
                                              CAT(pml_adj, SUF)(S, 1, 2, 1, iz, iy, ix, b6[i]), c, nc, inv_h);
                    }
                    if (pml_in(ix, nx, L)) {
                        CAT(pml_scatter, SUF)(vx, i, ix, nx, 1, 0,
                                              CAT(pml_adj, SUF)(S, 2, 0, 0, iz, iy, ix, b1[i]), c, nc, inv_h);
                        CAT(pml_scatter, SUF)(vz, i, ix, nx, 1, 1,
                                              CAT(pml_adj, SUF)(S, 2, 1, 1, iz, iy, ix, b5[i]), c, nc, inv_h);
                        CAT(pml_scatter, SUF)(vy, i, ix, nx, 1, 1,
                                              CAT(pml_adj, SUF)(S, 2, 2, 1, iz, iy, ix, b6[i]), c, nc, inv_h);
                    }
                }
    }
    for (int iz = 0; iz < nz; ++iz)
        for (int iy = 0; iy < ny; ++iy)
            for (int ix = 0; ix < nx; ++ix) {
                const int64_t i = ((int64_t)iz * ny + iy) * nx + ix;
                for (int j = 0; j < nc; ++j) {
                    const REAL cj = c[j] * inv_h;
                    if (ix >= j + 1 && ix < nx - j) {
                        vx[i + j] += cj * b1[i];
                        vx[i - (j + 1)] -= cj * b1[i];
                    }
                    if (iy >= j + 1 && iy < ny - j) {
                        vy[i + j * sy] += cj * b2[i];
                        vy[i - (j + 1) * sy] -= cj * b2[i];
                    }
                    if (iz >= j + 1 && iz < nz - j) {
                        vz[i + j * sz] += cj * b3[i];
                        vz[i - (j + 1) * sz] -= cj * b3[i];
                    }
                    if (iy >= j && iy < ny - j - 1) {
                        vz[i + (j + 1) * sy] += cj * b4[i];
                        vz[i - j * sy] -= cj * b4[i];
                    }
                    if (iz >= j && iz < nz - j - 1) {
                        vy[i + (j + 1) * sz] += cj * b4[i];
                        vy[i - j * sz] -= cj * b4[i];
                    }
                    if (ix >= j && ix < nx - j - 1) {
                        vz[i + (j + 1)] += cj * b5[i];
                        vz[i - j] -= cj * b5[i];
                    }
                    if (iz >= j && iz < nz - j - 1) {
                        vx[i + (j + 1) * sz] += cj * b5[i];
                        vx[i - j * sz] -= cj * b5[i];
                    }
                    if (ix >= j && ix < nx - j - 1) {
                        vy[i + (j + 1)] += cj * b6[i];
                        vy[i - j] -= cj * b6[i];
                    }
                    if (iy >= j && iy < ny - j - 1) {
                        vx[i + (j + 1) * sy] += cj * b6[i];
                        vx[i - j * sy] -= cj * b6[i];
                    }
                }
            }
    if (orc_adj_plant != 3)
        for (int I = 0; I < 6; ++I)
            for (int64_t i = 0; i < ncell; ++i) ls[I][i] *= damp ? damp[i] : (REAL)1;
    FP_EXIT
}

/* One reverse step n.  S->v must hold v^n (the forward state at the start of
 * step n); lv / ls hold the adjoints of the state after step n and are
 * overwritten with those of the state before it.  res is the trace adjoint
 * row dJ/dy_n (n_rec) or NULL; gC (n_mat * 36 doubles) accumulates dJ/dCt
 * when non-NULL; *gw receives dJ/dwavelet[n] when non-NULL.  gbar is 6 *
 * ncell scratch. */
void CAT(orc_adjoint_step, SUF)(const orc_shot *S, REAL *lv0, REAL *lv1, REAL *lv2, REAL *ls_flat,
                                const REAL *res, double *gC, double *gw, REAL *gbar)
{
    const int64_t ncell = (int64_t)S->nz * S->ny * S->nx;
    REAL *lv[3] = {lv0, lv1, lv2};
    REAL *ls[6];
    for (int I = 0; I < 6; ++I) ls[I] = ls_flat + I * ncell;
    if (res && S->record_mode != 0) CAT(adj_record, SUF)(S, lv, ls, res);
    CAT(adj_velocity, SUF)(S, lv, ls);
    if (res && S->record_mode == 0) CAT(adj_record, SUF)(S, lv, ls, res);
    if (gw) {
        const REAL *w = (const REAL *)S->src_w;
        double acc = 0;
        for (int k = 0; k < S->n_src; ++k) {
            const int64_t q = S->src_idx[k];
            acc += (double)w[k] * ((double)ls[0][q] + (double)ls[1][q] + (double)ls[2][q]);
        }
        *gw = acc;
    }
    CAT(adj_stress, SUF)(S, lv, ls, gC, gbar);
}

#undef CAT
#undef CAT_
