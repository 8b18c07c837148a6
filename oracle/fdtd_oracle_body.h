/*
 * CPU ORACLE body (test infrastructure only), instantiated for float and
 * double by fdtd_oracle.c.  See that file's header for the upstream lines each
 * step restates and for the floating-point convention.
 *
 * Every function that does arithmetic brackets its loops with FP_ENTER /
 * FP_EXIT, which for the float instantiation switch the calling thread and
 * each OpenMP worker to flush-to-zero / denormals-are-zero and restore the
 * previous mode on exit.
 */
#define CAT_(a, b) a##b
#define CAT(a, b) CAT_(a, b)

/* CPML correction of one derivative (no upstream counterpart): dd is the
 * derivative along axis a taken on its own (sum_j c_j * difference, with the
 * per-term edge guards of the plain sums; scaled by 1/h for strains, unscaled
 * for accelerations) at stagger st (0 integer point, 1 half point), for a
 * cell inside the band of axis a.  The memory variable k of the cell is
 * advanced in place,
 *   psi' = fma(b, psi, a_coef * dd),
 * and the correction returned is corr = fma(dd, 1/kappa - 1, psi'), so that
 * the stretched derivative dd/kappa + psi' equals dd + corr. */
static inline REAL CAT(pml_corr, SUF)(const orc_shot *S, int a, int k, int st, int iz, int iy,
                                      int ix, REAL dd)
{
    const int dims[3] = {S->nz, S->ny, S->nx}, idx[3] = {iz, iy, ix};
    const int n = dims[a], ia = idx[a];
    const REAL *pr = (const REAL *)S->cpml_prof[a] + (int64_t)st * 3 * n;
    const int64_t f = (int64_t)k * pml_vol(S, a) + pml_flat(S, a, iz, iy, ix);
    REAL *psi = (REAL *)S->cpml_psi[a];
    const REAL p = FMA(pr[ia], psi[f], pr[n + ia] * dd);
    psi[f] = p;
    return FMA(dd, pr[2 * n + ia] - (REAL)1, p);
}

/* Strain corrections ga[6] of a CPML band cell: for each band axis in the
 * order z, y, x, its three derivatives (z: Db_z vz -> g3, Df_z vy -> g4,
 * Df_z vx -> g5; y: Db_y vy -> g2, Df_y vz -> g4, Df_y vx -> g6; x: Db_x vx
 * -> g1, Df_x vz -> g5, Df_x vy -> g6; memory variables 0, 1, 2 of the
 * axis, staggers 0, 1, 1) each through pml_corr, added into the component's
 * correction in that order, starting from 0.  Returns the mask of the
 * components that received a correction (bit k: g_{k+1}). */
static inline int CAT(pml_strain_corr, SUF)(const orc_shot *S, int64_t i, int iz, int iy,
                                            int ix, REAL ga[6])
{
    int mask = 0;
    const int nz = S->nz, ny = S->ny, nx = S->nx, nc = S->nc, L = S->cpml_width;
    const int64_t sy = nx, sz = (int64_t)nx * ny;
    const REAL *vx = (const REAL *)S->v[0], *vy = (const REAL *)S->v[1],
               *vz = (const REAL *)S->v[2];
    REAL c[4];
    for (int j = 0; j < 4; ++j) c[j] = (REAL)S->coeffs[j];
    const REAL inv_h = (REAL)S->inv_h;
    for (int k = 0; k < 6; ++k) ga[k] = 0;
    if (pml_in(iz, nz, L)) {
        REAL d0 = 0, d1 = 0, d2 = 0;
        for (int j = 0; j < nc; ++j) {
            if (iz >= j + 1 && iz < nz - j) d0 = FMA(c[j], vz[i + j * sz] - vz[i - (j + 1) * sz], d0);
            if (iz >= j && iz < nz - j - 1) {
                d1 = FMA(c[j], vy[i + (j + 1) * sz] - vy[i - j * sz], d1);
                d2 = FMA(c[j], vx[i + (j + 1) * sz] - vx[i - j * sz], d2);
            }
        }
        ga[2] = ga[2] + CAT(pml_corr, SUF)(S, 0, 0, 0, iz, iy, ix, d0 * inv_h);
        ga[3] = ga[3] + CAT(pml_corr, SUF)(S, 0, 1, 1, iz, iy, ix, d1 * inv_h);
        ga[4] = ga[4] + CAT(pml_corr, SUF)(S, 0, 2, 1, iz, iy, ix, d2 * inv_h);
        mask |= 0x1c;
    }
    if (pml_in(iy, ny, L)) {
        REAL d0 = 0, d1 = 0, d2 = 0;
        for (int j = 0; j < nc; ++j) {
            if (iy >= j + 1 && iy < ny - j) d0 = FMA(c[j], vy[i + j * sy] - vy[i - (j + 1) * sy], d0);
            if (iy >= j && iy < ny - j - 1) {
                d1 = FMA(c[j], vz[i + (j + 1) * sy] - vz[i - j * sy], d1);
                d2 = FMA(c[j], vx[i + (j + 1) * sy] - vx[i - j * sy], d2);
            }
        }
        ga[1] = ga[1] + CAT(pml_corr, SUF)(S, 1, 0, 0, iz, iy, ix, d0 * inv_h);
        ga[3] = ga[3] + CAT(pml_corr, SUF)(S, 1, 1, 1, iz, iy, ix, d1 * inv_h);
        ga[5] = ga[5] + CAT(pml_corr, SUF)(S, 1, 2, 1, iz, iy, ix, d2 * inv_h);
        mask |= 0x2a;
    }
    if (pml_in(ix, nx, L)) {
        REAL d0 = 0, d1 = 0, d2 = 0;
        for (int j = 0; j < nc; ++j) {
            if (ix >= j + 1 && ix < nx - j) d0 = FMA(c[j], vx[i + j] - vx[i - (j + 1)], d0);
            if (ix >= j && ix < nx - j - 1) {
                d1 = FMA(c[j], vz[i + (j + 1)] - vz[i - j], d1);
                d2 = FMA(c[j], vy[i + (j + 1)] - vy[i - j], d2);
            }
        }
        ga[0] = ga[0] + CAT(pml_corr, SUF)(S, 2, 0, 0, iz, iy, ix, d0 * inv_h);
        ga[4] = ga[4] + CAT(pml_corr, SUF)(S, 2, 1, 1, iz, iy, ix, d1 * inv_h);
        ga[5] = ga[5] + CAT(pml_corr, SUF)(S, 2, 2, 1, iz, iy, ix, d2 * inv_h);
        mask |= 0x31;
    }
    return mask;
}

/* Acceleration corrections gv[3] of a CPML band cell (unscaled, like the
 * plain accumulators): for each band axis in the order z, y, x, its three
 * derivatives (z: D_z s5 -> ax, D_z s4 -> ay, D_z s3 -> az, staggers 0, 0,
 * 1; y: D_y s6 -> ax, D_y s2 -> ay, D_y s4 -> az, staggers 0, 1, 0; x:
 * D_x s1 -> ax, D_x s6 -> ay, D_x s5 -> az, staggers 1, 0, 0; memory
 * variables 3, 4, 5 of the axis) through pml_corr. */
static inline void CAT(pml_accel_corr, SUF)(const orc_shot *S, int64_t i, int iz, int iy,
                                            int ix, REAL gv[3])
{
    const int nz = S->nz, ny = S->ny, nx = S->nx, nc = S->nc, L = S->cpml_width;
    const int64_t sy = nx, sz = (int64_t)nx * ny;
    const REAL *s1 = (const REAL *)S->s[0], *s2 = (const REAL *)S->s[1],
               *s3 = (const REAL *)S->s[2], *s4 = (const REAL *)S->s[3],
               *s5 = (const REAL *)S->s[4], *s6 = (const REAL *)S->s[5];
    REAL c[4];
    for (int j = 0; j < 4; ++j) c[j] = (REAL)S->coeffs[j];
    gv[0] = gv[1] = gv[2] = 0;
    if (pml_in(iz, nz, L)) {
        REAL e0 = 0, e1 = 0, e2 = 0;
        for (int j = 0; j < nc; ++j) {
            if (iz >= j + 1 && iz < nz - j) {
                e0 = FMA(c[j], s5[i + j * sz] - s5[i - (j + 1) * sz], e0);
                e1 = FMA(c[j], s4[i + j * sz] - s4[i - (j + 1) * sz], e1);
            }
            if (iz >= j && iz < nz - j - 1) e2 = FMA(c[j], s3[i + (j + 1) * sz] - s3[i - j * sz], e2);
        }
        gv[0] = gv[0] + CAT(pml_corr, SUF)(S, 0, 3, 0, iz, iy, ix, e0);
        gv[1] = gv[1] + CAT(pml_corr, SUF)(S, 0, 4, 0, iz, iy, ix, e1);
        gv[2] = gv[2] + CAT(pml_corr, SUF)(S, 0, 5, 1, iz, iy, ix, e2);
    }
    if (pml_in(iy, ny, L)) {
        REAL e0 = 0, e1 = 0, e2 = 0;
        for (int j = 0; j < nc; ++j) {
            if (iy >= j + 1 && iy < ny - j) {
                e0 = FMA(c[j], s6[i + j * sy] - s6[i - (j + 1) * sy], e0);
                e2 = FMA(c[j], s4[i + j * sy] - s4[i - (j + 1) * sy], e2);
            }
            if (iy >= j && iy < ny - j - 1) e1 = FMA(c[j], s2[i + (j + 1) * sy] - s2[i - j * sy], e1);
        }
        gv[0] = gv[0] + CAT(pml_corr, SUF)(S, 1, 3, 0, iz, iy, ix, e0);
        gv[1] = gv[1] + CAT(pml_corr, SUF)(S, 1, 4, 1, iz, iy, ix, e1);
        gv[2] = gv[2] + CAT(pml_corr, SUF)(S, 1, 5, 0, iz, iy, ix, e2);
    }
    if (pml_in(ix, nx, L)) {
        REAL e0 = 0, e1 = 0, e2 = 0;
        for (int j = 0; j < nc; ++j) {
            if (ix >= j && ix < nx - j - 1) e0 = FMA(c[j], s1[i + (j + 1)] - s1[i - j], e0);
            if (ix >= j + 1 && ix < nx - j) {
                e1 = FMA(c[j], s6[i + j] - s6[i - (j + 1)], e1);
                e2 = FMA(c[j], s5[i + j] - s5[i - (j + 1)], e2);
            }
        }
        gv[0] = gv[0] + CAT(pml_corr, SUF)(S, 2, 3, 1, iz, iy, ix, e0);
        gv[1] = gv[1] + CAT(pml_corr, SUF)(S, 2, 4, 0, iz, iy, ix, e1);
        gv[2] = gv[2] + CAT(pml_corr, SUF)(S, 2, 5, 0, iz, iy, ix, e2);
    }
}

/* Strains g1..g6 at every cell, from v (sim/core/fdtd.py:1136-1164).
 * Each g is sum_j c_j * (difference) accumulated j-ascending with per-term
 * edge guards, then scaled by 1/h.  A CPML band cell (no upstream
 * counterpart) then adds its corrections, g_k = g_k + ga_k for each
 * component a band axis of the cell corrects (pml_strain_corr). */
static void CAT(strains, SUF)(const orc_shot *S)
{
    const int nz = S->nz, ny = S->ny, nx = S->nx, nc = S->nc;
    const int64_t sy = nx, sz = (int64_t)nx * ny, ncell = sz * nz;
    const REAL *vx = (const REAL *)S->v[0], *vy = (const REAL *)S->v[1],
               *vz = (const REAL *)S->v[2];
    REAL *g = (REAL *)S->g;
    REAL c[4];
    for (int j = 0; j < 4; ++j) c[j] = (REAL)S->coeffs[j];
    const REAL inv_h = (REAL)S->inv_h;
#pragma omp parallel
    {
        FP_ENTER
#pragma omp for collapse(2) schedule(static)
        for (int iz = 0; iz < nz; ++iz)
            for (int iy = 0; iy < ny; ++iy)
                for (int ix = 0; ix < nx; ++ix) {
                    const int64_t i = ((int64_t)iz * ny + iy) * nx + ix;
                    REAL a1 = 0, a2 = 0, a3 = 0, a4 = 0, a5 = 0, a6 = 0;
                    for (int j = 0; j < nc; ++j) {
                        const REAL cj = c[j];
                        if (ix >= j + 1 && ix < nx - j)
                            a1 = FMA(cj, vx[i + j] - vx[i - (j + 1)], a1);
                        if (iy >= j + 1 && iy < ny - j)
                            a2 = FMA(cj, vy[i + j * sy] - vy[i - (j + 1) * sy], a2);
                        if (iz >= j + 1 && iz < nz - j)
                            a3 = FMA(cj, vz[i + j * sz] - vz[i - (j + 1) * sz], a3);
                        if (iy >= j && iy < ny - j - 1)
                            a4 = FMA(cj, vz[i + (j + 1) * sy] - vz[i - j * sy], a4);
                        if (iz >= j && iz < nz - j - 1)
                            a4 = FMA(cj, vy[i + (j + 1) * sz] - vy[i - j * sz], a4);
                        if (ix >= j && ix < nx - j - 1)
                            a5 = FMA(cj, vz[i + (j + 1)] - vz[i - j], a5);
                        if (iz >= j && iz < nz - j - 1)
                            a5 = FMA(cj, vx[i + (j + 1) * sz] - vx[i - j * sz], a5);
                        if (ix >= j && ix < nx - j - 1)
                            a6 = FMA(cj, vy[i + (j + 1)] - vy[i - j], a6);
                        if (iy >= j && iy < ny - j - 1)
                            a6 = FMA(cj, vx[i + (j + 1) * sy] - vx[i - j * sy], a6);
                    }
                    REAL a[6] = {a1 * inv_h, a2 * inv_h, a3 * inv_h,
                                 a4 * inv_h, a5 * inv_h, a6 * inv_h};
                    if (pml_cell(S, iz, iy, ix)) {
                        REAL ga[6];
                        const int mask = CAT(pml_strain_corr, SUF)(S, i, iz, iy, ix, ga);
                        for (int k = 0; k < 6; ++k)
                            if (mask >> k & 1) a[k] = a[k] + ga[k];
                    }
                    for (int k = 0; k < 6; ++k) g[k * ncell + i] = a[k];
                }
        FP_EXIT
    }
}

/* Strain J moved from its own position to position p
 * (sim/core/fdtd.py:612-642): per axis in z, y, x order, d > 0 takes taps
 * {0,+1} and needs i <= n-2, d < 0 takes taps {-1,0} and needs i >= 1; the
 * taps are summed left to right in z-major order and scaled by 1/2 or 1/4.
 * The tap offsets and guards of each (p, J) pair come from move_table(). */
static inline REAL CAT(moved, SUF)(const REAL *gJ, const move_t *mv, const int idx3[3],
                                   const int dims[3], int64_t i)
{
    if (mv->ntap == 1) return gJ[i];
    for (int a = 0; a < 3; ++a) {
        if (mv->guard[a] > 0 && !(idx3[a] <= dims[a] - 2)) return (REAL)0;
        if (mv->guard[a] < 0 && !(idx3[a] >= 1)) return (REAL)0;
    }
    REAL acc = gJ[i + mv->off[0]];
    for (int t = 1; t < mv->ntap; ++t) acc = acc + gJ[i + mv->off[t]];
    return acc * (mv->ntap == 4 ? (REAL)0.25 : (REAL)0.5);
}

/* Stress half-step with deferred damping (sim/core/fdtd.py:643-657):
 * s_I <- fma(dt, sum_{J active} C_IJ[m] * q_{pos(I), J}, s_I * damp). */
static void CAT(stress, SUF)(const orc_shot *S)
{
    const int nz = S->nz, ny = S->ny, nx = S->nx;
    const int64_t ncell = (int64_t)nz * ny * nx;
    const REAL *g = (const REAL *)S->g;
    const REAL *Ct = (const REAL *)S->Ct;
    const REAL *damp = (const REAL *)S->damp;
    const REAL dt = (REAL)S->dt;
    const uint64_t act = S->act_mask;
    const int dims[3] = {nz, ny, nx};
    move_t mv[4][6];
    move_table(mv, nz, ny, nx);
    int any[6];
    for (int I = 0; I < 6; ++I) {
        any[I] = 0;
        for (int J = 0; J < 6; ++J)
            if (act >> (6 * I + J) & 1u) any[I] = 1;
    }
#pragma omp parallel
    {
        FP_ENTER
#pragma omp for collapse(2) schedule(static)
        for (int iz = 0; iz < nz; ++iz)
            for (int iy = 0; iy < ny; ++iy)
                for (int ix = 0; ix < nx; ++ix) {
                    const int64_t i = ((int64_t)iz * ny + iy) * nx + ix;
                    const int idx3[3] = {iz, iy, ix};
                    REAL q[4][6];
                    for (int p = 0; p < 4; ++p)
                        for (int J = 0; J < 6; ++J)
                            q[p][J] = CAT(moved, SUF)(g + J * ncell, &mv[p][J], idx3, dims, i);
                    const REAL *Cm = Ct + (int64_t)S->mats[i] * 36;
                    const REAL dmp = damp ? damp[i] : (REAL)1;
                    for (int I = 0; I < 6; ++I) {
                        const int p = VP[I];
                        REAL acc = 0;
                        for (int J = 0; J < 6; ++J)
                            if (act >> (6 * I + J) & 1u) acc = FMA(Cm[6 * I + J], q[p][J], acc);
                        REAL *sI = (REAL *)S->s[I];
                        REAL sv = sI[i] * dmp;
                        if (any[I]) sv = FMA(dt, acc, sv);
                        sI[i] = sv;
                    }
                }
        FP_EXIT
    }
}

/* Explosive source (sim/core/fdtd.py:1798-1800): s1..s3[k] += w[n] * w_k. */
static void CAT(source, SUF)(const orc_shot *S, REAL wn)
{
    const REAL *w = (const REAL *)S->src_w;
    FP_ENTER
    for (int I = 0; I < 3; ++I) {
        REAL *sI = (REAL *)S->s[I];
        for (int k = 0; k < S->n_src; ++k) sI[S->src_idx[k]] = sI[S->src_idx[k]] + wn * w[k];
    }
    FP_EXIT
}

/* Velocity half-step (sim/core/fdtd.py:1211-1244): j-interleaved sums in
 * x, y, z tap order; face density 0.5 (rho[m] + rho[m(+1)]) or rho[m] on the
 * last slice; v <- (v + dt/h * a / rho_face) * damp.  A CPML band cell (no
 * upstream counterpart) adds its corrections to the sums first: a = a + gv
 * (pml_accel_corr). */
static void CAT(velocity, SUF)(const orc_shot *S)
{
    const int nz = S->nz, ny = S->ny, nx = S->nx, nc = S->nc;
    const int64_t sy = nx, sz = (int64_t)nx * ny;
    const REAL *s1 = (const REAL *)S->s[0], *s2 = (const REAL *)S->s[1],
               *s3 = (const REAL *)S->s[2], *s4 = (const REAL *)S->s[3],
               *s5 = (const REAL *)S->s[4], *s6 = (const REAL *)S->s[5];
    REAL *vx = (REAL *)S->v[0], *vy = (REAL *)S->v[1], *vz = (REAL *)S->v[2];
    const REAL *rho = (const REAL *)S->rho_t;
    const REAL *damp = (const REAL *)S->damp;
    const int32_t *mats = S->mats;
    REAL c[4];
    for (int j = 0; j < 4; ++j) c[j] = (REAL)S->coeffs[j];
    const REAL dt_h = (REAL)S->dt_h;
    const REAL half = (REAL)0.5;
#pragma omp parallel
    {
        FP_ENTER
#pragma omp for collapse(2) schedule(static)
        for (int iz = 0; iz < nz; ++iz)
            for (int iy = 0; iy < ny; ++iy)
                for (int ix = 0; ix < nx; ++ix) {
                    const int64_t i = ((int64_t)iz * ny + iy) * nx + ix;
                    REAL ax = 0, ay = 0, az = 0;
                    for (int j = 0; j < nc; ++j) {
                        const REAL cj = c[j];
                        if (ix >= j && ix < nx - j - 1)
                            ax = FMA(cj, s1[i + (j + 1)] - s1[i - j], ax);
                        if (ix >= j + 1 && ix < nx - j) {
                            ay = FMA(cj, s6[i + j] - s6[i - (j + 1)], ay);
                            az = FMA(cj, s5[i + j] - s5[i - (j + 1)], az);
                        }
                        if (iy >= j && iy < ny - j - 1)
                            ay = FMA(cj, s2[i + (j + 1) * sy] - s2[i - j * sy], ay);
                        if (iy >= j + 1 && iy < ny - j) {
                            ax = FMA(cj, s6[i + j * sy] - s6[i - (j + 1) * sy], ax);
                            az = FMA(cj, s4[i + j * sy] - s4[i - (j + 1) * sy], az);
                        }
                        if (iz >= j && iz < nz - j - 1)
                            az = FMA(cj, s3[i + (j + 1) * sz] - s3[i - j * sz], az);
                        if (iz >= j + 1 && iz < nz - j) {
                            ax = FMA(cj, s5[i + j * sz] - s5[i - (j + 1) * sz], ax);
                            ay = FMA(cj, s4[i + j * sz] - s4[i - (j + 1) * sz], ay);
                        }
                    }
                    if (pml_cell(S, iz, iy, ix)) {
                        REAL gv[3];
                        CAT(pml_accel_corr, SUF)(S, i, iz, iy, ix, gv);
                        ax = ax + gv[0];
                        ay = ay + gv[1];
                        az = az + gv[2];
                    }
                    const REAL r0 = rho[mats[i]];
                    const REAL rvx = ix < nx - 1 ? half * (r0 + rho[mats[i + 1]]) : r0;
                    const REAL rvy = iy < ny - 1 ? half * (r0 + rho[mats[i + sy]]) : r0;
                    const REAL rvz = iz < nz - 1 ? half * (r0 + rho[mats[i + sz]]) : r0;
                    const REAL dmp = damp ? damp[i] : (REAL)1;
                    vx[i] = (vx[i] + dt_h * ax / rvx) * dmp;
                    vy[i] = (vy[i] + dt_h * ay / rvy) * dmp;
                    vz[i] = (vz[i] + dt_h * az / rvz) * dmp;
                }
        FP_EXIT
    }
}

/* Receiver groups (sim/core/fdtd.py:1803-1813): weighted sum of
 * -(s1+s2+s3)/3 (or one velocity component) over each group's cells,
 * accumulated in footprint order. */
static void CAT(record, SUF)(const orc_shot *S, REAL *row)
{
    const REAL *s1 = (const REAL *)S->s[0], *s2 = (const REAL *)S->s[1],
               *s3 = (const REAL *)S->s[2];
    const REAL *w = (const REAL *)S->rec_w;
    FP_ENTER
    for (int r = 0; r < S->n_rec; ++r) {
        REAL acc = 0;
        for (int k = S->rec_ptr[r]; k < S->rec_ptr[r + 1]; ++k) {
            const int64_t q = S->rec_idx[k];
            REAL val;
            if (S->record_mode == 0)
                val = -(s1[q] + s2[q] + s3[q]) / (REAL)3;
            else
                val = ((const REAL *)S->v[S->record_mode - 1])[q];
            acc = acc + val * w[k];
        }
        row[r] = acc;
    }
    FP_EXIT
}

/* Steps n0 .. n1-1 of the hot loop (sim/core/fdtd.py:1786-1815):
 * stress, source, velocity, record.  traces is (nt, n_rec), row n written. */
void CAT(orc_run, SUF)(const orc_shot *S, int n0, int n1, const double *wavelet,
                       REAL *traces)
{
    orc_shot Sn = *S;
    for (int n = n0; n < n1; ++n) {
        CAT(strains, SUF)(&Sn);
        CAT(stress, SUF)(&Sn);
        CAT(source, SUF)(&Sn, (REAL)wavelet[n]);
        CAT(velocity, SUF)(&Sn);
        if (traces) CAT(record, SUF)(&Sn, traces + (int64_t)n * Sn.n_rec);
    }
}

/* Individual half-steps, exposed for finer-grained checks. */
void CAT(orc_stress, SUF)(const orc_shot *S) { CAT(strains, SUF)(S); CAT(stress, SUF)(S); }
void CAT(orc_velocity, SUF)(const orc_shot *S) { CAT(velocity, SUF)(S); }
void CAT(orc_source, SUF)(const orc_shot *S, double wn) { CAT(source, SUF)(S, (REAL)wn); }
void CAT(orc_record, SUF)(const orc_shot *S, REAL *row) { CAT(record, SUF)(S, row); }

#undef CAT
#undef CAT_
