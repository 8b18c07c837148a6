// Shared device-side definitions of the gfx950 FDTD kernels.
//
// Staggering (half-cell offsets along z, y, x), as in the upstream solver
// (sim/core/fdtd.py:39-41): vx (0,0,+), vy (0,+,0), vz (+,0,0); s1..s3 at
// (0,0,0); s4 (yz) at (+,+,0); s5 (xz) at (+,0,+); s6 (xy) at (0,+,+).
// Strain component J lives at position kVP[J]; kPos gives each position's
// half-cell offsets in (z, y, x) order.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pefdtd.h"

namespace pe {

constexpr int kVP[6] = {0, 0, 0, 1, 2, 3};
constexpr int kPos[4][3] = {{0, 0, 0}, {1, 1, 0}, {1, 0, 1}, {0, 1, 1}};

// Kernel-side view of a pe_shot, passed by value.
template <typename R>
struct Shot {
    int nz, ny, nx, nc, n_mat;
    R c[4];
    R inv_h, dt, dt_h;
    uint64_t act;
    const int32_t *__restrict__ mats;
    const R *__restrict__ Ct;
    const R *__restrict__ rho;
    const R *__restrict__ damp;
    R *v[3];
    R *s[6];
    R *g;
    // CPML (no upstream counterpart): band half-width (0 = none), per-axis
    // (z, y, x) profiles real[2 staggers][b, a, 1/kappa][n], memory
    // variables real[6][band volume] (advanced in place) and, for the tiled
    // kernels, the corrections real[6][ncell] the band pre-pass writes.
    int pml;
    const R *prof[3];
    R *psi[3];
    R *corr;
};

template <typename R>
inline Shot<R> make_shot(const pe_shot *p)
{
    Shot<R> s;
    s.nz = p->nz;
    s.ny = p->ny;
    s.nx = p->nx;
    s.nc = p->nc;
    s.n_mat = p->n_mat;
    for (int j = 0; j < 4; ++j) s.c[j] = j < p->nc ? (R)p->coeffs[j] : (R)0;
    s.inv_h = (R)p->inv_h;
    s.dt = (R)p->dt;
    s.dt_h = (R)p->dt_h;
    s.act = p->act_mask;
    s.mats = p->mats;
    s.Ct = (const R *)p->Ct;
    s.rho = (const R *)p->rho_t;
    s.damp = (const R *)p->damp;
    for (int k = 0; k < 3; ++k) s.v[k] = (R *)p->v[k];
    for (int k = 0; k < 6; ++k) s.s[k] = (R *)p->s[k];
    s.g = (R *)p->g;
    s.pml = p->cpml_width;
    for (int a = 0; a < 3; ++a) {
        s.prof[a] = (const R *)p->cpml_prof[a];
        s.psi[a] = (R *)p->cpml_psi[a];
    }
    s.corr = (R *)p->cpml_corr;
    return s;
}

__device__ __forceinline__ float fma_r(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fma_r(double a, double b, double c) { return __builtin_fma(a, b, c); }

// ------------------------------------------------------------------ CPML
// Convolutional PML (no upstream counterpart).  Axis A (0 z, 1 y, 2 x) of
// length n has a band of L+1 cells at each end: i <= L or i >= n-1-L, band
// index i or i - (n - 2L - 2).  A cell in the band of any axis is the plain
// cell plus corrections: for each band axis (in the order z, y, x) its three
// derivatives dd are taken on their own (the plain sums' taps and per-term
// guards; strains scaled by 1/h, accelerations unscaled), memory variable K
// of the cell is advanced in place, psi' = fma(b, psi, a * dd), with the
// profile of the derivative's stagger ST (0 integer point, 1 half point),
// and corr = fma(dd, 1/kappa - 1, psi') is added into the correction of the
// component the derivative belongs to.  Then dd/kappa + psi' = dd + corr.
__device__ __forceinline__ bool pml_in(int i, int n, int L) { return i <= L || i >= n - 1 - L; }

template <typename R>
__device__ __forceinline__ bool pml_cell(const Shot<R> &S, int iz, int iy, int ix)
{
    const int L = S.pml;
    return L > 0 && (pml_in(iz, S.nz, L) || pml_in(iy, S.ny, L) || pml_in(ix, S.nx, L));
}

template <int A, int K, int ST, typename R>
__device__ __forceinline__ R pml_corr(const Shot<R> &S, int iz, int iy, int ix, R dd)
{
    const int L = S.pml, W = 2 * L + 2;
    const int n = A == 0 ? S.nz : (A == 1 ? S.ny : S.nx);
    const int ia = A == 0 ? iz : (A == 1 ? iy : ix);
    const int bi = ia <= L ? ia : ia - (n - W);
    size_t vol, f;
    if constexpr (A == 0) {
        vol = (size_t)W * S.ny * S.nx;
        f = ((size_t)bi * S.ny + iy) * S.nx + ix;
    } else if constexpr (A == 1) {
        vol = (size_t)S.nz * W * S.nx;
        f = ((size_t)iz * W + bi) * S.nx + ix;
    } else {
        vol = (size_t)S.nz * S.ny * W;
        f = ((size_t)iz * S.ny + iy) * W + bi;
    }
    const R *pr = S.prof[A] + (size_t)ST * 3 * n;
    R *psi = S.psi[A] + (size_t)K * vol + f;
    const R p = fma_r(pr[ia], *psi, pr[n + ia] * dd);
    *psi = p;
    return fma_r(dd, pr[2 * n + ia] - (R)1, p);
}

// Components of the strains a cell's bands correct (bit k: g_{k+1}): the z
// band g3, g4, g5; the y band g2, g4, g6; the x band g1, g5, g6.
constexpr int kPmlMaskZ = 0x1c, kPmlMaskY = 0x2a, kPmlMaskX = 0x31;

template <typename R>
__device__ __forceinline__ int pml_strain_mask(const Shot<R> &S, int iz, int iy, int ix)
{
    const int L = S.pml;
    return (pml_in(iz, S.nz, L) ? kPmlMaskZ : 0) | (pml_in(iy, S.ny, L) ? kPmlMaskY : 0) |
           (pml_in(ix, S.nx, L) ? kPmlMaskX : 0);
}

// Strain corrections ga[6] of the band cell (iz, iy, ix) at flat index i,
// from v: z band: Db_z vz -> g3, Df_z vy -> g4, Df_z vx -> g5; y band:
// Db_y vy -> g2, Df_y vz -> g4, Df_y vx -> g6; x band: Db_x vx -> g1,
// Df_x vz -> g5, Df_x vy -> g6 (memory variables 0, 1, 2; staggers 0, 1, 1).
// Only the components of pml_strain_mask are corrected (added to the plain
// strain); the others are left 0.
template <typename R>
__device__ __forceinline__ void pml_strain_corr(const Shot<R> &S, int i, int iz, int iy, int ix,
                                                R (&ga)[6])
{
    const int nz = S.nz, ny = S.ny, nx = S.nx, L = S.pml;
    const int sy = nx, sz = nx * ny;
    const R *__restrict__ vx = S.v[0];
    const R *__restrict__ vy = S.v[1];
    const R *__restrict__ vz = S.v[2];
#pragma unroll
    for (int k = 0; k < 6; ++k) ga[k] = 0;
    if (pml_in(iz, nz, L)) {
        R d0 = 0, d1 = 0, d2 = 0;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (j >= S.nc) break;
            const R cj = S.c[j];
            if (iz >= j + 1 && iz < nz - j) d0 = fma_r(cj, vz[i + j * sz] - vz[i - (j + 1) * sz], d0);
            if (iz >= j && iz < nz - j - 1) {
                d1 = fma_r(cj, vy[i + (j + 1) * sz] - vy[i - j * sz], d1);
                d2 = fma_r(cj, vx[i + (j + 1) * sz] - vx[i - j * sz], d2);
            }
        }
        ga[2] = ga[2] + pml_corr<0, 0, 0>(S, iz, iy, ix, d0 * S.inv_h);
        ga[3] = ga[3] + pml_corr<0, 1, 1>(S, iz, iy, ix, d1 * S.inv_h);
        ga[4] = ga[4] + pml_corr<0, 2, 1>(S, iz, iy, ix, d2 * S.inv_h);
    }
    if (pml_in(iy, ny, L)) {
        R d0 = 0, d1 = 0, d2 = 0;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (j >= S.nc) break;
            const R cj = S.c[j];
            if (iy >= j + 1 && iy < ny - j) d0 = fma_r(cj, vy[i + j * sy] - vy[i - (j + 1) * sy], d0);
            if (iy >= j && iy < ny - j - 1) {
                d1 = fma_r(cj, vz[i + (j + 1) * sy] - vz[i - j * sy], d1);
                d2 = fma_r(cj, vx[i + (j + 1) * sy] - vx[i - j * sy], d2);
            }
        }
        ga[1] = ga[1] + pml_corr<1, 0, 0>(S, iz, iy, ix, d0 * S.inv_h);
        ga[3] = ga[3] + pml_corr<1, 1, 1>(S, iz, iy, ix, d1 * S.inv_h);
        ga[5] = ga[5] + pml_corr<1, 2, 1>(S, iz, iy, ix, d2 * S.inv_h);
    }
    if (pml_in(ix, nx, L)) {
        R d0 = 0, d1 = 0, d2 = 0;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (j >= S.nc) break;
            const R cj = S.c[j];
            if (ix >= j + 1 && ix < nx - j) d0 = fma_r(cj, vx[i + j] - vx[i - (j + 1)], d0);
            if (ix >= j && ix < nx - j - 1) {
                d1 = fma_r(cj, vz[i + (j + 1)] - vz[i - j], d1);
                d2 = fma_r(cj, vy[i + (j + 1)] - vy[i - j], d2);
            }
        }
        ga[0] = ga[0] + pml_corr<2, 0, 0>(S, iz, iy, ix, d0 * S.inv_h);
        ga[4] = ga[4] + pml_corr<2, 1, 1>(S, iz, iy, ix, d1 * S.inv_h);
        ga[5] = ga[5] + pml_corr<2, 2, 1>(S, iz, iy, ix, d2 * S.inv_h);
    }
}

// Acceleration corrections gv[3] (unscaled, like the plain sums) of a band
// cell, from s: z band: D_z s5 -> ax, D_z s4 -> ay, D_z s3 -> az (staggers
// 0, 0, 1); y band: D_y s6 -> ax, D_y s2 -> ay, D_y s4 -> az (0, 1, 0); x
// band: D_x s1 -> ax, D_x s6 -> ay, D_x s5 -> az (1, 0, 0); memory
// variables 3, 4, 5.
template <typename R>
__device__ __forceinline__ void pml_accel_corr(const Shot<R> &S, int i, int iz, int iy, int ix,
                                               R (&gv)[3])
{
    const int nz = S.nz, ny = S.ny, nx = S.nx, L = S.pml;
    const int sy = nx, sz = nx * ny;
    const R *__restrict__ s1 = S.s[0];
    const R *__restrict__ s2 = S.s[1];
    const R *__restrict__ s3 = S.s[2];
    const R *__restrict__ s4 = S.s[3];
    const R *__restrict__ s5 = S.s[4];
    const R *__restrict__ s6 = S.s[5];
    gv[0] = gv[1] = gv[2] = 0;
    if (pml_in(iz, nz, L)) {
        R e0 = 0, e1 = 0, e2 = 0;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (j >= S.nc) break;
            const R cj = S.c[j];
            if (iz >= j + 1 && iz < nz - j) {
                e0 = fma_r(cj, s5[i + j * sz] - s5[i - (j + 1) * sz], e0);
                e1 = fma_r(cj, s4[i + j * sz] - s4[i - (j + 1) * sz], e1);
            }
            if (iz >= j && iz < nz - j - 1) e2 = fma_r(cj, s3[i + (j + 1) * sz] - s3[i - j * sz], e2);
        }
        gv[0] = gv[0] + pml_corr<0, 3, 0>(S, iz, iy, ix, e0);
        gv[1] = gv[1] + pml_corr<0, 4, 0>(S, iz, iy, ix, e1);
        gv[2] = gv[2] + pml_corr<0, 5, 1>(S, iz, iy, ix, e2);
    }
    if (pml_in(iy, ny, L)) {
        R e0 = 0, e1 = 0, e2 = 0;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (j >= S.nc) break;
            const R cj = S.c[j];
            if (iy >= j + 1 && iy < ny - j) {
                e0 = fma_r(cj, s6[i + j * sy] - s6[i - (j + 1) * sy], e0);
                e2 = fma_r(cj, s4[i + j * sy] - s4[i - (j + 1) * sy], e2);
            }
            if (iy >= j && iy < ny - j - 1) e1 = fma_r(cj, s2[i + (j + 1) * sy] - s2[i - j * sy], e1);
        }
        gv[0] = gv[0] + pml_corr<1, 3, 0>(S, iz, iy, ix, e0);
        gv[1] = gv[1] + pml_corr<1, 4, 1>(S, iz, iy, ix, e1);
        gv[2] = gv[2] + pml_corr<1, 5, 0>(S, iz, iy, ix, e2);
    }
    if (pml_in(ix, nx, L)) {
        R e0 = 0, e1 = 0, e2 = 0;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (j >= S.nc) break;
            const R cj = S.c[j];
            if (ix >= j && ix < nx - j - 1) e0 = fma_r(cj, s1[i + (j + 1)] - s1[i - j], e0);
            if (ix >= j + 1 && ix < nx - j) {
                e1 = fma_r(cj, s6[i + j] - s6[i - (j + 1)], e1);
                e2 = fma_r(cj, s5[i + j] - s5[i - (j + 1)], e2);
            }
        }
        gv[0] = gv[0] + pml_corr<2, 3, 1>(S, iz, iy, ix, e0);
        gv[1] = gv[1] + pml_corr<2, 4, 0>(S, iz, iy, ix, e1);
        gv[2] = gv[2] + pml_corr<2, 5, 0>(S, iz, iy, ix, e2);
    }
}

// The band cells of a grid enumerated as one index range (the pre-passes of
// the tiled kernels run one thread per band cell): first the 2L+2 z-band
// planes, then, plane by plane over the middle planes, the 2L+2 y-band rows
// and the x-band cells of the middle rows.  Every axis holds >= 2L+2 cells.
__host__ __device__ __forceinline__ long long pml_band_cells(int nz, int ny, int nx, int L)
{
    const long long W = 2 * (long long)L + 2;
    const long long ring = (long long)ny * nx - ((long long)ny - W) * ((long long)nx - W);
    return W * ny * nx + ((long long)nz - W) * ring;
}

__device__ __forceinline__ void pml_band_cell(int b, int nz, int ny, int nx, int L, int &iz,
                                              int &iy, int &ix)
{
    const int W = 2 * L + 2;
    const int plane = ny * nx;
    if (b < W * plane) {
        const int p = b / plane, r = b - p * plane;
        iz = p <= L ? p : p + (nz - W);
        iy = r / nx;
        ix = r - iy * nx;
        return;
    }
    b -= W * plane;
    const int ring = plane - (ny - W) * (nx - W);
    const int m = b / ring;
    int r = b - m * ring;
    iz = L + 1 + m;
    if (r < W * nx) {
        const int q = r / nx;
        iy = q <= L ? q : q + (ny - W);
        ix = r - q * nx;
    } else {
        r -= W * nx;
        const int q = r / W, c = r - q * W;
        iy = L + 1 + q;
        ix = c <= L ? c : c + (nx - W);
    }
}

// Workgroup barrier for LDS hand-offs only.  __syncthreads() also fences
// global memory, which makes every wave wait for its outstanding global
// loads (the prefetches of the next plane) at each barrier; the kernels
// exchange data between threads through LDS alone, so an LDS-scoped
// release/acquire around s_barrier is sufficient.
__device__ __forceinline__ void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Strain J moved onto position P at flat index i (sim/core/fdtd.py:612-642):
// per axis (z, y, x) with d = pos[P] - pos[J]: d > 0 taps {0,+1} and needs
// i <= n-2; d < 0 taps {-1,0} and needs i >= 1.  Taps are summed z-major,
// left to right, then scaled by 1/2 or 1/4; 0 if any guard fails.
template <int P, int J, typename R>
__device__ __forceinline__ R moved_direct(const R *__restrict__ gJ, int i, int iz, int iy,
                                          int ix, int nz, int ny, int nx, int sy, int sz)
{
    constexpr int src = kVP[J];
    if constexpr (src == P) {
        return gJ[i];
    } else {
        constexpr int dz = kPos[P][0] - kPos[src][0];
        constexpr int dy = kPos[P][1] - kPos[src][1];
        constexpr int dx = kPos[P][2] - kPos[src][2];
        bool ok = true;
        if constexpr (dz > 0) ok = ok && (iz <= nz - 2);
        if constexpr (dz < 0) ok = ok && (iz >= 1);
        if constexpr (dy > 0) ok = ok && (iy <= ny - 2);
        if constexpr (dy < 0) ok = ok && (iy >= 1);
        if constexpr (dx > 0) ok = ok && (ix <= nx - 2);
        if constexpr (dx < 0) ok = ok && (ix >= 1);
        if (!ok) return (R)0;
        constexpr int z0 = dz > 0 ? 0 : (dz < 0 ? -1 : 0), nzt = dz != 0 ? 2 : 1;
        constexpr int y0 = dy > 0 ? 0 : (dy < 0 ? -1 : 0), nyt = dy != 0 ? 2 : 1;
        constexpr int x0 = dx > 0 ? 0 : (dx < 0 ? -1 : 0), nxt = dx != 0 ? 2 : 1;
        R acc = (R)0;
        bool first = true;
#pragma unroll
        for (int a = 0; a < nzt; ++a)
#pragma unroll
            for (int b = 0; b < nyt; ++b)
#pragma unroll
                for (int c = 0; c < nxt; ++c) {
                    const R t = gJ[i + (z0 + a) * sz + (y0 + b) * sy + (x0 + c)];
                    acc = first ? t : acc + t;
                    first = false;
                }
        constexpr int ntap = nzt * nyt * nxt;
        return acc * (ntap == 4 ? (R)0.25 : (R)0.5);
    }
}

// One shot of a multi-shot batch (device table entry): the kernel view of the
// shot, its step count, per half-step tiling ([0] stress, [1] velocity) and
// first workgroup in the batched launch, and its source / receiver arrays.
template <typename R>
struct BatchEntry {
    Shot<R> S;
    int nt;
    // [0] stress, [1] velocity; [2] the CPML band pre-passes (blocks of
    // kPmlBlock band cells; 0 without a CPML)
    int tiles_x[2], tiles_y[2], zc_len[2], blk_off[3];
    int n_src, n_rec, mode;
    const int64_t *src_idx;
    const R *src_w;
    const double *wavelet;
    const int32_t *rec_ptr;
    const int64_t *rec_idx;
    const R *rec_w;
    R *traces;
};

// Launchers (defined in the kernel translation units).
template <typename R> hipError_t launch_stress_direct(const Shot<R> &s, hipStream_t st);
template <typename R> hipError_t launch_velocity_direct(const Shot<R> &s, hipStream_t st);
// Strains of s.v into the 6*ncell scratch s.g (first kernel of the direct stress half-step).
template <typename R> hipError_t launch_strain_direct(const Shot<R> &s, hipStream_t st);
// Densities up to this many materials are staged in LDS by the tiled
// velocity kernel (face-density lookups then wait on LDS only).
constexpr int kRhoLdsBytes = 16384;

// Tiled launchers: `variant` picks a tile shape (see tile_variants),
// `zchunk` the z planes per workgroup (0 = automatic).
// store_g: also write the strains of every cell into s.g (6 x ncell).
template <typename R>
hipError_t launch_stress_tiled(const Shot<R> &s, int variant, int zchunk, hipStream_t st,
                               bool store_g = false);
template <typename R>
hipError_t launch_velocity_tiled(const Shot<R> &s, int variant, int zchunk, hipStream_t st);
// Number of tile variants of a half-step (0 stress, 1 velocity) at a
// precision; fills (tx, ty, pad) of `variant` when it exists.
int tile_variants(int precision, int half, int variant, int *tx, int *ty, int *pad, int *fr);
// Tiling of one half-step of a grid; returns the workgroup count (-1: no
// such variant).
int tiled_geometry(int precision, int half, int variant, int zchunk, int nz, int ny, int nx,
                   int *tiles_x, int *tiles_y, int *zc_len);
// cp: some shot of the batch has a CPML (the kernel adds band corrections).
template <typename R>
hipError_t launch_stress_batched(int variant, const BatchEntry<R> *tab, const int32_t *blk, int nb,
                                 const int32_t *step, hipStream_t st, bool cp);
// max_pml: the widest CPML band of the batch (0: no CPML shot).
template <typename R>
hipError_t launch_velocity_batched(int variant, const BatchEntry<R> *tab, const int32_t *blk,
                                   int nb, const int32_t *step, hipStream_t st, int max_mat,
                                   int max_pml);
// CPML band pre-pass of a batch (strain corrections from v, before the
// stress kernel) over the band cells of every CPML shot; blk maps a block to
// its shot.
constexpr int kPmlBlock = 256;
template <typename R>
hipError_t launch_pml_corr_batched(int half, const BatchEntry<R> *tab, const int32_t *blk, int nb,
                                   const int32_t *step, hipStream_t st);
// Dense-table twin (kernels_direct.hip): half 0 = strains into s.g then the
// stress from the 36 planes Cflat; 1 = velocity from the face-density
// volumes rv; 2 = damping of all nine fields.
template <typename R>
hipError_t launch_dense_half(const Shot<R> &s, int half, const void *Cflat, uint64_t act,
                             const void *const rv[3], hipStream_t st);
// Source / record of every active shot (phase 0, after the stress
// half-step) and velocity-mode record plus step advance (phase 1, after the
// velocity half-step); one workgroup per shot.
template <typename R>
hipError_t launch_source_record_batched(const BatchEntry<R> *tab, int n_shot, const int32_t *step,
                                        int phase, hipStream_t st);
template <typename R>
hipError_t launch_source_record(const Shot<R> &s, const pe_shot *p, const double *wavelet,
                                int32_t n, bool do_source, R *traces, bool do_record,
                                hipStream_t st);

// Kernel-side view of a pe_adjoint, passed by value.
template <typename R>
struct Adj {
    R *lv[3];
    R *ls[6];
    R *work;
    R *wface;
    const R *proj;
    int hist_kind;
    int n_proj;
    int zbreak;
    const int32_t *run_ptr;
    double *run_acc;
    // CPML: adjoint memory variables (6 x band volume per axis) and the band
    // cells' derivative adjoints (3 x band volume per axis)
    R *psib[3];
    R *pw[3];
};

template <typename R>
inline Adj<R> make_adj(const pe_adjoint *a)
{
    Adj<R> A;
    for (int k = 0; k < 3; ++k) A.lv[k] = (R *)a->lv[k];
    for (int k = 0; k < 6; ++k) A.ls[k] = (R *)a->ls[k];
    A.work = (R *)a->work;
    A.wface = (R *)a->wface;
    A.hist_kind = a->hist_kind;
    A.proj = (const R *)a->proj;
    A.n_proj = a->n_proj;
    A.zbreak = a->zbreak;
    A.run_ptr = a->run_ptr;
    A.run_acc = a->run_acc;
    for (int k = 0; k < 3; ++k) {
        A.psib[k] = (R *)a->cpml_psib[k];
        A.pw[k] = (R *)a->cpml_work[k];
    }
    return A;
}

// Adjoint (kernels_adjoint.hip).
template <typename R> hipError_t store_velocity(const Shot<R> &s, R *dst, hipStream_t st);
template <typename R>
hipError_t adjoint_reset(const pe_shot *p, const pe_adjoint *a, hipStream_t st);
template <typename R>
hipError_t adjoint_run(const pe_shot *p, const pe_adjoint *a, int32_t n_hi, int32_t n_lo,
                       const void *vhist, const void *residual, double *gw, hipStream_t st);
hipError_t adjoint_finalize(const pe_shot *p, const pe_adjoint *a, double *grad, hipStream_t st);
// Tiled adjoint (kernels_adjoint_tiled.hip).  stage 0: D^T (damp_ls: apply
// the deferred stress-adjoint damping); 1: projected stiffness gradient on
// Sv (v = history slot); 2: M^T C^T then G^T.
template <typename R>
hipError_t adjoint_face_weights(const pe_shot *p, const pe_adjoint *a, hipStream_t st);
template <typename R>
hipError_t adjoint_step_tiled(const Shot<R> &S, const Shot<R> &Sv, const Adj<R> &A, int stage,
                              int damp_ls, hipStream_t st, const R *ghist = nullptr);
template <typename R>
hipError_t adjoint_scale_ls(const pe_shot *p, const pe_adjoint *a, hipStream_t st);

// Shot preparation (kernels_prep.hip).
hipError_t prep_rotate(const int32_t *src, int nx_s, int ny_s, int nz_s, double h_spec, double h,
                       double ca, double sa, int nd, int ndz, int m, int32_t *lab, hipStream_t st);
hipError_t prep_label_power(const double *seeds, const double *w, int n_seed, double h, double D,
                            int nd, int ndz, int m, int32_t *lab, hipStream_t st);
hipError_t prep_edt_damp(const int32_t *lab, int nz, int ny, int nx, double fluid_damp,
                         const float *pz, const float *py, const float *px, int32_t *s0,
                         int32_t *s1, float *damp, hipStream_t st);
template <typename R>
hipError_t prep_vmax(const int32_t *mats, int nz, int ny, int nx, const R *Ct, const R *rho,
                     double *partial, int nblocks, hipStream_t st);

}  // namespace pe
