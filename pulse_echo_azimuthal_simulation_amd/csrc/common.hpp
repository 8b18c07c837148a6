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
    int nz, ny, nx, nc;
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
};

template <typename R>
inline Shot<R> make_shot(const pe_shot *p)
{
    Shot<R> s;
    s.nz = p->nz;
    s.ny = p->ny;
    s.nx = p->nx;
    s.nc = p->nc;
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
    return s;
}

__device__ __forceinline__ float fma_r(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fma_r(double a, double b, double c) { return __builtin_fma(a, b, c); }

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

// Launchers (defined in the kernel translation units).
template <typename R> hipError_t launch_stress_direct(const Shot<R> &s, hipStream_t st);
template <typename R> hipError_t launch_velocity_direct(const Shot<R> &s, hipStream_t st);
// Tiled launchers: `variant` picks a tile shape (see tile_variants),
// `zchunk` the z planes per workgroup (0 = automatic).
template <typename R>
hipError_t launch_stress_tiled(const Shot<R> &s, int variant, int zchunk, hipStream_t st);
template <typename R>
hipError_t launch_velocity_tiled(const Shot<R> &s, int variant, int zchunk, hipStream_t st);
// Number of tile variants of a half-step (0 stress, 1 velocity) at a
// precision; fills (tx, ty, pad) of `variant` when it exists.
int tile_variants(int precision, int half, int variant, int *tx, int *ty, int *pad);
template <typename R>
hipError_t launch_source_record(const Shot<R> &s, const pe_shot *p, const double *wavelet,
                                int32_t n, bool do_source, R *traces, bool do_record,
                                hipStream_t st);

}  // namespace pe
