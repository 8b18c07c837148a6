// Building blocks shared by the z-marching tiled kernels (forward and
// adjoint): workgroup remapping, the LDS strain ring and its moved-strain
// reads, the strain stencils of one column, and launch helpers.
#pragma once

#include "common.hpp"

namespace pe {
namespace tiled {

constexpr int kXcd = 8;

// Element i of a field addressed by a 32-bit unsigned byte offset from the
// (wave-uniform) field base: the access takes the base in SGPRs and one
// offset VGPR, with no 64-bit address arithmetic per lane.  Valid while the
// field is smaller than 4 GiB (pe_check_shot enforces it for the tiled
// kernels).
template <typename T>
__device__ __forceinline__ T ldg(const T *base, int i)
{
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) +
                                        (uint32_t)i * (uint32_t)sizeof(T));
}
template <typename T>
__device__ __forceinline__ void stg(T *base, int i, T v)
{
    *reinterpret_cast<T *>(reinterpret_cast<char *>(base) + (uint32_t)i * (uint32_t)sizeof(T)) = v;
}

// Store of element i of an n-element field through a raw buffer resource,
// issued by every lane: a lane with ok == false aims at byte n * sizeof(T),
// past the resource's range, and the hardware drops its write.  Unlike a
// store under a branch, the instruction is issued on every path, so the
// vector-memory operations a loop body issues after a given load are a fixed
// count and a wait for that load at the next step can leave the stores in
// flight.  The empty asm makes the base look modified, so the resource is
// built at each store instead of holding four scalar registers per field
// across the caller's loop.  Needs n * sizeof(T) < 4 GiB, like ldg / stg.
constexpr int kBufferFlags = 0x00020000;   // gfx9 raw buffer: 32-bit data format

template <typename T>
__device__ __forceinline__ void stb(T *base, uint32_t n, int i, bool ok, T v)
{
    // (the base is wave-uniform; readfirstlane states it where the compiler
    // cannot see it, e.g. a base read from a batch table)
    const uint64_t a = reinterpret_cast<uint64_t>(base);
    uint64_t u = ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(a >> 32)) << 32) |
                 (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)a);
    __asm__ volatile("" : "+s"(u));
    base = reinterpret_cast<T *>(u);
    const __amdgpu_buffer_rsrc_t r =
        __builtin_amdgcn_make_buffer_rsrc(base, (short)0, (int)(n * (uint32_t)sizeof(T)),
                                          kBufferFlags);
    const int off = (int)((ok ? (uint32_t)i : n) * (uint32_t)sizeof(T));
    if constexpr (sizeof(T) == 4) {
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(int, v), r, off, 0, 0);
    } else {
        typedef int int2v __attribute__((ext_vector_type(2)));
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(int2v, v), r, off, 0, 0);
    }
}

// Remap a workgroup id so consecutive logical tiles land on the same XCD.
__device__ __forceinline__ int xcd_remap(int bid, int nb)
{
    const int per = nb / kXcd;
    if (bid >= per * kXcd) return bid;
    return (bid % kXcd) * per + bid / kXcd;
}

template <typename R>
__device__ __forceinline__ R ld(const R *p, int i, bool ok)
{
    return ok ? p[i] : (R)0;
}

// ------------------------------------------------------------------ stress
// FR: the frame of each v plane is fetched in the same step as the z rings'
// plane (5 planes before it is used) and parked in a per-thread LDS ring of
// 4 planes; without FR it is fetched one step before use.
template <typename R, int TX, int TY, int PAD, bool FR = false>
struct StressTile {
    static constexpr int GX = TX + 2, GY = TY + 2, NG = GX * GY;   // strain columns
    static constexpr int NT = (NG + 63) / 64 * 64;                 // threads
    // v plane tile; PAD extra words per row shift consecutive tile rows
    // onto disjoint LDS banks
    static constexpr int WX = GX + 8, VX = WX + PAD, VY = GY + 8, NV = VX * VY;
    static constexpr int NFRAME = 8 * (WX + GY);                   // v tile outside G
    static constexpr int NF = (NFRAME + NT - 1) / NT;              // frame entries per thread
    static constexpr int NRING = FR ? 4 * NF * 3 * NT : 0;         // frame ring words
    static constexpr size_t LDS = sizeof(R) * (3 * NV + 6 * 3 * NT + NRING);
    // with a CPML (float): the strain corrections of two planes, [2][6][NT],
    // for the tiles where that stays within 96 KiB (a 102 KiB request of this
    // kernel was refused at launch); the others load them directly
    static constexpr bool CORR_LDS =
        sizeof(R) == 4 && LDS + sizeof(R) * 12 * NT <= 96 * 1024;
    static constexpr size_t LDS_CP = LDS + (CORR_LDS ? sizeof(R) * 12 * NT : 0);
};

// Vector-memory wait for every load of this wave, the LDS-DMA loads
// included (s_waitcnt vmcnt(0); expcnt and lgkmcnt left at their maxima).
__device__ __forceinline__ void wait_vmem() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// The same wait with the last N (<= 15) vector-memory operations of this
// wave left in flight: the counter retires in issue order, so everything
// issued before them has completed (s_waitcnt vmcnt(N)).
template <int N>
__device__ __forceinline__ void wait_vmem_le()
{
    static_assert(N >= 0 && N <= 15, "vmcnt field");
    __builtin_amdgcn_s_waitcnt(0x0F70 | N);
}

// Moved strain q_{P,J} read from the LDS strain ring (see moved_direct in
// kernels_direct.hip for the rule), laid out [6][NS][NG].  sl[d+1] is the
// ring slot of plane z+d.  G = false drops the edge guards (the caller
// guarantees they all hold).
template <bool G, int P, int J, typename R, int GX, int NG, int NS = 3>
__device__ __forceinline__ R moved_lds(const R *__restrict__ gt, const int (&sl)[3], int c,
                                       int iz, int iy, int ix, int nz, int ny, int nx)
{
    const R *gJ = gt + J * NS * NG;
    constexpr int src = kVP[J];
    if constexpr (src == P) {
        return gJ[sl[1] * NG + c];
    } else {
        constexpr int dz = kPos[P][0] - kPos[src][0];
        constexpr int dy = kPos[P][1] - kPos[src][1];
        constexpr int dx = kPos[P][2] - kPos[src][2];
        bool ok = true;
        if constexpr (G) {
            if constexpr (dz > 0) ok = ok && (iz <= nz - 2);
            if constexpr (dz < 0) ok = ok && (iz >= 1);
            if constexpr (dy > 0) ok = ok && (iy <= ny - 2);
            if constexpr (dy < 0) ok = ok && (iy >= 1);
            if constexpr (dx > 0) ok = ok && (ix <= nx - 2);
            if constexpr (dx < 0) ok = ok && (ix >= 1);
        }
        constexpr int z0 = dz < 0 ? -1 : 0, nzt = dz != 0 ? 2 : 1;
        constexpr int y0 = dy < 0 ? -1 : 0, nyt = dy != 0 ? 2 : 1;
        constexpr int x0 = dx < 0 ? -1 : 0, nxt = dx != 0 ? 2 : 1;
        R acc = (R)0;
#pragma unroll
        for (int a = 0; a < nzt; ++a)
#pragma unroll
            for (int b = 0; b < nyt; ++b)
#pragma unroll
                for (int e = 0; e < nxt; ++e) {
                    const R t = gJ[sl[1 + z0 + a] * NG + c + (y0 + b) * GX + (x0 + e)];
                    acc = (a == 0 && b == 0 && e == 0) ? t : acc + t;
                }
        constexpr int ntap = nzt * nyt * nxt;
        const R q = acc * (ntap == 4 ? (R)0.25 : (R)0.5);
        if constexpr (G) return ok ? q : (R)0;
        return q;
    }
}

// One stress component: acc = sum over active J (ascending) of C_IJ q_J, then
// s = fma(dt, acc, s * damp) (stress_lab_u, sim/core/fdtd.py:643-657).
// FULL = true: every C_IJ of the table is active (no per-term test).
// Six consecutive row entries starting at an 8-byte (float) / 16-byte
// (double) aligned offset: three paired loads per lane.
__device__ __forceinline__ void load6(const float *p, float (&c)[6])
{
    const float2 *q = reinterpret_cast<const float2 *>(p);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float2 t = q[k];
        c[2 * k] = t.x;
        c[2 * k + 1] = t.y;
    }
}
__device__ __forceinline__ void load6(const double *p, double (&c)[6])
{
    const double2 *q = reinterpret_cast<const double2 *>(p);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double2 t = q[k];
        c[2 * k] = t.x;
        c[2 * k + 1] = t.y;
    }
}
template <typename R>
__device__ __forceinline__ void load6(const __attribute__((address_space(4))) R *p, R (&c)[6])
{
#pragma unroll
    for (int k = 0; k < 6; ++k) c[k] = p[k];
}

template <int I, bool FULL, typename R, typename Row>
__device__ __forceinline__ void stress_out(const Shot<R> &S, Row cm, const R (&q)[6], R dmp, R &sv,
                                           uint64_t act)
{
    R row[6];
    load6(cm + 6 * I, row);
    R acc = 0;
    bool any = FULL;
#pragma unroll
    for (int J = 0; J < 6; ++J) {
        if (FULL || ((act >> (6 * I + J)) & 1ull)) {
            acc = fma_r(row[J], q[J], acc);
            any = true;
        }
    }
    sv = sv * dmp;
    if (any) sv = fma_r(S.dt, acc, sv);
}

template <bool G, int P, typename R, int GX, int NG, int NS = 3>
__device__ __forceinline__ void gather_lds(const R *__restrict__ gt, const int (&sl)[3], int c,
                                           int iz, int iy, int ix, int nz, int ny, int nx,
                                           R (&q)[6])
{
    q[0] = moved_lds<G, P, 0, R, GX, NG, NS>(gt, sl, c, iz, iy, ix, nz, ny, nx);
    q[1] = moved_lds<G, P, 1, R, GX, NG, NS>(gt, sl, c, iz, iy, ix, nz, ny, nx);
    q[2] = moved_lds<G, P, 2, R, GX, NG, NS>(gt, sl, c, iz, iy, ix, nz, ny, nx);
    q[3] = moved_lds<G, P, 3, R, GX, NG, NS>(gt, sl, c, iz, iy, ix, nz, ny, nx);
    q[4] = moved_lds<G, P, 4, R, GX, NG, NS>(gt, sl, c, iz, iy, ix, nz, ny, nx);
    q[5] = moved_lds<G, P, 5, R, GX, NG, NS>(gt, sl, c, iz, iy, ix, nz, ny, nx);
}

// Strains g1..g6 of plane zc at one column (strains_zp statement order,
// sim/core/fdtd.py:1136-1164): x / y taps from the LDS plane tile (pointers
// at the column), z taps from the register rings (r?[k] = plane zc-4+k).
// G = false drops the per-term edge guards.
template <bool G, typename R, int VX>
__device__ __forceinline__ void strains_col(const Shot<R> &S, const R *tx_, const R *ty_,
                                            const R *tz_, const R (&rx)[9], const R (&ry)[9],
                                            const R (&rz)[9], int ix, int iy, int zc, R (&a)[6])
{
    const int nz = S.nz, ny = S.ny, nx = S.nx;
    R a1 = 0, a2 = 0, a3 = 0, a4 = 0, a5 = 0, a6 = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        if (j >= S.nc) break;
        const R cj = S.c[j];
        if (!G || (ix >= j + 1 && ix < nx - j)) a1 = fma_r(cj, tx_[j] - tx_[-(j + 1)], a1);
        if (!G || (iy >= j + 1 && iy < ny - j)) a2 = fma_r(cj, ty_[j * VX] - ty_[-(j + 1) * VX], a2);
        if (!G || (zc >= j + 1 && zc < nz - j)) a3 = fma_r(cj, rz[4 + j] - rz[3 - j], a3);
        if (!G || (iy >= j && iy < ny - j - 1)) a4 = fma_r(cj, tz_[(j + 1) * VX] - tz_[-j * VX], a4);
        if (!G || (zc >= j && zc < nz - j - 1)) a4 = fma_r(cj, ry[5 + j] - ry[4 - j], a4);
        if (!G || (ix >= j && ix < nx - j - 1)) a5 = fma_r(cj, tz_[j + 1] - tz_[-j], a5);
        if (!G || (zc >= j && zc < nz - j - 1)) a5 = fma_r(cj, rx[5 + j] - rx[4 - j], a5);
        if (!G || (ix >= j && ix < nx - j - 1)) a6 = fma_r(cj, ty_[j + 1] - ty_[-j], a6);
        if (!G || (iy >= j && iy < ny - j - 1)) a6 = fma_r(cj, tx_[(j + 1) * VX] - tx_[-j * VX], a6);
    }
    a[0] = a1 * S.inv_h;
    a[1] = a2 * S.inv_h;
    a[2] = a3 * S.inv_h;
    a[3] = a4 * S.inv_h;
    a[4] = a5 * S.inv_h;
    a[5] = a6 * S.inv_h;
}

// CPML profiles of the band cells in LDS, bp[((A*2 + st)*3 + c)*W + bi]
// (axis A, stagger st, c = b, a, 1/kappa; W = 2L+2 band indices), copied at
// kernel start by every thread of the workgroup.
template <typename R>
__device__ __forceinline__ void pml_profiles_to_lds(const Shot<R> &S, R *bp, int nthreads)
{
    const int L = S.pml, W = 2 * L + 2;
    for (int e = threadIdx.x; e < 18 * W; e += nthreads) {
        const int bi = e % W, c = (e / W) % 3, st = (e / (3 * W)) % 2, a = e / (6 * W);
        const int n = a == 0 ? S.nz : (a == 1 ? S.ny : S.nx);
        const int i = bi <= L ? bi : bi + (n - W);
        // select the axis without indexing the kernel argument (which would
        // move it to scratch memory)
        const R *pa = a == 0 ? S.prof[0] : (a == 1 ? S.prof[1] : S.prof[2]);
        bp[e] = pa[(st * 3 + c) * n + i];
    }
}

// pml_corr with the profile from the LDS table and the memory variable in a
// register (advanced in place; the caller stores it).
template <int A, int ST, typename R>
__device__ __forceinline__ R pml_corr_lds(const R *bp, int W, int bi, R &psi, R dd)
{
    const R *pr = bp + ((A * 2 + ST) * 3) * W + bi;
    const R p = fma_r(pr[0], psi, pr[W] * dd);
    psi = p;
    return fma_r(dd, pr[2 * W] - (R)1, p);
}

// Stiffness rows are read in place.  A wave whose active lanes share one
// material (the common case: grains span thousands of cells) reads its row
// through a wave-uniform pointer in the constant address space, which the
// compiler serves with scalar loads.
template <typename R>
using ConstRow = const __attribute__((address_space(4))) R *;

// ``act`` is the shot's active-entry mask (S.act), passed in so that a caller
// can hand over a copy the compiler does not hoist (see stress_tiled_body).
template <bool G, bool FULL, typename R, int GX, int NG, typename Row, int NS = 3>
__device__ __forceinline__ void stress_cell(const Shot<R> &S, const R *__restrict__ gt,
                                            const int (&sl)[3], int c, int iz, int iy, int ix,
                                            Row cm, R dmp, R (&sv)[6], uint64_t act)
{
    const int nz = S.nz, ny = S.ny, nx = S.nx;
    R q[6];
    gather_lds<G, 0, R, GX, NG, NS>(gt, sl, c, iz, iy, ix, nz, ny, nx, q);
    stress_out<0, FULL>(S, cm, q, dmp, sv[0], act);
    stress_out<1, FULL>(S, cm, q, dmp, sv[1], act);
    stress_out<2, FULL>(S, cm, q, dmp, sv[2], act);
    gather_lds<G, 1, R, GX, NG, NS>(gt, sl, c, iz, iy, ix, nz, ny, nx, q);
    stress_out<3, FULL>(S, cm, q, dmp, sv[3], act);
    gather_lds<G, 2, R, GX, NG, NS>(gt, sl, c, iz, iy, ix, nz, ny, nx, q);
    stress_out<4, FULL>(S, cm, q, dmp, sv[4], act);
    gather_lds<G, 3, R, GX, NG, NS>(gt, sl, c, iz, iy, ix, nz, ny, nx, q);
    stress_out<5, FULL>(S, cm, q, dmp, sv[5], act);
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Global loads are branch-free: every address is clamped into the grid and
// the values a clamped address returns are never used (the edge guards of
// the generic path exclude them; the fast path only runs where none occur).
// Waves per SIMD the register allocator must leave room for so that two
// workgroups of NT threads fit on a CU (float; one workgroup from 1024
// threads up); double keeps the compiler's choice (its larger register
// footprint would spill).
template <typename R, int NT> struct MinWaves { static constexpr int v = 1; };
template <int NT> struct MinWaves<float, NT> {
    static constexpr int v = NT >= 1024 ? NT / 256 : (2 * (NT / 64) + 3) / 4;
};

// ------------------------------------------------------------- launchers
// z planes per workgroup when the caller does not fix it: about a dozen
// workgroups per CU (measured best on 400x400x200), but no chunk shorter than
// 16 planes (each chunk refills its 9-plane rings).
inline int chunk_len(int nz, int tiles, int zchunk, int min_len = 16)
{
    if (zchunk > 0) return zchunk > nz ? nz : zchunk;
    const int target = 256 * 12;
    int chunks = (target + tiles - 1) / tiles;
    chunks = chunks < 1 ? 1 : chunks;
    int len = (nz + chunks - 1) / chunks;
    if (len < min_len) len = min_len;
    return len > nz ? nz : len;
}

template <typename K>
hipError_t allow_lds(K kernel, size_t lds)
{
    if (lds <= 65536) return hipSuccess;
    return hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
}


}  // namespace tiled
}  // namespace pe
