// Tiled z-marching kernels for gfx950 (PE_KERNEL_TILED).
//
// Each workgroup owns a TX x TY tile of (x, y) columns and marches z over a
// chunk of planes.  Every field byte the step needs crosses HBM once:
//  * z taps come from per-thread register rings (9 planes deep);
//  * x / y taps come from an LDS copy of the current plane with a 4-cell
//    frame (5 for the stress kernel), loaded once per plane;
//  * the stress kernel computes the strains of plane z+1 on the fly for the
//    tile plus a one-cell ring of columns and keeps planes z-1, z, z+1 of
//    them in LDS, so the moved (averaged) strains never touch HBM.  This
//    replaces the upstream strain scratch volume and its slab loop
//    (sim/core/fdtd.py:1102-1177, 1763-1797).
//
// The arithmetic of every cell is statement-for-statement that of the direct
// kernels (kernels_direct.hip), hence of the upstream kernels: results are
// bit-identical to PE_KERNEL_DIRECT.
//
// Launch: a 1-D grid of tiles x tiles_y x chunks workgroups.  Hardware hands
// consecutive workgroup ids to the 8 XCDs round robin; the id is remapped so
// that each XCD receives a contiguous run of neighbouring tiles and their
// shared halo planes stay in that XCD's L2.
#include "tiled_common.hpp"

namespace pe {

namespace {

using namespace tiled;

// Body of the stress kernel for logical tile t (tile x fastest, then tile
// y, then z-chunk) of shot S.  STORE_G: also write the strains g1..g6 of the
// tile's own columns into S.g (6 x ncell, the history the adjoint's
// stiffness gradient reads), each plane once.  FR: see StressTile.  CP:
// the shot may have a CPML; the strains of its band cells (tile, ring
// columns and chunk-edge planes alike) then get the corrections the band
// pre-pass wrote into S.corr, loaded after the step's first barrier (loaded
// with the step's prefetches they would not fit the register budget).
template <typename R, int TX, int TY, int PAD, bool STORE_G = false, bool FR = false,
          bool CP = false>
__device__ __forceinline__ void stress_tiled_body(const Shot<R> &S, int tiles_x, int tiles_y,
                                                  int zc_len, int t, int zlo, int zhi)
{
    using T = StressTile<R, TX, TY, PAD, FR>;
    constexpr int GX = T::GX, GY = T::GY, NG = T::NG, VX = T::VX, VY = T::VY, NV = T::NV;
    constexpr int NT = T::NT, NF = T::NF, WX = T::WX;
    extern __shared__ unsigned char lds_raw[];
    R *vt = reinterpret_cast<R *>(lds_raw);       // [3][VY][VX]
    R *gt = vt + 3 * NV;                           // [6][3][NT]: strains by thread
    R *fr = gt + 6 * 3 * NT;                       // FR: [4][NF][3][NT] frame ring
    R *cb = fr + T::NRING;                         // CP, CORR_LDS: [2][6][NT] corrections

    const int nz = S.nz, ny = S.ny, nx = S.nx;
    const int sy = nx, sz = nx * ny;
    const int tx = t % tiles_x, ty = (t / tiles_x) % tiles_y, chunk = t / (tiles_x * tiles_y);
    const int x0 = tx * TX, y0 = ty * TY;
    const int z0 = zlo + chunk * zc_len, z1 = min(zhi, z0 + zc_len);
    // every strain column at least 4 cells (every output 1 cell) from the x/y edges
    const bool inner_xy = x0 >= 5 && x0 + TX <= nx - 5 && y0 >= 5 && y0 + TY <= ny - 5;
    const bool full = S.act == 0xFFFFFFFFFull;
    const bool has_damp = S.damp != nullptr;

    const int tid = threadIdx.x;
    const bool gact = tid < NG;
    const int gx = gact ? tid % GX : 0, gy = gact ? tid / GX : 0;
    const int ix = x0 - 1 + gx, iy = y0 - 1 + gy;                  // column of this thread
    const bool col_in = gact && ix >= 0 && ix < nx && iy >= 0 && iy < ny;
    const int L = CP ? S.pml : 0;
    const bool band_xy = L > 0 && col_in && (pml_in(ix, nx, L) || pml_in(iy, ny, L));
    const int mask_xy = L > 0 ? (pml_in(iy, ny, L) ? kPmlMaskY : 0) |
                                    (pml_in(ix, nx, L) ? kPmlMaskX : 0) : 0;
    const bool core = col_in && gx >= 1 && gx <= TX && gy >= 1 && gy <= TY;
    const int col = clampi(iy, 0, ny - 1) * sy + clampi(ix, 0, nx - 1);
    // LDS slot of this column in each v plane tile; idle threads write the
    // far corner cell, which no stencil tap reads
    const int vc = gact ? (gy + 4) * VX + (gx + 4) : NV - 1;

    const R *__restrict__ vx = S.v[0];
    const R *__restrict__ vy = S.v[1];
    const R *__restrict__ vz = S.v[2];

    // this thread's entries of the v-tile frame (the cells outside the
    // strain columns): LDS slot (the unread far corner if none) and clamped
    // x-y offset
    int f_lds[NF], f_xy[NF];
#pragma unroll
    for (int k = 0; k < NF; ++k) {
        const int e = tid + k * NT;
        int r = 0, c = 0;
        if (e < 4 * WX) {
            r = e / WX;
            c = e % WX;
        } else if (e < 8 * WX) {
            r = VY - 4 + (e - 4 * WX) / WX;
            c = (e - 4 * WX) % WX;
        } else if (e < 8 * WX + 4 * GY) {
            r = 4 + (e - 8 * WX) / 4;
            c = (e - 8 * WX) % 4;
        } else {
            r = 4 + (e - 8 * WX - 4 * GY) / 4;
            c = GX + 4 + (e - 8 * WX - 4 * GY) % 4;
        }
        f_lds[k] = e < T::NFRAME ? r * VX + c : NV - 1;
        f_xy[k] = clampi(y0 - 5 + r, 0, ny - 1) * sy + clampi(x0 - 5 + c, 0, nx - 1);
    }

    // z rings: at the top of iteration zc, r?[k] holds plane zc - 5 + k and
    // p? holds plane zc + 4 (loaded one iteration earlier); the shift at the
    // top makes r?[k] = plane zc - 4 + k.
    // Frame: without FR, fq holds plane zc of this thread's frame entries
    // (fetched one step ahead).  With FR, fq holds plane zc + 4, fetched in
    // the same step as the ring's plane zc + 4 and so together with the
    // neighbouring tiles' ring loads of that plane (an L2 hit for one of
    // them); at the top of step zc it goes to ring slot zc & 3 right after
    // plane zc is read from that slot.  Each thread reads back only what it
    // wrote itself, so the ring needs no barrier.
    constexpr int FD = 1;
    R rx[9], ry[9], rz[9], px, py, pz, fq[FD][NF][3];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        const int i = clampi(z0 - 6 + k, 0, nz - 1) * sz + col;
        rx[k] = ldg(vx, i);
        ry[k] = ldg(vy, i);
        rz[k] = ldg(vz, i);
    }
    {
        const int i = clampi(z0 + 3, 0, nz - 1) * sz + col;
        px = ldg(vx, i);
        py = ldg(vy, i);
        pz = ldg(vz, i);
    }
    // output-plane operands of the next step: s, label, damping (only the
    // tile's own columns write a stress)
    R ns[6] = {0, 0, 0, 0, 0, 0}, nd = 1;
    int nm = 0;
    if (core) {
        const int io1 = clampi(z0 - 2, 0, nz - 1) * sz + col;
#pragma unroll
        for (int k = 0; k < 6; ++k) ns[k] = ldg(S.s[k], io1);
        nm = ldg(S.mats, io1);
        nd = has_damp ? ldg(S.damp, io1) : (R)1;
    }
    if constexpr (FR) {
        // planes z0-1 .. z0+2 into their ring slots, plane z0+3 into fq
#pragma unroll
        for (int d = 0; d < 5; ++d)
#pragma unroll
            for (int k = 0; k < NF; ++k) {
                const int i = clampi(z0 - 1 + d, 0, nz - 1) * sz + f_xy[k];
                const R a = ldg(vx, i), b = ldg(vy, i), c = ldg(vz, i);
                if (d < 4) {
                    R *sl = fr + ((((z0 - 1 + d) & 3) * NF + k) * 3) * NT + tid;
                    sl[0] = a;
                    sl[NT] = b;
                    sl[2 * NT] = c;
                } else {
                    fq[0][k][0] = a;
                    fq[0][k][1] = b;
                    fq[0][k][2] = c;
                }
            }
    } else {
#pragma unroll
        for (int d = 0; d < FD; ++d)
#pragma unroll
            for (int k = 0; k < NF; ++k) {
                const int i = clampi(z0 - 1 + d, 0, nz - 1) * sz + f_xy[k];
                fq[d][k][0] = ldg(vx, i);
                fq[d][k][1] = ldg(vy, i);
                fq[d][k][2] = ldg(vz, i);
            }
    }

    // CPML (float): the strain corrections of plane zp, which the band
    // pre-pass wrote into S.corr, fetched one step ahead straight into LDS
    // (slot [zp & 1][k][tid]; LDS-DMA loads take no registers) for the
    // components any lane of the wave needs.  Each lane reads back only its
    // own slots, after the wait at the top of the next step (and the wait
    // for in-flight LDS-DMA loads that the compiler puts before a barrier's
    // release fence).  Buffer loads rather than global_load_lds: a FLAT-
    // encoded load that writes LDS counts as both a vector-memory and an
    // LDS access, and the compiler then waits for both counters to drain
    // (vmcnt(0) lgkmcnt(0)) at the first dependency that follows it.
    constexpr bool CPD = CP && T::CORR_LDS;
    auto corr_dma = [&](int zp) {
        if constexpr (CPD) {
            if (L <= 0 || zp < 0 || zp >= nz) return;
            const bool zb = pml_in(zp, nz, L);
            const int lm = col_in && (band_xy || zb) ? mask_xy | (zb ? kPmlMaskZ : 0) : 0;
            const uint32_t n = (uint32_t)nz * (uint32_t)sz;
            const int off = (int)(((uint32_t)zp * (uint32_t)sz + (uint32_t)col) * 4u);
            R *dst = cb + (zp & 1) * 6 * NT + (tid & ~63);
#pragma unroll
            for (int k = 0; k < 6; ++k)
                if (__any(lm >> k & 1)) {
                    const __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc(
                        S.corr + (size_t)k * n, (short)0, (int)(n * 4u), kBufferFlags);
                    __builtin_amdgcn_raw_ptr_buffer_load_lds(
                        r, (__attribute__((address_space(3))) void *)(dst + k * NT), 4, off, 0, 0,
                        0);
                }
        }
    };
    corr_dma(z0 - 1);
    // the prologue's loads all complete before the loop: at the loop head the
    // compiler then knows that only the last step's stores can be in flight
    wait_vmem();

    for (int zc = z0 - 1; zc <= z1; ++zc) {
        const bool plane_in = zc >= 0 && zc < nz;
        // this plane's corrections are in LDS: everything but the previous
        // step's six output stores (issued after the corrections' DMA on
        // every path) has completed
        if constexpr (CPD) wait_vmem_le<6>();
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            rx[k] = rx[k + 1];
            ry[k] = ry[k + 1];
            rz[k] = rz[k + 1];
        }
        rx[8] = px;
        ry[8] = py;
        rz[8] = pz;
        // (1) plane zc of v into LDS: own column from the ring, frame from
        //     the delay line
        vt[0 * NV + vc] = rx[4];
        vt[1 * NV + vc] = ry[4];
        vt[2 * NV + vc] = rz[4];
        if constexpr (FR) {
#pragma unroll
            for (int k = 0; k < NF; ++k) {
                R *sl = fr + (((zc & 3) * NF + k) * 3) * NT + tid;
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    vt[c * NV + f_lds[k]] = sl[c * NT];
                    sl[c * NT] = fq[0][k][c];
                }
            }
        } else {
#pragma unroll
            for (int k = 0; k < NF; ++k) {
                vt[0 * NV + f_lds[k]] = fq[0][k][0];
                vt[1 * NV + f_lds[k]] = fq[0][k][1];
                vt[2 * NV + f_lds[k]] = fq[0][k][2];
            }
        }
#pragma unroll
        for (int d = 0; d + 1 < FD; ++d)
#pragma unroll
            for (int k = 0; k < NF; ++k) {
                fq[d][k][0] = fq[d + 1][k][0];
                fq[d][k][1] = fq[d + 1][k][1];
                fq[d][k][2] = fq[d + 1][k][2];
            }
        // this step's output-plane operands (loaded one step ahead); then the
        // loads for later steps in the order they are needed: the next
        // step's operands, the frame (plane zc+1, or zc+5 with FR), the
        // ring's plane zc+5
        const int zo = zc - 1;
        const bool out = core && zo >= z0 && zo < z1;
        const int io = clampi(zo, 0, nz - 1) * sz + col;
        R sv[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) sv[k] = ns[k];
        const int m = nm;
        const R dmp = nd;
        if (core) {
            const int io1 = clampi(zo + 1, 0, nz - 1) * sz + col;
#pragma unroll
            for (int k = 0; k < 6; ++k) ns[k] = ldg(S.s[k], io1);
            nm = ldg(S.mats, io1);
            nd = has_damp ? ldg(S.damp, io1) : (R)1;
        }
        {
            const int zf = clampi(zc + (FR ? 5 : FD), 0, nz - 1) * sz;
#pragma unroll
            for (int k = 0; k < NF; ++k) {
                fq[FD - 1][k][0] = ldg(vx, zf + f_xy[k]);
                fq[FD - 1][k][1] = ldg(vy, zf + f_xy[k]);
                fq[FD - 1][k][2] = ldg(vz, zf + f_xy[k]);
            }
        }
        {
            const int i = clampi(zc + 5, 0, nz - 1) * sz + col;
            px = ldg(vx, i);
            py = ldg(vy, i);
            pz = ldg(vz, i);
        }
        lds_barrier();

        // The guarded paths' column and table inputs, re-read every step:
        // left loop-invariant, the compiler hoists each per-lane edge mask
        // and table flag they derive out of the loop, and the ~140 scalar
        // registers that takes spill into vector-register lanes, whose
        // reloads then cost the fast path dozens of instructions a step.
        int gtid = tid;
        uint64_t gact = S.act;
        __asm__ volatile("" : "+v"(gtid), "+s"(gact));
        const int gix = x0 - 1 + gtid % GX, giy = y0 - 1 + gtid / GX;
        // (2) strains of plane zc at every column of the tile and its ring
        const int slot = (zc + 3) % 3;
        {
            R a[6] = {0, 0, 0, 0, 0, 0};
            if (inner_xy && zc >= 4 && zc <= nz - 5)
                strains_col<false, R, VX>(S, vt + vc, vt + NV + vc, vt + 2 * NV + vc, rx, ry, rz,
                                          ix, iy, zc, a);
            else if (col_in && plane_in)
                strains_col<true, R, VX>(S, vt + vc, vt + NV + vc, vt + 2 * NV + vc, rx, ry, rz,
                                         gix, giy, zc, a);
            // CPML band cell: this plane's strain corrections (the components
            // its bands correct)
            if (CP && L > 0 && col_in && plane_in && (band_xy || pml_in(zc, nz, L))) {
                const int mask = mask_xy | (pml_in(zc, nz, L) ? kPmlMaskZ : 0);
                if constexpr (CPD) {
                    const R *cs = cb + (zc & 1) * 6 * NT + tid;
#pragma unroll
                    for (int k = 0; k < 6; ++k)
                        if (mask >> k & 1) a[k] = a[k] + cs[k * NT];
                } else {
                    const size_t n = (size_t)nz * sz;
                    const int ig = zc * sz + col;
#pragma unroll
                    for (int k = 0; k < 6; ++k)
                        if (mask >> k & 1) a[k] = a[k] + ldg(S.corr + k * n, ig);
                }
            }
            R *gs = gt + slot * NT + tid;
#pragma unroll
            for (int k = 0; k < 6; ++k) gs[k * 3 * NT] = a[k];
            if constexpr (STORE_G) {
                if (core && zc >= z0 && zc < z1) {
                    // component bases are wave-uniform; the offset within one
                    // component fits 32 bits like every field
                    const size_t n = (size_t)nz * sz;
                    const int ig = zc * sz + col;
#pragma unroll
                    for (int k = 0; k < 6; ++k) stg(S.g + k * n, ig, a[k]);
                }
            }
        }
        lds_barrier();
        // the next plane's corrections: issued here, after this step's
        // reads of the other slot, so that the release fence of the next
        // barrier -- which waits for every LDS-DMA load still in flight --
        // finds them issued a stress stage and a loop latch earlier (issued
        // before the first barrier, they made it wait for the step's
        // prefetches too)
        if (zc < z1) corr_dma(zc + 1);

        // (3) stress of plane zc - 1 from the strain planes zc-2, zc-1, zc
        if (out) {
            const int sl[3] = {(zc + 1) % 3, (zc + 2) % 3, slot};
            const bool fast = full && inner_xy && zo >= 1 && zo <= nz - 2;   // wave-uniform
            const int mu = __builtin_amdgcn_readfirstlane(m);
            if (fast && __all(m == mu)) {
                // every lane of the wave in one grain: wave-uniform row
                ConstRow<R> cm = (ConstRow<R>)(S.Ct + (size_t)mu * 36);
                stress_cell<false, true, R, GX, NT>(S, gt, sl, tid, zo, iy, ix, cm, dmp, sv,
                                                    gact);
            } else if (fast) {
                stress_cell<false, true, R, GX, NT>(S, gt, sl, tid, zo, iy, ix,
                                                    S.Ct + (size_t)m * 36, dmp, sv, gact);
            } else {
                stress_cell<true, false, R, GX, NT>(S, gt, sl, tid, zo, giy, gix,
                                                    S.Ct + (size_t)m * 36, dmp, sv, gact);
            }
        }
        // With the CPML's correction DMA every lane issues the six stores
        // (stb drops those without an output): the loop body then issues a
        // fixed number of vector-memory operations after the DMA, and the
        // wait at the top of the next step leaves the stores in flight.
        // Without it the stores stay under the branch (measured: the buffer
        // resources cost the plain kernel more than the wait they save).
        if constexpr (CPD) {
            const uint32_t n = (uint32_t)nz * (uint32_t)sz;
#pragma unroll
            for (int k = 0; k < 6; ++k) stb(S.s[k], n, io, out, sv[k]);
        } else if (out) {
#pragma unroll
            for (int k = 0; k < 6; ++k) stg(S.s[k], io, sv[k]);
        }
    }
}

// Stress body with one barrier per step ("pipelined"): three stages of
// different planes run in the same step, each reading only what an earlier
// step wrote.  Step zc writes the v tile of plane zc (double-buffered),
// computes the strains of plane zc-1 from the tile written one step earlier
// into a 4-plane strain ring, and the stress of plane zc-3 from the strain
// planes zc-4 .. zc-2.  Same arithmetic as stress_tiled_body, bit for bit;
// plain tiles (no frame ring, no strain store, no CPML).  LDS:
// [2][3][NV] v tiles + [6][4][NT] strains.
template <typename R, int TX, int TY>
struct PipeTile {
    using T = StressTile<R, TX, TY, 0>;
    static constexpr size_t LDS = sizeof(R) * (2 * 3 * T::NV + 6 * 4 * T::NT);
};

template <typename R, int TX, int TY>
__device__ __forceinline__ void stress_pipe_body(const Shot<R> &S, int tiles_x, int tiles_y,
                                                 int zc_len, int t, int zlo, int zhi)
{
    using T = StressTile<R, TX, TY, 0>;
    constexpr int GX = T::GX, GY = T::GY, NG = T::NG, VX = T::VX, VY = T::VY, NV = T::NV;
    constexpr int NT = T::NT, NF = T::NF, WX = T::WX;
    extern __shared__ unsigned char lds_raw[];
    R *vt0 = reinterpret_cast<R *>(lds_raw);      // [2][3][VY][VX]
    R *gt = vt0 + 2 * 3 * NV;                     // [6][4][NT]: strains by thread

    const int nz = S.nz, ny = S.ny, nx = S.nx;
    const int sy = nx, sz = nx * ny;
    const int tx = t % tiles_x, ty = (t / tiles_x) % tiles_y, chunk = t / (tiles_x * tiles_y);
    const int x0 = tx * TX, y0 = ty * TY;
    const int z0 = zlo + chunk * zc_len, z1 = min(zhi, z0 + zc_len);
    const bool inner_xy = x0 >= 5 && x0 + TX <= nx - 5 && y0 >= 5 && y0 + TY <= ny - 5;
    const bool full = S.act == 0xFFFFFFFFFull;
    const bool has_damp = S.damp != nullptr;

    const int tid = threadIdx.x;
    const bool gact = tid < NG;
    const int gx = gact ? tid % GX : 0, gy = gact ? tid / GX : 0;
    const int ix = x0 - 1 + gx, iy = y0 - 1 + gy;
    const bool col_in = gact && ix >= 0 && ix < nx && iy >= 0 && iy < ny;
    const bool core = col_in && gx >= 1 && gx <= TX && gy >= 1 && gy <= TY;
    const int col = clampi(iy, 0, ny - 1) * sy + clampi(ix, 0, nx - 1);
    const int vc = gact ? (gy + 4) * VX + (gx + 4) : NV - 1;

    const R *__restrict__ vx = S.v[0];
    const R *__restrict__ vy = S.v[1];
    const R *__restrict__ vz = S.v[2];

    int f_lds[NF], f_xy[NF];
#pragma unroll
    for (int k = 0; k < NF; ++k) {
        const int e = tid + k * NT;
        int r = 0, c = 0;
        if (e < 4 * WX) {
            r = e / WX;
            c = e % WX;
        } else if (e < 8 * WX) {
            r = VY - 4 + (e - 4 * WX) / WX;
            c = (e - 4 * WX) % WX;
        } else if (e < 8 * WX + 4 * GY) {
            r = 4 + (e - 8 * WX) / 4;
            c = (e - 8 * WX) % 4;
        } else {
            r = 4 + (e - 8 * WX - 4 * GY) / 4;
            c = GX + 4 + (e - 8 * WX - 4 * GY) % 4;
        }
        f_lds[k] = e < T::NFRAME ? r * VX + c : NV - 1;
        f_xy[k] = clampi(y0 - 5 + r, 0, ny - 1) * sy + clampi(x0 - 5 + c, 0, nx - 1);
    }

    // z rings: after the shift at the top of step zc, r?[k] = plane zc-5+k
    // (strains of plane zc-1 use r?[0..8]; the tile of plane zc takes
    // r?[5]); p? = plane zc+3 before the shift.  fq = frame of plane zc.
    // ns/nm/nd = output operands of plane zc-3 (loaded one step ahead).
    R rx[9], ry[9], rz[9], px, py, pz, fq[NF][3];
    const int zs = z0 - 1;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        const int i = clampi(zs - 6 + k, 0, nz - 1) * sz + col;
        rx[k] = ldg(vx, i);
        ry[k] = ldg(vy, i);
        rz[k] = ldg(vz, i);
    }
    {
        const int i = clampi(zs + 3, 0, nz - 1) * sz + col;
        px = ldg(vx, i);
        py = ldg(vy, i);
        pz = ldg(vz, i);
    }
#pragma unroll
    for (int k = 0; k < NF; ++k) {
        const int i = clampi(zs, 0, nz - 1) * sz + f_xy[k];
        fq[k][0] = ldg(vx, i);
        fq[k][1] = ldg(vy, i);
        fq[k][2] = ldg(vz, i);
    }
    R ns[6] = {0, 0, 0, 0, 0, 0}, nd = 1;
    int nm = 0;

    for (int zc = zs; zc <= z1 + 2; ++zc) {
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            rx[k] = rx[k + 1];
            ry[k] = ry[k + 1];
            rz[k] = rz[k + 1];
        }
        rx[8] = px;
        ry[8] = py;
        rz[8] = pz;
        // (1) v tile of plane zc (planes zs .. z1)
        if (zc <= z1) {
            R *vt = vt0 + (zc & 1) * 3 * NV;
            vt[0 * NV + vc] = rx[5];
            vt[1 * NV + vc] = ry[5];
            vt[2 * NV + vc] = rz[5];
#pragma unroll
            for (int k = 0; k < NF; ++k) {
                vt[0 * NV + f_lds[k]] = fq[k][0];
                vt[1 * NV + f_lds[k]] = fq[k][1];
                vt[2 * NV + f_lds[k]] = fq[k][2];
            }
        }
        // this step's output operands (plane zc-3), then the loads for the
        // next step: operands of plane zc-2, frame of plane zc+1, ring plane
        // zc+4
        const int zo = zc - 3;
        const bool out = core && zo >= z0 && zo < z1;
        R sv[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) sv[k] = ns[k];
        const int m = nm;
        const R dmp = nd;
        if (core) {
            const int io1 = clampi(zo + 1, 0, nz - 1) * sz + col;
#pragma unroll
            for (int k = 0; k < 6; ++k) ns[k] = ldg(S.s[k], io1);
            nm = ldg(S.mats, io1);
            nd = has_damp ? ldg(S.damp, io1) : (R)1;
        }
        {
            const int zf = clampi(zc + 1, 0, nz - 1) * sz;
#pragma unroll
            for (int k = 0; k < NF; ++k) {
                fq[k][0] = ldg(vx, zf + f_xy[k]);
                fq[k][1] = ldg(vy, zf + f_xy[k]);
                fq[k][2] = ldg(vz, zf + f_xy[k]);
            }
        }
        {
            const int i = clampi(zc + 4, 0, nz - 1) * sz + col;
            px = ldg(vx, i);
            py = ldg(vy, i);
            pz = ldg(vz, i);
        }

        // (2) strains of plane q = zc-1 (planes zs .. z1) from the tile of
        //     the previous step
        const int q = zc - 1;
        if (q >= zs && q <= z1) {
            const R *vt = vt0 + (q & 1) * 3 * NV;
            const bool plane_in = q >= 0 && q < nz;
            R a[6] = {0, 0, 0, 0, 0, 0};
            if (inner_xy && q >= 4 && q <= nz - 5)
                strains_col<false, R, VX>(S, vt + vc, vt + NV + vc, vt + 2 * NV + vc, rx, ry, rz,
                                          ix, iy, q, a);
            else if (col_in && plane_in)
                strains_col<true, R, VX>(S, vt + vc, vt + NV + vc, vt + 2 * NV + vc, rx, ry, rz,
                                         ix, iy, q, a);
            R *gs = gt + (q & 3) * NT + tid;
#pragma unroll
            for (int k = 0; k < 6; ++k) gs[k * 4 * NT] = a[k];
        }

        // (3) stress of plane zo = zc-3 from the strain planes zo-1 .. zo+1,
        //     all written by earlier steps
        if (out) {
            const int sl[3] = {(zo - 1) & 3, zo & 3, (zo + 1) & 3};
            const int io = clampi(zo, 0, nz - 1) * sz + col;
            const bool fast = full && inner_xy && zo >= 1 && zo <= nz - 2;
            const int mu = __builtin_amdgcn_readfirstlane(m);
            if (fast && __all(m == mu)) {
                ConstRow<R> cm = (ConstRow<R>)(S.Ct + (size_t)mu * 36);
                stress_cell<false, true, R, GX, NT, ConstRow<R>, 4>(S, gt, sl, tid, zo, iy, ix, cm,
                                                                     dmp, sv, S.act);
            } else if (fast) {
                stress_cell<false, true, R, GX, NT, const R *, 4>(S, gt, sl, tid, zo, iy, ix,
                                                                   S.Ct + (size_t)m * 36, dmp, sv,
                                                                   S.act);
            } else {
                stress_cell<true, false, R, GX, NT, const R *, 4>(S, gt, sl, tid, zo, iy, ix,
                                                                   S.Ct + (size_t)m * 36, dmp, sv,
                                                                   S.act);
            }
#pragma unroll
            for (int k = 0; k < 6; ++k) stg(S.s[k], io, sv[k]);
        }
        lds_barrier();
    }
}

// ---------------------------------------------------------------- velocity
template <typename R, int TX, int TY>
struct VelTile {
    static constexpr int NT = TX * TY;
    static constexpr int VX = TX + 8, VY = TY + 8, NV = VX * VY;   // s plane tiles
    static constexpr int NFRAME = 8 * TY + 8 * TX;                 // x frame, then y frame
    static constexpr size_t LDS = sizeof(R) * 5 * NV;
};

// Velocity of one cell (vel_zp statement order, sim/core/fdtd.py:1211-1244).
template <bool G, typename R, int VX>
__device__ __forceinline__ void accel_col(const Shot<R> &S, const R *p1, const R *p2, const R *p4,
                                          const R *p5, const R *p6, const R (&r3)[9],
                                          const R (&r4)[9], const R (&r5)[9], int ix, int iy,
                                          int z, R &ax, R &ay, R &az)
{
    const int nz = S.nz, ny = S.ny, nx = S.nx;
    ax = 0;
    ay = 0;
    az = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        if (j >= S.nc) break;
        const R cj = S.c[j];
        if (!G || (ix >= j && ix < nx - j - 1)) ax = fma_r(cj, p1[j + 1] - p1[-j], ax);
        if (!G || (ix >= j + 1 && ix < nx - j)) {
            ay = fma_r(cj, p6[j] - p6[-(j + 1)], ay);
            az = fma_r(cj, p5[j] - p5[-(j + 1)], az);
        }
        if (!G || (iy >= j && iy < ny - j - 1)) ay = fma_r(cj, p2[(j + 1) * VX] - p2[-j * VX], ay);
        if (!G || (iy >= j + 1 && iy < ny - j)) {
            ax = fma_r(cj, p6[j * VX] - p6[-(j + 1) * VX], ax);
            az = fma_r(cj, p4[j * VX] - p4[-(j + 1) * VX], az);
        }
        if (!G || (z >= j && z < nz - j - 1)) az = fma_r(cj, r3[5 + j] - r3[4 - j], az);
        if (!G || (z >= j + 1 && z < nz - j)) {
            ax = fma_r(cj, r5[4 + j] - r5[3 - j], ax);
            ay = fma_r(cj, r4[4 + j] - r4[3 - j], ay);
        }
    }
}

// RHO_LDS: the density table sits in LDS after the five plane tiles, so the
// face-density lookups of the output stage wait on LDS only.  A global
// lookup there would make the wave wait for every prefetch issued before it
// (vector memory counters retire in order).
// CP: the shot may have a CPML.  Each velocity cell has one owner thread,
// which for a band cell forms the acceleration corrections itself (the taps
// of its band axes from the same LDS planes and z rings, profiles from an
// LDS table) and advances its memory variables 3..5 in place.
template <typename R, int TX, int TY, bool RHO_LDS = false, bool CP = false>
__device__ __forceinline__ void velocity_tiled_body(const Shot<R> &S, int tiles_x, int tiles_y,
                                                    int zc_len, int t, int zlo, int zhi)
{
    using T = VelTile<R, TX, TY>;
    constexpr int VX = T::VX, NV = T::NV, NT = T::NT, NF = (T::NFRAME + NT - 1) / NT;
    extern __shared__ unsigned char lds_raw[];
    R *st = reinterpret_cast<R *>(lds_raw);   // [5][VY][VX]: s1, s2, s4, s5, s6
    R *t1 = st + 0 * NV;
    R *t2 = st + 1 * NV;
    R *t4 = st + 2 * NV;
    R *t5 = st + 3 * NV;
    R *t6 = st + 4 * NV;
    const R *rho_tab = S.rho;
    if constexpr (RHO_LDS) {
        R *rl = st + 5 * NV;
        for (int k = threadIdx.x; k < S.n_mat; k += NT) rl[k] = S.rho[k];
        rho_tab = rl;                   // ordered by the loop's first barrier
    }
    // CP: band profiles after the density table (same ordering)
    R *bp = st + 5 * NV + (RHO_LDS ? S.n_mat : 0);
    if constexpr (CP)
        if (S.pml > 0) pml_profiles_to_lds(S, bp, NT);

    const int nz = S.nz, ny = S.ny, nx = S.nx;
    const int sy = nx, sz = nx * ny;
    const int tx = t % tiles_x, ty = (t / tiles_x) % tiles_y, chunk = t / (tiles_x * tiles_y);
    const int x0 = tx * TX, y0 = ty * TY;
    const int z0 = zlo + chunk * zc_len, z1 = min(zhi, z0 + zc_len);
    const bool inner_xy = x0 >= 4 && x0 + TX <= nx - 4 && y0 >= 4 && y0 + TY <= ny - 4;

    const int tid = threadIdx.x;
    const int lx = tid % TX, ly = tid / TX;
    const int ix = x0 + lx, iy = y0 + ly;
    const bool col_in = ix < nx && iy < ny;
    const int L = CP ? S.pml : 0, W = 2 * L + 2;
    const int col = clampi(iy, 0, ny - 1) * sy + clampi(ix, 0, nx - 1);
    const int vc = (ly + 4) * VX + (lx + 4);

    const R *__restrict__ s1 = S.s[0];
    const R *__restrict__ s2 = S.s[1];
    const R *__restrict__ s3 = S.s[2];
    const R *__restrict__ s4 = S.s[3];
    const R *__restrict__ s5 = S.s[4];
    const R *__restrict__ s6 = S.s[5];

    // this thread's frame entries: x frame (s1, s5, s6 beside the tile
    // rows) or y frame (s2, s4, s6 above / below the tile columns); LDS slot
    // (the far corner cell, which no tap reads, if none) and clamped x-y
    // offset.  Loads are branch-free: a clamped address only feeds terms the
    // edge guards exclude.
    int f_lds[NF], f_xy[NF];
    bool f_x[NF];
#pragma unroll
    for (int k = 0; k < NF; ++k) {
        const int e = tid + k * NT;
        int li = NV - 1, gx_ = 0, gy_ = 0;
        const bool xf = e < 8 * TY;
        if (xf) {
            const int r = e / 8, q = e % 8;
            const int c = q < 4 ? q : TX + q;
            gx_ = x0 - 4 + c;
            gy_ = y0 + r;
            li = (r + 4) * VX + c;
        } else if (e < T::NFRAME) {
            const int q = (e - 8 * TY) / TX, c = (e - 8 * TY) % TX;
            const int r = q < 4 ? q : TY + q;
            gx_ = x0 + c;
            gy_ = y0 - 4 + r;
            li = r * VX + c + 4;
        }
        f_lds[k] = li;
        f_x[k] = xf;
        f_xy[k] = clampi(gy_, 0, ny - 1) * sy + clampi(gx_, 0, nx - 1);
    }

    // z rings: at the top of iteration z, r?[k] holds plane z - 5 + k and
    // p? plane z + 4 (loaded one iteration earlier); after the shift at the
    // top r?[k] = plane z - 4 + k.  Registers for plane z: own column (s1,
    // s2, s6, v, labels, damp) and frame.
    R r3[9], r4[9], r5[9], p3, p4, p5, fv[NF][3];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        const int i = clampi(z0 - 5 + k, 0, nz - 1) * sz + col;
        r3[k] = ldg(s3, i);
        r4[k] = ldg(s4, i);
        r5[k] = ldg(s5, i);
    }
    {
        const int i = clampi(z0 + 4, 0, nz - 1) * sz + col;
        p3 = ldg(s3, i);
        p4 = ldg(s4, i);
        p5 = ldg(s5, i);
    }
    const bool has_damp = S.damp != nullptr;
    auto load_frame = [&](int z) {
        const int zo = clampi(z, 0, nz - 1) * sz;
#pragma unroll
        for (int k = 0; k < NF; ++k) {
            const int gi = zo + f_xy[k];
            fv[k][0] = (f_x[k] ? s1 : s2)[gi];
            fv[k][1] = (f_x[k] ? s5 : s4)[gi];
            fv[k][2] = ldg(s6, gi);
        }
    };
    R c1, c2, c6, cvx, cvy, cvz, cdmp;
    int cm0, cmx, cmy, cmz;
    auto load_col = [&](int z) {
        const int i = clampi(z, 0, nz - 1) * sz + col;
        c1 = ldg(s1, i);
        c2 = ldg(s2, i);
        c6 = ldg(s6, i);
        cvx = ldg(S.v[0], i);
        cvy = ldg(S.v[1], i);
        cvz = ldg(S.v[2], i);
        cdmp = has_damp ? ldg(S.damp, i) : (R)1;
        cm0 = ldg(S.mats, i);
        cmx = ldg(S.mats, ix < nx - 1 ? i + 1 : i);
        cmy = ldg(S.mats, iy < ny - 1 ? i + sy : i);
        cmz = ldg(S.mats, z < nz - 1 ? i + sz : i);
    };
    load_frame(z0);
    load_col(z0);

    auto step = [&](const int z) __attribute__((always_inline)) {
        const int i = z * sz + col;
        // the guarded and band paths' column inputs, re-derived every step
        // from a copy of the thread id the compiler cannot see through (else
        // it hoists their per-lane masks out of the loop and spills them;
        // see stress_tiled_body)
        int gtid = tid;
        __asm__ volatile("" : "+v"(gtid));
        const int gix = x0 + gtid % TX, giy = y0 + gtid / TX;
        const bool gcol = gix < nx && giy < ny;
        const bool gbx = L > 0 && gcol && pml_in(gix, nx, L);
        const bool gby = L > 0 && gcol && pml_in(giy, ny, L);
        const int gbix = gix <= L ? gix : gix - (nx - W), gbiy = giy <= L ? giy : giy - (ny - W);
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            r3[k] = r3[k + 1];
            r4[k] = r4[k + 1];
            r5[k] = r5[k + 1];
        }
        r3[8] = p3;
        r4[8] = p4;
        r5[8] = p5;
        // CPML band cell: memory variables 3..5 of its band axes, loaded
        // ahead of the step's prefetches
        const bool bz = CP && L > 0 && gcol && pml_in(z, nz, L);
        // (memory-variable offsets in 32 bits: the host caps each axis's
        // 6 x band-volume array under 4 GiB)
        // oz, oy, ox: offset of memory variable 3 of each band axis, formed
        // once here so that the stores below take adds only (re-derived
        // there, the compiler folded it into a 64-bit multiply-add whose
        // unused high half sat in a prefetch's destination register, and
        // waited for every prefetch of the next plane at that point; other
        // waves hid that wait: 0.644 ms per step before and after at
        // 400x400x200)
        R ps[9];
        int oz = 0, oy = 0, ox = 0, vz_ = 0, vy_ = 0, vx_ = 0;
        if constexpr (CP) {
            if (bz) {
                const int biz = z <= L ? z : z - (nz - W);
                vz_ = W * ny * nx;
                oz = 3 * vz_ + (biz * ny + giy) * nx + gix;
                __asm__ volatile("" : "+v"(oz));
#pragma unroll
                for (int k = 0; k < 3; ++k) ps[k] = ldg(S.psi[0], oz + k * vz_);
            }
            if (gby) {
                vy_ = nz * W * nx;
                oy = 3 * vy_ + (z * W + gbiy) * nx + gix;
                __asm__ volatile("" : "+v"(oy));
#pragma unroll
                for (int k = 0; k < 3; ++k) ps[3 + k] = ldg(S.psi[1], oy + k * vy_);
            }
            if (gbx) {
                vx_ = nz * ny * W;
                ox = 3 * vx_ + (z * ny + giy) * W + gbix;
                __asm__ volatile("" : "+v"(ox));
#pragma unroll
                for (int k = 0; k < 3; ++k) ps[6 + k] = ldg(S.psi[2], ox + k * vx_);
            }
        }
        // (1) plane z of s1, s2, s4, s5, s6 into LDS
        t1[vc] = c1;
        t2[vc] = c2;
        t6[vc] = c6;
        t4[vc] = r4[4];
        t5[vc] = r5[4];
#pragma unroll
        for (int k = 0; k < NF; ++k) {
            (f_x[k] ? t1 : t2)[f_lds[k]] = fv[k][0];
            (f_x[k] ? t5 : t4)[f_lds[k]] = fv[k][1];
            t6[f_lds[k]] = fv[k][2];
        }
        const R vxo = cvx, vyo = cvy, vzo = cvz, dmp = cdmp;
        const int m0 = cm0, mxp = cmx, myp = cmy, mzp = cmz;
        // prefetch plane z+1 and the ring's plane z+5
        load_frame(z + 1);
        load_col(z + 1);
        {
            const int i5 = clampi(z + 5, 0, nz - 1) * sz + col;
            p3 = ldg(s3, i5);
            p4 = ldg(s4, i5);
            p5 = ldg(s5, i5);
        }
        lds_barrier();

        R nv0 = 0, nv1 = 0, nv2 = 0;
        if (col_in) {
            R ax = 0, ay = 0, az = 0;
            if (inner_xy && z >= 4 && z <= nz - 5)
                accel_col<false, R, VX>(S, t1 + vc, t2 + vc, t4 + vc, t5 + vc, t6 + vc, r3, r4, r5,
                                        ix, iy, z, ax, ay, az);
            else
                accel_col<true, R, VX>(S, t1 + vc, t2 + vc, t4 + vc, t5 + vc, t6 + vc, r3, r4, r5,
                                       gix, giy, z, ax, ay, az);
            if constexpr (CP) {
                if (bz || gby || gbx) {
                    // corrections of the band axes z, y, x (pml_accel_corr)
                    const R *p1 = t1 + vc, *p2 = t2 + vc, *p4 = t4 + vc, *p5 = t5 + vc,
                            *p6 = t6 + vc;
                    R g0 = 0, g1 = 0, g2 = 0;
                    if (bz) {
                        const int biz = z <= L ? z : z - (nz - W);
                        R e0 = 0, e1 = 0, e2 = 0;
#pragma unroll
                        for (int j = 0; j < 4; ++j) {
                            if (j >= S.nc) break;
                            const R cj = S.c[j];
                            if (z >= j + 1 && z < nz - j) {
                                e0 = fma_r(cj, r5[4 + j] - r5[3 - j], e0);
                                e1 = fma_r(cj, r4[4 + j] - r4[3 - j], e1);
                            }
                            if (z >= j && z < nz - j - 1) e2 = fma_r(cj, r3[5 + j] - r3[4 - j], e2);
                        }
                        g0 = g0 + pml_corr_lds<0, 0>(bp, W, biz, ps[0], e0);
                        g1 = g1 + pml_corr_lds<0, 0>(bp, W, biz, ps[1], e1);
                        g2 = g2 + pml_corr_lds<0, 1>(bp, W, biz, ps[2], e2);
#pragma unroll
                        for (int k = 0; k < 3; ++k) stg(S.psi[0], oz + k * vz_, ps[k]);
                    }
                    if (gby) {
                        R e0 = 0, e1 = 0, e2 = 0;
#pragma unroll
                        for (int j = 0; j < 4; ++j) {
                            if (j >= S.nc) break;
                            const R cj = S.c[j];
                            if (giy >= j + 1 && giy < ny - j) {
                                e0 = fma_r(cj, p6[j * VX] - p6[-(j + 1) * VX], e0);
                                e2 = fma_r(cj, p4[j * VX] - p4[-(j + 1) * VX], e2);
                            }
                            if (giy >= j && giy < ny - j - 1)
                                e1 = fma_r(cj, p2[(j + 1) * VX] - p2[-j * VX], e1);
                        }
                        g0 = g0 + pml_corr_lds<1, 0>(bp, W, gbiy, ps[3], e0);
                        g1 = g1 + pml_corr_lds<1, 1>(bp, W, gbiy, ps[4], e1);
                        g2 = g2 + pml_corr_lds<1, 0>(bp, W, gbiy, ps[5], e2);
#pragma unroll
                        for (int k = 0; k < 3; ++k) stg(S.psi[1], oy + k * vy_, ps[3 + k]);
                    }
                    if (gbx) {
                        R e0 = 0, e1 = 0, e2 = 0;
#pragma unroll
                        for (int j = 0; j < 4; ++j) {
                            if (j >= S.nc) break;
                            const R cj = S.c[j];
                            if (gix >= j && gix < nx - j - 1) e0 = fma_r(cj, p1[j + 1] - p1[-j], e0);
                            if (gix >= j + 1 && gix < nx - j) {
                                e1 = fma_r(cj, p6[j] - p6[-(j + 1)], e1);
                                e2 = fma_r(cj, p5[j] - p5[-(j + 1)], e2);
                            }
                        }
                        g0 = g0 + pml_corr_lds<2, 1>(bp, W, gbix, ps[6], e0);
                        g1 = g1 + pml_corr_lds<2, 0>(bp, W, gbix, ps[7], e1);
                        g2 = g2 + pml_corr_lds<2, 0>(bp, W, gbix, ps[8], e2);
#pragma unroll
                        for (int k = 0; k < 3; ++k) stg(S.psi[2], ox + k * vx_, ps[6 + k]);
                    }
                    ax = ax + g0;
                    ay = ay + g1;
                    az = az + g2;
                }
            }
            const R tstep = S.dt_h;
            const R *rho = rho_tab;
            const R r0 = rho[m0];
            const R rvx = ix < nx - 1 ? (R)0.5 * (r0 + rho[mxp]) : r0;
            const R rvy = iy < ny - 1 ? (R)0.5 * (r0 + rho[myp]) : r0;
            const R rvz = z < nz - 1 ? (R)0.5 * (r0 + rho[mzp]) : r0;
            nv0 = (vxo + tstep * ax / rvx) * dmp;
            nv1 = (vyo + tstep * ay / rvy) * dmp;
            nv2 = (vzo + tstep * az / rvz) * dmp;
        }
        // every lane issues the three stores (stb drops the lanes outside
        // the grid): the step's vector-memory count after its prefetches is
        // then fixed, and the next step's wait for them leaves the stores in
        // flight (under the branch it was vmcnt(0); 0.678 -> 0.668 ms with
        // the CPML at 400x400x200, 0.463 -> 0.461 with the sponge)
        {
            const uint32_t n = (uint32_t)nz * (uint32_t)sz;
            stb(S.v[0], n, i, col_in, nv0);
            stb(S.v[1], n, i, col_in, nv1);
            stb(S.v[2], n, i, col_in, nv2);
        }
        lds_barrier();
    };
    // two planes per iteration: the loads one plane issues feed the next
    // without the register moves of a loop latch, each of which waits for
    // its load (measured 0.481 -> 0.457 ms at 400x400x200, 64x8 tiles; every
    // variant stays within the registers of its occupancy, the CPML ones at
    // 107-115, where the 64x8 tile went from 0.750 to 0.700 ms)
    int z = z0;
    for (; z + 1 < z1; z += 2) {
        step(z);
        step(z + 1);
    }
    for (; z < z1; ++z) step(z);
}

// ------------------------------------------------------------- launchers
template <typename R, int TX, int TY, int PAD, bool STORE_G = false, bool FR = false,
          bool CP = false>
__global__ __launch_bounds__((StressTile<R, TX, TY, PAD>::NT),
                             (MinWaves<R, StressTile<R, TX, TY, PAD>::NT>::v)) void
k_stress_tiled(Shot<R> S, int tiles_x, int tiles_y, int zc_len, int zlo, int zhi)
{
    const int nb = tiles_x * tiles_y * ((zhi - zlo + zc_len - 1) / zc_len);
    const int t = xcd_remap(blockIdx.x, nb);
    stress_tiled_body<R, TX, TY, PAD, STORE_G, FR, CP>(S, tiles_x, tiles_y, zc_len, t, zlo, zhi);
}

template <typename R, int TX, int TY, bool RHO_LDS = false, bool CP = false>
__global__ __launch_bounds__(TX *TY, (MinWaves<R, TX * TY>::v)) void k_velocity_tiled(
    Shot<R> S, int tiles_x, int tiles_y, int zc_len, int zlo, int zhi)
{
    const int nb = tiles_x * tiles_y * ((zhi - zlo + zc_len - 1) / zc_len);
    const int t = xcd_remap(blockIdx.x, nb);
    velocity_tiled_body<R, TX, TY, RHO_LDS, CP>(S, tiles_x, tiles_y, zc_len, t, zlo, zhi);
}

template <typename R, int TX, int TY>
__global__ __launch_bounds__((StressTile<R, TX, TY, 0>::NT),
                             (MinWaves<R, StressTile<R, TX, TY, 0>::NT>::v)) void
k_stress_pipe(Shot<R> S, int tiles_x, int tiles_y, int zc_len, int zlo, int zhi)
{
    const int nb = tiles_x * tiles_y * ((zhi - zlo + zc_len - 1) / zc_len);
    stress_pipe_body<R, TX, TY>(S, tiles_x, tiles_y, zc_len, xcd_remap(blockIdx.x, nb), zlo, zhi);
}

// CPML band pre-pass: one thread per band cell (pml_band_cell order), which
// advances the cell's memory variables and writes its corrections into
// S.corr -- HALF 0 the six strain corrections (from v, before the stress
// kernel), HALF 1 the three acceleration corrections (from s, before the
// velocity kernel).  Each memory variable has one owner, so it is advanced
// in place.
// pml_strain_corr with every load of an axis issued before its arithmetic:
// the taps at clamped indices (a tap outside the grid only feeds a term the
// edge guards drop), the memory variables and the profile values, so that a
// band cell waits for its loads once per axis.  Same arithmetic.
template <int A, typename R>
__device__ __forceinline__ void pml_strain_axis(const Shot<R> &S, int i, int iz, int iy, int ix,
                                                R (&ga)[6])
{
    constexpr int K0 = A == 0 ? 2 : (A == 1 ? 1 : 0);     // Db component
    constexpr int K1 = A == 0 ? 3 : (A == 1 ? 3 : 4);     // first Df component
    constexpr int K2 = A == 0 ? 4 : 5;                    // second Df component
    const int n = A == 0 ? S.nz : (A == 1 ? S.ny : S.nx);
    const int ia = A == 0 ? iz : (A == 1 ? iy : ix);
    const int st = A == 0 ? S.nx * S.ny : (A == 1 ? S.nx : 1);
    // Db on the component along A, Df on the other two (z: vz | vy, vx;
    // y: vy | vz, vx; x: vx | vz, vy)
    const R *__restrict__ fb = S.v[A == 0 ? 2 : (A == 1 ? 1 : 0)];
    const R *__restrict__ f1 = S.v[A == 2 ? 2 : (A == 1 ? 2 : 1)];
    const R *__restrict__ f2 = S.v[A == 2 ? 1 : 0];
    // pml_corr<A, 0, 0>, <A, 1, 1>, <A, 2, 1> written out with their memory
    // variables and profile values loaded together with the taps (a memory
    // variable's store would otherwise keep the next one's load behind it).
    // Their addresses are formed, and the loads issued, ahead of the taps:
    // formed after them, the compiler folded the offsets into 64-bit
    // multiply-adds whose unused high halves sat in tap destination
    // registers, so these loads waited for most taps to return (eight waves
    // per SIMD hid it: 215 us per launch at 400x400x200 before and after).
    // Offsets in 32 bits (the host caps each axis's 6 x band-volume array
    // under 4 GiB).
    const int L = S.pml, W = 2 * L + 2;
    const int bi = ia <= L ? ia : ia - (n - W);
    int vol, f;
    if constexpr (A == 0) {
        vol = W * S.ny * S.nx;
        f = (bi * S.ny + iy) * S.nx + ix;
    } else if constexpr (A == 1) {
        vol = S.nz * W * S.nx;
        f = (iz * W + bi) * S.nx + ix;
    } else {
        vol = S.nz * S.ny * W;
        f = (iz * S.ny + iy) * W + bi;
    }
    R *__restrict__ psi = S.psi[A];
    const R *pr0 = S.prof[A], *pr1 = S.prof[A] + 3 * n;
    const R q0 = psi[f], q1 = psi[f + vol], q2 = psi[f + 2 * vol];
    const R b0 = pr0[ia], a0 = pr0[n + ia], k0 = pr0[2 * n + ia];
    const R b1 = pr1[ia], a1 = pr1[n + ia], k1 = pr1[2 * n + ia];
    // taps from the cell's own address (an index i + d * st folds into the
    // same kind of 64-bit multiply-add as above)
    const R *__restrict__ cb = fb + i;
    const R *__restrict__ c1 = f1 + i;
    const R *__restrict__ c2 = f2 + i;
    R tb[8], t1[8], t2[8];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        // Db taps ia+j, ia-(j+1); Df taps ia+j+1, ia-j
        const int pb = min(max(ia + j, 0), n - 1), mb = min(max(ia - (j + 1), 0), n - 1);
        const int pf = min(max(ia + j + 1, 0), n - 1), mf = min(max(ia - j, 0), n - 1);
        tb[2 * j] = cb[(pb - ia) * st];
        tb[2 * j + 1] = cb[(mb - ia) * st];
        t1[2 * j] = c1[(pf - ia) * st];
        t1[2 * j + 1] = c1[(mf - ia) * st];
        t2[2 * j] = c2[(pf - ia) * st];
        t2[2 * j + 1] = c2[(mf - ia) * st];
    }
    R d0 = 0, d1 = 0, d2 = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        if (j >= S.nc) break;
        const R cj = S.c[j];
        if (ia >= j + 1 && ia < n - j) d0 = fma_r(cj, tb[2 * j] - tb[2 * j + 1], d0);
        if (ia >= j && ia < n - j - 1) {
            d1 = fma_r(cj, t1[2 * j] - t1[2 * j + 1], d1);
            d2 = fma_r(cj, t2[2 * j] - t2[2 * j + 1], d2);
        }
    }
    const R dd0 = d0 * S.inv_h, dd1 = d1 * S.inv_h, dd2 = d2 * S.inv_h;
    const R p0 = fma_r(b0, q0, a0 * dd0), p1 = fma_r(b1, q1, a1 * dd1),
            p2 = fma_r(b1, q2, a1 * dd2);
    psi[f] = p0;
    psi[f + vol] = p1;
    psi[f + 2 * vol] = p2;
    ga[K0] = ga[K0] + fma_r(dd0, k0 - (R)1, p0);
    ga[K1] = ga[K1] + fma_r(dd1, k1 - (R)1, p1);
    ga[K2] = ga[K2] + fma_r(dd2, k1 - (R)1, p2);
}

template <typename R>
__device__ __forceinline__ void pml_strain_corr_hoisted(const Shot<R> &S, int i, int iz, int iy,
                                                        int ix, R (&ga)[6])
{
    const int L = S.pml;
#pragma unroll
    for (int k = 0; k < 6; ++k) ga[k] = 0;
    if (pml_in(iz, S.nz, L)) pml_strain_axis<0>(S, i, iz, iy, ix, ga);
    if (pml_in(iy, S.ny, L)) pml_strain_axis<1>(S, i, iz, iy, ix, ga);
    if (pml_in(ix, S.nx, L)) pml_strain_axis<2>(S, i, iz, iy, ix, ga);
}

template <typename R, int HALF>
__device__ __forceinline__ void pml_corr_body(const Shot<R> &S, int b)
{
    if (b >= pml_band_cells(S.nz, S.ny, S.nx, S.pml)) return;
    int iz, iy, ix;
    pml_band_cell(b, S.nz, S.ny, S.nx, S.pml, iz, iy, ix);
    const int sz = S.nx * S.ny;
    const int i = iz * sz + iy * S.nx + ix;
    const size_t n = (size_t)S.nz * sz;
    if constexpr (HALF == 0) {
        R ga[6];
        pml_strain_corr_hoisted(S, i, iz, iy, ix, ga);
        const int mask = pml_strain_mask(S, iz, iy, ix);
#pragma unroll
        for (int k = 0; k < 6; ++k)
            if (mask >> k & 1) S.corr[k * n + i] = ga[k];
    } else {
        R gv[3];
        pml_accel_corr(S, i, iz, iy, ix, gv);
#pragma unroll
        for (int k = 0; k < 3; ++k) S.corr[k * n + i] = gv[k];
    }
}

template <typename R, int HALF>
__global__ __launch_bounds__(kPmlBlock) void k_pml_corr(Shot<R> S)
{
    pml_corr_body<R, HALF>(S, blockIdx.x * kPmlBlock + threadIdx.x);
}

// Stress half-step band pre-pass over the band cells that k_pml_corr_zslab
// leaves: those of the z slabs also in the y or x band, and all the others.
template <typename R>
__global__ __launch_bounds__(kPmlBlock) void k_pml_corr_rest(Shot<R> S)
{
    const int b = blockIdx.x * kPmlBlock + threadIdx.x;
    const int plane = S.ny * S.nx;
    if (b < (2 * S.pml + 2) * plane) {
        const int r = b % plane, iy = r / S.nx, ix = r - iy * S.nx;
        if (!pml_in(iy, S.ny, S.pml) && !pml_in(ix, S.nx, S.pml)) return;
    }
    pml_corr_body<R, 0>(S, b);
}

// The z slabs of the stress half-step pre-pass (band cells 0 .. W*ny*nx of
// pml_band_cell), columns outside the y and x bands: one thread per (y, x)
// column, slab and run of kZslabRun planes, marching over them with register
// rings of the z taps, so that a cell loads one plane of vz, vy, vx instead
// of its 24 z taps.  The arithmetic is pml_strain_axis<0>'s (the same taps,
// guards and operation order).  Needs nz >= 2 (L + 1).
constexpr int kZslabRun = 8;

template <typename R>
__global__ __launch_bounds__(kPmlBlock) void k_pml_corr_zslab(Shot<R> S)
{
    const int nz = S.nz, ny = S.ny, nx = S.nx, L = S.pml, W = 2 * L + 2;
    const int plane = ny * nx;
    const int runs = (L + 1 + kZslabRun - 1) / kZslabRun;   // runs per slab
    const int t = blockIdx.x * kPmlBlock + threadIdx.x;
    if (t >= 2 * runs * plane) return;
    const int u = t / plane, r = t - u * plane;
    const int slab = u & 1, q0r = (u >> 1) * kZslabRun, q1r = min(L + 1, q0r + kZslabRun);
    const int iy = r / nx, ix = r - iy * nx;
    if (pml_in(iy, ny, L) || pml_in(ix, nx, L)) return;   // k_pml_corr_rest's
    const int zb = (slab ? nz - 1 - L : 0) + q0r;  // first plane of the run
    const int vol = W * plane;                     // one memory variable
    const size_t n = (size_t)nz * plane;
    const R *__restrict__ vz = S.v[2] + r;
    const R *__restrict__ vy = S.v[1] + r;
    const R *__restrict__ vx = S.v[0] + r;
    const R *pr0 = S.prof[0], *pr1 = S.prof[0] + 3 * nz;
    auto pz = [&](int z) { return min(max(z, 0), nz - 1) * plane; };
    // rz[k] = vz at plane iz - 4 + k (Db taps iz+j, iz-(j+1)); ry, rx at
    // plane iz - 3 + k (Df taps iz+j+1, iz-j); planes clamped to the grid
    R rz[8], ry[8], rx[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        rz[k] = vz[pz(zb - 4 + k)];
        ry[k] = vy[pz(zb - 3 + k)];
        rx[k] = vx[pz(zb - 3 + k)];
    }
#pragma unroll 1
    for (int q = q0r; q < q1r; ++q) {
        const int iz = zb + (q - q0r);
        const int i = iz * plane + r;
        const int f = (slab * (L + 1) + q) * plane + r;
        R *__restrict__ psi = S.psi[0];
        const R q0 = psi[f], q1 = psi[f + vol], q2 = psi[f + 2 * vol];
        const R b0 = pr0[iz], a0 = pr0[nz + iz], k0 = pr0[2 * nz + iz];
        const R b1 = pr1[iz], a1 = pr1[nz + iz], k1 = pr1[2 * nz + iz];
        R d0 = 0, d1 = 0, d2 = 0;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (j >= S.nc) break;
            const R cj = S.c[j];
            if (iz >= j + 1 && iz < nz - j) d0 = fma_r(cj, rz[4 + j] - rz[3 - j], d0);
            if (iz >= j && iz < nz - j - 1) {
                d1 = fma_r(cj, ry[4 + j] - ry[3 - j], d1);
                d2 = fma_r(cj, rx[4 + j] - rx[3 - j], d2);
            }
        }
        const R dd0 = d0 * S.inv_h, dd1 = d1 * S.inv_h, dd2 = d2 * S.inv_h;
        const R p0 = fma_r(b0, q0, a0 * dd0), p1 = fma_r(b1, q1, a1 * dd1),
                p2 = fma_r(b1, q2, a1 * dd2);
        psi[f] = p0;
        psi[f + vol] = p1;
        psi[f + 2 * vol] = p2;
        R ga[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) ga[k] = 0;
        ga[2] = ga[2] + fma_r(dd0, k0 - (R)1, p0);
        ga[3] = ga[3] + fma_r(dd1, k1 - (R)1, p1);
        ga[4] = ga[4] + fma_r(dd2, k1 - (R)1, p2);
        // (pml_strain_mask: the z band's components)
        S.corr[2 * n + i] = ga[2];
        S.corr[3 * n + i] = ga[3];
        S.corr[4 * n + i] = ga[4];
        // next plane: shift the rings, load plane iz + 4 of vz and iz + 5
        // of vy, vx
#pragma unroll
        for (int k = 0; k < 7; ++k) {
            rz[k] = rz[k + 1];
            ry[k] = ry[k + 1];
            rx[k] = rx[k + 1];
        }
        rz[7] = vz[pz(iz + 4)];
        ry[7] = vy[pz(iz + 5)];
        rx[7] = vx[pz(iz + 5)];
    }
}

// Multi-shot kernels: the workgroups of every shot of a batch in one launch.
// blk_shot maps a (remapped) workgroup id to its shot; a shot whose step
// counter has reached its step count exits at once.
template <typename R, int TX, int TY, int PAD, bool FR = false, bool CP = false>
__global__ __launch_bounds__((StressTile<R, TX, TY, PAD>::NT),
                             (MinWaves<R, StressTile<R, TX, TY, PAD>::NT>::v)) void
k_stress_batched(const BatchEntry<R> *__restrict__ tab, const int32_t *__restrict__ blk_shot,
                 int nb_total, const int32_t *__restrict__ step)
{
    const int g = xcd_remap(blockIdx.x, nb_total);
    const int s = blk_shot[g];
    const BatchEntry<R> *e = tab + s;
    if (step[s] >= e->nt) return;
    const Shot<R> &S = e->S;
    stress_tiled_body<R, TX, TY, PAD, false, FR, CP>(S, e->tiles_x[0], e->tiles_y[0],
                                                     e->zc_len[0], g - e->blk_off[0], 0, S.nz);
}

template <typename R, int TX, int TY>
__global__ __launch_bounds__((StressTile<R, TX, TY, 0>::NT),
                             (MinWaves<R, StressTile<R, TX, TY, 0>::NT>::v)) void
k_stress_pipe_batched(const BatchEntry<R> *__restrict__ tab, const int32_t *__restrict__ blk_shot,
                      int nb_total, const int32_t *__restrict__ step)
{
    const int g = xcd_remap(blockIdx.x, nb_total);
    const int s = blk_shot[g];
    const BatchEntry<R> *e = tab + s;
    if (step[s] >= e->nt) return;
    const Shot<R> S = e->S;
    stress_pipe_body<R, TX, TY>(S, e->tiles_x[0], e->tiles_y[0], e->zc_len[0], g - e->blk_off[0],
                                0, S.nz);
}

template <typename R, int TX, int TY, bool RHO_LDS = false, bool CP = false>
__global__ __launch_bounds__(TX *TY, (MinWaves<R, TX * TY>::v)) void k_velocity_batched(
    const BatchEntry<R> *__restrict__ tab, const int32_t *__restrict__ blk_shot, int nb_total,
    const int32_t *__restrict__ step)
{
    const int g = xcd_remap(blockIdx.x, nb_total);
    const int s = blk_shot[g];
    const BatchEntry<R> *e = tab + s;
    if (step[s] >= e->nt) return;
    const Shot<R> &S = e->S;
    velocity_tiled_body<R, TX, TY, RHO_LDS, CP>(S, e->tiles_x[1], e->tiles_y[1], e->zc_len[1],
                                                g - e->blk_off[1], 0, S.nz);
}

// Band pre-pass of a batch: blk_shot maps a block to its CPML shot, whose
// blocks start at blk_off[2].
template <typename R, int HALF>
__global__ __launch_bounds__(kPmlBlock) void k_pml_corr_batched(
    const BatchEntry<R> *__restrict__ tab, const int32_t *__restrict__ blk_shot,
    const int32_t *__restrict__ step)
{
    const int s = blk_shot[blockIdx.x];
    const BatchEntry<R> *e = tab + s;
    if (step[s] >= e->nt) return;
    pml_corr_body<R, HALF>(e->S, (blockIdx.x - e->blk_off[2]) * kPmlBlock + threadIdx.x);
}

template <typename R>
hipError_t pml_corr_launch(const Shot<R> &s, int half, hipStream_t st)
{
    const long long cells = pml_band_cells(s.nz, s.ny, s.nx, s.pml);
    const int nb = (int)((cells + kPmlBlock - 1) / kPmlBlock);
    const int W = 2 * s.pml + 2;
    if (half == 0 && s.nz >= W) {
        // the z slabs' columns outside the y and x bands marched, then the
        // rest of the band cell by cell
        const long long thr =
            2LL * ((s.pml + 1 + kZslabRun - 1) / kZslabRun) * s.ny * s.nx;
        hipLaunchKernelGGL((k_pml_corr_zslab<R>), dim3((int)((thr + kPmlBlock - 1) / kPmlBlock)),
                           dim3(kPmlBlock), 0, st, s);
        hipLaunchKernelGGL((k_pml_corr_rest<R>), dim3(nb), dim3(kPmlBlock), 0, st, s);
    } else if (half == 0)
        hipLaunchKernelGGL((k_pml_corr<R, 0>), dim3(nb), dim3(kPmlBlock), 0, st, s);
    else
        hipLaunchKernelGGL((k_pml_corr<R, 1>), dim3(nb), dim3(kPmlBlock), 0, st, s);
    return hipGetLastError();
}

template <typename R, int TX, int TY, int PAD, bool FR = false>
hipError_t stress_launch(const Shot<R> &s, int zchunk, hipStream_t st, bool store_g = false)
{
    using T = StressTile<R, TX, TY, PAD, FR>;
    const int tiles_x = (s.nx + TX - 1) / TX, tiles_y = (s.ny + TY - 1) / TY;
    const size_t lds = s.pml > 0 ? T::LDS_CP : T::LDS;
    auto run = [&](auto kern) -> hipError_t {
        const int zl = chunk_len(s.nz, tiles_x * tiles_y, zchunk);
        const int nb = tiles_x * tiles_y * ((s.nz + zl - 1) / zl);
        hipError_t e = allow_lds(kern, lds);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(kern, dim3(nb), dim3(T::NT), lds, st, s, tiles_x, tiles_y, zl, 0,
                           s.nz);
        return hipGetLastError();
    };
    if (s.pml > 0) {
        hipError_t e = pml_corr_launch(s, 0, st);
        if (e != hipSuccess) return e;
        // with the strain history of the adjoint: the strains with their
        // band corrections
        if (store_g) return run(k_stress_tiled<R, TX, TY, PAD, true, FR, true>);
        return run(k_stress_tiled<R, TX, TY, PAD, false, FR, true>);
    }
    if (store_g) return run(k_stress_tiled<R, TX, TY, PAD, true, FR>);
    return run(k_stress_tiled<R, TX, TY, PAD, false, FR>);
}

// Pipelined stress tile (stress_pipe_body); the strain-history store and
// CPML shots run the plain kernel of the same tile.
template <typename R, int TX, int TY>
hipError_t stress_pipe_launch(const Shot<R> &s, int zchunk, hipStream_t st, bool store_g)
{
    if (store_g || s.pml > 0) return stress_launch<R, TX, TY, 0>(s, zchunk, st, store_g);
    using T = StressTile<R, TX, TY, 0>;
    const int tiles_x = (s.nx + TX - 1) / TX, tiles_y = (s.ny + TY - 1) / TY;
    const int zl = chunk_len(s.nz, tiles_x * tiles_y, zchunk);
    const int nb = tiles_x * tiles_y * ((s.nz + zl - 1) / zl);
    const size_t lds = PipeTile<R, TX, TY>::LDS;
    hipError_t e = allow_lds(k_stress_pipe<R, TX, TY>, lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_stress_pipe<R, TX, TY>), dim3(nb), dim3(T::NT), lds, st, s, tiles_x,
                       tiles_y, zl, 0, s.nz);
    return hipGetLastError();
}

template <typename R, int TX, int TY>
hipError_t velocity_launch(const Shot<R> &s, int zchunk, hipStream_t st)
{
    using T = VelTile<R, TX, TY>;
    const int tiles_x = (s.nx + TX - 1) / TX, tiles_y = (s.ny + TY - 1) / TY;
    const size_t rho_bytes = sizeof(R) * (size_t)s.n_mat;
    const bool rho_lds = rho_bytes <= (size_t)kRhoLdsBytes;
    // CPML: the band profile table after the densities
    const size_t lds = T::LDS + (rho_lds ? rho_bytes : 0) +
                       (s.pml > 0 ? sizeof(R) * 18 * (2 * (size_t)s.pml + 2) : 0);
    auto run = [&](auto kern) -> hipError_t {
        const int zl = chunk_len(s.nz, tiles_x * tiles_y, zchunk);
        const int nb = tiles_x * tiles_y * ((s.nz + zl - 1) / zl);
        hipError_t e = allow_lds(kern, lds);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(kern, dim3(nb), dim3(T::NT), lds, st, s, tiles_x, tiles_y, zl, 0,
                           s.nz);
        return hipGetLastError();
    };
    if (s.pml > 0) {
        if (rho_lds) return run(k_velocity_tiled<R, TX, TY, true, true>);
        return run(k_velocity_tiled<R, TX, TY, false, true>);
    }
    if (rho_lds) return run(k_velocity_tiled<R, TX, TY, true>);
    return run(k_velocity_tiled<R, TX, TY>);
}

template <typename R, int TX, int TY, int PAD, bool FR = false>
hipError_t stress_batched_launch(const BatchEntry<R> *tab, const int32_t *blk, int nb,
                                 const int32_t *step, hipStream_t st, bool cp)
{
    using T = StressTile<R, TX, TY, PAD, FR>;
    const size_t lds = cp ? T::LDS_CP : T::LDS;
    auto go = [&](auto kern) -> hipError_t {
        hipError_t e = allow_lds(kern, lds);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(kern, dim3(nb), dim3(T::NT), lds, st, tab, blk, nb, step);
        return hipGetLastError();
    };
    return cp ? go(k_stress_batched<R, TX, TY, PAD, FR, true>)
              : go(k_stress_batched<R, TX, TY, PAD, FR, false>);
}

template <typename R, int TX, int TY>
hipError_t stress_pipe_batched_launch(const BatchEntry<R> *tab, const int32_t *blk, int nb,
                                      const int32_t *step, hipStream_t st)
{
    using T = StressTile<R, TX, TY, 0>;
    const size_t lds = PipeTile<R, TX, TY>::LDS;
    hipError_t e = allow_lds(k_stress_pipe_batched<R, TX, TY>, lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_stress_pipe_batched<R, TX, TY>), dim3(nb), dim3(T::NT), lds, st, tab,
                       blk, nb, step);
    return hipGetLastError();
}

// max_pml: the widest CPML band of the batch (sizes the LDS profile table;
// 0 = no CPML shot).
template <typename R, int TX, int TY>
hipError_t velocity_batched_launch(const BatchEntry<R> *tab, const int32_t *blk, int nb,
                                   const int32_t *step, hipStream_t st, int max_mat, int max_pml)
{
    using T = VelTile<R, TX, TY>;
    const size_t rho_bytes = sizeof(R) * (size_t)max_mat;
    const bool rho_lds = rho_bytes <= (size_t)kRhoLdsBytes;
    const bool cp = max_pml > 0;
    const size_t lds = T::LDS + (rho_lds ? rho_bytes : 0) +
                       (cp ? sizeof(R) * 18 * (2 * (size_t)max_pml + 2) : 0);
    auto go = [&](auto kern) -> hipError_t {
        hipError_t e = allow_lds(kern, lds);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(kern, dim3(nb), dim3(T::NT), lds, st, tab, blk, nb, step);
        return hipGetLastError();
    };
    if (rho_lds)
        return cp ? go(k_velocity_batched<R, TX, TY, true, true>)
                  : go(k_velocity_batched<R, TX, TY, true, false>);
    return cp ? go(k_velocity_batched<R, TX, TY, false, true>)
              : go(k_velocity_batched<R, TX, TY, false, false>);
}

// Tile variants (TX, TY, LDS row padding); variant 0 is the default.
// fr: 1 = the stress kernel's frame ring (StressTile), 2 = the one-barrier
// pipelined stress body (stress_pipe_body).
struct TileInfo {
    int tx, ty, pad, fr;
};
constexpr TileInfo kStressF32[] = {{32, 13, 24, 0}, {62, 6, 0, 0},  {30, 14, 24, 0},
                                   {32, 16, 24, 0}, {32, 8, 24, 0}, {16, 16, 0, 0},
                                   {32, 16, 0, 0},  {30, 14, 24, 1}, {32, 13, 24, 1},
                                   {30, 14, 0, 1},  {62, 14, 0, 0}, {62, 14, 0, 1},
                                   {32, 13, 0, 2},  {30, 14, 0, 2}, {40, 10, 24, 1},
                                   {40, 10, 24, 0}};
constexpr TileInfo kStressF64[] = {{32, 6, 0, 0}, {32, 4, 0, 0}, {16, 8, 0, 0}, {32, 6, 24, 0},
                                   {32, 6, 0, 2}};
constexpr TileInfo kVelF32[] = {{64, 8, 0, 0},  {32, 16, 0, 0}, {32, 8, 0, 0},
                                {64, 4, 0, 0},  {128, 4, 0, 0}, {32, 32, 0, 0}};
constexpr TileInfo kVelF64[] = {{32, 8, 0, 0}, {64, 4, 0, 0}, {32, 16, 0, 0}, {32, 4, 0, 0}};

}  // namespace

int tile_variants(int precision, int half, int variant, int *tx, int *ty, int *pad, int *fr)
{
    const TileInfo *t = nullptr;
    int n = 0;
    if (half == 0 && precision == PE_F32) { t = kStressF32; n = sizeof kStressF32 / sizeof *t; }
    if (half == 0 && precision == PE_F64) { t = kStressF64; n = sizeof kStressF64 / sizeof *t; }
    if (half == 1 && precision == PE_F32) { t = kVelF32; n = sizeof kVelF32 / sizeof *t; }
    if (half == 1 && precision == PE_F64) { t = kVelF64; n = sizeof kVelF64 / sizeof *t; }
    if (variant >= 0 && variant < n) {
        *tx = t[variant].tx;
        *ty = t[variant].ty;
        *pad = t[variant].pad;
        *fr = t[variant].fr;
    }
    return n;
}

int tiled_geometry(int precision, int half, int variant, int zchunk, int nz, int ny, int nx,
                   int *tiles_x, int *tiles_y, int *zc_len)
{
    int tx = 0, ty = 0, pad = 0, fr = 0;
    const int n = tile_variants(precision, half, variant, &tx, &ty, &pad, &fr);
    if (variant < 0 || variant >= n) return -1;
    *tiles_x = (nx + tx - 1) / tx;
    *tiles_y = (ny + ty - 1) / ty;
    *zc_len = chunk_len(nz, *tiles_x * *tiles_y, zchunk);
    return *tiles_x * *tiles_y * ((nz + *zc_len - 1) / *zc_len);
}

template <>
hipError_t launch_stress_batched<float>(int v, const BatchEntry<float> *tab, const int32_t *blk,
                                        int nb, const int32_t *step, hipStream_t st, bool cp)
{
    switch (v) {
    case 1: return stress_batched_launch<float, 62, 6, 0>(tab, blk, nb, step, st, cp);
    case 2: return stress_batched_launch<float, 30, 14, 24>(tab, blk, nb, step, st, cp);
    case 3: return stress_batched_launch<float, 32, 16, 24>(tab, blk, nb, step, st, cp);
    case 4: return stress_batched_launch<float, 32, 8, 24>(tab, blk, nb, step, st, cp);
    case 5: return stress_batched_launch<float, 16, 16, 0>(tab, blk, nb, step, st, cp);
    case 6: return stress_batched_launch<float, 32, 16, 0>(tab, blk, nb, step, st, cp);
    case 7: return stress_batched_launch<float, 30, 14, 24, true>(tab, blk, nb, step, st, cp);
    case 8: return stress_batched_launch<float, 32, 13, 24, true>(tab, blk, nb, step, st, cp);
    case 9: return stress_batched_launch<float, 30, 14, 0, true>(tab, blk, nb, step, st, cp);
    case 10: return stress_batched_launch<float, 62, 14, 0>(tab, blk, nb, step, st, cp);
    case 11: return stress_batched_launch<float, 62, 14, 0, true>(tab, blk, nb, step, st, cp);
    // the pipelined body has no CPML path: a batch with CPML shots runs the
    // plain kernel of the same tile, as a single shot does
    case 12:
        return cp ? stress_batched_launch<float, 32, 13, 0>(tab, blk, nb, step, st, cp)
                  : stress_pipe_batched_launch<float, 32, 13>(tab, blk, nb, step, st);
    case 13:
        return cp ? stress_batched_launch<float, 30, 14, 0>(tab, blk, nb, step, st, cp)
                  : stress_pipe_batched_launch<float, 30, 14>(tab, blk, nb, step, st);
    case 14: return stress_batched_launch<float, 40, 10, 24, true>(tab, blk, nb, step, st, cp);
    case 15: return stress_batched_launch<float, 40, 10, 24>(tab, blk, nb, step, st, cp);
    default: return stress_batched_launch<float, 32, 13, 24>(tab, blk, nb, step, st, cp);
    }
}

template <>
hipError_t launch_stress_batched<double>(int v, const BatchEntry<double> *tab, const int32_t *blk,
                                         int nb, const int32_t *step, hipStream_t st, bool cp)
{
    switch (v) {
    case 1: return stress_batched_launch<double, 32, 4, 0>(tab, blk, nb, step, st, cp);
    case 2: return stress_batched_launch<double, 16, 8, 0>(tab, blk, nb, step, st, cp);
    case 3: return stress_batched_launch<double, 32, 6, 24>(tab, blk, nb, step, st, cp);
    case 4:
        return cp ? stress_batched_launch<double, 32, 6, 0>(tab, blk, nb, step, st, cp)
                  : stress_pipe_batched_launch<double, 32, 6>(tab, blk, nb, step, st);
    default: return stress_batched_launch<double, 32, 6, 0>(tab, blk, nb, step, st, cp);
    }
}

template <>
hipError_t launch_velocity_batched<float>(int v, const BatchEntry<float> *tab, const int32_t *blk,
                                          int nb, const int32_t *step, hipStream_t st,
                                          int max_mat, int cp)
{
    switch (v) {
    case 1: return velocity_batched_launch<float, 32, 16>(tab, blk, nb, step, st, max_mat, cp);
    case 2: return velocity_batched_launch<float, 32, 8>(tab, blk, nb, step, st, max_mat, cp);
    case 3: return velocity_batched_launch<float, 64, 4>(tab, blk, nb, step, st, max_mat, cp);
    case 4: return velocity_batched_launch<float, 128, 4>(tab, blk, nb, step, st, max_mat, cp);
    case 5: return velocity_batched_launch<float, 32, 32>(tab, blk, nb, step, st, max_mat, cp);
    default: return velocity_batched_launch<float, 64, 8>(tab, blk, nb, step, st, max_mat, cp);
    }
}

template <>
hipError_t launch_velocity_batched<double>(int v, const BatchEntry<double> *tab,
                                           const int32_t *blk, int nb, const int32_t *step,
                                           hipStream_t st, int max_mat, int cp)
{
    switch (v) {
    case 1: return velocity_batched_launch<double, 64, 4>(tab, blk, nb, step, st, max_mat, cp);
    case 2: return velocity_batched_launch<double, 32, 16>(tab, blk, nb, step, st, max_mat, cp);
    case 3: return velocity_batched_launch<double, 32, 4>(tab, blk, nb, step, st, max_mat, cp);
    default: return velocity_batched_launch<double, 32, 8>(tab, blk, nb, step, st, max_mat, cp);
    }
}

template <typename R>
hipError_t launch_pml_corr_batched(int half, const BatchEntry<R> *tab, const int32_t *blk, int nb,
                                   const int32_t *step, hipStream_t st)
{
    if (nb <= 0) return hipSuccess;
    if (half == 0)
        hipLaunchKernelGGL((k_pml_corr_batched<R, 0>), dim3(nb), dim3(kPmlBlock), 0, st, tab, blk,
                           step);
    else
        hipLaunchKernelGGL((k_pml_corr_batched<R, 1>), dim3(nb), dim3(kPmlBlock), 0, st, tab, blk,
                           step);
    return hipGetLastError();
}

template hipError_t launch_pml_corr_batched<float>(int, const BatchEntry<float> *, const int32_t *,
                                                   int, const int32_t *, hipStream_t);
template hipError_t launch_pml_corr_batched<double>(int, const BatchEntry<double> *,
                                                    const int32_t *, int, const int32_t *,
                                                    hipStream_t);

template <>
hipError_t launch_stress_tiled<float>(const Shot<float> &s, int v, int zc, hipStream_t st,
                                      bool g)
{
    switch (v) {
    case 1: return stress_launch<float, 62, 6, 0>(s, zc, st, g);
    case 2: return stress_launch<float, 30, 14, 24>(s, zc, st, g);
    case 3: return stress_launch<float, 32, 16, 24>(s, zc, st, g);
    case 4: return stress_launch<float, 32, 8, 24>(s, zc, st, g);
    case 5: return stress_launch<float, 16, 16, 0>(s, zc, st, g);
    case 6: return stress_launch<float, 32, 16, 0>(s, zc, st, g);
    case 7: return stress_launch<float, 30, 14, 24, true>(s, zc, st, g);
    case 8: return stress_launch<float, 32, 13, 24, true>(s, zc, st, g);
    case 9: return stress_launch<float, 30, 14, 0, true>(s, zc, st, g);
    case 10: return stress_launch<float, 62, 14, 0>(s, zc, st, g);
    case 11: return stress_launch<float, 62, 14, 0, true>(s, zc, st, g);
    case 12: return stress_pipe_launch<float, 32, 13>(s, zc, st, g);
    case 13: return stress_pipe_launch<float, 30, 14>(s, zc, st, g);
    case 14: return stress_launch<float, 40, 10, 24, true>(s, zc, st, g);
    case 15: return stress_launch<float, 40, 10, 24>(s, zc, st, g);
    default: return stress_launch<float, 32, 13, 24>(s, zc, st, g);
    }
}

template <>
hipError_t launch_stress_tiled<double>(const Shot<double> &s, int v, int zc, hipStream_t st,
                                       bool g)
{
    switch (v) {
    case 1: return stress_launch<double, 32, 4, 0>(s, zc, st, g);
    case 2: return stress_launch<double, 16, 8, 0>(s, zc, st, g);
    case 3: return stress_launch<double, 32, 6, 24>(s, zc, st, g);
    case 4: return stress_pipe_launch<double, 32, 6>(s, zc, st, g);
    default: return stress_launch<double, 32, 6, 0>(s, zc, st, g);
    }
}

template <>
hipError_t launch_velocity_tiled<float>(const Shot<float> &s, int v, int zc, hipStream_t st)
{
    switch (v) {
    case 1: return velocity_launch<float, 32, 16>(s, zc, st);
    case 2: return velocity_launch<float, 32, 8>(s, zc, st);
    case 3: return velocity_launch<float, 64, 4>(s, zc, st);
    case 4: return velocity_launch<float, 128, 4>(s, zc, st);
    case 5: return velocity_launch<float, 32, 32>(s, zc, st);
    default: return velocity_launch<float, 64, 8>(s, zc, st);
    }
}

template <>
hipError_t launch_velocity_tiled<double>(const Shot<double> &s, int v, int zc, hipStream_t st)
{
    switch (v) {
    case 1: return velocity_launch<double, 64, 4>(s, zc, st);
    case 2: return velocity_launch<double, 32, 16>(s, zc, st);
    case 3: return velocity_launch<double, 32, 4>(s, zc, st);
    default: return velocity_launch<double, 32, 8>(s, zc, st);
    }
}

}  // namespace pe
