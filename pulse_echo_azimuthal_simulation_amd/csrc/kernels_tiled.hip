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
#include <mutex>

#include "tiled_common.hpp"

namespace pe {

namespace {

using namespace tiled;

// Body of the stress kernel for logical tile t (tile x fastest, then tile
// y, then z-chunk) of shot S.  STORE_G: also write the strains g1..g6 of the
// tile's own columns into S.g (6 x ncell, the history the adjoint's
// stiffness gradient reads), each plane once.  FR: see StressTile.  PM: the
// CPML band kernel of the tile (band_kernel; 0 = no band path compiled), band
// columns take strains_col_pml.
template <typename R, int TX, int TY, int PAD, bool STORE_G = false, bool FR = false,
          int PM = 0>
__device__ __forceinline__ void stress_tiled_body(const Shot<R> &S, int tiles_x, int tiles_y,
                                                  int zc_len, int t, int zlo, int zhi)
{
    using T = StressTile<R, TX, TY, PAD, FR>;
    constexpr bool PML = PM != 0;
    constexpr int GX = T::GX, GY = T::GY, NG = T::NG, VX = T::VX, VY = T::VY, NV = T::NV;
    constexpr int NT = T::NT, NF = T::NF, WX = T::WX;
    extern __shared__ unsigned char lds_raw[];
    R *vt = reinterpret_cast<R *>(lds_raw);       // [3][VY][VX]
    R *gt = vt + 3 * NV;                           // [6][3][NT]: strains by thread
    R *fr = gt + 6 * 3 * NT;                       // FR: [4][NF][3][NT] frame ring
    R *bp = fr + T::NRING;                         // PML: band profiles (pml_profiles_to_lds)
    if constexpr (PML) pml_profiles_to_lds(S, bp, NT);

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
    const bool pml_xy = ((PM & 4) && pml_in(ix, nx, S.pml)) || ((PM & 2) && pml_in(iy, ny, S.pml));
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

    for (int zc = z0 - 1; zc <= z1; ++zc) {
        const bool plane_in = zc >= 0 && zc < nz;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            rx[k] = rx[k + 1];
            ry[k] = ry[k + 1];
            rz[k] = rz[k + 1];
        }
        rx[8] = px;
        ry[8] = py;
        rz[8] = pz;
        // CPML band: this plane's memory variables, loaded ahead of the
        // step's prefetches
        R ps[9];
        bool band = false;
        if constexpr (PML) {
            band = pml_xy || ((PM & 1) && pml_in(zc, nz, S.pml));
            if (band && col_in && plane_in) pml_psi_load<PM>(S, zc, iy, ix, 0, ps);
        }
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

        // (2) strains of plane zc at every column of the tile and its ring
        const int slot = (zc + 3) % 3;
        {
            R a[6] = {0, 0, 0, 0, 0, 0};
            if (band) {
                // CPML band: only the owner of the cell advances its memory
                // variables (ring columns and chunk-edge planes are also
                // computed by neighbouring workgroups)
                if constexpr (PML)
                    if (col_in && plane_in)
                        strains_col_pml<R, VX, PM>(S, bp, ps, vt + vc, vt + NV + vc, vt + 2 * NV + vc,
                                               rx, ry, rz, ix, iy, zc,
                                               core && zc >= z0 && zc < z1, a);
            } else if (inner_xy && zc >= 4 && zc <= nz - 5)
                strains_col<false, R, VX>(S, vt + vc, vt + NV + vc, vt + 2 * NV + vc, rx, ry, rz,
                                          ix, iy, zc, a);
            else if (col_in && plane_in)
                strains_col<true, R, VX>(S, vt + vc, vt + NV + vc, vt + 2 * NV + vc, rx, ry, rz,
                                         ix, iy, zc, a);
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

        // (3) stress of plane zc - 1 from the strain planes zc-2, zc-1, zc
        if (out) {
            const int sl[3] = {(zc + 1) % 3, (zc + 2) % 3, slot};
            const bool fast = full && inner_xy && zo >= 1 && zo <= nz - 2;
            const int mu = __builtin_amdgcn_readfirstlane(m);
            if (fast && __all(m == mu)) {
                // every lane of the wave in one grain: wave-uniform row
                ConstRow<R> cm = (ConstRow<R>)(S.Ct + (size_t)mu * 36);
                stress_cell<false, true, R, GX, NT>(S, gt, sl, tid, zo, iy, ix, cm, dmp, sv);
            } else if (fast) {
                stress_cell<false, true, R, GX, NT>(S, gt, sl, tid, zo, iy, ix,
                                                    S.Ct + (size_t)m * 36, dmp, sv);
            } else {
                stress_cell<true, false, R, GX, NT>(S, gt, sl, tid, zo, iy, ix,
                                                    S.Ct + (size_t)m * 36, dmp, sv);
            }
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
                                                                     dmp, sv);
            } else if (fast) {
                stress_cell<false, true, R, GX, NT, const R *, 4>(S, gt, sl, tid, zo, iy, ix,
                                                                   S.Ct + (size_t)m * 36, dmp, sv);
            } else {
                stress_cell<true, false, R, GX, NT, const R *, 4>(S, gt, sl, tid, zo, iy, ix,
                                                                   S.Ct + (size_t)m * 36, dmp, sv);
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

// Accelerations of one column at a CPML band cell (see pml_d): the taps of
// accel_col per derivative, scaled by 1/h, through pml_t (memory variables
// 3..5 in ps); summed x, y, z per component.  Every velocity cell has one
// owner.  M: see pml_psi_load.
template <typename R, int VX, int M = 7>
__device__ __forceinline__ void accel_col_pml(const Shot<R> &S, const R *bp, const R (&ps)[9],
                                              const R *p1, const R *p2, const R *p4,
                                              const R *p5, const R *p6, const R (&r3)[9],
                                              const R (&r4)[9], const R (&r5)[9], int ix, int iy,
                                              int z, R &ax, R &ay, R &az)
{
    const int nz = S.nz, ny = S.ny, nx = S.nx;
    R e[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        if (j >= S.nc) break;
        const R cj = S.c[j];
        if (ix >= j && ix < nx - j - 1) e[0] = fma_r(cj, p1[j + 1] - p1[-j], e[0]);
        if (ix >= j + 1 && ix < nx - j) {
            e[3] = fma_r(cj, p6[j] - p6[-(j + 1)], e[3]);
            e[6] = fma_r(cj, p5[j] - p5[-(j + 1)], e[6]);
        }
        if (iy >= j && iy < ny - j - 1) e[4] = fma_r(cj, p2[(j + 1) * VX] - p2[-j * VX], e[4]);
        if (iy >= j + 1 && iy < ny - j) {
            e[1] = fma_r(cj, p6[j * VX] - p6[-(j + 1) * VX], e[1]);
            e[7] = fma_r(cj, p4[j * VX] - p4[-(j + 1) * VX], e[7]);
        }
        if (z >= j && z < nz - j - 1) e[8] = fma_r(cj, r3[5 + j] - r3[4 - j], e[8]);
        if (z >= j + 1 && z < nz - j) {
            e[2] = fma_r(cj, r5[4 + j] - r5[3 - j], e[2]);
            e[5] = fma_r(cj, r4[4 + j] - r4[3 - j], e[5]);
        }
    }
#pragma unroll
    for (int t = 0; t < 9; ++t) e[t] = e[t] * S.inv_h;
    ax = pml_t<2, 3, 1, M>(S, bp, z, iy, ix, e[0], ps, true) +
         pml_t<1, 3, 0, M>(S, bp, z, iy, ix, e[1], ps, true) +
         pml_t<0, 3, 0, M>(S, bp, z, iy, ix, e[2], ps, true);
    ay = pml_t<2, 4, 0, M>(S, bp, z, iy, ix, e[3], ps, true) +
         pml_t<1, 4, 1, M>(S, bp, z, iy, ix, e[4], ps, true) +
         pml_t<0, 4, 0, M>(S, bp, z, iy, ix, e[5], ps, true);
    az = pml_t<2, 5, 0, M>(S, bp, z, iy, ix, e[6], ps, true) +
         pml_t<1, 5, 0, M>(S, bp, z, iy, ix, e[7], ps, true) +
         pml_t<0, 5, 1, M>(S, bp, z, iy, ix, e[8], ps, true);
}

// RHO_LDS: the density table sits in LDS after the five plane tiles, so the
// face-density lookups of the output stage wait on LDS only.  A global
// lookup there would make the wave wait for every prefetch issued before it
// (vector memory counters retire in order).
// PM: the CPML band kernel of the tile, as in stress_tiled_body.
template <typename R, int TX, int TY, bool RHO_LDS = false, int PM = 0>
__device__ __forceinline__ void velocity_tiled_body(const Shot<R> &S, int tiles_x, int tiles_y,
                                                    int zc_len, int t, int zlo, int zhi)
{
    using T = VelTile<R, TX, TY>;
    constexpr bool PML = PM != 0;
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
    // PML: band profiles after the density table
    R *bp = st + 5 * NV + (RHO_LDS ? S.n_mat : 0);
    if constexpr (PML) pml_profiles_to_lds(S, bp, NT);

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
    const bool pml_xy = ((PM & 4) && pml_in(ix, nx, S.pml)) || ((PM & 2) && pml_in(iy, ny, S.pml));
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
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            r3[k] = r3[k + 1];
            r4[k] = r4[k + 1];
            r5[k] = r5[k + 1];
        }
        r3[8] = p3;
        r4[8] = p4;
        r5[8] = p5;
        // CPML band: this plane's memory variables, loaded ahead of the
        // step's prefetches
        R ps[9];
        bool band = false;
        if constexpr (PML) {
            band = pml_xy || ((PM & 1) && pml_in(z, nz, S.pml));
            if (band && col_in) pml_psi_load<PM>(S, z, iy, ix, 3, ps);
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

        if (col_in) {
            R ax = 0, ay = 0, az = 0;
            // a CPML band cell's accelerations are in units of 1/h already
            R tstep = S.dt_h;
            if (band) {
                if constexpr (PML) {
                    accel_col_pml<R, VX, PM>(S, bp, ps, t1 + vc, t2 + vc, t4 + vc, t5 + vc, t6 + vc,
                                         r3, r4, r5, ix, iy, z, ax, ay, az);
                    tstep = S.dt;
                }
            } else if (inner_xy && z >= 4 && z <= nz - 5)
                accel_col<false, R, VX>(S, t1 + vc, t2 + vc, t4 + vc, t5 + vc, t6 + vc, r3, r4, r5,
                                        ix, iy, z, ax, ay, az);
            else
                accel_col<true, R, VX>(S, t1 + vc, t2 + vc, t4 + vc, t5 + vc, t6 + vc, r3, r4, r5,
                                       ix, iy, z, ax, ay, az);
            const R *rho = rho_tab;
            const R r0 = rho[m0];
            const R rvx = ix < nx - 1 ? (R)0.5 * (r0 + rho[mxp]) : r0;
            const R rvy = iy < ny - 1 ? (R)0.5 * (r0 + rho[myp]) : r0;
            const R rvz = z < nz - 1 ? (R)0.5 * (r0 + rho[mzp]) : r0;
            stg(S.v[0], i, (vxo + tstep * ax / rvx) * dmp);
            stg(S.v[1], i, (vyo + tstep * ay / rvy) * dmp);
            stg(S.v[2], i, (vzo + tstep * az / rvz) * dmp);
        }
        lds_barrier();
    };
    // two planes per iteration: the loads one plane issues feed the next
    // without the register moves of a loop latch, each of which waits for
    // its load (measured 0.481 -> 0.457 ms at 400x400x200, 64x8 tiles; the
    // kernel stays within the registers of its occupancy)
    int z = z0;
    for (; z + 1 < z1; z += 2) {
        step(z);
        step(z + 1);
    }
    if (z < z1) step(z);
}

// ------------------------------------------------------------- launchers
// Single-shot kernels.  With a CPML boundary each tile is given to one of
// three kernels by the bands it meets: 0 (none: the plain kernel), 1 (a z
// band only) and 7 (an x or y band, maybe also a z band); the template
// argument PM is the mask of axes whose memory variables the kernel compiles
// (bit 0 z, bit 1 y, bit 2 x).  Each kernel exits for the tiles of the
// others, so launches of the same geometry over the same planes split the
// tiles between them.  Kernel 7 may use twice the registers (its band path
// needs them); kernel 1 keeps the plain kernel's register bound, and so its
// occupancy.  (Kernels for the x and y bands apart, 3 and 5, were measured
// slower: their launches run too few tiles to fill the device.)
inline __device__ int band_kernel(bool x, bool y, bool z)
{
    return x || y ? 7 : (z ? 1 : 0);
}

template <typename R, int TX, int TY>
__device__ __forceinline__ int stress_band_kernel(const Shot<R> &S, int t, int tiles_x,
                                                  int tiles_y, int zc_len, int zlo, int zhi)
{
    // strain columns x0-1 .. x0+TX, y0-1 .. y0+TY; planes z0-1 .. z1
    const int L = S.pml;
    const int x0 = (t % tiles_x) * TX, y0 = ((t / tiles_x) % tiles_y) * TY;
    const int z0 = zlo + (t / (tiles_x * tiles_y)) * zc_len, z1 = min(zhi, z0 + zc_len);
    return band_kernel(x0 - 1 <= L || x0 + TX >= S.nx - 1 - L,
                       y0 - 1 <= L || y0 + TY >= S.ny - 1 - L, z0 - 1 <= L || z1 >= S.nz - 1 - L);
}

template <typename R, int TX, int TY>
__device__ __forceinline__ int velocity_band_kernel(const Shot<R> &S, int t, int tiles_x,
                                                    int tiles_y, int zc_len, int zlo, int zhi)
{
    const int L = S.pml;
    const int x0 = (t % tiles_x) * TX, y0 = ((t / tiles_x) % tiles_y) * TY;
    const int z0 = zlo + (t / (tiles_x * tiles_y)) * zc_len, z1 = min(zhi, z0 + zc_len);
    return band_kernel(x0 <= L || x0 + TX - 1 >= S.nx - 1 - L,
                       y0 <= L || y0 + TY - 1 >= S.ny - 1 - L, z0 <= L || z1 - 1 >= S.nz - 1 - L);
}

// Register bound of band kernel PM: the plain kernel's (MinWaves) but for
// kernel 7.
template <typename R, int NT, int PM> struct ClassWaves {
    static constexpr int v = PM == 7 ? 1 : MinWaves<R, NT>::v;
};

template <typename R, int TX, int TY, int PAD, bool STORE_G = false, bool FR = false, int PM = 0>
__global__ __launch_bounds__((StressTile<R, TX, TY, PAD>::NT),
                             (ClassWaves<R, StressTile<R, TX, TY, PAD>::NT, PM>::v)) void
k_stress_tiled(Shot<R> S, int tiles_x, int tiles_y, int zc_len, int zlo, int zhi)
{
    const int nb = tiles_x * tiles_y * ((zhi - zlo + zc_len - 1) / zc_len);
    const int t = xcd_remap(blockIdx.x, nb);
    if (S.pml > 0 &&
        stress_band_kernel<R, TX, TY>(S, t, tiles_x, tiles_y, zc_len, zlo, zhi) != PM)
        return;
    stress_tiled_body<R, TX, TY, PAD, STORE_G, FR, PM>(S, tiles_x, tiles_y, zc_len, t, zlo, zhi);
}

template <typename R, int TX, int TY, bool RHO_LDS = false, int PM = 0>
__global__ __launch_bounds__(TX *TY, (ClassWaves<R, TX * TY, PM>::v)) void k_velocity_tiled(
    Shot<R> S, int tiles_x, int tiles_y, int zc_len, int zlo, int zhi)
{
    const int nb = tiles_x * tiles_y * ((zhi - zlo + zc_len - 1) / zc_len);
    const int t = xcd_remap(blockIdx.x, nb);
    if (S.pml > 0 &&
        velocity_band_kernel<R, TX, TY>(S, t, tiles_x, tiles_y, zc_len, zlo, zhi) != PM)
        return;
    velocity_tiled_body<R, TX, TY, RHO_LDS, PM>(S, tiles_x, tiles_y, zc_len, t, zlo, zhi);
}

template <typename R, int TX, int TY>
__global__ __launch_bounds__((StressTile<R, TX, TY, 0>::NT),
                             (MinWaves<R, StressTile<R, TX, TY, 0>::NT>::v)) void
k_stress_pipe(Shot<R> S, int tiles_x, int tiles_y, int zc_len, int zlo, int zhi)
{
    const int nb = tiles_x * tiles_y * ((zhi - zlo + zc_len - 1) / zc_len);
    stress_pipe_body<R, TX, TY>(S, tiles_x, tiles_y, zc_len, xcd_remap(blockIdx.x, nb), zlo, zhi);
}

// Multi-shot kernels: the workgroups of every shot of a batch in one launch.
// blk_shot maps a (remapped) workgroup id to its shot; a shot whose step
// counter has reached its step count exits at once.  A batch with CPML shots
// runs the band kernels PM = 0, 1, 7 one after another over every
// workgroup; a tile of a CPML shot runs in the kernel of its band class
// (stress_band_kernel) and a sponge shot's tiles in kernel 0.  The memory
// variables' buffer parity is that of the shot's own step counter (the entry
// holds the shot once per parity, so the kernel reads it in place: a
// modified local copy would be kept in scratch memory).
template <typename R, int TX, int TY, int PAD, bool FR = false, int PM = 0>
__global__ __launch_bounds__((StressTile<R, TX, TY, PAD>::NT),
                             (ClassWaves<R, StressTile<R, TX, TY, PAD>::NT, PM>::v)) void
k_stress_batched(const BatchEntry<R> *__restrict__ tab, const int32_t *__restrict__ blk_shot,
                 int nb_total, const int32_t *__restrict__ step)
{
    const int g = xcd_remap(blockIdx.x, nb_total);
    const int s = blk_shot[g];
    const BatchEntry<R> *e = tab + s;
    const int n = step[s];
    if (n >= e->nt) return;
    const Shot<R> &S = (PM != 0 && (n & 1)) ? e->S_odd : e->S;
    const int t = g - e->blk_off[0];
    if (S.pml > 0) {
        if (stress_band_kernel<R, TX, TY>(S, t, e->tiles_x[0], e->tiles_y[0], e->zc_len[0], 0,
                                          S.nz) != PM)
            return;
    } else if (PM != 0) {
        return;
    }
    stress_tiled_body<R, TX, TY, PAD, false, FR, PM>(S, e->tiles_x[0], e->tiles_y[0],
                                                     e->zc_len[0], t, 0, S.nz);
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

template <typename R, int TX, int TY, bool RHO_LDS = false, int PM = 0>
__global__ __launch_bounds__(TX *TY, (ClassWaves<R, TX * TY, PM>::v)) void k_velocity_batched(
    const BatchEntry<R> *__restrict__ tab, const int32_t *__restrict__ blk_shot, int nb_total,
    const int32_t *__restrict__ step)
{
    const int g = xcd_remap(blockIdx.x, nb_total);
    const int s = blk_shot[g];
    const BatchEntry<R> *e = tab + s;
    const int n = step[s];
    if (n >= e->nt) return;
    const Shot<R> &S = (PM != 0 && (n & 1)) ? e->S_odd : e->S;
    const int t = g - e->blk_off[1];
    if (S.pml > 0) {
        if (velocity_band_kernel<R, TX, TY>(S, t, e->tiles_x[1], e->tiles_y[1], e->zc_len[1], 0,
                                            S.nz) != PM)
            return;
    } else if (PM != 0) {
        return;
    }
    velocity_tiled_body<R, TX, TY, RHO_LDS, PM>(S, e->tiles_x[1], e->tiles_y[1], e->zc_len[1], t,
                                                0, S.nz);
}

// A second stream per device (default priority, created once) on which a
// CPML half-step runs its full band kernel (tiles in an x or y band) while
// the plain and z-band launches run on the caller's stream; the sets of
// tiles write disjoint cells.  Calls that
// overlap on one device share it: a caller holds `use` from the fork record
// to the join wait, so host threads enqueueing at the same time cannot
// interleave their fork/join records (the lock covers enqueueing only).
struct SideStream {
    hipStream_t s = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    std::mutex use;
};

inline hipError_t side_stream(SideStream **out)
{
    static std::mutex mu;
    static SideStream per_dev[64];
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= 64) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(mu);
    SideStream &ss = per_dev[dev];
    if (!ss.s) {
        e = hipStreamCreateWithFlags(&ss.s, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&ss.fork, hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&ss.join, hipEventDisableTiming);
        if (e != hipSuccess) return e;
    }
    *out = &ss;
    return hipSuccess;
}

// CPML z ranges: the middle planes [m0, m1), whose tiles lie in a band only
// through x or y, and the two slabs that contain the z bands.  A stress tile
// of planes z0..z1-1 computes strains on z0-1 .. z1, so its middle range is
// [L+2, nz-L-2); a velocity tile touches only its own planes: [L+1, nz-L-1).
inline void pml_zranges(int nz, int L, int half, int &m0, int &m1)
{
    m0 = half == 0 ? L + 2 : L + 1;
    m1 = half == 0 ? nz - L - 2 : nz - L - 1;
    if (m1 <= m0) m0 = m1 = nz;                  // no band-free middle
}

template <typename R, int TX, int TY, int PAD, bool FR = false, bool PML = false>
hipError_t stress_launch(const Shot<R> &s, int zchunk, hipStream_t st, bool store_g = false)
{
    using T = StressTile<R, TX, TY, PAD, FR>;
    const int tiles_x = (s.nx + TX - 1) / TX, tiles_y = (s.ny + TY - 1) / TY;
    auto run = [&](auto kern, size_t lds, int lo, int hi) -> hipError_t {
        if (hi <= lo) return hipSuccess;
        const int zl = chunk_len(hi - lo, tiles_x * tiles_y, zchunk);
        const int nb = tiles_x * tiles_y * ((hi - lo + zl - 1) / zl);
        hipError_t e = allow_lds(kern, lds);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(kern, dim3(nb), dim3(T::NT), lds, st, s, tiles_x, tiles_y, zl, lo, hi);
        return hipGetLastError();
    };
    if constexpr (PML) {
        if (store_g) return hipErrorInvalidValue;   // the strain history serves the adjoint only
        const size_t lds_pml = T::LDS + sizeof(R) * 18 * (2 * (size_t)s.pml + 2);
        int m0, m1;
        pml_zranges(s.nz, s.pml, 0, m0, m1);
        // the caller's stream takes the band-free middle tiles (kernel 0)
        // and then the z-only tiles of the two slabs that hold the z bands
        // (kernel 1); concurrently the side stream takes the tiles in an x
        // or y band (kernel 7) in one launch over every plane
        const int lo_end = m0 < m1 ? m0 : s.nz;
        SideStream *ss = nullptr;
        hipError_t e = side_stream(&ss);
        if (e != hipSuccess) return e;
        std::lock_guard<std::mutex> hold(ss->use);
        e = hipEventRecord(ss->fork, st);
        if (e == hipSuccess) e = hipStreamWaitEvent(ss->s, ss->fork, 0);
        if (e == hipSuccess) e = run(k_stress_tiled<R, TX, TY, PAD, false, FR, 0>, T::LDS, m0, m1);
        if (e == hipSuccess) e = run(k_stress_tiled<R, TX, TY, PAD, false, FR, 1>, lds_pml, 0, lo_end);
        if (e == hipSuccess && m0 < m1)
            e = run(k_stress_tiled<R, TX, TY, PAD, false, FR, 1>, lds_pml, m1, s.nz);
        const hipStream_t main = st;
        st = ss->s;
        if (e == hipSuccess) e = run(k_stress_tiled<R, TX, TY, PAD, false, FR, 7>, lds_pml, 0, s.nz);
        st = main;
        if (e == hipSuccess) e = hipEventRecord(ss->join, ss->s);
        if (e == hipSuccess) e = hipStreamWaitEvent(st, ss->join, 0);
        return e;
    }
    if (store_g) return run(k_stress_tiled<R, TX, TY, PAD, true, FR>, T::LDS, 0, s.nz);
    return run(k_stress_tiled<R, TX, TY, PAD, false, FR>, T::LDS, 0, s.nz);
}

// Pipelined stress tile (stress_pipe_body); the strain-history store runs
// the plain kernel of the same tile.
template <typename R, int TX, int TY>
hipError_t stress_pipe_launch(const Shot<R> &s, int zchunk, hipStream_t st, bool store_g)
{
    if (store_g) return stress_launch<R, TX, TY, 0>(s, zchunk, st, true);
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

template <typename R, int TX, int TY, bool PML = false>
hipError_t velocity_launch(const Shot<R> &s, int zchunk, hipStream_t st)
{
    using T = VelTile<R, TX, TY>;
    const int tiles_x = (s.nx + TX - 1) / TX, tiles_y = (s.ny + TY - 1) / TY;
    auto run = [&](auto kern, size_t lds, int lo, int hi) -> hipError_t {
        if (hi <= lo) return hipSuccess;
        const int zl = chunk_len(hi - lo, tiles_x * tiles_y, zchunk);
        const int nb = tiles_x * tiles_y * ((hi - lo + zl - 1) / zl);
        hipError_t e = allow_lds(kern, lds);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(kern, dim3(nb), dim3(T::NT), lds, st, s, tiles_x, tiles_y, zl, lo, hi);
        return hipGetLastError();
    };
    const size_t rho_bytes = sizeof(R) * (size_t)s.n_mat;
    const bool rho_lds = rho_bytes <= (size_t)kRhoLdsBytes;
    const size_t lds = T::LDS + (rho_lds ? rho_bytes : 0);
    if constexpr (PML) {
        const size_t lds_pml = lds + sizeof(R) * 18 * (2 * (size_t)s.pml + 2);
        int m0, m1;
        pml_zranges(s.nz, s.pml, 1, m0, m1);
        const int lo_end = m0 < m1 ? m0 : s.nz;
        // the launch sequence of stress_launch: kernels 0 and 1 on the
        // caller's stream, kernel 7 on the side stream
        auto seq = [&](auto k0, auto k1, auto k7) -> hipError_t {
            SideStream *ss = nullptr;
            hipError_t e = side_stream(&ss);
            if (e != hipSuccess) return e;
            std::lock_guard<std::mutex> hold(ss->use);
            e = hipEventRecord(ss->fork, st);
            if (e == hipSuccess) e = hipStreamWaitEvent(ss->s, ss->fork, 0);
            if (e == hipSuccess) e = run(k0, lds, m0, m1);
            if (e == hipSuccess) e = run(k1, lds_pml, 0, lo_end);
            if (e == hipSuccess && m0 < m1) e = run(k1, lds_pml, m1, s.nz);
            const hipStream_t main = st;
            st = ss->s;
            if (e == hipSuccess) e = run(k7, lds_pml, 0, s.nz);
            st = main;
            if (e == hipSuccess) e = hipEventRecord(ss->join, ss->s);
            if (e == hipSuccess) e = hipStreamWaitEvent(st, ss->join, 0);
            return e;
        };
        if (rho_lds)
            return seq(k_velocity_tiled<R, TX, TY, true, 0>, k_velocity_tiled<R, TX, TY, true, 1>,
                       k_velocity_tiled<R, TX, TY, true, 7>);
        return seq(k_velocity_tiled<R, TX, TY, false, 0>, k_velocity_tiled<R, TX, TY, false, 1>,
                   k_velocity_tiled<R, TX, TY, false, 7>);
    }
    if (rho_lds) return run(k_velocity_tiled<R, TX, TY, true>, lds, 0, s.nz);
    return run(k_velocity_tiled<R, TX, TY>, lds, 0, s.nz);
}

template <typename R, int TX, int TY, int PAD, bool FR = false>
hipError_t stress_batched_launch(const BatchEntry<R> *tab, const int32_t *blk, int nb,
                                 const int32_t *step, hipStream_t st)
{
    using T = StressTile<R, TX, TY, PAD, FR>;
    hipError_t e = allow_lds(k_stress_batched<R, TX, TY, PAD, FR>, T::LDS);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_stress_batched<R, TX, TY, PAD, FR>), dim3(nb), dim3(T::NT), T::LDS, st,
                       tab, blk, nb, step);
    return hipGetLastError();
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

template <typename R, int TX, int TY>
hipError_t velocity_batched_launch(const BatchEntry<R> *tab, const int32_t *blk, int nb,
                                   const int32_t *step, hipStream_t st, int max_mat)
{
    using T = VelTile<R, TX, TY>;
    const size_t rho_bytes = sizeof(R) * (size_t)max_mat;
    if (rho_bytes <= (size_t)kRhoLdsBytes) {
        const size_t lds = T::LDS + rho_bytes;
        hipError_t e = allow_lds(k_velocity_batched<R, TX, TY, true>, lds);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL((k_velocity_batched<R, TX, TY, true>), dim3(nb), dim3(T::NT), lds, st,
                           tab, blk, nb, step);
        return hipGetLastError();
    }
    hipError_t e = allow_lds(k_velocity_batched<R, TX, TY>, T::LDS);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_velocity_batched<R, TX, TY>), dim3(nb), dim3(T::NT), T::LDS, st, tab,
                       blk, nb, step);
    return hipGetLastError();
}

// CPML batches: the three band kernels over every workgroup, variant-0
// tiles; max_pml sizes the LDS band-profile table for the widest band.  As
// for a single shot, kernels 0 and 1 run on the caller's stream and kernel 7
// concurrently on the side stream (fork / join by events, which a graph
// capture of the caller's stream follows).
template <typename R, int NT, typename K0, typename K1, typename K7>
hipError_t batched_pml_fork(K0 k0, K1 k1, K7 k7, size_t lds, size_t lds_pml,
                            const BatchEntry<R> *tab, const int32_t *blk, int nb,
                            const int32_t *step, hipStream_t st)
{
    auto go = [&](auto kern, size_t l, hipStream_t q) -> hipError_t {
        hipError_t e = allow_lds(kern, l);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(kern, dim3(nb), dim3(NT), l, q, tab, blk, nb, step);
        return hipGetLastError();
    };
    SideStream *ss = nullptr;
    hipError_t e = side_stream(&ss);
    if (e != hipSuccess) return e;
    std::lock_guard<std::mutex> hold(ss->use);
    e = hipEventRecord(ss->fork, st);
    if (e == hipSuccess) e = hipStreamWaitEvent(ss->s, ss->fork, 0);
    if (e == hipSuccess) e = go(k0, lds, st);
    if (e == hipSuccess) e = go(k1, lds_pml, st);
    if (e == hipSuccess) e = go(k7, lds_pml, ss->s);
    if (e == hipSuccess) e = hipEventRecord(ss->join, ss->s);
    if (e == hipSuccess) e = hipStreamWaitEvent(st, ss->join, 0);
    return e;
}

template <typename R, int TX, int TY, int PAD>
hipError_t stress_batched_pml_launch(const BatchEntry<R> *tab, const int32_t *blk, int nb,
                                     const int32_t *step, hipStream_t st, int max_pml)
{
    using T = StressTile<R, TX, TY, PAD>;
    const size_t lds_pml = T::LDS + sizeof(R) * 18 * (2 * (size_t)max_pml + 2);
    return batched_pml_fork<R, T::NT>(k_stress_batched<R, TX, TY, PAD, false, 0>,
                                      k_stress_batched<R, TX, TY, PAD, false, 1>,
                                      k_stress_batched<R, TX, TY, PAD, false, 7>, T::LDS, lds_pml,
                                      tab, blk, nb, step, st);
}

template <typename R, int TX, int TY>
hipError_t velocity_batched_pml_launch(const BatchEntry<R> *tab, const int32_t *blk, int nb,
                                       const int32_t *step, hipStream_t st, int max_mat,
                                       int max_pml)
{
    using T = VelTile<R, TX, TY>;
    const size_t rho_bytes = sizeof(R) * (size_t)max_mat;
    const bool rho_lds = rho_bytes <= (size_t)kRhoLdsBytes;
    const size_t lds = T::LDS + (rho_lds ? rho_bytes : 0);
    const size_t lds_pml = lds + sizeof(R) * 18 * (2 * (size_t)max_pml + 2);
    if (rho_lds)
        return batched_pml_fork<R, T::NT>(k_velocity_batched<R, TX, TY, true, 0>,
                                          k_velocity_batched<R, TX, TY, true, 1>,
                                          k_velocity_batched<R, TX, TY, true, 7>, lds, lds_pml,
                                          tab, blk, nb, step, st);
    return batched_pml_fork<R, T::NT>(k_velocity_batched<R, TX, TY, false, 0>,
                                      k_velocity_batched<R, TX, TY, false, 1>,
                                      k_velocity_batched<R, TX, TY, false, 7>, lds, lds_pml, tab,
                                      blk, nb, step, st);
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
                                   {32, 13, 0, 2},  {30, 14, 0, 2}};
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
                                        int nb, const int32_t *step, hipStream_t st)
{
    switch (v) {
    case 1: return stress_batched_launch<float, 62, 6, 0>(tab, blk, nb, step, st);
    case 2: return stress_batched_launch<float, 30, 14, 24>(tab, blk, nb, step, st);
    case 3: return stress_batched_launch<float, 32, 16, 24>(tab, blk, nb, step, st);
    case 4: return stress_batched_launch<float, 32, 8, 24>(tab, blk, nb, step, st);
    case 5: return stress_batched_launch<float, 16, 16, 0>(tab, blk, nb, step, st);
    case 6: return stress_batched_launch<float, 32, 16, 0>(tab, blk, nb, step, st);
    case 7: return stress_batched_launch<float, 30, 14, 24, true>(tab, blk, nb, step, st);
    case 8: return stress_batched_launch<float, 32, 13, 24, true>(tab, blk, nb, step, st);
    case 9: return stress_batched_launch<float, 30, 14, 0, true>(tab, blk, nb, step, st);
    case 10: return stress_batched_launch<float, 62, 14, 0>(tab, blk, nb, step, st);
    case 11: return stress_batched_launch<float, 62, 14, 0, true>(tab, blk, nb, step, st);
    case 12: return stress_pipe_batched_launch<float, 32, 13>(tab, blk, nb, step, st);
    case 13: return stress_pipe_batched_launch<float, 30, 14>(tab, blk, nb, step, st);
    default: return stress_batched_launch<float, 32, 13, 24>(tab, blk, nb, step, st);
    }
}

template <>
hipError_t launch_stress_batched<double>(int v, const BatchEntry<double> *tab, const int32_t *blk,
                                         int nb, const int32_t *step, hipStream_t st)
{
    switch (v) {
    case 1: return stress_batched_launch<double, 32, 4, 0>(tab, blk, nb, step, st);
    case 2: return stress_batched_launch<double, 16, 8, 0>(tab, blk, nb, step, st);
    case 3: return stress_batched_launch<double, 32, 6, 24>(tab, blk, nb, step, st);
    case 4: return stress_pipe_batched_launch<double, 32, 6>(tab, blk, nb, step, st);
    default: return stress_batched_launch<double, 32, 6, 0>(tab, blk, nb, step, st);
    }
}

template <>
hipError_t launch_velocity_batched<float>(int v, const BatchEntry<float> *tab, const int32_t *blk,
                                          int nb, const int32_t *step, hipStream_t st,
                                          int max_mat)
{
    switch (v) {
    case 1: return velocity_batched_launch<float, 32, 16>(tab, blk, nb, step, st, max_mat);
    case 2: return velocity_batched_launch<float, 32, 8>(tab, blk, nb, step, st, max_mat);
    case 3: return velocity_batched_launch<float, 64, 4>(tab, blk, nb, step, st, max_mat);
    case 4: return velocity_batched_launch<float, 128, 4>(tab, blk, nb, step, st, max_mat);
    case 5: return velocity_batched_launch<float, 32, 32>(tab, blk, nb, step, st, max_mat);
    default: return velocity_batched_launch<float, 64, 8>(tab, blk, nb, step, st, max_mat);
    }
}

template <>
hipError_t launch_velocity_batched<double>(int v, const BatchEntry<double> *tab,
                                           const int32_t *blk, int nb, const int32_t *step,
                                           hipStream_t st, int max_mat)
{
    switch (v) {
    case 1: return velocity_batched_launch<double, 64, 4>(tab, blk, nb, step, st, max_mat);
    case 2: return velocity_batched_launch<double, 32, 16>(tab, blk, nb, step, st, max_mat);
    case 3: return velocity_batched_launch<double, 32, 4>(tab, blk, nb, step, st, max_mat);
    default: return velocity_batched_launch<double, 32, 8>(tab, blk, nb, step, st, max_mat);
    }
}

template <>
hipError_t launch_stress_batched_pml<float>(const BatchEntry<float> *tab, const int32_t *blk,
                                            int nb, const int32_t *step, hipStream_t st,
                                            int max_pml)
{
    return stress_batched_pml_launch<float, 32, 13, 24>(tab, blk, nb, step, st, max_pml);
}

template <>
hipError_t launch_stress_batched_pml<double>(const BatchEntry<double> *tab, const int32_t *blk,
                                             int nb, const int32_t *step, hipStream_t st,
                                             int max_pml)
{
    return stress_batched_pml_launch<double, 32, 6, 0>(tab, blk, nb, step, st, max_pml);
}

template <>
hipError_t launch_velocity_batched_pml<float>(const BatchEntry<float> *tab, const int32_t *blk,
                                              int nb, const int32_t *step, hipStream_t st,
                                              int max_mat, int max_pml)
{
    return velocity_batched_pml_launch<float, 64, 8>(tab, blk, nb, step, st, max_mat, max_pml);
}

template <>
hipError_t launch_velocity_batched_pml<double>(const BatchEntry<double> *tab, const int32_t *blk,
                                               int nb, const int32_t *step, hipStream_t st,
                                               int max_mat, int max_pml)
{
    return velocity_batched_pml_launch<double, 32, 8>(tab, blk, nb, step, st, max_mat, max_pml);
}

template <>
hipError_t launch_stress_tiled<float>(const Shot<float> &s, int v, int zc, hipStream_t st,
                                      bool g)
{
    // CPML shots run the variant-0 tile shape whatever the variant (6-wave
    // band tiles at 168 VGPRs measured 1.5-1.8x slower)
    if (s.pml > 0) return stress_launch<float, 32, 13, 24, false, true>(s, zc, st, g);
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
    default: return stress_launch<float, 32, 13, 24>(s, zc, st, g);
    }
}

template <>
hipError_t launch_stress_tiled<double>(const Shot<double> &s, int v, int zc, hipStream_t st,
                                       bool g)
{
    if (s.pml > 0) return stress_launch<double, 32, 6, 0, false, true>(s, zc, st, g);
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
    // CPML shots: the band path is compiled for variants 0, 2 and 3 (the
    // 4-wave tiles fit three band workgroups per CU); others run variant 0
    if (s.pml > 0) {
        if (v == 2) return velocity_launch<float, 32, 8, true>(s, zc, st);
        if (v == 3) return velocity_launch<float, 64, 4, true>(s, zc, st);
        return velocity_launch<float, 64, 8, true>(s, zc, st);
    }
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
    if (s.pml > 0) return velocity_launch<double, 32, 8, true>(s, zc, st);
    switch (v) {
    case 1: return velocity_launch<double, 64, 4>(s, zc, st);
    case 2: return velocity_launch<double, 32, 16>(s, zc, st);
    case 3: return velocity_launch<double, 32, 4>(s, zc, st);
    default: return velocity_launch<double, 32, 8>(s, zc, st);
    }
}

}  // namespace pe
