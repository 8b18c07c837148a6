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
#include "common.hpp"

namespace pe {

namespace {

constexpr int kXcd = 8;

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
template <typename R, int TX, int TY>
struct StressTile {
    static constexpr int GX = TX + 2, GY = TY + 2, NG = GX * GY;   // strain columns
    static constexpr int NT = (NG + 63) / 64 * 64;                 // threads
    static constexpr int VX = GX + 8, VY = GY + 8, NV = VX * VY;   // v plane tile
    static constexpr int NFRAME = 8 * (VX + GY);                   // v tile outside G
    static constexpr size_t LDS = sizeof(R) * (3 * NV + 6 * 3 * NG);
};

// Moved strain q_{P,J} read from the LDS strain ring (see moved_direct in
// kernels_direct.hip for the rule).  sl[d+1] is the ring slot of plane z+d.
template <int P, int J, typename R, int GX, int NG>
__device__ __forceinline__ R moved_lds(const R *__restrict__ gt, const int (&sl)[3], int c,
                                       int iz, int iy, int ix, int nz, int ny, int nx)
{
    const R *gJ = gt + J * 3 * NG;
    constexpr int src = kVP[J];
    if constexpr (src == P) {
        return gJ[sl[1] * NG + c];
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
        return ok ? acc * (ntap == 4 ? (R)0.25 : (R)0.5) : (R)0;
    }
}

template <int I, typename R>
__device__ __forceinline__ void stress_out(const Shot<R> &S, const R *__restrict__ Cm,
                                           const R (&q)[6], R dmp, R &sv)
{
    R acc = 0;
    bool any = false;
#pragma unroll
    for (int J = 0; J < 6; ++J) {
        if ((S.act >> (6 * I + J)) & 1ull) {
            acc = fma_r(Cm[6 * I + J], q[J], acc);
            any = true;
        }
    }
    sv = sv * dmp;
    if (any) sv = fma_r(S.dt, acc, sv);
}

template <int P, typename R, int GX, int NG>
__device__ __forceinline__ void gather_lds(const R *__restrict__ gt, const int (&sl)[3], int c,
                                           int iz, int iy, int ix, int nz, int ny, int nx,
                                           R (&q)[6])
{
    q[0] = moved_lds<P, 0, R, GX, NG>(gt, sl, c, iz, iy, ix, nz, ny, nx);
    q[1] = moved_lds<P, 1, R, GX, NG>(gt, sl, c, iz, iy, ix, nz, ny, nx);
    q[2] = moved_lds<P, 2, R, GX, NG>(gt, sl, c, iz, iy, ix, nz, ny, nx);
    q[3] = moved_lds<P, 3, R, GX, NG>(gt, sl, c, iz, iy, ix, nz, ny, nx);
    q[4] = moved_lds<P, 4, R, GX, NG>(gt, sl, c, iz, iy, ix, nz, ny, nx);
    q[5] = moved_lds<P, 5, R, GX, NG>(gt, sl, c, iz, iy, ix, nz, ny, nx);
}

template <typename R, int TX, int TY>
__global__ __launch_bounds__((StressTile<R, TX, TY>::NT)) void k_stress_tiled(Shot<R> S, int tiles_x,
                                                                           int tiles_y, int zc_len)
{
    using T = StressTile<R, TX, TY>;
    constexpr int GX = T::GX, GY = T::GY, NG = T::NG, VX = T::VX, VY = T::VY, NV = T::NV;
    extern __shared__ unsigned char lds_raw[];
    R *vt = reinterpret_cast<R *>(lds_raw);       // [3][VY][VX]
    R *gt = vt + 3 * NV;                           // [6][3][GY][GX]

    const int nz = S.nz, ny = S.ny, nx = S.nx;
    const int sy = nx, sz = nx * ny;
    const int nb = tiles_x * tiles_y * ((nz + zc_len - 1) / zc_len);
    const int t = xcd_remap(blockIdx.x, nb);
    const int tx = t % tiles_x, ty = (t / tiles_x) % tiles_y, chunk = t / (tiles_x * tiles_y);
    const int x0 = tx * TX, y0 = ty * TY;
    const int z0 = chunk * zc_len, z1 = min(nz, z0 + zc_len);

    const int tid = threadIdx.x;
    const bool gact = tid < NG;
    const int gx = gact ? tid % GX : 0, gy = gact ? tid / GX : 0;
    const int ix = x0 - 1 + gx, iy = y0 - 1 + gy;                  // column of this thread
    const bool col_in = gact && ix >= 0 && ix < nx && iy >= 0 && iy < ny;
    const bool core = col_in && gx >= 1 && gx <= TX && gy >= 1 && gy <= TY;
    const int col = col_in ? iy * sy + ix : 0;
    const int vc = (gy + 4) * VX + (gx + 4);                        // this column in vt

    const R *__restrict__ vx = S.v[0];
    const R *__restrict__ vy = S.v[1];
    const R *__restrict__ vz = S.v[2];

    // z rings: r?[k] = v at plane zc - 4 + k
    R rx[9], ry[9], rz[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        const int z = z0 - 5 + k;
        const bool ok = col_in && z >= 0 && z < nz;
        rx[k] = ld(vx, col + z * sz, ok);
        ry[k] = ld(vy, col + z * sz, ok);
        rz[k] = ld(vz, col + z * sz, ok);
    }

    for (int zc = z0 - 1; zc <= z1; ++zc) {
        const bool plane_in = zc >= 0 && zc < nz;
        // (1) plane zc of v into LDS: own column from the ring, frame from HBM/L2
        if (gact) {
            vt[0 * NV + vc] = rx[4];
            vt[1 * NV + vc] = ry[4];
            vt[2 * NV + vc] = rz[4];
        }
        for (int e = tid; e < T::NFRAME; e += T::NT) {
            int r, c;
            if (e < 4 * VX) {
                r = e / VX;
                c = e % VX;
            } else if (e < 8 * VX) {
                r = VY - 4 + (e - 4 * VX) / VX;
                c = (e - 4 * VX) % VX;
            } else if (e < 8 * VX + 4 * GY) {
                r = 4 + (e - 8 * VX) / 4;
                c = (e - 8 * VX) % 4;
            } else {
                r = 4 + (e - 8 * VX - 4 * GY) / 4;
                c = VX - 4 + (e - 8 * VX - 4 * GY) % 4;
            }
            const int gy_ = y0 - 5 + r, gx_ = x0 - 5 + c;
            const bool ok = plane_in && gy_ >= 0 && gy_ < ny && gx_ >= 0 && gx_ < nx;
            const int gi = zc * sz + gy_ * sy + gx_;
            vt[0 * NV + r * VX + c] = ld(vx, gi, ok);
            vt[1 * NV + r * VX + c] = ld(vy, gi, ok);
            vt[2 * NV + r * VX + c] = ld(vz, gi, ok);
        }
        // prefetch of the ring's next plane, consumed after the stress update
        const int zn = zc + 5;
        const bool okn = col_in && zn >= 0 && zn < nz;
        const R nxv = ld(vx, col + zn * sz, okn);
        const R nyv = ld(vy, col + zn * sz, okn);
        const R nzv = ld(vz, col + zn * sz, okn);
        // loads of the output plane's stress, label and damping
        const int zo = zc - 1;
        const bool out = core && zo >= z0 && zo < z1;
        const int io = zo * sz + col;
        R sv[6];
        int m = 0;
        R dmp = (R)1;
        if (out) {
#pragma unroll
            for (int k = 0; k < 6; ++k) sv[k] = S.s[k][io];
            m = S.mats[io];
            if (S.damp) dmp = S.damp[io];
        }
        __syncthreads();

        // (2) strains of plane zc at every column of the tile and its ring
        const int slot = (zc + 3) % 3;
        if (gact) {
            R a1 = 0, a2 = 0, a3 = 0, a4 = 0, a5 = 0, a6 = 0;
            if (col_in && plane_in) {
                const R *tx_ = vt + 0 * NV + vc;
                const R *ty_ = vt + 1 * NV + vc;
                const R *tz_ = vt + 2 * NV + vc;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    if (j >= S.nc) break;
                    const R cj = S.c[j];
                    if (ix >= j + 1 && ix < nx - j) a1 = fma_r(cj, tx_[j] - tx_[-(j + 1)], a1);
                    if (iy >= j + 1 && iy < ny - j)
                        a2 = fma_r(cj, ty_[j * VX] - ty_[-(j + 1) * VX], a2);
                    if (zc >= j + 1 && zc < nz - j) a3 = fma_r(cj, rz[4 + j] - rz[3 - j], a3);
                    if (iy >= j && iy < ny - j - 1)
                        a4 = fma_r(cj, tz_[(j + 1) * VX] - tz_[-j * VX], a4);
                    if (zc >= j && zc < nz - j - 1) a4 = fma_r(cj, ry[5 + j] - ry[4 - j], a4);
                    if (ix >= j && ix < nx - j - 1) a5 = fma_r(cj, tz_[j + 1] - tz_[-j], a5);
                    if (zc >= j && zc < nz - j - 1) a5 = fma_r(cj, rx[5 + j] - rx[4 - j], a5);
                    if (ix >= j && ix < nx - j - 1) a6 = fma_r(cj, ty_[j + 1] - ty_[-j], a6);
                    if (iy >= j && iy < ny - j - 1)
                        a6 = fma_r(cj, tx_[(j + 1) * VX] - tx_[-j * VX], a6);
                }
                a1 = a1 * S.inv_h;
                a2 = a2 * S.inv_h;
                a3 = a3 * S.inv_h;
                a4 = a4 * S.inv_h;
                a5 = a5 * S.inv_h;
                a6 = a6 * S.inv_h;
            }
            R *gs = gt + slot * NG + tid;
            gs[0 * 3 * NG] = a1;
            gs[1 * 3 * NG] = a2;
            gs[2 * 3 * NG] = a3;
            gs[3 * 3 * NG] = a4;
            gs[4 * 3 * NG] = a5;
            gs[5 * 3 * NG] = a6;
        }
        __syncthreads();

        // (3) stress of plane zc - 1 from the strain planes zc-2, zc-1, zc
        if (out) {
            const int sl[3] = {(zc + 1) % 3, (zc + 2) % 3, slot};
            const R *__restrict__ Cm = S.Ct + (size_t)m * 36;
            R q[6];
            gather_lds<0, R, GX, NG>(gt, sl, tid, zo, iy, ix, nz, ny, nx, q);
            stress_out<0>(S, Cm, q, dmp, sv[0]);
            stress_out<1>(S, Cm, q, dmp, sv[1]);
            stress_out<2>(S, Cm, q, dmp, sv[2]);
            gather_lds<1, R, GX, NG>(gt, sl, tid, zo, iy, ix, nz, ny, nx, q);
            stress_out<3>(S, Cm, q, dmp, sv[3]);
            gather_lds<2, R, GX, NG>(gt, sl, tid, zo, iy, ix, nz, ny, nx, q);
            stress_out<4>(S, Cm, q, dmp, sv[4]);
            gather_lds<3, R, GX, NG>(gt, sl, tid, zo, iy, ix, nz, ny, nx, q);
            stress_out<5>(S, Cm, q, dmp, sv[5]);
#pragma unroll
            for (int k = 0; k < 6; ++k) S.s[k][io] = sv[k];
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            rx[k] = rx[k + 1];
            ry[k] = ry[k + 1];
            rz[k] = rz[k + 1];
        }
        rx[8] = nxv;
        ry[8] = nyv;
        rz[8] = nzv;
    }
}

// ---------------------------------------------------------------- velocity
template <typename R, int TX, int TY>
struct VelTile {
    static constexpr int NT = TX * TY;
    static constexpr int VX = TX + 8, VY = TY + 8, NV = VX * VY;   // s plane tiles
    static constexpr size_t LDS = sizeof(R) * 5 * NV;
};

template <typename R, int TX, int TY>
__global__ __launch_bounds__(TX *TY) void k_velocity_tiled(Shot<R> S, int tiles_x, int tiles_y,
                                                          int zc_len)
{
    using T = VelTile<R, TX, TY>;
    constexpr int VX = T::VX, NV = T::NV;
    extern __shared__ unsigned char lds_raw[];
    R *st = reinterpret_cast<R *>(lds_raw);   // [5][VY][VX]: s1, s2, s4, s5, s6
    R *t1 = st + 0 * NV;
    R *t2 = st + 1 * NV;
    R *t4 = st + 2 * NV;
    R *t5 = st + 3 * NV;
    R *t6 = st + 4 * NV;

    const int nz = S.nz, ny = S.ny, nx = S.nx;
    const int sy = nx, sz = nx * ny;
    const int nb = tiles_x * tiles_y * ((nz + zc_len - 1) / zc_len);
    const int t = xcd_remap(blockIdx.x, nb);
    const int tx = t % tiles_x, ty = (t / tiles_x) % tiles_y, chunk = t / (tiles_x * tiles_y);
    const int x0 = tx * TX, y0 = ty * TY;
    const int z0 = chunk * zc_len, z1 = min(nz, z0 + zc_len);

    const int tid = threadIdx.x;
    const int lx = tid % TX, ly = tid / TX;
    const int ix = x0 + lx, iy = y0 + ly;
    const bool col_in = ix < nx && iy < ny;
    const int col = col_in ? iy * sy + ix : 0;
    const int vc = (ly + 4) * VX + (lx + 4);

    const R *__restrict__ s1 = S.s[0];
    const R *__restrict__ s2 = S.s[1];
    const R *__restrict__ s3 = S.s[2];
    const R *__restrict__ s4 = S.s[3];
    const R *__restrict__ s5 = S.s[4];
    const R *__restrict__ s6 = S.s[5];

    // z rings: r?[k] = s? at plane z - 4 + k
    R r3[9], r4[9], r5[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        const int z = z0 - 4 + k;
        const bool ok = col_in && z >= 0 && z < nz;
        r3[k] = ld(s3, col + z * sz, ok);
        r4[k] = ld(s4, col + z * sz, ok);
        r5[k] = ld(s5, col + z * sz, ok);
    }
    int mnext = col_in && z0 < nz ? S.mats[z0 * sz + col] : 0;

    for (int z = z0; z < z1; ++z) {
        const int i = z * sz + col;
        // (1) plane z of s1, s2, s4, s5, s6 into LDS
        if (col_in) {
            t1[vc] = s1[i];
            t2[vc] = s2[i];
            t6[vc] = s6[i];
        } else {
            t1[vc] = (R)0;
            t2[vc] = (R)0;
            t6[vc] = (R)0;
        }
        t4[vc] = r4[4];
        t5[vc] = r5[4];
        // x frame (4 columns each side of the tile rows): s1, s5, s6
        for (int e = tid; e < 8 * TY; e += T::NT) {
            const int r = e / 8, k = e % 8;
            const int c = k < 4 ? k : TX + k;             // 0..3 and TX+4..TX+7
            const int gx_ = x0 - 4 + c, gy_ = y0 + r;
            const bool ok = gx_ >= 0 && gx_ < nx && gy_ < ny;
            const int gi = z * sz + gy_ * sy + gx_;
            const int li = (r + 4) * VX + c;
            t1[li] = ld(s1, gi, ok);
            t5[li] = ld(s5, gi, ok);
            t6[li] = ld(s6, gi, ok);
        }
        // y frame (4 rows above and below the tile columns): s2, s4, s6
        for (int e = tid; e < 8 * TX; e += T::NT) {
            const int k = e / TX, c = e % TX;
            const int r = k < 4 ? k : TY + k;             // 0..3 and TY+4..TY+7
            const int gx_ = x0 + c, gy_ = y0 - 4 + r;
            const bool ok = gx_ < nx && gy_ >= 0 && gy_ < ny;
            const int gi = z * sz + gy_ * sy + gx_;
            const int li = r * VX + c + 4;
            t2[li] = ld(s2, gi, ok);
            t4[li] = ld(s4, gi, ok);
            t6[li] = ld(s6, gi, ok);
        }
        // per-cell operands of plane z
        const int m0 = mnext;
        const int zn = z + 1;
        mnext = col_in && zn < nz ? S.mats[zn * sz + col] : 0;
        const int mxp = col_in && ix < nx - 1 ? S.mats[i + 1] : 0;
        const int myp = col_in && iy < ny - 1 ? S.mats[i + sy] : 0;
        const R dmp = col_in && S.damp ? S.damp[i] : (R)1;
        R vxo = (R)0, vyo = (R)0, vzo = (R)0;
        if (col_in) {
            vxo = S.v[0][i];
            vyo = S.v[1][i];
            vzo = S.v[2][i];
        }
        const int z5 = z + 5;
        const bool ok5 = col_in && z5 < nz;
        const R n3 = ld(s3, col + z5 * sz, ok5);
        const R n4 = ld(s4, col + z5 * sz, ok5);
        const R n5 = ld(s5, col + z5 * sz, ok5);
        __syncthreads();

        if (col_in) {
            R ax = 0, ay = 0, az = 0;
            const R *p1 = t1 + vc, *p2 = t2 + vc, *p4 = t4 + vc, *p5 = t5 + vc, *p6 = t6 + vc;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                if (j >= S.nc) break;
                const R cj = S.c[j];
                if (ix >= j && ix < nx - j - 1) ax = fma_r(cj, p1[j + 1] - p1[-j], ax);
                if (ix >= j + 1 && ix < nx - j) {
                    ay = fma_r(cj, p6[j] - p6[-(j + 1)], ay);
                    az = fma_r(cj, p5[j] - p5[-(j + 1)], az);
                }
                if (iy >= j && iy < ny - j - 1) ay = fma_r(cj, p2[(j + 1) * VX] - p2[-j * VX], ay);
                if (iy >= j + 1 && iy < ny - j) {
                    ax = fma_r(cj, p6[j * VX] - p6[-(j + 1) * VX], ax);
                    az = fma_r(cj, p4[j * VX] - p4[-(j + 1) * VX], az);
                }
                if (z >= j && z < nz - j - 1) az = fma_r(cj, r3[5 + j] - r3[4 - j], az);
                if (z >= j + 1 && z < nz - j) {
                    ax = fma_r(cj, r5[4 + j] - r5[3 - j], ax);
                    ay = fma_r(cj, r4[4 + j] - r4[3 - j], ay);
                }
            }
            const R *__restrict__ rho = S.rho;
            const R r0 = rho[m0];
            const R rvx = ix < nx - 1 ? (R)0.5 * (r0 + rho[mxp]) : r0;
            const R rvy = iy < ny - 1 ? (R)0.5 * (r0 + rho[myp]) : r0;
            const R rvz = z < nz - 1 ? (R)0.5 * (r0 + rho[mnext]) : r0;
            S.v[0][i] = (vxo + S.dt_h * ax / rvx) * dmp;
            S.v[1][i] = (vyo + S.dt_h * ay / rvy) * dmp;
            S.v[2][i] = (vzo + S.dt_h * az / rvz) * dmp;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            r3[k] = r3[k + 1];
            r4[k] = r4[k + 1];
            r5[k] = r5[k + 1];
        }
        r3[8] = n3;
        r4[8] = n4;
        r5[8] = n5;
    }
}

// ------------------------------------------------------------- launchers
template <typename R> struct Tiles;
template <> struct Tiles<float> {
    static constexpr int STX = 32, STY = 16, VTX = 64, VTY = 4;
};
template <> struct Tiles<double> {
    static constexpr int STX = 32, STY = 6, VTX = 64, VTY = 4;
};

// z planes per workgroup: enough workgroups to cover the chip about four
// times over, but no chunk shorter than 16 planes (ring refill cost).
inline int chunk_len(int nz, int tiles)
{
    const int target = 256 * 4;
    int chunks = (target + tiles - 1) / tiles;
    chunks = chunks < 1 ? 1 : chunks;
    int len = (nz + chunks - 1) / chunks;
    if (len < 16) len = 16;
    return len > nz ? nz : len;
}

}  // namespace

template <typename R>
hipError_t launch_stress_tiled(const Shot<R> &s, hipStream_t st)
{
    constexpr int TX = Tiles<R>::STX, TY = Tiles<R>::STY;
    using T = StressTile<R, TX, TY>;
    const int tiles_x = (s.nx + TX - 1) / TX, tiles_y = (s.ny + TY - 1) / TY;
    const int zl = chunk_len(s.nz, tiles_x * tiles_y);
    const int nb = tiles_x * tiles_y * ((s.nz + zl - 1) / zl);
    hipLaunchKernelGGL((k_stress_tiled<R, TX, TY>), dim3(nb), dim3(T::NT), T::LDS, st, s, tiles_x,
                       tiles_y, zl);
    return hipGetLastError();
}

template <typename R>
hipError_t launch_velocity_tiled(const Shot<R> &s, hipStream_t st)
{
    constexpr int TX = Tiles<R>::VTX, TY = Tiles<R>::VTY;
    using T = VelTile<R, TX, TY>;
    const int tiles_x = (s.nx + TX - 1) / TX, tiles_y = (s.ny + TY - 1) / TY;
    const int zl = chunk_len(s.nz, tiles_x * tiles_y);
    const int nb = tiles_x * tiles_y * ((s.nz + zl - 1) / zl);
    hipLaunchKernelGGL((k_velocity_tiled<R, TX, TY>), dim3(nb), dim3(T::NT), T::LDS, st, s,
                       tiles_x, tiles_y, zl);
    return hipGetLastError();
}

template hipError_t launch_stress_tiled<float>(const Shot<float> &, hipStream_t);
template hipError_t launch_stress_tiled<double>(const Shot<double> &, hipStream_t);
template hipError_t launch_velocity_tiled<float>(const Shot<float> &, hipStream_t);
template hipError_t launch_velocity_tiled<double>(const Shot<double> &, hipStream_t);

}  // namespace pe
