// Tiled z-marching adjoint kernels (pe_adjoint.kernel = PE_KERNEL_TILED).
//
// Same reverse step as kernels_adjoint.hip (which states the transposed
// stencils and moves), with every field streamed once per kernel: x / y taps
// from LDS plane tiles, z taps from per-thread register rings, as in the
// forward tiled kernels.  Reverse step n:
//
//   k_adj_velocity_tiled   ls = [d *] ls + D^T e,  e_a = w_a * lv_a with the
//                          static face weights w_a = d * dt_h / rho_face_a;
//                          the "d *" is the stress-adjoint damping of the
//                          previous reverse step, deferred to here
//   k_adj_grad_tiled       per-run sums of ls_I * q_{VP[I],J}(v^n) (forward
//                          stress-kernel strain machinery on v^n), projected
//                          on the material rows P_p when the run ends
//   k_adj_gbar_tiled       gbar = M^T (dt C^T ls): every cell forms its 24
//                          products C^T ls from its own row once per plane;
//                          the 18 that neighbours read sit in a 3-plane LDS
//                          ring
//   k_adj_strain_t_tiled   lv = d * lv + inv_h * G^T gbar (the data layout of
//                          the forward velocity kernel)
//
// Summation orders differ from the direct adjoint kernels (exact arithmetic
// is the same), so results agree to rounding.
#include "tiled_common.hpp"

namespace pe {

namespace {

using namespace tiled;

constexpr int kPw = 256;

// Transposed staggered differences summed over j (see kernels_adjoint.hip):
// FWD: sum_j c_j ([k >= 2j+1] u(k-j-1) - [k <= n-2j-2] u(k+j))
// else sum_j c_j ([k >= 2j+1] u(k-j)   - [k <= n-2j-2] u(k+j+1)).
// p points at u(k); stride between taps.  G = false: every term is kept.
template <bool G, bool FWD, typename R>
__device__ __forceinline__ R tdiff(const Shot<R> &S, const R *p, int stride, int k, int n)
{
    R acc = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        if (j >= S.nc) break;
        const int lo = FWD ? j + 1 : j, hi = FWD ? j : j + 1;
        R t = 0;
        if (!G || k >= 2 * j + 1) t = p[-lo * stride];
        if (!G || k <= n - 2 * j - 2) t = t - p[hi * stride];
        acc = fma_r(S.c[j], t, acc);
    }
    return acc;
}

// Same on a 9-deep register ring, r[4] = u(k).
template <bool G, bool FWD, typename R>
__device__ __forceinline__ R tdiff_ring(const Shot<R> &S, const R (&r)[9], int k, int n)
{
    R acc = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        if (j >= S.nc) break;
        const int lo = FWD ? j + 1 : j, hi = FWD ? j : j + 1;
        R t = 0;
        if (!G || k >= 2 * j + 1) t = r[4 - lo];
        if (!G || k <= n - 2 * j - 2) t = t - r[4 + hi];
        acc = fma_r(S.c[j], t, acc);
    }
    return acc;
}

// Interior test of the transposed stencils: every term kept needs
// 2*nc-1 <= k <= n - 2*nc along the axis.
__device__ __forceinline__ bool t_inner(int k0, int k1, int n) { return k0 >= 7 && k1 <= n - 8; }

// ----------------------------------------------------------- face weights
template <typename R>
__global__ __launch_bounds__(kPw) void k_face_weights(Shot<R> S, R *__restrict__ w)
{
    const int nz = S.nz, ny = S.ny, nx = S.nx, sy = nx, sz = nx * ny;
    const int n = sz * nz;
    const int i = blockIdx.x * kPw + threadIdx.x;
    if (i >= n) return;
    const int ix = i % nx, iy = (i / nx) % ny, iz = i / sz;
    const R r0 = S.rho[S.mats[i]];
    const R rvx = ix < nx - 1 ? (R)0.5 * (r0 + S.rho[S.mats[i + 1]]) : r0;
    const R rvy = iy < ny - 1 ? (R)0.5 * (r0 + S.rho[S.mats[i + sy]]) : r0;
    const R rvz = iz < nz - 1 ? (R)0.5 * (r0 + S.rho[S.mats[i + sz]]) : r0;
    const R f = (S.damp ? S.damp[i] : (R)1) * S.dt_h;
    w[i] = f / rvx;
    w[(size_t)n + i] = f / rvy;
    w[2 * (size_t)n + i] = f / rvz;
}

template <typename R>
__global__ __launch_bounds__(kPw) void k_scale_ls(Shot<R> S, Adj<R> A)
{
    const int n = S.nz * S.ny * S.nx;
    const int i = blockIdx.x * kPw + threadIdx.x;
    if (i >= n || !S.damp) return;
    const R d = S.damp[i];
#pragma unroll
    for (int I = 0; I < 6; ++I) A.ls[I][i] = A.ls[I][i] * d;
}

// ------------------------------------------------- D^T (adjoint velocity)
template <typename R, int TX, int TY>
struct AdjVelTile {
    static constexpr int NT = TX * TY;
    static constexpr int VX = TX + 8, VY = TY + 8, NV = VX * VY;
    static constexpr int NFRAME = 8 * TY + 8 * TX;                 // x frame, then y frame
    static constexpr size_t LDS = sizeof(R) * 3 * NV;
};

template <typename R, int TX, int TY>
__global__ __launch_bounds__(TX *TY, (MinWaves<R, TX * TY>::v)) void k_adj_velocity_tiled(
    Shot<R> S, Adj<R> A, int tiles_x, int tiles_y, int zc_len, int damp_ls)
{
    using T = AdjVelTile<R, TX, TY>;
    constexpr int VX = T::VX, NV = T::NV, NT = T::NT, NF = (T::NFRAME + NT - 1) / NT;
    extern __shared__ unsigned char lds_raw[];
    R *ex = reinterpret_cast<R *>(lds_raw);
    R *ey = ex + NV;
    R *ez = ey + NV;

    const int nz = S.nz, ny = S.ny, nx = S.nx;
    const int sy = nx, sz = nx * ny;
    const size_t ncell = (size_t)sz * nz;
    const int nb = tiles_x * tiles_y * ((nz + zc_len - 1) / zc_len);
    const int t = xcd_remap(blockIdx.x, nb);
    const int tx = t % tiles_x, ty = (t / tiles_x) % tiles_y, chunk = t / (tiles_x * tiles_y);
    const int x0 = tx * TX, y0 = ty * TY;
    const int z0 = chunk * zc_len, z1 = min(nz, z0 + zc_len);
    const bool inner_xy = t_inner(x0, x0 + TX - 1, nx) && t_inner(y0, y0 + TY - 1, ny);

    const int tid = threadIdx.x;
    const int lx = tid % TX, ly = tid / TX;
    const int ix = x0 + lx, iy = y0 + ly;
    const bool col_in = ix < nx && iy < ny;
    const int col = clampi(iy, 0, ny - 1) * sy + clampi(ix, 0, nx - 1);
    const int vc = (ly + 4) * VX + (lx + 4);
    const R *__restrict__ w = A.wface;

    int f_lds[NF], f_xy[NF];
#pragma unroll
    for (int k = 0; k < NF; ++k) {
        const int e = tid + k * NT;
        int li = NV - 1, gx_ = 0, gy_ = 0;
        if (e < 8 * TY) {
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
        f_xy[k] = clampi(gy_, 0, ny - 1) * sy + clampi(gx_, 0, nx - 1);
    }
    auto e_at = [&](int gi, int a) { return w[(size_t)a * ncell + gi] * A.lv[a][gi]; };

    R rx[9], ry[9], rz[9], px, py, pz, fe[NF][3];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        const int i = clampi(z0 - 5 + k, 0, nz - 1) * sz + col;
        rx[k] = e_at(i, 0);
        ry[k] = e_at(i, 1);
        rz[k] = e_at(i, 2);
    }
    {
        const int i = clampi(z0 + 4, 0, nz - 1) * sz + col;
        px = e_at(i, 0);
        py = e_at(i, 1);
        pz = e_at(i, 2);
    }
    const bool has_damp = S.damp != nullptr && damp_ls;
    auto load_frame = [&](int z) {
        const int zo = clampi(z, 0, nz - 1) * sz;
#pragma unroll
        for (int k = 0; k < NF; ++k) {
            const int gi = zo + f_xy[k];
            fe[k][0] = e_at(gi, 0);
            fe[k][1] = e_at(gi, 1);
            fe[k][2] = e_at(gi, 2);
        }
    };
    R cls[6], cd;
    auto load_col = [&](int z) {
        const int i = clampi(z, 0, nz - 1) * sz + col;
#pragma unroll
        for (int k = 0; k < 6; ++k) cls[k] = A.ls[k][i];
        cd = has_damp ? S.damp[i] : (R)1;
    };
    load_frame(z0);
    load_col(z0);

    for (int z = z0; z < z1; ++z) {
        const int i = z * sz + col;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            rx[k] = rx[k + 1];
            ry[k] = ry[k + 1];
            rz[k] = rz[k + 1];
        }
        rx[8] = px;
        ry[8] = py;
        rz[8] = pz;
        ex[vc] = rx[4];
        ey[vc] = ry[4];
        ez[vc] = rz[4];
#pragma unroll
        for (int k = 0; k < NF; ++k) {
            ex[f_lds[k]] = fe[k][0];
            ey[f_lds[k]] = fe[k][1];
            ez[f_lds[k]] = fe[k][2];
        }
        R sv[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) sv[k] = cls[k];
        const R d = cd;
        load_frame(z + 1);
        load_col(z + 1);
        {
            const int i5 = clampi(z + 5, 0, nz - 1) * sz + col;
            px = e_at(i5, 0);
            py = e_at(i5, 1);
            pz = e_at(i5, 2);
        }
        lds_barrier();
        R nl[6] = {0, 0, 0, 0, 0, 0};
        if (col_in) {
            R o[6];
            const bool in = inner_xy && z >= 7 && z <= nz - 8;
#define PE_DT(G)                                                                                  \
    o[0] = tdiff<G, true>(S, ex + vc, 1, ix, nx);                                                \
    o[1] = tdiff<G, true>(S, ey + vc, VX, iy, ny);                                               \
    o[2] = tdiff_ring<G, true>(S, rz, z, nz);                                                    \
    o[3] = tdiff_ring<G, false>(S, ry, z, nz) + tdiff<G, false>(S, ez + vc, VX, iy, ny);         \
    o[4] = tdiff_ring<G, false>(S, rx, z, nz) + tdiff<G, false>(S, ez + vc, 1, ix, nx);          \
    o[5] = tdiff<G, false>(S, ex + vc, VX, iy, ny) + tdiff<G, false>(S, ey + vc, 1, ix, nx);
            if (in) {
                PE_DT(false)
            } else {
                PE_DT(true)
            }
#undef PE_DT
#pragma unroll
            for (int k = 0; k < 6; ++k) nl[k] = sv[k] * d + o[k];
        }
        // every lane issues the stores (stb drops the lanes outside the
        // grid), so the next step's wait for its prefetches leaves them in
        // flight
#pragma unroll
        for (int k = 0; k < 6; ++k) stb(A.ls[k], (uint32_t)ncell, i, col_in, nl[k]);
        lds_barrier();
    }
}

// ------------------------------------------- G^T (adjoint of the strains)
template <typename R, int TX, int TY>
struct AdjStrTile {
    static constexpr int NT = TX * TY;
    static constexpr int VX = TX + 8, VY = TY + 8, NV = VX * VY;
    static constexpr int NFRAME = 8 * TY + 8 * TX;
    static constexpr size_t LDS = sizeof(R) * 5 * NV;
};

template <typename R, int TX, int TY>
__global__ __launch_bounds__(TX *TY, (MinWaves<R, TX * TY>::v)) void k_adj_strain_t_tiled(
    Shot<R> S, Adj<R> A, int tiles_x, int tiles_y, int zc_len)
{
    using T = AdjStrTile<R, TX, TY>;
    constexpr int VX = T::VX, NV = T::NV, NT = T::NT, NF = (T::NFRAME + NT - 1) / NT;
    extern __shared__ unsigned char lds_raw[];
    R *t1 = reinterpret_cast<R *>(lds_raw);   // gbar1, 2, 4, 5, 6 plane tiles
    R *t2 = t1 + NV;
    R *t4 = t2 + NV;
    R *t5 = t4 + NV;
    R *t6 = t5 + NV;

    const int nz = S.nz, ny = S.ny, nx = S.nx;
    const int sy = nx, sz = nx * ny;
    const size_t ncell = (size_t)sz * nz;
    const int nb = tiles_x * tiles_y * ((nz + zc_len - 1) / zc_len);
    const int t = xcd_remap(blockIdx.x, nb);
    const int tx = t % tiles_x, ty = (t / tiles_x) % tiles_y, chunk = t / (tiles_x * tiles_y);
    const int x0 = tx * TX, y0 = ty * TY;
    const int z0 = chunk * zc_len, z1 = min(nz, z0 + zc_len);
    const bool inner_xy = t_inner(x0, x0 + TX - 1, nx) && t_inner(y0, y0 + TY - 1, ny);

    const int tid = threadIdx.x;
    const int lx = tid % TX, ly = tid / TX;
    const int ix = x0 + lx, iy = y0 + ly;
    const bool col_in = ix < nx && iy < ny;
    const int col = clampi(iy, 0, ny - 1) * sy + clampi(ix, 0, nx - 1);
    const int vc = (ly + 4) * VX + (lx + 4);
    const R *__restrict__ g1 = A.work;
    const R *__restrict__ g2 = A.work + ncell;
    const R *__restrict__ g3 = A.work + 2 * ncell;
    const R *__restrict__ g4 = A.work + 3 * ncell;
    const R *__restrict__ g5 = A.work + 4 * ncell;
    const R *__restrict__ g6 = A.work + 5 * ncell;

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

    R r3[9], r4[9], r5[9], p3, p4, p5, fv[NF][3];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        const int i = clampi(z0 - 5 + k, 0, nz - 1) * sz + col;
        r3[k] = g3[i];
        r4[k] = g4[i];
        r5[k] = g5[i];
    }
    {
        const int i = clampi(z0 + 4, 0, nz - 1) * sz + col;
        p3 = g3[i];
        p4 = g4[i];
        p5 = g5[i];
    }
    const bool has_damp = S.damp != nullptr;
    auto load_frame = [&](int z) {
        const int zo = clampi(z, 0, nz - 1) * sz;
#pragma unroll
        for (int k = 0; k < NF; ++k) {
            const int gi = zo + f_xy[k];
            fv[k][0] = (f_x[k] ? g1 : g2)[gi];
            fv[k][1] = (f_x[k] ? g5 : g4)[gi];
            fv[k][2] = g6[gi];
        }
    };
    R c1, c2, c6, clx, cly, clz, cd;
    auto load_col = [&](int z) {
        const int i = clampi(z, 0, nz - 1) * sz + col;
        c1 = g1[i];
        c2 = g2[i];
        c6 = g6[i];
        clx = A.lv[0][i];
        cly = A.lv[1][i];
        clz = A.lv[2][i];
        cd = has_damp ? S.damp[i] : (R)1;
    };
    load_frame(z0);
    load_col(z0);

    for (int z = z0; z < z1; ++z) {
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
        const R lx0 = clx, ly0 = cly, lz0 = clz, d = cd;
        load_frame(z + 1);
        load_col(z + 1);
        {
            const int i5 = clampi(z + 5, 0, nz - 1) * sz + col;
            p3 = g3[i5];
            p4 = g4[i5];
            p5 = g5[i5];
        }
        lds_barrier();
        R nl[3] = {0, 0, 0};
        if (col_in) {
            R gx, gy, gz;
            const bool in = inner_xy && z >= 7 && z <= nz - 8;
#define PE_GT(G)                                                                                  \
    gx = tdiff<G, false>(S, t1 + vc, 1, ix, nx) + tdiff_ring<G, true>(S, r5, z, nz) +             \
         tdiff<G, true>(S, t6 + vc, VX, iy, ny);                                                 \
    gy = tdiff<G, false>(S, t2 + vc, VX, iy, ny) + tdiff_ring<G, true>(S, r4, z, nz) +            \
         tdiff<G, true>(S, t6 + vc, 1, ix, nx);                                                  \
    gz = tdiff_ring<G, false>(S, r3, z, nz) + tdiff<G, true>(S, t4 + vc, VX, iy, ny) +            \
         tdiff<G, true>(S, t5 + vc, 1, ix, nx);
            if (in) {
                PE_GT(false)
            } else {
                PE_GT(true)
            }
#undef PE_GT
            nl[0] = lx0 * d + S.inv_h * gx;
            nl[1] = ly0 * d + S.inv_h * gy;
            nl[2] = lz0 * d + S.inv_h * gz;
        }
        // (stores issued by every lane, as in k_adj_velocity_tiled)
#pragma unroll
        for (int k = 0; k < 3; ++k) stb(A.lv[k], (uint32_t)ncell, i, col_in, nl[k]);
        lds_barrier();
    }
}

// ------------------------------------------- M^T C^T (moved-stress adjoint)
// u_{P,J} = sum over I at position P of C_IJ[m] * ls_I at every cell: 24
// values, 6 of them at J's own position (used by the cell itself) and 18
// that neighbouring cells read through the transposed moves.  Each cell of
// the tile and its one-cell ring computes its u once per plane from its own
// row; the 18 moved values live in a 3-plane LDS ring, slot kUIdx[P][J].
constexpr int kUIdx[4][6] = {{-1, -1, -1, 0, 1, 2},
                             {3, 4, 5, -1, 6, 7},
                             {8, 9, 10, 11, -1, 12},
                             {13, 14, 15, 16, 17, -1}};

template <typename R, int TX, int TY>
struct GbarTile {
    static constexpr int GX = TX + 2, GY = TY + 2, NG = GX * GY;
    static constexpr int NT = (NG + 63) / 64 * 64;
    static constexpr size_t LDS = sizeof(R) * 18 * 3 * NG;
};

// u of one cell: own[J] = u_{pos(J),J}; mv[kUIdx[P][J]] = u_{P,J}, P != pos(J).
// Inactive table entries are zero in every row.
template <typename R>
__device__ __forceinline__ void cell_u(const R (&row)[36], const R (&l)[6], R (&own)[6],
                                       R (&mv)[18])
{
#pragma unroll
    for (int J = 0; J < 6; ++J) {
        const R uc = fma_r(row[12 + J], l[2], fma_r(row[6 + J], l[1], row[J] * l[0]));
        const R u3 = row[18 + J] * l[3], u4 = row[24 + J] * l[4], u5 = row[30 + J] * l[5];
        const R uP[4] = {uc, u3, u4, u5};
#pragma unroll
        for (int P = 0; P < 4; ++P) {
            if (P == kVP[J])
                own[J] = uP[P];
            else
                mv[kUIdx[P][J]] = uP[P];
        }
    }
}

template <typename R>
__device__ __forceinline__ void load_row(const R *p, R (&row)[36])
{
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        R r6[6];
        load6(p + 6 * k, r6);
#pragma unroll
        for (int q = 0; q < 6; ++q) row[6 * k + q] = r6[q];
    }
}

// Transposed move of u_{P,J} onto the output cell: the taps k - o of the
// forward move (P <- pos(J)) whose source cell passes the move guards.
template <bool G, int P, int J, typename R, int GX, int NG>
__device__ __forceinline__ R move_t(const Shot<R> &S, const R *ut, const int (&sl)[3], int c,
                                    int kz, int ky, int kx)
{
    constexpr int src = kVP[J];
    constexpr int dz = kPos[P][0] - kPos[src][0];
    constexpr int dy = kPos[P][1] - kPos[src][1];
    constexpr int dx = kPos[P][2] - kPos[src][2];
    constexpr int z0 = dz < 0 ? -1 : 0, nzt = dz != 0 ? 2 : 1;
    constexpr int y0 = dy < 0 ? -1 : 0, nyt = dy != 0 ? 2 : 1;
    constexpr int x0 = dx < 0 ? -1 : 0, nxt = dx != 0 ? 2 : 1;
    constexpr int v = kUIdx[P][J];
    R acc = 0;
#pragma unroll
    for (int a = 0; a < nzt; ++a)
#pragma unroll
        for (int b = 0; b < nyt; ++b)
#pragma unroll
            for (int e = 0; e < nxt; ++e) {
                const int oz = z0 + a, oy = y0 + b, ox = x0 + e;
                if constexpr (G) {
                    auto ok = [](int d, int cc, int n) {
                        return d > 0 ? (cc >= 0 && cc <= n - 2)
                                     : (d < 0 ? (cc >= 1 && cc <= n - 1) : true);
                    };
                    if (!ok(dz, kz - oz, S.nz) || !ok(dy, ky - oy, S.ny) || !ok(dx, kx - ox, S.nx))
                        continue;
                }
                acc += ut[(sl[1 - oz] * 18 + v) * NG + c - oy * GX - ox];
            }
    constexpr int ntap = nzt * nyt * nxt;
    return acc * (ntap == 4 ? (R)0.25 : (R)0.5);
}

template <bool G, int J, typename R, int GX, int NG>
__device__ __forceinline__ R gbar_j(const Shot<R> &S, const R *ut, const int (&sl)[3], int c,
                                    const R (&own)[6], int kz, int ky, int kx)
{
    constexpr int pj = kVP[J];
    R acc = own[J];
    if constexpr (pj != 0) acc += move_t<G, 0, J, R, GX, NG>(S, ut, sl, c, kz, ky, kx);
    if constexpr (pj != 1) acc += move_t<G, 1, J, R, GX, NG>(S, ut, sl, c, kz, ky, kx);
    if constexpr (pj != 2) acc += move_t<G, 2, J, R, GX, NG>(S, ut, sl, c, kz, ky, kx);
    if constexpr (pj != 3) acc += move_t<G, 3, J, R, GX, NG>(S, ut, sl, c, kz, ky, kx);
    return S.dt * acc;
}

template <typename R, int TX, int TY>
__global__ __launch_bounds__((GbarTile<R, TX, TY>::NT)) void k_adj_gbar_tiled(
    Shot<R> S, Adj<R> A, int tiles_x, int tiles_y, int zc_len)
{
    using T = GbarTile<R, TX, TY>;
    constexpr int GX = T::GX, NG = T::NG;
    extern __shared__ unsigned char lds_raw[];
    R *ut = reinterpret_cast<R *>(lds_raw);                       // [3][18][NG]

    const int nz = S.nz, ny = S.ny, nx = S.nx;
    const int sy = nx, sz = nx * ny;
    const size_t ncell = (size_t)sz * nz;
    const int nb = tiles_x * tiles_y * ((nz + zc_len - 1) / zc_len);
    const int t = xcd_remap(blockIdx.x, nb);
    const int tx = t % tiles_x, ty = (t / tiles_x) % tiles_y, chunk = t / (tiles_x * tiles_y);
    const int x0 = tx * TX, y0 = ty * TY;
    const int z0 = chunk * zc_len, z1 = min(nz, z0 + zc_len);
    const bool inner_xy = x0 >= 1 && x0 + TX <= nx - 1 && y0 >= 1 && y0 + TY <= ny - 1;

    const int tid = threadIdx.x;
    const bool gact = tid < NG;
    const int gx = gact ? tid % GX : 0, gy = gact ? tid / GX : 0;
    const int ix = x0 - 1 + gx, iy = y0 - 1 + gy;
    const bool core = gact && gx >= 1 && gx <= TX && gy >= 1 && gy <= TY && ix < nx && iy < ny;
    const int col = clampi(iy, 0, ny - 1) * sy + clampi(ix, 0, nx - 1);
    const int c = gact ? tid : NG - 1;

    auto at = [&](int z) { return clampi(z, 0, nz - 1) * sz + col; };
    auto load_ls = [&](int z, R (&l)[6]) {
        const int i = at(z);
#pragma unroll
        for (int k = 0; k < 6; ++k) l[k] = A.ls[k][i];
    };
    auto put = [&](int z, const R (&mv)[18]) {
        const int s = (z + 3) % 3;
        if (gact) {
#pragma unroll
            for (int v = 0; v < 18; ++v) ut[(s * 18 + v) * NG + c] = mv[v];
        }
    };
    R own_cur[6], own_nxt[6];
    {
        R row[36], l[6], mv[18], dummy[6];
        load_row(S.Ct + (size_t)S.mats[at(z0 - 1)] * 36, row);
        load_ls(z0 - 1, l);
        cell_u(row, l, dummy, mv);
        put(z0 - 1, mv);
        load_row(S.Ct + (size_t)S.mats[at(z0)] * 36, row);
        load_ls(z0, l);
        cell_u(row, l, own_nxt, mv);
        put(z0, mv);
    }
    // operands of plane z+1 (row, ls) and the label of plane z+2, one
    // iteration ahead
    R nrow[36], nls[6];
    int nm2;
    load_row(S.Ct + (size_t)S.mats[at(z0 + 1)] * 36, nrow);
    load_ls(z0 + 1, nls);
    nm2 = S.mats[at(z0 + 2)];

    for (int z = z0; z < z1; ++z) {
#pragma unroll
        for (int k = 0; k < 6; ++k) own_cur[k] = own_nxt[k];
        {
            R mv[18];
            cell_u(nrow, nls, own_nxt, mv);
            put(z + 1, mv);
        }
        load_row(S.Ct + (size_t)nm2 * 36, nrow);
        load_ls(z + 2, nls);
        nm2 = S.mats[at(z + 3)];
        lds_barrier();
        const int i = z * sz + col;
        R o[6] = {0, 0, 0, 0, 0, 0};
        if (core) {
            const int sl[3] = {(z + 2) % 3, (z + 3) % 3, (z + 4) % 3};   // planes z-1, z, z+1
            const bool in = inner_xy && z >= 1 && z <= nz - 2;
#define PE_GB(G)                                                                                  \
    o[0] = gbar_j<G, 0, R, GX, NG>(S, ut, sl, c, own_cur, z, iy, ix);                            \
    o[1] = gbar_j<G, 1, R, GX, NG>(S, ut, sl, c, own_cur, z, iy, ix);                            \
    o[2] = gbar_j<G, 2, R, GX, NG>(S, ut, sl, c, own_cur, z, iy, ix);                            \
    o[3] = gbar_j<G, 3, R, GX, NG>(S, ut, sl, c, own_cur, z, iy, ix);                            \
    o[4] = gbar_j<G, 4, R, GX, NG>(S, ut, sl, c, own_cur, z, iy, ix);                            \
    o[5] = gbar_j<G, 5, R, GX, NG>(S, ut, sl, c, own_cur, z, iy, ix);
            if (in) {
                PE_GB(false)
            } else {
                PE_GB(true)
            }
#undef PE_GB
        }
        // (stores issued by every lane, as in k_adj_velocity_tiled)
#pragma unroll
        for (int k = 0; k < 6; ++k) stb(A.work + (size_t)k * ncell, (uint32_t)ncell, i, core, o[k]);
        lds_barrier();
    }
}

// ------------------------------------------------------ stiffness gradient
// The forward stress kernel's strain machinery on v^n (Shot::v = the
// history slot).  Per output cell the 36 products ls_I * q_{VP[I],J}
// accumulate in registers over the cells of one material run; when the run
// ends the run's sums are projected on the material's rows P_p and dt * (P_p
// . sums) is added to the run's accumulators (double).
template <bool G, typename R, int GX, int NG>
__device__ __forceinline__ void grad_cell(const Shot<R> &S, const R *__restrict__ gt,
                                          const int (&sl)[3], int c, int iz, int iy, int ix,
                                          const R (&l)[6], R (&g)[36])
{
    const int nz = S.nz, ny = S.ny, nx = S.nx;
    R q[6];
    gather_lds<G, 0, R, GX, NG>(gt, sl, c, iz, iy, ix, nz, ny, nx, q);
#pragma unroll
    for (int J = 0; J < 6; ++J) {
        g[J] = fma_r(l[0], q[J], g[J]);
        g[6 + J] = fma_r(l[1], q[J], g[6 + J]);
        g[12 + J] = fma_r(l[2], q[J], g[12 + J]);
    }
    gather_lds<G, 1, R, GX, NG>(gt, sl, c, iz, iy, ix, nz, ny, nx, q);
#pragma unroll
    for (int J = 0; J < 6; ++J) g[18 + J] = fma_r(l[3], q[J], g[18 + J]);
    gather_lds<G, 2, R, GX, NG>(gt, sl, c, iz, iy, ix, nz, ny, nx, q);
#pragma unroll
    for (int J = 0; J < 6; ++J) g[24 + J] = fma_r(l[4], q[J], g[24 + J]);
    gather_lds<G, 3, R, GX, NG>(gt, sl, c, iz, iy, ix, nz, ny, nx, q);
#pragma unroll
    for (int J = 0; J < 6; ++J) g[30 + J] = fma_r(l[5], q[J], g[30 + J]);
}

// Double precision keeps the per-cell form (36 double accumulators would
// not fit the register file): t_p = sum_I ls_I sum_J P_p[m][6I+J] q_J per
// cell, added to the run's double accumulators directly.
template <int I, typename R>
__device__ __forceinline__ R proj_row(const R *pr, const R (&q)[6])
{
    R row[6];
    load6(pr + 6 * I, row);
    R acc = 0;
#pragma unroll
    for (int J = 0; J < 6; ++J) acc = fma_r(row[J], q[J], acc);
    return acc;
}

template <bool G, typename R, int GX, int NG>
__device__ __forceinline__ R proj_cell(const Shot<R> &S, const R *__restrict__ gt,
                                       const int (&sl)[3], int c, int iz, int iy, int ix,
                                       const R *pr, const R (&l)[6])
{
    const int nz = S.nz, ny = S.ny, nx = S.nx;
    R q[6];
    gather_lds<G, 0, R, GX, NG>(gt, sl, c, iz, iy, ix, nz, ny, nx, q);
    R t = l[0] * proj_row<0>(pr, q);
    t = fma_r(l[1], proj_row<1>(pr, q), t);
    t = fma_r(l[2], proj_row<2>(pr, q), t);
    gather_lds<G, 1, R, GX, NG>(gt, sl, c, iz, iy, ix, nz, ny, nx, q);
    t = fma_r(l[3], proj_row<3>(pr, q), t);
    gather_lds<G, 2, R, GX, NG>(gt, sl, c, iz, iy, ix, nz, ny, nx, q);
    t = fma_r(l[4], proj_row<4>(pr, q), t);
    gather_lds<G, 3, R, GX, NG>(gt, sl, c, iz, iy, ix, nz, ny, nx, q);
    t = fma_r(l[5], proj_row<5>(pr, q), t);
    return t;
}

template <int NP, typename R>
__device__ __forceinline__ void grad_flush(const R *__restrict__ proj, int m, double dt,
                                           R (&g)[36], double *__restrict__ dst)
{
#pragma unroll
    for (int p = 0; p < NP; ++p) {
        const R *pr = proj + ((size_t)m * NP + p) * 36;
        double t = 0.0;
#pragma unroll
        for (int k = 0; k < 36; ++k) t += (double)pr[k] * (double)g[k];
        dst[p] += dt * t;
    }
#pragma unroll
    for (int k = 0; k < 36; ++k) g[k] = 0;
}

template <typename R, int TX, int TY, int PAD, int NP>
__global__ __launch_bounds__((StressTile<R, TX, TY, PAD>::NT),
                             (MinWaves<R, StressTile<R, TX, TY, PAD>::NT>::v)) void k_adj_grad_tiled(
    Shot<R> S, Adj<R> A, int tiles_x, int tiles_y, int zc_len)
{
    using T = StressTile<R, TX, TY, PAD>;
    constexpr int GX = T::GX, GY = T::GY, NG = T::NG, VX = T::VX, VY = T::VY, NV = T::NV;
    constexpr int NT = T::NT, NF = (T::NFRAME + NT - 1) / NT, WX = T::WX;
    extern __shared__ unsigned char lds_raw[];
    R *vt = reinterpret_cast<R *>(lds_raw);
    R *gt = vt + 3 * NV;

    const int nz = S.nz, ny = S.ny, nx = S.nx;
    const int sy = nx, sz = nx * ny;
    const int nchunk = (nz + zc_len - 1) / zc_len;
    const int nb = tiles_x * tiles_y * nchunk;
    const int t = xcd_remap(blockIdx.x, nb);
    const int tx = t % tiles_x, ty = (t / tiles_x) % tiles_y, chunk = t / (tiles_x * tiles_y);
    const int x0 = tx * TX, y0 = ty * TY;
    const int z0 = chunk * zc_len, z1 = min(nz, z0 + zc_len);
    const bool inner_xy = x0 >= 5 && x0 + TX <= nx - 5 && y0 >= 5 && y0 + TY <= ny - 5;

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
        int r = 0, cc = 0;
        if (e < 4 * WX) {
            r = e / WX;
            cc = e % WX;
        } else if (e < 8 * WX) {
            r = VY - 4 + (e - 4 * WX) / WX;
            cc = (e - 4 * WX) % WX;
        } else if (e < 8 * WX + 4 * GY) {
            r = 4 + (e - 8 * WX) / 4;
            cc = (e - 8 * WX) % 4;
        } else {
            r = 4 + (e - 8 * WX - 4 * GY) / 4;
            cc = GX + 4 + (e - 8 * WX - 4 * GY) % 4;
        }
        f_lds[k] = e < T::NFRAME ? r * VX + cc : NV - 1;
        f_xy[k] = clampi(y0 - 5 + r, 0, ny - 1) * sy + clampi(x0 - 5 + cc, 0, nx - 1);
    }

    R rx[9], ry[9], rz[9], px, py, pz, fq[NF][3];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        const int i = clampi(z0 - 6 + k, 0, nz - 1) * sz + col;
        rx[k] = vx[i];
        ry[k] = vy[i];
        rz[k] = vz[i];
    }
    {
        const int i = clampi(z0 + 3, 0, nz - 1) * sz + col;
        px = vx[i];
        py = vy[i];
        pz = vz[i];
    }
#pragma unroll
    for (int k = 0; k < NF; ++k) {
        const int i = clampi(z0 - 1, 0, nz - 1) * sz + f_xy[k];
        fq[k][0] = vx[i];
        fq[k][1] = vy[i];
        fq[k][2] = vz[i];
    }
    R nl[6];
    int nm;
    {
        const int io1 = clampi(z0 - 2, 0, nz - 1) * sz + col;
#pragma unroll
        for (int k = 0; k < 6; ++k) nl[k] = A.ls[k][io1];
        nm = S.mats[io1];
    }
    // run bookkeeping of this thread's column chunk
    int run = core ? A.run_ptr[(size_t)col * nchunk + chunk] : 0;
    int prev = -1;
    constexpr bool kRunSums = sizeof(R) == 4;
    constexpr int NG36 = kRunSums ? 36 : 1;
    R g[NG36];
    double acc[NP];
#pragma unroll
    for (int k = 0; k < NG36; ++k) g[k] = 0;
#pragma unroll
    for (int p = 0; p < NP; ++p) acc[p] = 0.0;
    const double dt = (double)S.dt;

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
        vt[0 * NV + vc] = rx[4];
        vt[1 * NV + vc] = ry[4];
        vt[2 * NV + vc] = rz[4];
#pragma unroll
        for (int k = 0; k < NF; ++k) {
            vt[0 * NV + f_lds[k]] = fq[k][0];
            vt[1 * NV + f_lds[k]] = fq[k][1];
            vt[2 * NV + f_lds[k]] = fq[k][2];
        }
        const int zo = zc - 1;
        const bool out = core && zo >= z0 && zo < z1;
        R l[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) l[k] = nl[k];
        const int m = nm;
        {
            const int io1 = clampi(zo + 1, 0, nz - 1) * sz + col;
#pragma unroll
            for (int k = 0; k < 6; ++k) nl[k] = A.ls[k][io1];
            nm = S.mats[io1];
        }
        {
            const int zf = clampi(zc + 1, 0, nz - 1) * sz;
#pragma unroll
            for (int k = 0; k < NF; ++k) {
                fq[k][0] = vx[zf + f_xy[k]];
                fq[k][1] = vy[zf + f_xy[k]];
                fq[k][2] = vz[zf + f_xy[k]];
            }
        }
        {
            const int i = clampi(zc + 5, 0, nz - 1) * sz + col;
            px = vx[i];
            py = vy[i];
            pz = vz[i];
        }
        lds_barrier();

        const int slot = (zc + 3) % 3;
        {
            R a[6] = {0, 0, 0, 0, 0, 0};
            if (inner_xy && zc >= 4 && zc <= nz - 5)
                strains_col<false, R, VX>(S, vt + vc, vt + NV + vc, vt + 2 * NV + vc, rx, ry, rz,
                                          ix, iy, zc, a);
            else if (col_in && plane_in)
                strains_col<true, R, VX>(S, vt + vc, vt + NV + vc, vt + 2 * NV + vc, rx, ry, rz,
                                         ix, iy, zc, a);
            R *gs = gt + slot * NT + tid;
#pragma unroll
            for (int k = 0; k < 6; ++k) gs[k * 3 * NT] = a[k];
        }
        lds_barrier();

        if (out) {
            if (prev >= 0 && m != prev) {
                if constexpr (kRunSums) {
                    grad_flush<NP>(A.proj, prev, dt, g, A.run_acc + (size_t)run * NP);
                } else {
#pragma unroll
                    for (int p = 0; p < NP; ++p) {
                        A.run_acc[(size_t)run * NP + p] += acc[p];
                        acc[p] = 0.0;
                    }
                }
                ++run;
            }
            prev = m;
            const int sl[3] = {(zc + 1) % 3, (zc + 2) % 3, slot};
            const bool fast = inner_xy && zo >= 1 && zo <= nz - 2;
            if constexpr (kRunSums) {
                if (fast)
                    grad_cell<false, R, GX, NT>(S, gt, sl, tid, zo, iy, ix, l, g);
                else
                    grad_cell<true, R, GX, NT>(S, gt, sl, tid, zo, iy, ix, l, g);
            } else {
#pragma unroll
                for (int p = 0; p < NP; ++p) {
                    const R *pr = A.proj + ((size_t)m * NP + p) * 36;
                    const R tp = fast ? proj_cell<false, R, GX, NT>(S, gt, sl, tid, zo, iy, ix, pr, l)
                                      : proj_cell<true, R, GX, NT>(S, gt, sl, tid, zo, iy, ix, pr, l);
                    acc[p] += dt * (double)tp;
                }
            }
        }
    }
    if (core && prev >= 0) {
        if constexpr (kRunSums) {
            grad_flush<NP>(A.proj, prev, dt, g, A.run_acc + (size_t)run * NP);
        } else {
#pragma unroll
            for (int p = 0; p < NP; ++p) A.run_acc[(size_t)run * NP + p] += acc[p];
        }
    }
}

// Stiffness gradient from a stored strain history (PE_HIST_STRAIN): the
// strains of v^n come from memory, so only their moves are formed, from a
// 3-plane LDS ring of the tile and its one-cell ring (the layout of the
// stress kernel's strain ring); per-run sums as in k_adj_grad_tiled.
template <typename R, int TX, int TY>
struct GradGTile {
    static constexpr int GX = TX + 2, GY = TY + 2, NG = GX * GY;
    static constexpr int NT = (NG + 63) / 64 * 64;
    static constexpr size_t LDS = sizeof(R) * 6 * 3 * NG;
};

template <typename R, int TX, int TY, int NP>
__global__ __launch_bounds__((GradGTile<R, TX, TY>::NT)) void k_adj_grad_g_tiled(
    Shot<R> S, Adj<R> A, const R *__restrict__ gh, int tiles_x, int tiles_y, int zc_len)
{
    using T = GradGTile<R, TX, TY>;
    constexpr int GX = T::GX, NG = T::NG;
    extern __shared__ unsigned char lds_raw[];
    R *gt = reinterpret_cast<R *>(lds_raw);                       // [6][3][NG]

    const int nz = S.nz, ny = S.ny, nx = S.nx;
    const int sy = nx, sz = nx * ny;
    const size_t ncell = (size_t)sz * nz;
    const int nchunk = (nz + zc_len - 1) / zc_len;
    const int nb = tiles_x * tiles_y * nchunk;
    const int t = xcd_remap(blockIdx.x, nb);
    const int tx = t % tiles_x, ty = (t / tiles_x) % tiles_y, chunk = t / (tiles_x * tiles_y);
    const int x0 = tx * TX, y0 = ty * TY;
    const int z0 = chunk * zc_len, z1 = min(nz, z0 + zc_len);
    const bool inner_xy = x0 >= 1 && x0 + TX <= nx - 1 && y0 >= 1 && y0 + TY <= ny - 1;

    const int tid = threadIdx.x;
    const bool gact = tid < NG;
    const int gx = gact ? tid % GX : 0, gy = gact ? tid / GX : 0;
    const int ix = x0 - 1 + gx, iy = y0 - 1 + gy;
    const bool core = gact && gx >= 1 && gx <= TX && gy >= 1 && gy <= TY && ix < nx && iy < ny;
    const int col = clampi(iy, 0, ny - 1) * sy + clampi(ix, 0, nx - 1);
    const int c = gact ? tid : NG - 1;
    auto at = [&](int z) { return (size_t)clampi(z, 0, nz - 1) * sz + col; };

    R ng[6];
    auto load_g = [&](int z) {
        const size_t i = at(z);
#pragma unroll
        for (int k = 0; k < 6; ++k) ng[k] = gh[k * ncell + i];
    };
    auto put = [&](int z) {
        const int s = (z + 3) % 3;
        if (gact) {
#pragma unroll
            for (int k = 0; k < 6; ++k) gt[(k * 3 + s) * NG + c] = ng[k];
        }
    };
    load_g(z0 - 1);
    put(z0 - 1);
    load_g(z0);
    put(z0);
    load_g(z0 + 1);
    R nl[6];
    int nm;
    auto load_out = [&](int z) {
        const size_t i = at(z);
#pragma unroll
        for (int k = 0; k < 6; ++k) nl[k] = A.ls[k][i];
        nm = S.mats[i];
    };
    load_out(z0);

    int run = core ? A.run_ptr[(size_t)col * nchunk + chunk] : 0;
    int prev = -1;
    constexpr bool kRunSums = sizeof(R) == 4;
    constexpr int NG36 = kRunSums ? 36 : 1;
    R g[NG36];
    double acc[NP];
#pragma unroll
    for (int k = 0; k < NG36; ++k) g[k] = 0;
#pragma unroll
    for (int p = 0; p < NP; ++p) acc[p] = 0.0;
    const double dt = (double)S.dt;

    for (int z = z0; z < z1; ++z) {
        put(z + 1);
        load_g(z + 2);
        R l[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) l[k] = nl[k];
        const int m = nm;
        load_out(z + 1);
        lds_barrier();
        if (core) {
            if (prev >= 0 && m != prev) {
                if constexpr (kRunSums) {
                    grad_flush<NP>(A.proj, prev, dt, g, A.run_acc + (size_t)run * NP);
                } else {
#pragma unroll
                    for (int p = 0; p < NP; ++p) {
                        A.run_acc[(size_t)run * NP + p] += acc[p];
                        acc[p] = 0.0;
                    }
                }
                ++run;
            }
            prev = m;
            const int sl[3] = {(z + 2) % 3, (z + 3) % 3, (z + 4) % 3};   // planes z-1, z, z+1
            const bool fast = inner_xy && z >= 1 && z <= nz - 2;
            if constexpr (kRunSums) {
                if (fast)
                    grad_cell<false, R, GX, NG>(S, gt, sl, c, z, iy, ix, l, g);
                else
                    grad_cell<true, R, GX, NG>(S, gt, sl, c, z, iy, ix, l, g);
            } else {
#pragma unroll
                for (int p = 0; p < NP; ++p) {
                    const R *pr = A.proj + ((size_t)m * NP + p) * 36;
                    const R tp = fast ? proj_cell<false, R, GX, NG>(S, gt, sl, c, z, iy, ix, pr, l)
                                      : proj_cell<true, R, GX, NG>(S, gt, sl, c, z, iy, ix, pr, l);
                    acc[p] += dt * (double)tp;
                }
            }
        }
        lds_barrier();
    }
    if (core && prev >= 0) {
        if constexpr (kRunSums) {
            grad_flush<NP>(A.proj, prev, dt, g, A.run_acc + (size_t)run * NP);
        } else {
#pragma unroll
            for (int p = 0; p < NP; ++p) A.run_acc[(size_t)run * NP + p] += acc[p];
        }
    }
}

inline unsigned blocks_of(size_t n, int b) { return (unsigned)((n + b - 1) / b); }

// Shortest z-chunk of the reverse kernels when the chunk length is chosen
// for occupancy (small grids: more workgroups beat fewer ring refills).
constexpr int kAdjMinChunk = 4;

template <typename R, int TX, int TY>
hipError_t adj_velocity_launch(const Shot<R> &s, const Adj<R> &a, int damp_ls, hipStream_t st)
{
    using T = AdjVelTile<R, TX, TY>;
    const int tiles_x = (s.nx + TX - 1) / TX, tiles_y = (s.ny + TY - 1) / TY;
    const int zl = chunk_len(s.nz, tiles_x * tiles_y, 0, kAdjMinChunk);
    const int nb = tiles_x * tiles_y * ((s.nz + zl - 1) / zl);
    hipError_t e = allow_lds(k_adj_velocity_tiled<R, TX, TY>, T::LDS);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_adj_velocity_tiled<R, TX, TY>), dim3(nb), dim3(T::NT), T::LDS, st, s, a,
                       tiles_x, tiles_y, zl, damp_ls);
    return hipGetLastError();
}

template <typename R, int TX, int TY>
hipError_t adj_strain_t_launch(const Shot<R> &s, const Adj<R> &a, hipStream_t st)
{
    using T = AdjStrTile<R, TX, TY>;
    const int tiles_x = (s.nx + TX - 1) / TX, tiles_y = (s.ny + TY - 1) / TY;
    const int zl = chunk_len(s.nz, tiles_x * tiles_y, 0, kAdjMinChunk);
    const int nb = tiles_x * tiles_y * ((s.nz + zl - 1) / zl);
    hipError_t e = allow_lds(k_adj_strain_t_tiled<R, TX, TY>, T::LDS);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_adj_strain_t_tiled<R, TX, TY>), dim3(nb), dim3(T::NT), T::LDS, st, s,
                       a, tiles_x, tiles_y, zl);
    return hipGetLastError();
}

template <typename R, int TX, int TY>
hipError_t adj_gbar_launch(const Shot<R> &s, const Adj<R> &a, hipStream_t st)
{
    using T = GbarTile<R, TX, TY>;
    const int tiles_x = (s.nx + TX - 1) / TX, tiles_y = (s.ny + TY - 1) / TY;
    const int zl = chunk_len(s.nz, tiles_x * tiles_y, 0, kAdjMinChunk);
    const int nb = tiles_x * tiles_y * ((s.nz + zl - 1) / zl);
    hipError_t e = allow_lds(k_adj_gbar_tiled<R, TX, TY>, T::LDS);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_adj_gbar_tiled<R, TX, TY>), dim3(nb), dim3(T::NT), T::LDS, st, s, a,
                       tiles_x, tiles_y, zl);
    return hipGetLastError();
}

template <typename R, int TX, int TY, int PAD, int NP>
hipError_t adj_grad_launch(const Shot<R> &s, const Adj<R> &a, hipStream_t st)
{
    using T = StressTile<R, TX, TY, PAD>;
    const int tiles_x = (s.nx + TX - 1) / TX, tiles_y = (s.ny + TY - 1) / TY;
    const int zl = a.zbreak > 0 ? (a.zbreak < s.nz ? a.zbreak : s.nz) : s.nz;
    const int nb = tiles_x * tiles_y * ((s.nz + zl - 1) / zl);
    hipError_t e = allow_lds(k_adj_grad_tiled<R, TX, TY, PAD, NP>, T::LDS);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_adj_grad_tiled<R, TX, TY, PAD, NP>), dim3(nb), dim3(T::NT), T::LDS, st,
                       s, a, tiles_x, tiles_y, zl);
    return hipGetLastError();
}

template <typename R, int TX, int TY, int NP>
hipError_t adj_grad_g_launch(const Shot<R> &s, const Adj<R> &a, const R *gh, hipStream_t st)
{
    using T = GradGTile<R, TX, TY>;
    const int tiles_x = (s.nx + TX - 1) / TX, tiles_y = (s.ny + TY - 1) / TY;
    const int zl = a.zbreak > 0 ? (a.zbreak < s.nz ? a.zbreak : s.nz) : s.nz;
    const int nb = tiles_x * tiles_y * ((s.nz + zl - 1) / zl);
    hipError_t e = allow_lds(k_adj_grad_g_tiled<R, TX, TY, NP>, T::LDS);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_adj_grad_g_tiled<R, TX, TY, NP>), dim3(nb), dim3(T::NT), T::LDS, st, s,
                       a, gh, tiles_x, tiles_y, zl);
    return hipGetLastError();
}

template <typename R, int NP>
hipError_t grad_tiled(const Shot<R> &s, const Adj<R> &a, hipStream_t st, const R *gh)
{
    if (gh) return adj_grad_g_launch<R, 32, 8, NP>(s, a, gh, st);
    if constexpr (sizeof(R) == 4)
        return adj_grad_launch<R, 32, 8, 24, NP>(s, a, st);
    else
        return adj_grad_launch<R, 32, 6, 0, NP>(s, a, st);
}

}  // namespace

template <typename R>
hipError_t adjoint_face_weights(const pe_shot *p, const pe_adjoint *a, hipStream_t st)
{
    const Shot<R> S = make_shot<R>(p);
    const size_t n = (size_t)p->nz * p->ny * p->nx;
    hipLaunchKernelGGL(k_face_weights<R>, dim3(blocks_of(n, kPw)), dim3(kPw), 0, st, S,
                       (R *)a->wface);
    return hipGetLastError();
}

// One tiled reverse step's kernels after D^T and the record adjoint:
// gradient (when accumulating), M^T C^T, G^T.
template <typename R>
hipError_t adjoint_step_tiled(const Shot<R> &S, const Shot<R> &Sv, const Adj<R> &A, int stage,
                              int damp_ls, hipStream_t st, const R *gh)
{
    constexpr bool f32 = sizeof(R) == 4;
    // D^T and G^T: 32 x 8 tiles in both precisions (float32 at 80x160x160,
    // 4000 steps: gradient 1.427 s with 64 x 8, 1.402 / 1.407 s with 32 x 8
    // for D^T / G^T alone; 64 x 4 and 32 x 16 no better)
    if (stage == 0) return adj_velocity_launch<R, 32, 8>(S, A, damp_ls, st);
    if (stage == 1) {
        switch (A.n_proj) {
        case 1: return grad_tiled<R, 1>(Sv, A, st, gh);
        case 2: return grad_tiled<R, 2>(Sv, A, st, gh);
        case 3: return grad_tiled<R, 3>(Sv, A, st, gh);
        default: return grad_tiled<R, 4>(Sv, A, st, gh);
        }
    }
    // stage 2: M^T then G^T; stage 3: M^T alone; stage 4: G^T alone (a CPML
    // shot's band cells read gbar in between)
    if (stage == 4) return adj_strain_t_launch<R, 32, 8>(S, A, st);
    hipError_t e = f32 ? adj_gbar_launch<R, 32, 8>(S, A, st) : adj_gbar_launch<R, 32, 4>(S, A, st);
    if (e != hipSuccess || stage == 3) return e;
    return adj_strain_t_launch<R, 32, 8>(S, A, st);
}

template <typename R>
hipError_t adjoint_scale_ls(const pe_shot *p, const pe_adjoint *a, hipStream_t st)
{
    const Shot<R> S = make_shot<R>(p);
    const Adj<R> A = make_adj<R>(a);
    const size_t n = (size_t)p->nz * p->ny * p->nx;
    hipLaunchKernelGGL(k_scale_ls<R>, dim3(blocks_of(n, kPw)), dim3(kPw), 0, st, S, A);
    return hipGetLastError();
}

template hipError_t adjoint_face_weights<float>(const pe_shot *, const pe_adjoint *, hipStream_t);
template hipError_t adjoint_face_weights<double>(const pe_shot *, const pe_adjoint *, hipStream_t);
template hipError_t adjoint_step_tiled<float>(const Shot<float> &, const Shot<float> &,
                                              const Adj<float> &, int, int, hipStream_t,
                                              const float *);
template hipError_t adjoint_step_tiled<double>(const Shot<double> &, const Shot<double> &,
                                               const Adj<double> &, int, int, hipStream_t,
                                               const double *);
template hipError_t adjoint_scale_ls<float>(const pe_shot *, const pe_adjoint *, hipStream_t);
template hipError_t adjoint_scale_ls<double>(const pe_shot *, const pe_adjoint *, hipStream_t);

}  // namespace pe
