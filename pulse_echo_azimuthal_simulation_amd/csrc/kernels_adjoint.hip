// Reverse (adjoint) step of the label-indexed, deferred-damped FDTD step:
// the exact discrete transpose of one forward step, gathered per cell.
//
// The upstream repository's only exact adjoint is the vendored array-path
// one (vendor/ringfwi/elastic3d.py:348-463, reverse loop :386-461); these
// kernels give the same reverse-mode derivative for the label-path step of
// sim/core/fdtd.py::forward_fused_labels (stencils :1102-1257, moved
// strains :612-642, deferred damping :643-657, source :1798-1800, record
// :1803-1813).
//
// Forward step n from (v^n, s^n):
//   A  s <- s * d + dt * C[m] M G v        B  s1..3[src] += w[n] * w_k
//   C  y_n = R s (pressure)                 D  v <- (v + dt_h * D s / rho_face) * d
//   E  y_n = R_v v (velocity record modes, after D)
// Reverse step n on the adjoints (lv, ls) of (v, s):
//   E^T                      k_adj_record (velocity modes)
//   D^T: ls += D^T e,        k_adj_velocity; e_a = dt_h * d * lv_a / rho_face_a,
//        lv <- d * lv        (the scaling of lv is deferred to k_adj_strain_t)
//   C^T, B^T                 k_adj_record (pressure mode; wavelet gradient)
//   A^T: grad += dt ls q(v^n)  k_adj_grad (strains of v^n in the scratch)
//        gbar = M^T (dt C^T ls)  k_adj_gbar
//        lv = d * lv + G^T gbar, ls <- d * ls   k_adj_strain_t
// With a CPML (no upstream counterpart) the state also holds the memory
// variables; k_adj_pml_pre runs before D^T and before k_adj_strain_t, one
// thread per band cell, and the two gather kernels add the transposed band
// derivatives near the bands.  The gradient then reads the strain history
// (the strains depend on the forward memory variables).
//
// Transposed stencils.  With the forward differences
//   F_j f(i) = f(i+j+1) - f(i-j)  (kept iff j <= i <= n-j-2)
//   B_j f(i) = f(i+j) - f(i-j-1)  (kept iff j+1 <= i <= n-j-1)
// the transposes are, at cell k,
//   F_j^T u(k) = [k >= 2j+1] u(k-j-1) - [k <= n-2j-2] u(k+j)
//   B_j^T u(k) = [k >= 2j+1] u(k-j)   - [k <= n-2j-2] u(k+j+1).
// A moved strain q_{P,J}(i) = f * sum_t g_J(i + o_t) (kept iff its guards
// hold at i) transposes to gbar_J(k) += f * sum_t u(k - o_t) over the taps
// whose source cell k - o_t satisfies the guards.
//
// Stiffness gradient, deterministic: a thread walks one (y, x) column in z;
// the column is cut into runs of one material (also cut at multiples of
// zbreak when set), and each run owns one row of run_acc that only this
// thread updates, once per run per step, in step order.  The final per
// material sums walk each material's runs in ascending order.
#include "common.hpp"

namespace pe {

namespace {

constexpr int kBlock = 256;
constexpr int kColBlock = 64;
constexpr int kSrBlock = 1024;

// e_a at cell i: dt_h * (d * lv_a) / rho_face_a, with the forward's face
// density rule (mean with the +a neighbour, the last slice keeps rho[m]).
template <int A, typename R>
__device__ __forceinline__ R e_val(const Shot<R> &S, const Adj<R> &Ad, int i, int iz, int iy,
                                   int ix)
{
    const R dmp = S.damp ? S.damp[i] : (R)1;
    const R t = Ad.lv[A][i] * dmp;
    const R r0 = S.rho[S.mats[i]];
    R rv;
    if constexpr (A == 0)
        rv = ix < S.nx - 1 ? (R)0.5 * (r0 + S.rho[S.mats[i + 1]]) : r0;
    else if constexpr (A == 1)
        rv = iy < S.ny - 1 ? (R)0.5 * (r0 + S.rho[S.mats[i + S.nx]]) : r0;
    else
        rv = iz < S.nz - 1 ? (R)0.5 * (r0 + S.rho[S.mats[i + S.nx * S.ny]]) : r0;
    return S.dt_h * t / rv;
}

// Axis helpers: AX 0 = z, 1 = y, 2 = x.
template <int AX>
__device__ __forceinline__ int coord(int iz, int iy, int ix)
{
    return AX == 0 ? iz : (AX == 1 ? iy : ix);
}

template <int AX, typename R>
__device__ __forceinline__ int extent(const Shot<R> &S)
{
    return AX == 0 ? S.nz : (AX == 1 ? S.ny : S.nx);
}

template <int AX, typename R>
__device__ __forceinline__ int stride(const Shot<R> &S)
{
    return AX == 0 ? S.nx * S.ny : (AX == 1 ? S.nx : 1);
}

// c_j * F_j^T (FWD = true) or c_j * B_j^T of the field f (a callable
// f(i, iz, iy, ix)) along axis AX at cell k, summed over j.
template <bool FWD, int AX, typename R, typename Fn>
__device__ __forceinline__ R diff_t(const Shot<R> &S, Fn f, int k, int kz, int ky, int kx)
{
    const int n = extent<AX>(S), st = stride<AX>(S), kc = coord<AX>(kz, ky, kx);
    R acc = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        if (j >= S.nc) break;
        const int lo = FWD ? j + 1 : j;       // u(k - lo) when k >= 2j+1
        const int hi = FWD ? j : j + 1;       // u(k + hi) when k <= n-2j-2
        R t = 0;
        if (kc >= 2 * j + 1) {
            const int c = kc - lo;
            t = f(k - lo * st, AX == 0 ? c : kz, AX == 1 ? c : ky, AX == 2 ? c : kx);
        }
        if (kc <= n - 2 * j - 2) {
            const int c = kc + hi;
            t = t - f(k + hi * st, AX == 0 ? c : kz, AX == 1 ? c : ky, AX == 2 ? c : kx);
        }
        acc = fma_r(S.c[j], t, acc);
    }
    return acc;
}

// ------------------------------------------------------------------ CPML
// A band cell's correction of one derivative dd (pml_corr in common.hpp) is
// psi' = b psi + a dd, corr = (1/kappa - 1) dd + psi'.  Given the adjoint
// ebar of the strain (acceleration) it corrects and the adjoint pbar of psi'
// (A.psib, carried from the later step), with P = ebar + pbar: the adjoint of
// dd is (1/kappa - 1) ebar + a P, and that of psi is b P (stored over pbar).
// No upstream counterpart.
template <int AX, typename R>
__device__ __forceinline__ int band_vol(const Shot<R> &S)
{
    const int W = 2 * S.pml + 2;
    return AX == 0 ? W * S.ny * S.nx : (AX == 1 ? S.nz * W * S.nx : S.nz * S.ny * W);
}

template <int AX, typename R>
__device__ __forceinline__ int band_flat(const Shot<R> &S, int iz, int iy, int ix)
{
    const int L = S.pml, W = 2 * L + 2;
    if constexpr (AX == 0) return ((iz <= L ? iz : iz - (S.nz - W)) * S.ny + iy) * S.nx + ix;
    else if constexpr (AX == 1) return (iz * W + (iy <= L ? iy : iy - (S.ny - W))) * S.nx + ix;
    else return (iz * S.ny + iy) * W + (ix <= L ? ix : ix - (S.nx - W));
}

// The adjoint of the corrected derivatives of band axis AX at one cell
// (memory variables K0 .. K0+2, staggers S0, S1, S2; e0..e2 their ebar):
// P = ebar + pbar, the adjoint memory variable b P stored over pbar, the
// adjoint of dd, (1/kappa - 1) ebar + a P, into A.pw[AX][0..2][band cell].
// Every load of the axis (adjoint memory variables, profile values) comes
// ahead of its arithmetic and stores, so that the three wait for memory once.
template <int AX, int K0, int S0, int S1, int S2, typename R>
__device__ __forceinline__ void pml_adj_axis(const Shot<R> &S, const Adj<R> &A, int iz, int iy,
                                             int ix, R e0, R e1, R e2)
{
    const int n = extent<AX>(S), ia = coord<AX>(iz, iy, ix);
    const size_t v = (size_t)band_vol<AX>(S);
    const int f = band_flat<AX>(S, iz, iy, ix);
    R *pb = A.psib[AX] + (size_t)K0 * v + f;
    R *d = A.pw[AX] + f;
    const R *p0 = S.prof[AX] + (size_t)S0 * 3 * n;
    const R *p1 = S.prof[AX] + (size_t)S1 * 3 * n;
    const R *p2 = S.prof[AX] + (size_t)S2 * 3 * n;
    const R q0 = pb[0], q1 = pb[v], q2 = pb[2 * v];
    const R b0 = p0[ia], a0 = p0[n + ia], k0 = p0[2 * n + ia];
    const R b1 = p1[ia], a1 = p1[n + ia], k1 = p1[2 * n + ia];
    const R b2 = p2[ia], a2 = p2[n + ia], k2 = p2[2 * n + ia];
    const R P0 = e0 + q0, P1 = e1 + q1, P2 = e2 + q2;
    pb[0] = b0 * P0;
    pb[v] = b1 * P1;
    pb[2 * v] = b2 * P2;
    d[0] = (k0 - (R)1) * e0 + a0 * P0;
    d[v] = (k1 - (R)1) * e1 + a1 * P1;
    d[2 * v] = (k2 - (R)1) * e2 + a2 * P2;
}

// One thread per band cell (pml_band_cell order): the adjoints of the
// cell's band derivatives into A.pw[axis][k][band cell], k the derivative of
// that axis, and its adjoint memory variables advanced.  HALF 1: the
// velocity half-step's (memory variables 3..5), from the adjoints e of the
// accelerations; HALF 0: the strain half-step's (0..2), from gbar (A.work).
template <typename R, int HALF>
__global__ __launch_bounds__(kBlock) void k_adj_pml_pre(Shot<R> S, Adj<R> A)
{
    const long long b = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (b >= pml_band_cells(S.nz, S.ny, S.nx, S.pml)) return;
    int iz, iy, ix;
    pml_band_cell((int)b, S.nz, S.ny, S.nx, S.pml, iz, iy, ix);
    const int L = S.pml;
    const int i = (iz * S.ny + iy) * S.nx + ix;
    R t0, t1, t2, u0, u1, u2, w0, w1, w2;    // the corrected quantities' adjoints per axis
    if constexpr (HALF == 1) {
        // z band: D_z s5 -> ax, D_z s4 -> ay, D_z s3 -> az; y band: D_y s6,
        // D_y s2, D_y s4; x band: D_x s1, D_x s6, D_x s5
        const R ex = e_val<0>(S, A, i, iz, iy, ix), ey = e_val<1>(S, A, i, iz, iy, ix),
                ez = e_val<2>(S, A, i, iz, iy, ix);
        t0 = u0 = w0 = ex;
        t1 = u1 = w1 = ey;
        t2 = u2 = w2 = ez;
    } else {
        // z band: Db_z vz -> g3, Df_z vy -> g4, Df_z vx -> g5; y band: Db_y
        // vy -> g2, Df_y vz -> g4, Df_y vx -> g6; x band: Db_x vx -> g1, Df_x
        // vz -> g5, Df_x vy -> g6
        const size_t n = (size_t)S.nz * S.ny * S.nx;
        const R *g = A.work;
        t0 = g[2 * n + i];
        t1 = g[3 * n + i];
        t2 = g[4 * n + i];
        u0 = g[1 * n + i];
        u1 = g[3 * n + i];
        u2 = g[5 * n + i];
        w0 = g[0 * n + i];
        w1 = g[4 * n + i];
        w2 = g[5 * n + i];
    }
    if (pml_in(iz, S.nz, L)) {
        if constexpr (HALF == 1) pml_adj_axis<0, 3, 0, 0, 1>(S, A, iz, iy, ix, t0, t1, t2);
        else pml_adj_axis<0, 0, 0, 1, 1>(S, A, iz, iy, ix, t0, t1, t2);
    }
    if (pml_in(iy, S.ny, L)) {
        if constexpr (HALF == 1) pml_adj_axis<1, 3, 0, 1, 0>(S, A, iz, iy, ix, u0, u1, u2);
        else pml_adj_axis<1, 0, 0, 1, 1>(S, A, iz, iy, ix, u0, u1, u2);
    }
    if (pml_in(ix, S.nx, L)) {
        if constexpr (HALF == 1) pml_adj_axis<2, 3, 1, 0, 0>(S, A, iz, iy, ix, w0, w1, w2);
        else pml_adj_axis<2, 0, 0, 1, 1>(S, A, iz, iy, ix, w0, w1, w2);
    }
}

// The transposed stencil of diff_t applied to the derivative adjoints of
// axis AX, component K (A.pw, taken as a field over the cells that is 0
// outside the band), with every tap loaded before the arithmetic: a tap
// outside the band (or the grid) reads the cell at band index 0 of its row
// and is then dropped, so the wave waits for its loads once instead of once
// per guarded tap.  The terms and their guards are diff_t's.
template <bool FWD, int AX, int K, typename R>
__device__ __forceinline__ R band_diff_t(const Shot<R> &S, const Adj<R> &A, int kz, int ky, int kx)
{
    const int n = extent<AX>(S), kc = coord<AX>(kz, ky, kx), L = S.pml, W = 2 * L + 2;
    const R *pw = A.pw[AX] + (size_t)K * band_vol<AX>(S);
    auto flat = [&](int c) {
        const int b = (c >= 0 && c < n && pml_in(c, n, L)) ? (c <= L ? c : c - (n - W)) : 0;
        if constexpr (AX == 0) return (b * S.ny + ky) * S.nx + kx;
        else if constexpr (AX == 1) return (kz * W + b) * S.nx + kx;
        else return (kz * S.ny + ky) * W + b;
    };
    R vlo[4], vhi[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        vlo[j] = pw[flat(kc - (FWD ? j + 1 : j))];
        vhi[j] = pw[flat(kc + (FWD ? j : j + 1))];
    }
    R acc = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        if (j >= S.nc) break;
        const int lo = FWD ? j + 1 : j, hi = FWD ? j : j + 1;
        R t = 0;
        if (kc >= 2 * j + 1) t = pml_in(kc - lo, n, L) ? vlo[j] : (R)0;
        if (kc <= n - 2 * j - 2) t = t - (pml_in(kc + hi, n, L) ? vhi[j] : (R)0);
        acc = fma_r(S.c[j], t, acc);
    }
    return acc;
}

// Is cell (kz, ky, kx) within the taps' reach (4 cells) of the band of axis AX?
template <int AX, typename R>
__device__ __forceinline__ bool near_band(const Shot<R> &S, int kz, int ky, int kx)
{
    const int n = extent<AX>(S), c = coord<AX>(kz, ky, kx), L = S.pml;
    return L > 0 && (c <= L + 4 || c >= n - 1 - L - 4);
}

// Transposed band derivatives of the velocity half-step at cell k, added to
// the stress adjoint sums d1..d6: z band D_z s5, D_z s4 (backward
// differences), D_z s3 (forward); y band D_y s6 (b), D_y s2 (f), D_y s4 (b);
// x band D_x s1 (f), D_x s6, D_x s5 (b).
template <typename R>
__device__ __forceinline__ void vel_band_t(const Shot<R> &S, const Adj<R> &A, int k, int kz,
                                           int ky, int kx, R &d1, R &d2, R &d3, R &d4, R &d5,
                                           R &d6)
{
    if (near_band<0>(S, kz, ky, kx)) {
        d5 = d5 + band_diff_t<false, 0, 0>(S, A, kz, ky, kx);
        d4 = d4 + band_diff_t<false, 0, 1>(S, A, kz, ky, kx);
        d3 = d3 + band_diff_t<true, 0, 2>(S, A, kz, ky, kx);
    }
    if (near_band<1>(S, kz, ky, kx)) {
        d6 = d6 + band_diff_t<false, 1, 0>(S, A, kz, ky, kx);
        d2 = d2 + band_diff_t<true, 1, 1>(S, A, kz, ky, kx);
        d4 = d4 + band_diff_t<false, 1, 2>(S, A, kz, ky, kx);
    }
    if (near_band<2>(S, kz, ky, kx)) {
        d1 = d1 + band_diff_t<true, 2, 0>(S, A, kz, ky, kx);
        d6 = d6 + band_diff_t<false, 2, 1>(S, A, kz, ky, kx);
        d5 = d5 + band_diff_t<false, 2, 2>(S, A, kz, ky, kx);
    }
}

// Transposed band derivatives of the strain half-step at cell k, added to
// the velocity adjoint sums (before the 1/h scale): z band Db_z vz, Df_z vy,
// Df_z vx; y band Db_y vy, Df_y vz, Df_y vx; x band Db_x vx, Df_x vz, Df_x
// vy.
template <typename R>
__device__ __forceinline__ void strain_band_t(const Shot<R> &S, const Adj<R> &A, int k, int kz,
                                              int ky, int kx, R &gx, R &gy, R &gz)
{
    if (near_band<0>(S, kz, ky, kx)) {
        gz = gz + band_diff_t<false, 0, 0>(S, A, kz, ky, kx);
        gy = gy + band_diff_t<true, 0, 1>(S, A, kz, ky, kx);
        gx = gx + band_diff_t<true, 0, 2>(S, A, kz, ky, kx);
    }
    if (near_band<1>(S, kz, ky, kx)) {
        gy = gy + band_diff_t<false, 1, 0>(S, A, kz, ky, kx);
        gz = gz + band_diff_t<true, 1, 1>(S, A, kz, ky, kx);
        gx = gx + band_diff_t<true, 1, 2>(S, A, kz, ky, kx);
    }
    if (near_band<2>(S, kz, ky, kx)) {
        gx = gx + band_diff_t<false, 2, 0>(S, A, kz, ky, kx);
        gz = gz + band_diff_t<true, 2, 1>(S, A, kz, ky, kx);
        gy = gy + band_diff_t<true, 2, 2>(S, A, kz, ky, kx);
    }
}

// The tiled family's CPML band terms, as separate gathers after its own
// kernels (whose updates they add to): HALF 1 into ls after D^T, HALF 0
// into lv (times 1/h) after G^T.  One thread per cell; cells beyond the
// taps' reach of every band exit at once.
template <typename R, int HALF>
__global__ __launch_bounds__(kBlock) void k_adj_pml_gather(Shot<R> S, Adj<R> A)
{
    const int nz = S.nz, ny = S.ny, nx = S.nx, sz = nx * ny;
    const int k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= sz * nz) return;
    const int kx = k % nx, ky = (k / nx) % ny, kz = k / sz;
    if (!near_band<0>(S, kz, ky, kx) && !near_band<1>(S, kz, ky, kx) &&
        !near_band<2>(S, kz, ky, kx))
        return;
    if constexpr (HALF == 1) {
        R d[6] = {0, 0, 0, 0, 0, 0};
        vel_band_t(S, A, k, kz, ky, kx, d[0], d[1], d[2], d[3], d[4], d[5]);
#pragma unroll
        for (int I = 0; I < 6; ++I) A.ls[I][k] += d[I];
    } else {
        R gx = 0, gy = 0, gz = 0;
        strain_band_t(S, A, k, kz, ky, kx, gx, gy, gz);
        A.lv[0][k] += S.inv_h * gx;
        A.lv[1][k] += S.inv_h * gy;
        A.lv[2][k] += S.inv_h * gz;
    }
}

// D^T: ls_I(k) += transposed stress-divergence stencils of e (and, near a
// CPML band, those of the band cells' derivative adjoints).
template <typename R>
__global__ __launch_bounds__(kBlock) void k_adj_velocity_direct(Shot<R> S, Adj<R> A)
{
    const int nz = S.nz, ny = S.ny, nx = S.nx, sz = nx * ny;
    const int ncell = sz * nz;
    const int k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= ncell) return;
    const int kx = k % nx, ky = (k / nx) % ny, kz = k / sz;
    auto ex = [&](int i, int iz, int iy, int ix) { return e_val<0>(S, A, i, iz, iy, ix); };
    auto ey = [&](int i, int iz, int iy, int ix) { return e_val<1>(S, A, i, iz, iy, ix); };
    auto ez = [&](int i, int iz, int iy, int ix) { return e_val<2>(S, A, i, iz, iy, ix); };
    R d1 = diff_t<true, 2>(S, ex, k, kz, ky, kx);
    R d2 = diff_t<true, 1>(S, ey, k, kz, ky, kx);
    R d3 = diff_t<true, 0>(S, ez, k, kz, ky, kx);
    R d4 = diff_t<false, 0>(S, ey, k, kz, ky, kx) + diff_t<false, 1>(S, ez, k, kz, ky, kx);
    R d5 = diff_t<false, 0>(S, ex, k, kz, ky, kx) + diff_t<false, 2>(S, ez, k, kz, ky, kx);
    R d6 = diff_t<false, 1>(S, ex, k, kz, ky, kx) + diff_t<false, 2>(S, ey, k, kz, ky, kx);
    vel_band_t(S, A, k, kz, ky, kx, d1, d2, d3, d4, d5, d6);
    A.ls[0][k] += d1;
    A.ls[1][k] += d2;
    A.ls[2][k] += d3;
    A.ls[3][k] += d4;
    A.ls[4][k] += d5;
    A.ls[5][k] += d6;
}

// u_{P,J}(i) = sum over I at position P with C_IJ active of dt * C_IJ[m(i)] * ls_I(i).
template <int P, int J, typename R>
__device__ __forceinline__ R u_val(const Shot<R> &S, const Adj<R> &A, int i)
{
    const R *Cm = S.Ct + (size_t)S.mats[i] * 36;
    R acc = 0;
#pragma unroll
    for (int I = 0; I < 6; ++I) {
        if (kVP[I] != P) continue;
        if (!((S.act >> (6 * I + J)) & 1ull)) continue;
        acc += S.dt * Cm[6 * I + J] * A.ls[I][i];
    }
    return acc;
}

// Guard of a moved strain at source coordinate c along an axis moved by d.
__device__ __forceinline__ bool move_ok(int d, int c, int n)
{
    return d > 0 ? (c >= 0 && c <= n - 2) : (d < 0 ? (c >= 1 && c <= n - 1) : true);
}

// Contribution of position P to gbar_J(k): the transposed move of u_{P,J}.
template <int P, int J, typename R>
__device__ __forceinline__ R moved_t(const Shot<R> &S, const Adj<R> &A, int k, int kz, int ky,
                                     int kx)
{
    constexpr int src = kVP[J];
    if constexpr (src == P) {
        return u_val<P, J>(S, A, k);
    } else {
        const int sy = S.nx, sz = S.nx * S.ny;
        constexpr int dz = kPos[P][0] - kPos[src][0];
        constexpr int dy = kPos[P][1] - kPos[src][1];
        constexpr int dx = kPos[P][2] - kPos[src][2];
        constexpr int z0 = dz < 0 ? -1 : 0, nzt = dz != 0 ? 2 : 1;
        constexpr int y0 = dy < 0 ? -1 : 0, nyt = dy != 0 ? 2 : 1;
        constexpr int x0 = dx < 0 ? -1 : 0, nxt = dx != 0 ? 2 : 1;
        R acc = 0;
#pragma unroll
        for (int a = 0; a < nzt; ++a)
#pragma unroll
            for (int b = 0; b < nyt; ++b)
#pragma unroll
                for (int c = 0; c < nxt; ++c) {
                    const int oz = z0 + a, oy = y0 + b, ox = x0 + c;
                    const int iz = kz - oz, iy = ky - oy, ix = kx - ox;
                    if (!move_ok(dz, iz, S.nz) || !move_ok(dy, iy, S.ny) || !move_ok(dx, ix, S.nx))
                        continue;
                    acc += u_val<P, J>(S, A, k - oz * sz - oy * sy - ox);
                }
        constexpr int ntap = nzt * nyt * nxt;
        return acc * (ntap == 4 ? (R)0.25 : (R)0.5);
    }
}

template <int J, typename R>
__device__ __forceinline__ R gbar_j(const Shot<R> &S, const Adj<R> &A, int k, int kz, int ky,
                                    int kx)
{
    return moved_t<0, J>(S, A, k, kz, ky, kx) + moved_t<1, J>(S, A, k, kz, ky, kx) +
           moved_t<2, J>(S, A, k, kz, ky, kx) + moved_t<3, J>(S, A, k, kz, ky, kx);
}

// gbar = M^T (dt C^T ls), into the scratch `work`.
template <typename R>
__global__ __launch_bounds__(kBlock) void k_adj_gbar_direct(Shot<R> S, Adj<R> A)
{
    const int nz = S.nz, ny = S.ny, nx = S.nx, sz = nx * ny;
    const int ncell = sz * nz;
    const int k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= ncell) return;
    const int kx = k % nx, ky = (k / nx) % ny, kz = k / sz;
    const size_t n = (size_t)ncell;
    A.work[0 * n + k] = gbar_j<0>(S, A, k, kz, ky, kx);
    A.work[1 * n + k] = gbar_j<1>(S, A, k, kz, ky, kx);
    A.work[2 * n + k] = gbar_j<2>(S, A, k, kz, ky, kx);
    A.work[3 * n + k] = gbar_j<3>(S, A, k, kz, ky, kx);
    A.work[4 * n + k] = gbar_j<4>(S, A, k, kz, ky, kx);
    A.work[5 * n + k] = gbar_j<5>(S, A, k, kz, ky, kx);
}

// lv = d * lv + inv_h * G^T gbar ; ls <- d * ls.
template <typename R>
__global__ __launch_bounds__(kBlock) void k_adj_strain_t_direct(Shot<R> S, Adj<R> A)
{
    const int nz = S.nz, ny = S.ny, nx = S.nx, sz = nx * ny;
    const int ncell = sz * nz;
    const int k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= ncell) return;
    const int kx = k % nx, ky = (k / nx) % ny, kz = k / sz;
    const size_t n = (size_t)ncell;
    const R *w = A.work;
    auto b1 = [&](int i, int, int, int) { return w[0 * n + i]; };
    auto b2 = [&](int i, int, int, int) { return w[1 * n + i]; };
    auto b3 = [&](int i, int, int, int) { return w[2 * n + i]; };
    auto b4 = [&](int i, int, int, int) { return w[3 * n + i]; };
    auto b5 = [&](int i, int, int, int) { return w[4 * n + i]; };
    auto b6 = [&](int i, int, int, int) { return w[5 * n + i]; };
    R gx = diff_t<false, 2>(S, b1, k, kz, ky, kx) + diff_t<true, 0>(S, b5, k, kz, ky, kx) +
           diff_t<true, 1>(S, b6, k, kz, ky, kx);
    R gy = diff_t<false, 1>(S, b2, k, kz, ky, kx) + diff_t<true, 0>(S, b4, k, kz, ky, kx) +
           diff_t<true, 2>(S, b6, k, kz, ky, kx);
    R gz = diff_t<false, 0>(S, b3, k, kz, ky, kx) + diff_t<true, 1>(S, b4, k, kz, ky, kx) +
           diff_t<true, 2>(S, b5, k, kz, ky, kx);
    strain_band_t(S, A, k, kz, ky, kx, gx, gy, gz);
    const R dmp = S.damp ? S.damp[k] : (R)1;
    A.lv[0][k] = A.lv[0][k] * dmp + S.inv_h * gx;
    A.lv[1][k] = A.lv[1][k] * dmp + S.inv_h * gy;
    A.lv[2][k] = A.lv[2][k] * dmp + S.inv_h * gz;
#pragma unroll
    for (int I = 0; I < 6; ++I) A.ls[I][k] = A.ls[I][k] * dmp;
}

// Stiffness gradient of one step: per (y, x) column, walking z; the
// strains of v^n sit in `work`.  NP = 0 accumulates all 36 entries
// dt * ls_I * q_{VP[I],J}; NP > 0 accumulates their projections on the
// per-material rows proj[m][p][0..35].
template <typename R, int NP>
__global__ __launch_bounds__(kColBlock) void k_adj_grad_direct(Shot<R> S, Adj<R> A,
                                                              const R *__restrict__ g)
{
    constexpr int K = NP == 0 ? 36 : NP;
    const int nz = S.nz, ny = S.ny, nx = S.nx, sy = nx, sz = nx * ny;
    const int col = blockIdx.x * kColBlock + threadIdx.x;
    if (col >= sz) return;
    const int ix = col % nx, iy = col / nx;
    const size_t n = (size_t)sz * nz;
    const double dt = (double)S.dt;
    double acc[K];
#pragma unroll
    for (int k = 0; k < K; ++k) acc[k] = 0.0;
    const int nchunk = A.zbreak > 0 ? (nz + A.zbreak - 1) / A.zbreak : 1;
    int run = A.run_ptr[(size_t)col * nchunk];
    int prev = S.mats[col];
    for (int iz = 0; iz < nz; ++iz) {
        const int i = iz * sz + col;
        const int m = S.mats[i];
        if (iz > 0 && (m != prev || (A.zbreak > 0 && iz % A.zbreak == 0))) {
            double *dst = A.run_acc + (size_t)run * K;
#pragma unroll
            for (int k = 0; k < K; ++k) {
                dst[k] += acc[k];
                acc[k] = 0.0;
            }
            ++run;
            prev = m;
        }
        R q[4][6];
#define PE_Q(P)                                                                                 \
    q[P][0] = moved_direct<P, 0>(g + 0 * n, i, iz, iy, ix, nz, ny, nx, sy, sz);                 \
    q[P][1] = moved_direct<P, 1>(g + 1 * n, i, iz, iy, ix, nz, ny, nx, sy, sz);                 \
    q[P][2] = moved_direct<P, 2>(g + 2 * n, i, iz, iy, ix, nz, ny, nx, sy, sz);                 \
    q[P][3] = moved_direct<P, 3>(g + 3 * n, i, iz, iy, ix, nz, ny, nx, sy, sz);                 \
    q[P][4] = moved_direct<P, 4>(g + 4 * n, i, iz, iy, ix, nz, ny, nx, sy, sz);                 \
    q[P][5] = moved_direct<P, 5>(g + 5 * n, i, iz, iy, ix, nz, ny, nx, sy, sz);
        PE_Q(0)
        PE_Q(1)
        PE_Q(2)
        PE_Q(3)
#undef PE_Q
#pragma unroll
        for (int I = 0; I < 6; ++I) {
            const double l = dt * (double)A.ls[I][i];
#pragma unroll
            for (int J = 0; J < 6; ++J) {
                const double t = l * (double)q[kVP[I]][J];
                if constexpr (NP == 0) {
                    acc[6 * I + J] += t;
                } else {
                    const R *pr = A.proj + (size_t)m * NP * 36;
#pragma unroll
                    for (int p = 0; p < NP; ++p) acc[p] += (double)pr[p * 36 + 6 * I + J] * t;
                }
            }
        }
    }
    double *dst = A.run_acc + (size_t)run * K;
#pragma unroll
    for (int k = 0; k < K; ++k) dst[k] += acc[k];
}

// Record / source transposes, one workgroup.  Phase 1 (scatter): for each
// receiver group r in order, its cells (unique within the group) get
// pressure: ls_1..3 -= res[r] * w_k / 3, or velocity: lv_a += res[r] * w_k.
// Phase 2 (gw != NULL): gw[n] = sum_k w_k (ls1 + ls2 + ls3)[src_k] in
// double, lanes in footprint order then a fixed butterfly.
template <typename R>
__global__ __launch_bounds__(kSrBlock) void k_adj_record(
    Adj<R> A, int n, int n_rec, const int32_t *__restrict__ rec_ptr,
    const int64_t *__restrict__ rec_idx, const R *__restrict__ rec_w, int mode,
    const R *__restrict__ res, int n_src, const int64_t *__restrict__ src_idx,
    const R *__restrict__ src_w, double *__restrict__ gw)
{
    if (res) {
        const R *row = res + (size_t)n * n_rec;
        for (int r = 0; r < n_rec; ++r) {
            const R rr = row[r];
            for (int k = rec_ptr[r] + (int)threadIdx.x; k < rec_ptr[r + 1]; k += kSrBlock) {
                const int64_t q = rec_idx[k];
                if (mode == PE_RECORD_PRESSURE) {
                    const R a = rr * rec_w[k] / (R)3;
                    A.ls[0][q] -= a;
                    A.ls[1][q] -= a;
                    A.ls[2][q] -= a;
                } else {
                    A.lv[mode - 1][q] += rr * rec_w[k];
                }
            }
            __syncthreads();
        }
    }
    if (gw && threadIdx.x < 64) {
        const int lane = threadIdx.x;
        double acc = 0.0;
        for (int k = lane; k < n_src; k += 64) {
            const int64_t q = src_idx[k];
            acc += (double)src_w[k] * ((double)A.ls[0][q] + (double)A.ls[1][q] + (double)A.ls[2][q]);
        }
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) acc += __shfl_xor(acc, off, 64);
        if (lane == 0) gw[n] = acc;
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_fill_zero(T *__restrict__ p, size_t count)
{
    const size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < count) p[i] = (T)0;
}

// dst[k*n + i] = v[k][i], k < 3: the velocity at the start of a step, for
// the adjoint's history.
template <typename R>
__global__ __launch_bounds__(kBlock) void k_store_v(R *__restrict__ dst, const R *__restrict__ v0,
                                                    const R *__restrict__ v1,
                                                    const R *__restrict__ v2, int n)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    dst[i] = v0[i];
    dst[(size_t)n + i] = v1[i];
    dst[2 * (size_t)n + i] = v2[i];
}

// grad[m][k] = sum of run_acc over the runs of material m.  One workgroup
// per material: thread t sums runs t, t+256, ... in ascending order, then a
// fixed tree over the 256 partials (deterministic).
__global__ __launch_bounds__(kBlock) void k_adj_finalize(int K, const int32_t *__restrict__ mat_ptr,
                                                         const int32_t *__restrict__ mat_runs,
                                                         const double *__restrict__ run_acc,
                                                         double *__restrict__ grad)
{
    __shared__ double part[kBlock];
    const int m = blockIdx.x;
    const int r0 = mat_ptr[m], r1 = mat_ptr[m + 1];
    for (int k = 0; k < K; ++k) {
        double s = 0.0;
        for (int r = r0 + (int)threadIdx.x; r < r1; r += kBlock)
            s += run_acc[(size_t)mat_runs[r] * K + k];
        part[threadIdx.x] = s;
        __syncthreads();
        for (int w = kBlock / 2; w >= 1; w >>= 1) {
            if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) grad[(size_t)m * K + k] = part[0];
        __syncthreads();
    }
}

inline unsigned blocks_of(size_t n, int b) { return (unsigned)((n + b - 1) / b); }

template <typename T>
hipError_t fill_zero(T *p, size_t count, hipStream_t st)
{
    if (!p || count == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_zero<T>, dim3(blocks_of(count, kBlock)), dim3(kBlock), 0, st, p,
                       count);
    return hipGetLastError();
}

}  // namespace

template <typename R>
hipError_t store_velocity(const Shot<R> &s, R *dst, hipStream_t st)
{
    const int ncell = s.nz * s.ny * s.nx;
    hipLaunchKernelGGL(k_store_v<R>, dim3(blocks_of(ncell, kBlock)), dim3(kBlock), 0, st, dst,
                       s.v[0], s.v[1], s.v[2], ncell);
    return hipGetLastError();
}

template <typename R>
hipError_t adjoint_reset(const pe_shot *p, const pe_adjoint *a, hipStream_t st)
{
    const size_t ncell = (size_t)p->nz * p->ny * p->nx;
    hipError_t e = hipSuccess;
    for (int k = 0; k < 3 && e == hipSuccess; ++k) e = fill_zero((R *)a->lv[k], ncell, st);
    for (int k = 0; k < 6 && e == hipSuccess; ++k) e = fill_zero((R *)a->ls[k], ncell, st);
    const int K = a->n_proj > 0 ? a->n_proj : 36;
    if (e == hipSuccess && a->run_acc) e = fill_zero(a->run_acc, (size_t)a->n_runs * K, st);
    if (p->cpml_width > 0) {
        const size_t W = 2 * (size_t)p->cpml_width + 2;
        const size_t vols[3] = {W * p->ny * p->nx, (size_t)p->nz * W * p->nx,
                                (size_t)p->nz * p->ny * W};
        for (int k = 0; k < 3 && e == hipSuccess; ++k)
            e = fill_zero((R *)a->cpml_psib[k], 6 * vols[k], st);
    }
    if (e == hipSuccess && a->kernel == PE_KERNEL_TILED) e = adjoint_face_weights<R>(p, a, st);
    return e;
}

template <typename R>
hipError_t adjoint_run(const pe_shot *p, const pe_adjoint *a, int32_t n_hi, int32_t n_lo,
                       const void *vhist, const void *residual, double *gw, hipStream_t st)
{
    Shot<R> S = make_shot<R>(p);
    const Adj<R> A = make_adj<R>(a);
    const int ncell = p->nz * p->ny * p->nx;
    const int ncol = p->ny * p->nx;
    const size_t nband = S.pml > 0 ? (size_t)pml_band_cells(S.nz, S.ny, S.nx, S.pml) : 0;
    const R *res = (const R *)residual;
    const bool vel_mode = p->record_mode != PE_RECORD_PRESSURE;
    const bool rec = res && p->n_rec > 0;
    Shot<R> Sv = S;
    Sv.g = A.work;
    hipError_t e = hipSuccess;
    const bool tiled = a->kernel == PE_KERNEL_TILED;
    for (int32_t n = n_hi - 1; n >= n_lo && e == hipSuccess; --n) {
        if (rec && vel_mode) {
            hipLaunchKernelGGL(k_adj_record<R>, dim3(1), dim3(kSrBlock), 0, st, A, n, p->n_rec,
                               p->rec_ptr, p->rec_idx, (const R *)p->rec_w, p->record_mode, res,
                               0, (const int64_t *)nullptr, (const R *)nullptr, (double *)nullptr);
        }
        if (tiled) {
            if (S.pml > 0)
                hipLaunchKernelGGL((k_adj_pml_pre<R, 1>), dim3(blocks_of(nband, kBlock)),
                                   dim3(kBlock), 0, st, S, A);
            // the first reverse step of a call starts from consistent adjoints;
            // later ones apply the previous step's deferred damping of ls
            e = adjoint_step_tiled<R>(S, Sv, A, 0, n != n_hi - 1, st);
            if (e != hipSuccess) break;
            if (S.pml > 0)
                hipLaunchKernelGGL((k_adj_pml_gather<R, 1>), dim3(blocks_of(ncell, kBlock)),
                                   dim3(kBlock), 0, st, S, A);
        } else {
            if (S.pml > 0)
                hipLaunchKernelGGL((k_adj_pml_pre<R, 1>), dim3(blocks_of(nband, kBlock)),
                                   dim3(kBlock), 0, st, S, A);
            hipLaunchKernelGGL(k_adj_velocity_direct<R>, dim3(blocks_of(ncell, kBlock)),
                               dim3(kBlock), 0, st, S, A);
        }
        const bool pres = rec && !vel_mode;
        const bool want_gw = gw && p->n_src > 0;
        if (pres || want_gw) {
            hipLaunchKernelGGL(k_adj_record<R>, dim3(1), dim3(kSrBlock), 0, st, A, n, p->n_rec,
                               p->rec_ptr, p->rec_idx, (const R *)p->rec_w, p->record_mode,
                               pres ? res : (const R *)nullptr, p->n_src, p->src_idx,
                               (const R *)p->src_w, want_gw ? gw : (double *)nullptr);
        }
        const bool gh = a->hist_kind == PE_HIST_STRAIN;
        const R *slot = nullptr;
        if (a->run_acc) {
            slot = (const R *)vhist + (size_t)(n - n_lo) * (gh ? 6 : 3) * ncell;
            if (!gh)
                for (int k = 0; k < 3; ++k) Sv.v[k] = (R *)(slot + (size_t)k * ncell);
        }
        if (a->run_acc && tiled && a->n_proj > 0) {
            e = adjoint_step_tiled<R>(S, Sv, A, 1, 0, st, gh ? slot : nullptr);
            if (e != hipSuccess) break;
        } else if (a->run_acc) {
            const R *g = gh ? slot : A.work;
            if (!gh) {
                e = launch_strain_direct<R>(Sv, st);
                if (e != hipSuccess) break;
            }
            const dim3 gb(blocks_of(ncol, kColBlock)), tb(kColBlock);
            switch (a->n_proj) {
            case 0: hipLaunchKernelGGL((k_adj_grad_direct<R, 0>), gb, tb, 0, st, S, A, g); break;
            case 1: hipLaunchKernelGGL((k_adj_grad_direct<R, 1>), gb, tb, 0, st, S, A, g); break;
            case 2: hipLaunchKernelGGL((k_adj_grad_direct<R, 2>), gb, tb, 0, st, S, A, g); break;
            case 3: hipLaunchKernelGGL((k_adj_grad_direct<R, 3>), gb, tb, 0, st, S, A, g); break;
            default: hipLaunchKernelGGL((k_adj_grad_direct<R, 4>), gb, tb, 0, st, S, A, g); break;
            }
        }
        if (tiled && S.pml > 0) {
            // M^T, the band cells' strain adjoints, G^T, the band gathers
            e = adjoint_step_tiled<R>(S, Sv, A, 3, 0, st);
            if (e != hipSuccess) break;
            hipLaunchKernelGGL((k_adj_pml_pre<R, 0>), dim3(blocks_of(nband, kBlock)),
                               dim3(kBlock), 0, st, S, A);
            e = adjoint_step_tiled<R>(S, Sv, A, 4, 0, st);
            if (e != hipSuccess) break;
            hipLaunchKernelGGL((k_adj_pml_gather<R, 0>), dim3(blocks_of(ncell, kBlock)),
                               dim3(kBlock), 0, st, S, A);
        } else if (tiled) {
            e = adjoint_step_tiled<R>(S, Sv, A, 2, 0, st);
            if (e != hipSuccess) break;
        } else {
            hipLaunchKernelGGL(k_adj_gbar_direct<R>, dim3(blocks_of(ncell, kBlock)), dim3(kBlock),
                               0, st, S, A);
            if (S.pml > 0)
                hipLaunchKernelGGL((k_adj_pml_pre<R, 0>), dim3(blocks_of(nband, kBlock)),
                                   dim3(kBlock), 0, st, S, A);
            hipLaunchKernelGGL(k_adj_strain_t_direct<R>, dim3(blocks_of(ncell, kBlock)),
                               dim3(kBlock), 0, st, S, A);
        }
        e = hipGetLastError();
    }
    // the tiled family defers the last step's damping of ls
    if (e == hipSuccess && tiled && n_hi > n_lo) e = adjoint_scale_ls<R>(p, a, st);
    return e;
}

hipError_t adjoint_finalize(const pe_shot *p, const pe_adjoint *a, double *grad, hipStream_t st)
{
    const int K = a->n_proj > 0 ? a->n_proj : 36;
    hipLaunchKernelGGL(k_adj_finalize, dim3(p->n_mat), dim3(kBlock), 0, st, K, a->mat_ptr,
                       a->mat_runs, (const double *)a->run_acc, grad);
    return hipGetLastError();
}

template hipError_t store_velocity<float>(const Shot<float> &, float *, hipStream_t);
template hipError_t store_velocity<double>(const Shot<double> &, double *, hipStream_t);
template hipError_t adjoint_reset<float>(const pe_shot *, const pe_adjoint *, hipStream_t);
template hipError_t adjoint_reset<double>(const pe_shot *, const pe_adjoint *, hipStream_t);
template hipError_t adjoint_run<float>(const pe_shot *, const pe_adjoint *, int32_t, int32_t,
                                       const void *, const void *, double *, hipStream_t);
template hipError_t adjoint_run<double>(const pe_shot *, const pe_adjoint *, int32_t, int32_t,
                                        const void *, const void *, double *, hipStream_t);

}  // namespace pe
