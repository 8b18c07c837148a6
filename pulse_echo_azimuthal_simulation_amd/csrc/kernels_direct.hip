// One-thread-per-cell kernels (PE_KERNEL_DIRECT) and the source/receiver
// kernel shared by both kernel families.
//
// The direct family mirrors the upstream two-kernel shape (strains into a
// scratch volume, then stress; sim/core/fdtd.py:1102-1177, 594-689) and the
// velocity update (sim/core/fdtd.py:1179-1257).  It is the simplest correct
// form of the step and serves as the cross-check for the tiled family.
//
// Arithmetic follows the upstream statement order exactly: every
// "acc += c * x" is one fused multiply-add; the stress update is
// fma(dt, acc, s * damp); the velocity update is (v + (dt_h * a) / rho_face)
// * damp.  Built with -ffp-contract=off so nothing else is fused, and with
// f32 subnormals flushed, which reproduces the upstream float32 archives.
#include "common.hpp"

namespace pe {

namespace {

constexpr int kBlock = 256;

// PML: the shot has a CPML boundary (band cells add pml_strain_corr).
template <typename R, bool PML = false>
__global__ __launch_bounds__(kBlock) void k_strain_direct(Shot<R> S)
{
    const int nz = S.nz, ny = S.ny, nx = S.nx;
    const int sy = nx, sz = nx * ny;
    const int ncell = sz * nz;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= ncell) return;
    const int ix = i % nx, iy = (i / nx) % ny, iz = i / sz;
    const R *__restrict__ vx = S.v[0];
    const R *__restrict__ vy = S.v[1];
    const R *__restrict__ vz = S.v[2];
    R a1 = 0, a2 = 0, a3 = 0, a4 = 0, a5 = 0, a6 = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        if (j >= S.nc) break;
        const R cj = S.c[j];
        if (ix >= j + 1 && ix < nx - j) a1 = fma_r(cj, vx[i + j] - vx[i - (j + 1)], a1);
        if (iy >= j + 1 && iy < ny - j) a2 = fma_r(cj, vy[i + j * sy] - vy[i - (j + 1) * sy], a2);
        if (iz >= j + 1 && iz < nz - j) a3 = fma_r(cj, vz[i + j * sz] - vz[i - (j + 1) * sz], a3);
        if (iy >= j && iy < ny - j - 1) a4 = fma_r(cj, vz[i + (j + 1) * sy] - vz[i - j * sy], a4);
        if (iz >= j && iz < nz - j - 1) a4 = fma_r(cj, vy[i + (j + 1) * sz] - vy[i - j * sz], a4);
        if (ix >= j && ix < nx - j - 1) a5 = fma_r(cj, vz[i + (j + 1)] - vz[i - j], a5);
        if (iz >= j && iz < nz - j - 1) a5 = fma_r(cj, vx[i + (j + 1) * sz] - vx[i - j * sz], a5);
        if (ix >= j && ix < nx - j - 1) a6 = fma_r(cj, vy[i + (j + 1)] - vy[i - j], a6);
        if (iy >= j && iy < ny - j - 1) a6 = fma_r(cj, vx[i + (j + 1) * sy] - vx[i - j * sy], a6);
    }
    R a[6] = {a1 * S.inv_h, a2 * S.inv_h, a3 * S.inv_h, a4 * S.inv_h, a5 * S.inv_h, a6 * S.inv_h};
    if constexpr (PML) {
        if (pml_cell(S, iz, iy, ix)) {
            R ga[6];
            pml_strain_corr(S, i, iz, iy, ix, ga);
            const int mask = pml_strain_mask(S, iz, iy, ix);
#pragma unroll
            for (int k = 0; k < 6; ++k)
                if (mask >> k & 1) a[k] = a[k] + ga[k];
        }
    }
    R *__restrict__ g = S.g;
    const size_t n = (size_t)ncell;
#pragma unroll
    for (int k = 0; k < 6; ++k) g[k * n + i] = a[k];
}

// Stress component I at position P = kVP[I]: acc = sum over active J of
// C_IJ * q_{P,J} (J ascending), then s = fma(dt, acc, s * damp).
template <int I, typename R>
__device__ __forceinline__ void stress_component(const Shot<R> &S, const R *__restrict__ Cm,
                                                 const R (&q)[6], R dmp, int i)
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
    R *__restrict__ sI = S.s[I];
    R sv = sI[i] * dmp;
    if (any) sv = fma_r(S.dt, acc, sv);
    sI[i] = sv;
}

template <int P, typename R>
__device__ __forceinline__ void gather_q(const Shot<R> &S, R (&q)[6], int i, int iz, int iy,
                                         int ix)
{
    const int nz = S.nz, ny = S.ny, nx = S.nx, sy = nx, sz = nx * ny;
    const size_t n = (size_t)nz * sz;
    const R *g = S.g;
    q[0] = moved_direct<P, 0>(g + 0 * n, i, iz, iy, ix, nz, ny, nx, sy, sz);
    q[1] = moved_direct<P, 1>(g + 1 * n, i, iz, iy, ix, nz, ny, nx, sy, sz);
    q[2] = moved_direct<P, 2>(g + 2 * n, i, iz, iy, ix, nz, ny, nx, sy, sz);
    q[3] = moved_direct<P, 3>(g + 3 * n, i, iz, iy, ix, nz, ny, nx, sy, sz);
    q[4] = moved_direct<P, 4>(g + 4 * n, i, iz, iy, ix, nz, ny, nx, sy, sz);
    q[5] = moved_direct<P, 5>(g + 5 * n, i, iz, iy, ix, nz, ny, nx, sy, sz);
}

template <typename R>
__global__ __launch_bounds__(kBlock) void k_stress_direct(Shot<R> S)
{
    const int nz = S.nz, ny = S.ny, nx = S.nx;
    const int sz = nx * ny;
    const int ncell = sz * nz;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= ncell) return;
    const int ix = i % nx, iy = (i / nx) % ny, iz = i / sz;
    const R *__restrict__ Cm = S.Ct + (size_t)S.mats[i] * 36;
    const R dmp = S.damp ? S.damp[i] : (R)1;
    R q[6];
    gather_q<0>(S, q, i, iz, iy, ix);
    stress_component<0>(S, Cm, q, dmp, i);
    stress_component<1>(S, Cm, q, dmp, i);
    stress_component<2>(S, Cm, q, dmp, i);
    gather_q<1>(S, q, i, iz, iy, ix);
    stress_component<3>(S, Cm, q, dmp, i);
    gather_q<2>(S, q, i, iz, iy, ix);
    stress_component<4>(S, Cm, q, dmp, i);
    gather_q<3>(S, q, i, iz, iy, ix);
    stress_component<5>(S, Cm, q, dmp, i);
}

template <typename R, bool PML = false>
__global__ __launch_bounds__(kBlock) void k_velocity_direct(Shot<R> S)
{
    const int nz = S.nz, ny = S.ny, nx = S.nx;
    const int sy = nx, sz = nx * ny;
    const int ncell = sz * nz;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= ncell) return;
    const int ix = i % nx, iy = (i / nx) % ny, iz = i / sz;
    const R *__restrict__ s1 = S.s[0];
    const R *__restrict__ s2 = S.s[1];
    const R *__restrict__ s3 = S.s[2];
    const R *__restrict__ s4 = S.s[3];
    const R *__restrict__ s5 = S.s[4];
    const R *__restrict__ s6 = S.s[5];
    R ax = 0, ay = 0, az = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        if (j >= S.nc) break;
        const R cj = S.c[j];
        if (ix >= j && ix < nx - j - 1) ax = fma_r(cj, s1[i + (j + 1)] - s1[i - j], ax);
        if (ix >= j + 1 && ix < nx - j) {
            ay = fma_r(cj, s6[i + j] - s6[i - (j + 1)], ay);
            az = fma_r(cj, s5[i + j] - s5[i - (j + 1)], az);
        }
        if (iy >= j && iy < ny - j - 1) ay = fma_r(cj, s2[i + (j + 1) * sy] - s2[i - j * sy], ay);
        if (iy >= j + 1 && iy < ny - j) {
            ax = fma_r(cj, s6[i + j * sy] - s6[i - (j + 1) * sy], ax);
            az = fma_r(cj, s4[i + j * sy] - s4[i - (j + 1) * sy], az);
        }
        if (iz >= j && iz < nz - j - 1) az = fma_r(cj, s3[i + (j + 1) * sz] - s3[i - j * sz], az);
        if (iz >= j + 1 && iz < nz - j) {
            ax = fma_r(cj, s5[i + j * sz] - s5[i - (j + 1) * sz], ax);
            ay = fma_r(cj, s4[i + j * sz] - s4[i - (j + 1) * sz], ay);
        }
    }
    if constexpr (PML) {
        if (pml_cell(S, iz, iy, ix)) {
            R gv[3];
            pml_accel_corr(S, i, iz, iy, ix, gv);
            ax = ax + gv[0];
            ay = ay + gv[1];
            az = az + gv[2];
        }
    }
    const R *__restrict__ rho = S.rho;
    const int32_t *__restrict__ mats = S.mats;
    const R r0 = rho[mats[i]];
    const R rvx = ix < nx - 1 ? (R)0.5 * (r0 + rho[mats[i + 1]]) : r0;
    const R rvy = iy < ny - 1 ? (R)0.5 * (r0 + rho[mats[i + sy]]) : r0;
    const R rvz = iz < nz - 1 ? (R)0.5 * (r0 + rho[mats[i + sz]]) : r0;
    const R dmp = S.damp ? S.damp[i] : (R)1;
    R *__restrict__ vx = S.v[0];
    R *__restrict__ vy = S.v[1];
    R *__restrict__ vz = S.v[2];
    vx[i] = (vx[i] + S.dt_h * ax / rvx) * dmp;
    vy[i] = (vy[i] + S.dt_h * ay / rvy) * dmp;
    vz[i] = (vz[i] + S.dt_h * az / rvz) * dmp;
}

// ------------------------------------------------------------ dense twin
// The dense-table form of the step (forward_fused, sim/core/fdtd.py:831-921,
// kernels stress_acc and vel_update, sim/core/fdtd.py:392-583): stiffness
// from 36 per-cell planes Cflat[(6I+J)*ncell + i], face densities from three
// per-cell volumes, no damping inside the half-steps (k_damp_dense applies
// it to all nine fields after the velocity update).  Same statement order as
// the label kernels, so where damp == 1 the two forms are bit-identical (the
// upstream's bit-equivalence gate, sim/core/validate_labels.py:66-100).
template <typename R>
__global__ __launch_bounds__(kBlock) void k_stress_dense(Shot<R> S, const R *__restrict__ Cflat,
                                                         uint64_t act)
{
    const int nz = S.nz, ny = S.ny, nx = S.nx;
    const int sz = nx * ny;
    const int ncell = sz * nz;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= ncell) return;
    const int ix = i % nx, iy = (i / nx) % ny, iz = i / sz;
    const size_t n = (size_t)ncell;
    R q[6];
#pragma unroll
    for (int I = 0; I < 6; ++I) {
        if (I == 0) gather_q<0>(S, q, i, iz, iy, ix);
        if (I == 3) gather_q<1>(S, q, i, iz, iy, ix);
        if (I == 4) gather_q<2>(S, q, i, iz, iy, ix);
        if (I == 5) gather_q<3>(S, q, i, iz, iy, ix);
        R acc = 0;
        bool any = false;
#pragma unroll
        for (int J = 0; J < 6; ++J) {
            if ((act >> (6 * I + J)) & 1ull) {
                acc = fma_r(Cflat[(size_t)(6 * I + J) * n + i], q[J], acc);
                any = true;
            }
        }
        if (any) S.s[I][i] = fma_r(S.dt, acc, S.s[I][i]);
    }
}

template <typename R>
__global__ __launch_bounds__(kBlock) void k_velocity_dense(Shot<R> S, const R *__restrict__ rvx,
                                                           const R *__restrict__ rvy,
                                                           const R *__restrict__ rvz)
{
    const int nz = S.nz, ny = S.ny, nx = S.nx;
    const int sy = nx, sz = nx * ny;
    const int ncell = sz * nz;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= ncell) return;
    const int ix = i % nx, iy = (i / nx) % ny, iz = i / sz;
    const R *__restrict__ s1 = S.s[0];
    const R *__restrict__ s2 = S.s[1];
    const R *__restrict__ s3 = S.s[2];
    const R *__restrict__ s4 = S.s[3];
    const R *__restrict__ s5 = S.s[4];
    const R *__restrict__ s6 = S.s[5];
    R ax = 0, ay = 0, az = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        if (j >= S.nc) break;
        const R cj = S.c[j];
        if (ix >= j && ix < nx - j - 1) ax = fma_r(cj, s1[i + (j + 1)] - s1[i - j], ax);
        if (ix >= j + 1 && ix < nx - j) {
            ay = fma_r(cj, s6[i + j] - s6[i - (j + 1)], ay);
            az = fma_r(cj, s5[i + j] - s5[i - (j + 1)], az);
        }
        if (iy >= j && iy < ny - j - 1) ay = fma_r(cj, s2[i + (j + 1) * sy] - s2[i - j * sy], ay);
        if (iy >= j + 1 && iy < ny - j) {
            ax = fma_r(cj, s6[i + j * sy] - s6[i - (j + 1) * sy], ax);
            az = fma_r(cj, s4[i + j * sy] - s4[i - (j + 1) * sy], az);
        }
        if (iz >= j && iz < nz - j - 1) az = fma_r(cj, s3[i + (j + 1) * sz] - s3[i - j * sz], az);
        if (iz >= j + 1 && iz < nz - j) {
            ax = fma_r(cj, s5[i + j * sz] - s5[i - (j + 1) * sz], ax);
            ay = fma_r(cj, s4[i + j * sz] - s4[i - (j + 1) * sz], ay);
        }
    }
    S.v[0][i] = S.v[0][i] + S.dt_h * ax / rvx[i];
    S.v[1][i] = S.v[1][i] + S.dt_h * ay / rvy[i];
    S.v[2][i] = S.v[2][i] + S.dt_h * az / rvz[i];
}

template <typename R>
__global__ __launch_bounds__(kBlock) void k_damp_dense(Shot<R> S)
{
    const int ncell = S.nx * S.ny * S.nz;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= ncell || !S.damp) return;
    const R d = S.damp[i];
#pragma unroll
    for (int k = 0; k < 3; ++k) S.v[k][i] = S.v[k][i] * d;
#pragma unroll
    for (int k = 0; k < 6; ++k) S.s[k][i] = S.s[k][i] * d;
}

// Source injection and receiver recording in one single-workgroup launch.
// Phase 1 adds (real)wavelet[n] * w_k to s1..s3 at the (unique) footprint
// cells (sim/core/fdtd.py:1798-1800).  Phase 2, after a barrier, forms each
// receiver group's weighted sum of -(s1+s2+s3)/3 or of one velocity
// component (sim/core/fdtd.py:1803-1813).  With at least one group per wave,
// wave w takes groups w, w+nw, ...; lane l sums cells l, l+64, ... in order;
// the 64 partials are combined by a fixed butterfly.  With fewer groups all
// 1024 threads share each group.  Either way the result is deterministic.
constexpr int kSrBlock = 1024;

template <typename R>
__device__ __forceinline__ void source_record_body(
    const Shot<R> &S, int n_src, const int64_t *__restrict__ src_idx, const R *__restrict__ src_w,
    const double *__restrict__ wavelet, int n, int do_source, int n_rec,
    const int32_t *__restrict__ rec_ptr, const int64_t *__restrict__ rec_idx,
    const R *__restrict__ rec_w, int mode, R *__restrict__ traces, int do_record)
{
    if (do_source) {
        const R wn = (R)wavelet[n];
        for (int k = threadIdx.x; k < n_src; k += kSrBlock) {
            const int64_t q = src_idx[k];
            const R add = wn * src_w[k];
            S.s[0][q] = S.s[0][q] + add;
            S.s[1][q] = S.s[1][q] + add;
            S.s[2][q] = S.s[2][q] + add;
        }
        __syncthreads();
    }
    if (!do_record) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    constexpr int nwave = kSrBlock / 64;
    auto cell = [&](int k) {
        const int64_t q = rec_idx[k];
        R val;
        if (mode == PE_RECORD_PRESSURE)
            val = -(S.s[0][q] + S.s[1][q] + S.s[2][q]) / (R)3;
        else
            val = S.v[mode - 1][q];
        return val * rec_w[k];
    };
    if (n_rec >= nwave) {
        for (int r = wave; r < n_rec; r += nwave) {
            R acc = 0;
            for (int k = rec_ptr[r] + lane; k < rec_ptr[r + 1]; k += 64) acc = acc + cell(k);
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) acc = acc + __shfl_xor(acc, off, 64);
            if (lane == 0) traces[(size_t)n * n_rec + r] = acc;
        }
        return;
    }
    // few groups (a piston): every wave sums a share of each group's cells;
    // thread t takes cells t, t+1024, ...; lanes by a fixed butterfly, waves
    // in order
    __shared__ R part[nwave];
    for (int r = 0; r < n_rec; ++r) {
        R acc = 0;
        for (int k = rec_ptr[r] + (int)threadIdx.x; k < rec_ptr[r + 1]; k += kSrBlock)
            acc = acc + cell(k);
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) acc = acc + __shfl_xor(acc, off, 64);
        if (lane == 0) part[wave] = acc;
        __syncthreads();
        if (threadIdx.x == 0) {
            R t = part[0];
            for (int w = 1; w < nwave; ++w) t = t + part[w];
            traces[(size_t)n * n_rec + r] = t;
        }
        __syncthreads();
    }
}

template <typename R>
__global__ __launch_bounds__(kSrBlock) void k_source_record(
    Shot<R> S, int n_src, const int64_t *__restrict__ src_idx, const R *__restrict__ src_w,
    const double *__restrict__ wavelet, int n, int do_source, int n_rec,
    const int32_t *__restrict__ rec_ptr, const int64_t *__restrict__ rec_idx,
    const R *__restrict__ rec_w, int mode, R *__restrict__ traces, int do_record)
{
    source_record_body<R>(S, n_src, src_idx, src_w, wavelet, n, do_source, n_rec, rec_ptr, rec_idx,
                          rec_w, mode, traces, do_record);
}

// Batched: workgroup s serves shot s at its own step.  Phase 0 (after the
// stress half-step): source, and the pressure record.  Phase 1 (after the
// velocity half-step): velocity-mode record, then the shot's step counter
// advances (the last kernel of a batched step to read it).
template <typename R>
__global__ __launch_bounds__(kSrBlock) void k_source_record_batched(
    const BatchEntry<R> *__restrict__ tab, int32_t *__restrict__ step, int phase)
{
    const int s = blockIdx.x;
    const BatchEntry<R> *e = tab + s;
    const int n = step[s];
    if (n >= e->nt) return;
    const Shot<R> S = e->S;
    const bool pres = e->mode == PE_RECORD_PRESSURE;
    const bool rec = e->traces != nullptr && e->n_rec > 0;
    if (phase == 0) {
        if (e->n_src > 0 || (pres && rec))
            source_record_body<R>(S, e->n_src, e->src_idx, e->src_w, e->wavelet, n,
                                  e->n_src > 0, e->n_rec, e->rec_ptr, e->rec_idx, e->rec_w,
                                  e->mode, e->traces, pres && rec);
        return;
    }
    if (!pres && rec)
        source_record_body<R>(S, 0, nullptr, nullptr, nullptr, n, 0, e->n_rec, e->rec_ptr,
                              e->rec_idx, e->rec_w, e->mode, e->traces, 1);
    __syncthreads();
    if (threadIdx.x == 0) step[s] = n + 1;
}

inline unsigned blocks_for(int ncell) { return (unsigned)((ncell + kBlock - 1) / kBlock); }

}  // namespace

template <typename R>
hipError_t launch_stress_direct(const Shot<R> &s, hipStream_t st)
{
    const int ncell = s.nz * s.ny * s.nx;
    if (s.pml > 0)
        hipLaunchKernelGGL((k_strain_direct<R, true>), dim3(blocks_for(ncell)), dim3(kBlock), 0, st, s);
    else
        hipLaunchKernelGGL(k_strain_direct<R>, dim3(blocks_for(ncell)), dim3(kBlock), 0, st, s);
    hipLaunchKernelGGL(k_stress_direct<R>, dim3(blocks_for(ncell)), dim3(kBlock), 0, st, s);
    return hipGetLastError();
}

template <typename R>
hipError_t launch_strain_direct(const Shot<R> &s, hipStream_t st)
{
    const int ncell = s.nz * s.ny * s.nx;
    if (s.pml > 0)
        hipLaunchKernelGGL((k_strain_direct<R, true>), dim3(blocks_for(ncell)), dim3(kBlock), 0, st, s);
    else
        hipLaunchKernelGGL(k_strain_direct<R>, dim3(blocks_for(ncell)), dim3(kBlock), 0, st, s);
    return hipGetLastError();
}

template <typename R>
hipError_t launch_velocity_direct(const Shot<R> &s, hipStream_t st)
{
    const int ncell = s.nz * s.ny * s.nx;
    if (s.pml > 0)
        hipLaunchKernelGGL((k_velocity_direct<R, true>), dim3(blocks_for(ncell)), dim3(kBlock), 0, st, s);
    else
        hipLaunchKernelGGL(k_velocity_direct<R>, dim3(blocks_for(ncell)), dim3(kBlock), 0, st, s);
    return hipGetLastError();
}

template <typename R>
hipError_t launch_dense_half(const Shot<R> &s, int half, const void *Cflat, uint64_t act,
                             const void *const rv[3], hipStream_t st)
{
    const int ncell = s.nx * s.ny * s.nz;
    const dim3 grid(blocks_for(ncell)), block(kBlock);
    if (half == 0) {
        hipLaunchKernelGGL(k_strain_direct<R>, grid, block, 0, st, s);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_stress_dense<R>, grid, block, 0, st, s, (const R *)Cflat, act);
    } else if (half == 1) {
        hipLaunchKernelGGL(k_velocity_dense<R>, grid, block, 0, st, s, (const R *)rv[0],
                           (const R *)rv[1], (const R *)rv[2]);
    } else {
        hipLaunchKernelGGL(k_damp_dense<R>, grid, block, 0, st, s);
    }
    return hipGetLastError();
}

template <typename R>
hipError_t launch_source_record(const Shot<R> &s, const pe_shot *p, const double *wavelet,
                                int32_t n, bool do_source, R *traces, bool do_record,
                                hipStream_t st)
{
    if (!(do_source && p->n_src > 0) && !(do_record && p->n_rec > 0)) return hipSuccess;
    hipLaunchKernelGGL(k_source_record<R>, dim3(1), dim3(kSrBlock), 0, st, s, p->n_src,
                       p->src_idx, (const R *)p->src_w, wavelet, n,
                       (int)(do_source && p->n_src > 0), p->n_rec, p->rec_ptr, p->rec_idx,
                       (const R *)p->rec_w, p->record_mode, traces,
                       (int)(do_record && p->n_rec > 0));
    return hipGetLastError();
}

template <typename R>
hipError_t launch_source_record_batched(const BatchEntry<R> *tab, int n_shot, const int32_t *step,
                                        int phase, hipStream_t st)
{
    hipLaunchKernelGGL(k_source_record_batched<R>, dim3(n_shot), dim3(kSrBlock), 0, st, tab,
                       (int32_t *)step, phase);
    return hipGetLastError();
}

template hipError_t launch_source_record_batched<float>(const BatchEntry<float> *, int,
                                                        const int32_t *, int, hipStream_t);
template hipError_t launch_source_record_batched<double>(const BatchEntry<double> *, int,
                                                         const int32_t *, int, hipStream_t);
template hipError_t launch_dense_half<float>(const Shot<float> &, int, const void *, uint64_t,
                                             const void *const[3], hipStream_t);
template hipError_t launch_dense_half<double>(const Shot<double> &, int, const void *, uint64_t,
                                              const void *const[3], hipStream_t);
template hipError_t launch_stress_direct<float>(const Shot<float> &, hipStream_t);
template hipError_t launch_stress_direct<double>(const Shot<double> &, hipStream_t);
template hipError_t launch_strain_direct<float>(const Shot<float> &, hipStream_t);
template hipError_t launch_strain_direct<double>(const Shot<double> &, hipStream_t);
template hipError_t launch_velocity_direct<float>(const Shot<float> &, hipStream_t);
template hipError_t launch_velocity_direct<double>(const Shot<double> &, hipStream_t);
template hipError_t launch_source_record<float>(const Shot<float> &, const pe_shot *,
                                                const double *, int32_t, bool, float *, bool,
                                                hipStream_t);
template hipError_t launch_source_record<double>(const Shot<double> &, const pe_shot *,
                                                 const double *, int32_t, bool, double *, bool,
                                                 hipStream_t);

}  // namespace pe
