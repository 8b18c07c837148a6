// Per-azimuth shot preparation on the GPU (the "GPU specimen" of a sweep).
//
// The upstream sweep prepares every azimuth on the host
// (sim/pipeline/sweep_runner.py:331-341): nearest-neighbour re-grid of the
// rotated disk (sim/core/rotation_test.py:77-91, 125-132), the couplant
// damping mask exp(-fluid_damp * EDT) (scipy distance_transform_edt), the
// damping volume sponge x mask (sim/core/fdtd.py:1693-1700) and the stable
// dt (sim/core/fdtd.py:1606-1615).  These kernels compute the same arrays
// from device-resident inputs:
//
//  * k_rotate_nn: the re-grid, every index computed with the host's float64
//    statements in the same order (no contraction), np.rint = rint
//    (round half to even), truncating casts: bit-identical labels;
//  * exact Euclidean distance transform by three separable passes of the
//    1-D lower envelope min_q g(q) + (p - q)^2 over squared integer
//    distances (exact), then exp(-fluid_damp * sqrt(d2)) in float64 cast
//    to float32, times the float32 sponge product (z * y) * x;
//  * k_vmax: max over cells of max(C11, C22, C33)[m] / min(rho, face
//    densities) in float64 (a max is exact in any order): the host takes the
//    square root and forms dt.
#include "common.hpp"

namespace pe {

namespace {

constexpr int kB = 256;
constexpr int32_t kInf = 1 << 29;

__global__ __launch_bounds__(kB) void k_rotate_nn(const int32_t *__restrict__ src, int nx_s,
                                                  int ny_s, int nz_s, double h_spec, double h,
                                                  double ca, double sa, int nd, int ndz, int m,
                                                  int nzg, int nxg, int32_t *__restrict__ lab)
{
    const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
    const int64_t n = (int64_t)nzg * nxg * nxg;
    if (i >= n) return;
    const int x = (int)(i % nxg), y = (int)((i / nxg) % nxg), z = (int)(i / ((int64_t)nxg * nxg));
    const int ci = x - m, cj = y - m, cz = z - m;
    int32_t out = -1;
    if (ci >= 0 && ci < nd && cj >= 0 && cj < nd && cz >= 0 && cz < ndz) {
        const double D = (double)nx_s * h_spec;
        const double X = ((double)ci + 0.5) * h - D / 2.0;
        const double Y = ((double)cj + 0.5) * h - D / 2.0;
        const double Xr = ca * X - sa * Y;
        const double Yr = sa * X + ca * Y;
        const double c0 = (double)(nx_s - 1) / 2.0;
        const long gi = (long)__builtin_rint(Xr / h_spec + c0);
        const long gj = (long)__builtin_rint(Yr / h_spec + c0);
        double fz = ((double)cz + 0.5) * h / h_spec;
        fz = fz < 0.0 ? 0.0 : (fz > (double)(nz_s - 1) ? (double)(nz_s - 1) : fz);
        const long gz = (long)fz;
        if (gi >= 0 && gi < nx_s && gj >= 0 && gj < nx_s)
            out = src[((int64_t)gi * ny_s + gj) * nz_s + gz];
    }
    lab[i] = out;
}

// Pass along x: one workgroup per row (z, y); f = 0 on specimen cells,
// kInf on couplant; out(p) = min_q f(q) + (p - q)^2.
__global__ __launch_bounds__(kB) void k_edt_x(const int32_t *__restrict__ lab, int nx,
                                              int32_t *__restrict__ out)
{
    extern __shared__ int32_t row[];
    const int64_t base = (int64_t)blockIdx.x * nx;
    for (int q = threadIdx.x; q < nx; q += kB) row[q] = lab[base + q] >= 0 ? 0 : kInf;
    __syncthreads();
    for (int p = threadIdx.x; p < nx; p += kB) {
        int32_t best = kInf;
        for (int q = 0; q < nx; ++q) {
            const int32_t d = p - q;
            const int32_t v = row[q] + d * d;
            best = v < best ? v : best;
        }
        out[base + p] = best;
    }
}

// Pass along y (axis stride nx, lines of length ny) or z (stride nx*ny,
// length nz): a workgroup takes 16 adjacent x columns of one line set and
// stages them in LDS as [len][16].
constexpr int kCols = 16;

__global__ __launch_bounds__(kB) void k_edt_line(const int32_t *__restrict__ in, int nx, int ny,
                                                 int nz, int axis, int32_t *__restrict__ out)
{
    extern __shared__ int32_t tile[];
    const int len = axis == 1 ? ny : nz;
    const int xb = (nx + kCols - 1) / kCols;
    const int bx = blockIdx.x % xb, other = blockIdx.x / xb;   // other: z (y pass) or y (z pass)
    const int x0 = bx * kCols;
    const int64_t sline = axis == 1 ? nx : (int64_t)nx * ny;
    const int64_t base = axis == 1 ? (int64_t)other * nx * ny : (int64_t)other * nx;
    const int c = threadIdx.x % kCols, r0 = threadIdx.x / kCols;
    constexpr int rstep = kB / kCols;
    const bool col_ok = x0 + c < nx;
    for (int q = r0; q < len; q += rstep)
        tile[q * kCols + c] = col_ok ? in[base + q * sline + x0 + c] : kInf;
    __syncthreads();
    if (!col_ok) return;
    for (int p = r0; p < len; p += rstep) {
        int32_t best = kInf;
        for (int q = 0; q < len; ++q) {
            const int32_t d = p - q;
            const int32_t g = tile[q * kCols + c];
            const int32_t v = g >= kInf ? kInf : g + d * d;
            best = v < best ? v : best;
        }
        out[base + p * sline + x0 + c] = best;
    }
}

// damp = ((pz[z] * py[y]) * px[x]) * (float)exp(-fluid_damp * sqrt(d2)).
__global__ __launch_bounds__(kB) void k_damp(const int32_t *__restrict__ d2, int nz, int ny,
                                             int nx, double fluid_damp,
                                             const float *__restrict__ pz,
                                             const float *__restrict__ py,
                                             const float *__restrict__ px,
                                             float *__restrict__ damp)
{
    const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
    const int64_t n = (int64_t)nz * ny * nx;
    if (i >= n) return;
    const int x = (int)(i % nx), y = (int)((i / nx) % ny), z = (int)(i / ((int64_t)nx * ny));
    const double dist = sqrt((double)d2[i]);
    const float mask = (float)exp(-fluid_damp * dist);
    const float sp = (pz[z] * py[y]) * px[x];
    damp[i] = sp * mask;
}

template <typename R>
__global__ __launch_bounds__(kB) void k_vmax(const int32_t *__restrict__ mats, int nz, int ny,
                                             int nx, const R *__restrict__ Ct,
                                             const R *__restrict__ rho,
                                             double *__restrict__ partial)
{
    __shared__ double red[kB];
    const int64_t n = (int64_t)nz * ny * nx;
    double best = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x; i < n; i += (int64_t)gridDim.x * kB) {
        const int x = (int)(i % nx), y = (int)((i / nx) % ny), z = (int)(i / ((int64_t)nx * ny));
        const int m = mats[i];
        const R *c = Ct + (size_t)m * 36;
        R cm = c[0];
        cm = c[7] > cm ? c[7] : cm;
        cm = c[14] > cm ? c[14] : cm;
        const double r0 = (double)rho[m];
        double rmin = r0;
        if (z < nz - 1) rmin = fmin(rmin, 0.5 * (r0 + (double)rho[mats[i + (int64_t)nx * ny]]));
        if (y < ny - 1) rmin = fmin(rmin, 0.5 * (r0 + (double)rho[mats[i + nx]]));
        if (x < nx - 1) rmin = fmin(rmin, 0.5 * (r0 + (double)rho[mats[i + 1]]));
        const double v = (double)cm / rmin;
        best = v > best ? v : best;
    }
    red[threadIdx.x] = best;
    __syncthreads();
    for (int w = kB / 2; w >= 1; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// Single rasterisation: the power diagram (Laguerre; Voronoi for zero
// weights) evaluated at the solver points.  Solver column (i, j) of plane z
// sits at (u_i, u_j, zc_z), u_i = (i + 0.5) h - D/2, zc_z = (z - (ndz-1)/2) h;
// inside the disk (u_i^2 + u_j^2 <= (D/2)^2) it takes the seed s minimising
// |x - s|^2 - w_s over all seeds (float64, (dx^2 + dy^2) + dz^2, lowest
// index on ties), outside -1.  Seeds already rotated by the caller
// (rotation_test.py:108-124).
__global__ __launch_bounds__(kB) void k_label_power(const double *__restrict__ seeds,
                                                    const double *__restrict__ w, int n_seed,
                                                    double h, double D, int nd, int ndz, int m,
                                                    int nzg, int nxg, int32_t *__restrict__ lab)
{
    const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
    const int64_t n = (int64_t)nzg * nxg * nxg;
    if (i >= n) return;
    const int x = (int)(i % nxg), y = (int)((i / nxg) % nxg), z = (int)(i / ((int64_t)nxg * nxg));
    const int ci = x - m, cj = y - m, cz = z - m;
    int32_t out = -1;
    if (ci >= 0 && ci < nd && cj >= 0 && cj < nd && cz >= 0 && cz < ndz) {
        const double X = ((double)ci + 0.5) * h - D / 2.0;
        const double Y = ((double)cj + 0.5) * h - D / 2.0;
        const double Z = ((double)cz - (double)(ndz - 1) / 2.0) * h;
        const double R = 0.5 * D;
        if (X * X + Y * Y <= R * R) {
            double best = 0.0;
            int arg = -1;
            for (int s = 0; s < n_seed; ++s) {
                const double dx = X - seeds[3 * s], dy = Y - seeds[3 * s + 1],
                             dz = Z - seeds[3 * s + 2];
                const double d = (dx * dx + dy * dy) + dz * dz - w[s];
                if (arg < 0 || d < best) {
                    best = d;
                    arg = s;
                }
            }
            out = arg;
        }
    }
    lab[i] = out;
}

inline unsigned nblk(int64_t n) { return (unsigned)((n + kB - 1) / kB); }

}  // namespace

hipError_t prep_rotate(const int32_t *src, int nx_s, int ny_s, int nz_s, double h_spec, double h,
                       double ca, double sa, int nd, int ndz, int m, int32_t *lab, hipStream_t st)
{
    const int nzg = ndz + 2 * m, nxg = nd + 2 * m;
    const int64_t n = (int64_t)nzg * nxg * nxg;
    hipLaunchKernelGGL(k_rotate_nn, dim3(nblk(n)), dim3(kB), 0, st, src, nx_s, ny_s, nz_s, h_spec,
                       h, ca, sa, nd, ndz, m, nzg, nxg, lab);
    return hipGetLastError();
}

hipError_t prep_label_power(const double *seeds, const double *w, int n_seed, double h, double D,
                            int nd, int ndz, int m, int32_t *lab, hipStream_t st)
{
    const int nzg = ndz + 2 * m, nxg = nd + 2 * m;
    const int64_t n = (int64_t)nzg * nxg * nxg;
    hipLaunchKernelGGL(k_label_power, dim3(nblk(n)), dim3(kB), 0, st, seeds, w, n_seed, h, D, nd,
                       ndz, m, nzg, nxg, lab);
    return hipGetLastError();
}

hipError_t prep_edt_damp(const int32_t *lab, int nz, int ny, int nx, double fluid_damp,
                         const float *pz, const float *py, const float *px, int32_t *s0,
                         int32_t *s1, float *damp, hipStream_t st)
{
    hipLaunchKernelGGL(k_edt_x, dim3((unsigned)((int64_t)nz * ny)), dim3(kB),
                       sizeof(int32_t) * nx, st, lab, nx, s0);
    const int xb = (nx + kCols - 1) / kCols;
    hipLaunchKernelGGL(k_edt_line, dim3((unsigned)((int64_t)xb * nz)), dim3(kB),
                       sizeof(int32_t) * kCols * ny, st, s0, nx, ny, nz, 1, s1);
    hipLaunchKernelGGL(k_edt_line, dim3((unsigned)((int64_t)xb * ny)), dim3(kB),
                       sizeof(int32_t) * kCols * nz, st, s1, nx, ny, nz, 2, s0);
    const int64_t n = (int64_t)nz * ny * nx;
    hipLaunchKernelGGL(k_damp, dim3(nblk(n)), dim3(kB), 0, st, s0, nz, ny, nx, fluid_damp, pz, py,
                       px, damp);
    return hipGetLastError();
}

template <typename R>
hipError_t prep_vmax(const int32_t *mats, int nz, int ny, int nx, const R *Ct, const R *rho,
                     double *partial, int nblocks, hipStream_t st)
{
    hipLaunchKernelGGL(k_vmax<R>, dim3(nblocks), dim3(kB), 0, st, mats, nz, ny, nx, Ct, rho,
                       partial);
    return hipGetLastError();
}

template hipError_t prep_vmax<float>(const int32_t *, int, int, int, const float *, const float *,
                                     double *, int, hipStream_t);
template hipError_t prep_vmax<double>(const int32_t *, int, int, int, const double *,
                                      const double *, double *, int, hipStream_t);

}  // namespace pe
