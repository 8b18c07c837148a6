// C ABI of the gfx950 FDTD solver: argument checking and kernel dispatch.
// Declarations and the upstream code each entry point replaces are in
// include/pefdtd.h.
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "common.hpp"

namespace {

thread_local char g_err[512] = "";

int fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

int hip_fail(hipError_t e, const char *what)
{
    return fail(PE_EHIP, "%s: %s", what, hipGetErrorString(e));
}

template <typename R>
int stress_impl(const pe_shot *p, hipStream_t st)
{
    const pe::Shot<R> s = pe::make_shot<R>(p);
    hipError_t e = p->kernel == PE_KERNEL_DIRECT ? pe::launch_stress_direct<R>(s, st)
                                                 : pe::launch_stress_tiled<R>(s, p->tile[0], p->zchunk[0], st);
    return e == hipSuccess ? PE_OK : hip_fail(e, "stress launch");
}

template <typename R>
int velocity_impl(const pe_shot *p, hipStream_t st)
{
    const pe::Shot<R> s = pe::make_shot<R>(p);
    hipError_t e = p->kernel == PE_KERNEL_DIRECT ? pe::launch_velocity_direct<R>(s, st)
                                                 : pe::launch_velocity_tiled<R>(s, p->tile[1], p->zchunk[1], st);
    return e == hipSuccess ? PE_OK : hip_fail(e, "velocity launch");
}

template <typename R>
int source_record_impl(const pe_shot *p, const double *wavelet, int32_t n, bool src,
                       void *traces, bool rec, hipStream_t st)
{
    const pe::Shot<R> s = pe::make_shot<R>(p);
    hipError_t e = pe::launch_source_record<R>(s, p, wavelet, n, src, (R *)traces, rec, st);
    return e == hipSuccess ? PE_OK : hip_fail(e, "source/record launch");
}

// hist: NULL, or a per-step history written at the start of every step n
// into slot n - n0: hist_kind 0 = v (3 x ncell), 1 = the strains the stress
// half-step computes from v (6 x ncell).
template <typename R>
int run_impl(const pe_shot *p, int32_t n0, int32_t n1, const double *wavelet, void *traces,
             hipStream_t st, void *vhist = nullptr, int hist_kind = 0)
{
    pe::Shot<R> s = pe::make_shot<R>(p);
    const bool direct = p->kernel == PE_KERNEL_DIRECT;
    // Pressure is read from s, which the velocity half-step does not touch,
    // so source and record share one launch before the velocity update.
    const bool rec_before = p->record_mode == PE_RECORD_PRESSURE;
    const size_t ncell = (size_t)p->nz * p->ny * p->nx;
    for (int32_t n = n0; n < n1; ++n) {
        hipError_t e;
        if (vhist && hist_kind == 0) {
            e = pe::store_velocity<R>(s, (R *)vhist + (size_t)(n - n0) * 3 * ncell, st);
            if (e != hipSuccess) return hip_fail(e, "velocity store launch");
        }
        if (vhist && hist_kind == 1) {
            // the strain scratch of this step is the history slot itself
            pe::Shot<R> sg = s;
            sg.g = (R *)vhist + (size_t)(n - n0) * 6 * ncell;
            e = direct ? pe::launch_stress_direct<R>(sg, st)
                       : pe::launch_stress_tiled<R>(sg, p->tile[0], p->zchunk[0], st, true);
        } else {
            e = direct ? pe::launch_stress_direct<R>(s, st) : pe::launch_stress_tiled<R>(s, p->tile[0], p->zchunk[0], st);
        }
        if (e != hipSuccess) return hip_fail(e, "stress launch");
        e = pe::launch_source_record<R>(s, p, wavelet, n, true, (R *)traces,
                                        rec_before && traces != nullptr, st);
        if (e != hipSuccess) return hip_fail(e, "source/record launch");
        e = direct ? pe::launch_velocity_direct<R>(s, st) : pe::launch_velocity_tiled<R>(s, p->tile[1], p->zchunk[1], st);
        if (e != hipSuccess) return hip_fail(e, "velocity launch");
        if (!rec_before && traces != nullptr) {
            e = pe::launch_source_record<R>(s, p, wavelet, n, false, (R *)traces, true, st);
            if (e != hipSuccess) return hip_fail(e, "record launch");
        }
    }
    return PE_OK;
}

// Same sequence as run_impl with a hipEvent pair around every half-step
// kernel; per-kernel time is summed over the run.  Events are recorded in
// batches of kBatch steps, each batch read back after it completes.
template <typename R>
int run_timed_impl(const pe_shot *p, int32_t n0, int32_t n1, const double *wavelet,
                   void *traces, hipStream_t st, double *ms)
{
    constexpr int kBatch = 128;
    pe::Shot<R> s = pe::make_shot<R>(p);
    const bool direct = p->kernel == PE_KERNEL_DIRECT;
    const bool rec_before = p->record_mode == PE_RECORD_PRESSURE;
    hipEvent_t ev[kBatch][4];
    for (int b = 0; b < kBatch; ++b)
        for (int k = 0; k < 4; ++k)
            if (hipEventCreate(&ev[b][k]) != hipSuccess) return fail(PE_EHIP, "hipEventCreate");
    ms[0] = ms[1] = ms[2] = 0.0;
    int rc = PE_OK;
    for (int32_t b0 = n0; b0 < n1 && rc == PE_OK; b0 += kBatch) {
        const int nb = (n1 - b0) < kBatch ? (n1 - b0) : kBatch;
        for (int b = 0; b < nb && rc == PE_OK; ++b) {
            const int32_t n = b0 + b;
            (void)hipEventRecord(ev[b][0], st);
            hipError_t e = direct ? pe::launch_stress_direct<R>(s, st) : pe::launch_stress_tiled<R>(s, p->tile[0], p->zchunk[0], st);
            (void)hipEventRecord(ev[b][1], st);
            if (e == hipSuccess)
                e = pe::launch_source_record<R>(s, p, wavelet, n, true, (R *)traces,
                                                rec_before && traces != nullptr, st);
            (void)hipEventRecord(ev[b][2], st);
            if (e == hipSuccess)
                e = direct ? pe::launch_velocity_direct<R>(s, st) : pe::launch_velocity_tiled<R>(s, p->tile[1], p->zchunk[1], st);
            (void)hipEventRecord(ev[b][3], st);
            if (e == hipSuccess && !rec_before && traces != nullptr)
                e = pe::launch_source_record<R>(s, p, wavelet, n, false, (R *)traces, true, st);
            if (e != hipSuccess) rc = hip_fail(e, "timed step launch");
        }
        hipError_t e = hipEventSynchronize(ev[nb - 1][3]);
        if (e != hipSuccess && rc == PE_OK) rc = hip_fail(e, "hipEventSynchronize");
        for (int b = 0; b < nb && rc == PE_OK; ++b) {
            float t01 = 0, t12 = 0, t23 = 0;
            (void)hipEventElapsedTime(&t01, ev[b][0], ev[b][1]);
            (void)hipEventElapsedTime(&t12, ev[b][1], ev[b][2]);
            (void)hipEventElapsedTime(&t23, ev[b][2], ev[b][3]);
            ms[0] += t01;
            ms[1] += t23;
            ms[2] += t12;
        }
    }
    for (int b = 0; b < kBatch; ++b)
        for (int k = 0; k < 4; ++k) (void)hipEventDestroy(ev[b][k]);
    return rc;
}

// Layout of a batch table: entries (256-byte aligned block), then the
// workgroup -> shot maps of the stress and velocity launches and of the
// CPML band pre-passes (nb[2] blocks, 0 without CPML shots).
struct BatchLayout {
    size_t entry_bytes, blk_s_off, blk_v_off, blk_p_off, total;
    int64_t nb[3];
};

int64_t pml_blocks(const pe_shot *p)
{
    if (p->cpml_width <= 0) return 0;
    const long long cells = pe::pml_band_cells(p->nz, p->ny, p->nx, p->cpml_width);
    return (cells + pe::kPmlBlock - 1) / pe::kPmlBlock;
}

int batch_layout(const pe_shot *shots, int32_t n, BatchLayout *L)
{
    if (!shots || n < 1) return fail(PE_EINVAL, "a batch needs at least one shot");
    const int prec = shots[0].precision;
    L->nb[0] = L->nb[1] = L->nb[2] = 0;
    for (int32_t i = 0; i < n; ++i) {
        const pe_shot *p = &shots[i];
        int rc = pe_check_shot(p);
        if (rc) return fail(rc, "shot %d: %s", i, pe_last_error());
        if (p->precision != prec) return fail(PE_EINVAL, "shot %d: precision differs from shot 0", i);
        if (p->kernel != PE_KERNEL_TILED)
            return fail(PE_EINVAL, "shot %d: batches run the tiled kernels only", i);
        L->nb[2] += pml_blocks(p);
        for (int h = 0; h < 2; ++h) {
            int tx, ty, zl;
            const int nb = pe::tiled_geometry(prec, h, shots[0].tile[h], p->zchunk[h], p->nz, p->ny,
                                              p->nx, &tx, &ty, &zl);
            if (nb < 0) return fail(PE_EINVAL, "shot %d: tile variant", i);
            L->nb[h] += nb;
        }
    }
    if (L->nb[0] >= ((int64_t)1 << 31) || L->nb[1] >= ((int64_t)1 << 31) ||
        L->nb[2] >= ((int64_t)1 << 31))
        return fail(PE_ERANGE, "batch exceeds 2^31 workgroups");
    const size_t esz = prec == PE_F32 ? sizeof(pe::BatchEntry<float>) : sizeof(pe::BatchEntry<double>);
    L->entry_bytes = (esz * n + 255) / 256 * 256;
    L->blk_s_off = L->entry_bytes;
    L->blk_v_off = L->blk_s_off + ((size_t)L->nb[0] * 4 + 255) / 256 * 256;
    L->blk_p_off = L->blk_v_off + ((size_t)L->nb[1] * 4 + 255) / 256 * 256;
    L->total = L->blk_p_off + (size_t)L->nb[2] * 4;
    return PE_OK;
}

template <typename R>
int batch_setup_impl(pe_batch *b, const BatchLayout &L)
{
    std::vector<unsigned char> host(L.total, 0);
    auto *ent = reinterpret_cast<pe::BatchEntry<R> *>(host.data());
    auto *blk_s = reinterpret_cast<int32_t *>(host.data() + L.blk_s_off);
    auto *blk_v = reinterpret_cast<int32_t *>(host.data() + L.blk_v_off);
    auto *blk_p = reinterpret_cast<int32_t *>(host.data() + L.blk_p_off);
    int64_t off[3] = {0, 0, 0};
    for (int32_t i = 0; i < b->n_shot; ++i) {
        const pe_shot *p = &b->shots[i];
        pe::BatchEntry<R> &e = ent[i];
        e.S = pe::make_shot<R>(p);
        e.nt = b->nt[i];
        const int64_t np = pml_blocks(p);
        e.blk_off[2] = (int)off[2];
        for (int64_t k = 0; k < np; ++k) blk_p[off[2] + k] = i;
        off[2] += np;
        for (int h = 0; h < 2; ++h) {
            const int nb = pe::tiled_geometry(p->precision, h, b->shots[0].tile[h], p->zchunk[h],
                                              p->nz, p->ny, p->nx, &e.tiles_x[h], &e.tiles_y[h],
                                              &e.zc_len[h]);
            e.blk_off[h] = (int)off[h];
            int32_t *blk = h == 0 ? blk_s : blk_v;
            for (int k = 0; k < nb; ++k) blk[off[h] + k] = i;
            off[h] += nb;
        }
        e.n_src = p->n_src;
        e.n_rec = p->n_rec;
        e.mode = p->record_mode;
        e.src_idx = p->src_idx;
        e.src_w = (const R *)p->src_w;
        e.wavelet = b->wavelet[i];
        e.rec_ptr = p->rec_ptr;
        e.rec_idx = p->rec_idx;
        e.rec_w = (const R *)p->rec_w;
        e.traces = b->traces ? (R *)b->traces[i] : nullptr;
    }
    hipError_t e = hipMemcpy(b->table, host.data(), L.total, hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(e, "batch table upload");
    std::vector<int32_t> zeros(b->n_shot, 0);
    e = hipMemcpy(b->step, zeros.data(), sizeof(int32_t) * b->n_shot, hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(e, "batch step reset");
    b->nb[0] = L.nb[0];
    b->nb[1] = L.nb[1];
    b->nb[2] = L.nb[2];
    return PE_OK;
}

template <typename R>
int batch_step_impl(const pe_batch *b, const BatchLayout &L, int32_t steps, hipStream_t st)
{
    const auto *tab = reinterpret_cast<const pe::BatchEntry<R> *>(b->table);
    const auto *blk_s = reinterpret_cast<const int32_t *>((const char *)b->table + L.blk_s_off);
    const auto *blk_v = reinterpret_cast<const int32_t *>((const char *)b->table + L.blk_v_off);
    const auto *blk_p = reinterpret_cast<const int32_t *>((const char *)b->table + L.blk_p_off);
    const int vs = b->shots[0].tile[0];
    const int vv = b->shots[0].tile[1];
    int max_mat = 1, max_pml = 0;
    for (int32_t i = 0; i < b->n_shot; ++i) {
        max_mat = std::max(max_mat, (int)b->shots[i].n_mat);
        max_pml = std::max(max_pml, (int)b->shots[i].cpml_width);
    }
    // (the one-barrier pipelined stress body has no CPML corrections: with
    // CPML shots launch_stress_batched runs the plain kernel of its tile)
    const int np = (int)L.nb[2];
    for (int32_t k = 0; k < steps; ++k) {
        hipError_t e = pe::launch_pml_corr_batched<R>(0, tab, blk_p, np, b->step, st);
        if (e == hipSuccess)
            e = pe::launch_stress_batched<R>(vs, tab, blk_s, (int)L.nb[0], b->step, st, np > 0);
        if (e == hipSuccess) e = pe::launch_source_record_batched<R>(tab, b->n_shot, b->step, 0, st);
        if (e == hipSuccess)
            e = pe::launch_velocity_batched<R>(vv, tab, blk_v, (int)L.nb[1], b->step, st, max_mat,
                                               max_pml);
        if (e == hipSuccess) e = pe::launch_source_record_batched<R>(tab, b->n_shot, b->step, 1, st);
        if (e != hipSuccess) return hip_fail(e, "batched step launch");
    }
    return PE_OK;
}

int batch_check(const pe_batch *b, BatchLayout *L)
{
    if (!b) return fail(PE_EINVAL, "batch is NULL");
    int rc = batch_layout(b->shots, b->n_shot, L);
    if (rc) return rc;
    if (!b->wavelet || !b->nt || !b->table || !b->step)
        return fail(PE_EINVAL, "batch needs wavelet, nt, table and step");
    for (int32_t i = 0; i < b->n_shot; ++i) {
        if (b->nt[i] < 0) return fail(PE_EINVAL, "shot %d: negative step count", i);
        if (!b->wavelet[i] && b->shots[i].n_src > 0) return fail(PE_EINVAL, "shot %d: wavelet is NULL", i);
    }
    return PE_OK;
}

// Library-owned device memory of pe_batch_setup_noalloc: the batch table
// and the step counters of the last batch set up that way on each device.
// Host code never uses the array's name as an address: the per-device
// address comes from hipGetSymbolAddress, taken once per device with that
// device current, and only that address is passed to kernels and copies.
__device__ __align__(256) unsigned char scratch[PE_BATCH_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

int scratch_address(void **out)
{
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices) return fail(PE_ERANGE, "device %d beyond %d", dev, kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *ptr = nullptr;
        e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
        if (!ptr) return fail(PE_EHIP, "hipGetSymbolAddress(scratch) returned NULL");
        g_scratch_ptr[dev] = ptr;
    }
    *out = g_scratch_ptr[dev];
    return PE_OK;
}

}  // namespace

extern "C" {

int pe_abi_version(void) { return PE_ABI_VERSION; }

const char *pe_last_error(void) { return g_err; }

int pe_check_shot(const pe_shot *p)
{
    if (!p) return fail(PE_EINVAL, "shot descriptor is NULL");
    if (p->nz < 1 || p->ny < 1 || p->nx < 1)
        return fail(PE_EINVAL, "grid (%d, %d, %d) must be positive", p->nz, p->ny, p->nx);
    const int64_t ncell = (int64_t)p->nz * p->ny * p->nx;
    if (ncell >= ((int64_t)1 << 31))
        return fail(PE_ERANGE, "grid of %lld cells exceeds the 2^31-1 cell limit",
                    (long long)ncell);
    if (p->nc < 1 || p->nc > 4)
        return fail(PE_EINVAL,
                    "nc = %d stencil coefficients given: the kernels support order <= 8 "
                    "(1..4 coefficients)", p->nc);
    if (p->precision != PE_F32 && p->precision != PE_F64)
        return fail(PE_EINVAL, "precision %d is neither PE_F32 nor PE_F64", p->precision);
    if (p->kernel != PE_KERNEL_DIRECT && p->kernel != PE_KERNEL_TILED)
        return fail(PE_EINVAL, "unknown kernel family %d", p->kernel);
    // the tiled kernels address a field by a 32-bit byte offset
    const int64_t rbytes = p->precision == PE_F64 ? 8 : 4;
    if (p->kernel == PE_KERNEL_TILED && ncell * rbytes >= ((int64_t)1 << 32))
        return fail(PE_ERANGE,
                    "grid of %lld cells: the tiled kernels need each field under 4 GiB "
                    "(%lld bytes); use PE_KERNEL_DIRECT",
                    (long long)ncell, (long long)(ncell * rbytes));
    if (p->n_mat < 1) return fail(PE_EINVAL, "n_mat = %d must be >= 1", p->n_mat);
    if (!p->mats || !p->Ct || !p->rho_t) return fail(PE_EINVAL, "mats, Ct and rho_t are required");
    for (int k = 0; k < 3; ++k)
        if (!p->v[k]) return fail(PE_EINVAL, "velocity field %d is NULL", k);
    for (int k = 0; k < 6; ++k)
        if (!p->s[k]) return fail(PE_EINVAL, "stress field %d is NULL", k + 1);
    if (p->kernel == PE_KERNEL_DIRECT && !p->g)
        return fail(PE_EINVAL, "PE_KERNEL_DIRECT needs the 6*ncell strain scratch g");
    if (p->n_src < 0 || (p->n_src > 0 && (!p->src_idx || !p->src_w)))
        return fail(PE_EINVAL, "source footprint arrays missing");
    if (p->n_rec < 0 || (p->n_rec > 0 && (!p->rec_ptr || !p->rec_idx || !p->rec_w)))
        return fail(PE_EINVAL, "receiver arrays missing");
    for (int h = 0; h < 2; ++h) {
        int tx, ty, pad, fr;
        const int n = pe::tile_variants(p->precision, h, 0, &tx, &ty, &pad, &fr);
        if (p->tile[h] < 0 || p->tile[h] >= n)
            return fail(PE_EINVAL, "tile variant %d of half-step %d out of range [0, %d)",
                        p->tile[h], h, n);
        if (p->zchunk[h] < 0) return fail(PE_EINVAL, "zchunk must be >= 0");
    }
    if (p->record_mode < PE_RECORD_PRESSURE || p->record_mode > PE_RECORD_VZ)
        return fail(PE_EINVAL, "record mode %d unknown", p->record_mode);
    if (p->cpml_width < 0) return fail(PE_EINVAL, "cpml_width %d < 0", p->cpml_width);
    if (p->cpml_width > 0) {
        const int dims[3] = {p->nz, p->ny, p->nx};
        for (int a = 0; a < 3; ++a) {
            if (dims[a] < 2 * p->cpml_width + 2)
                return fail(PE_EINVAL, "CPML width %d needs every axis >= %d cells (axis %d has %d)",
                            p->cpml_width, 2 * p->cpml_width + 2, a, dims[a]);
            if (!p->cpml_prof[a] || !p->cpml_psi[a])
                return fail(PE_EINVAL, "CPML profile or memory variables of axis %d are NULL", a);
        }
        if (p->kernel == PE_KERNEL_TILED && !p->cpml_corr)
            return fail(PE_EINVAL, "a tiled CPML shot needs cpml_corr (6 x ncell)");
        // the tiled velocity kernel addresses the memory variables by 32-bit
        // byte offsets
        const int64_t W = 2 * (int64_t)p->cpml_width + 2;
        const int64_t vol[3] = {W * p->ny * p->nx, (int64_t)p->nz * W * p->nx,
                                (int64_t)p->nz * p->ny * W};
        for (int a = 0; a < 3; ++a)
            if (p->kernel == PE_KERNEL_TILED && 6 * vol[a] * rbytes >= ((int64_t)1 << 32))
                return fail(PE_ERANGE,
                            "CPML memory variables of axis %d take %lld bytes: the tiled "
                            "kernels need them under 4 GiB; use PE_KERNEL_DIRECT",
                            a, (long long)(6 * vol[a] * rbytes));
    }
    return PE_OK;
}

int pe_stress(const pe_shot *p, void *stream)
{
    int rc = pe_check_shot(p);
    if (rc) return rc;
    return p->precision == PE_F32 ? stress_impl<float>(p, (hipStream_t)stream)
                                  : stress_impl<double>(p, (hipStream_t)stream);
}

int pe_velocity(const pe_shot *p, void *stream)
{
    int rc = pe_check_shot(p);
    if (rc) return rc;
    return p->precision == PE_F32 ? velocity_impl<float>(p, (hipStream_t)stream)
                                  : velocity_impl<double>(p, (hipStream_t)stream);
}

int pe_source(const pe_shot *p, const double *wavelet, int32_t n, void *stream)
{
    int rc = pe_check_shot(p);
    if (rc) return rc;
    if (!wavelet) return fail(PE_EINVAL, "wavelet is NULL");
    if (n < 0) return fail(PE_EINVAL, "step %d is negative", n);
    return p->precision == PE_F32
               ? source_record_impl<float>(p, wavelet, n, true, nullptr, false, (hipStream_t)stream)
               : source_record_impl<double>(p, wavelet, n, true, nullptr, false, (hipStream_t)stream);
}

int pe_record(const pe_shot *p, void *traces, int32_t n, void *stream)
{
    int rc = pe_check_shot(p);
    if (rc) return rc;
    if (!traces) return fail(PE_EINVAL, "traces is NULL");
    if (n < 0) return fail(PE_EINVAL, "step %d is negative", n);
    return p->precision == PE_F32
               ? source_record_impl<float>(p, nullptr, n, false, traces, true, (hipStream_t)stream)
               : source_record_impl<double>(p, nullptr, n, false, traces, true, (hipStream_t)stream);
}

int pe_run(const pe_shot *p, int32_t n0, int32_t n1, const double *wavelet, void *traces,
           void *stream)
{
    int rc = pe_check_shot(p);
    if (rc) return rc;
    if (n0 < 0 || n1 < n0) return fail(PE_EINVAL, "bad step range [%d, %d)", n0, n1);
    if (!wavelet && p->n_src > 0) return fail(PE_EINVAL, "wavelet is NULL");
    return p->precision == PE_F32
               ? run_impl<float>(p, n0, n1, wavelet, traces, (hipStream_t)stream)
               : run_impl<double>(p, n0, n1, wavelet, traces, (hipStream_t)stream);
}

int pe_run_store(const pe_shot *p, int32_t n0, int32_t n1, const double *wavelet, void *traces,
                 void *vhist, void *stream)
{
    int rc = pe_check_shot(p);
    if (rc) return rc;
    if (n0 < 0 || n1 < n0) return fail(PE_EINVAL, "bad step range [%d, %d)", n0, n1);
    if (!wavelet && p->n_src > 0) return fail(PE_EINVAL, "wavelet is NULL");
    if (!vhist) return fail(PE_EINVAL, "vhist is NULL");
    return p->precision == PE_F32
               ? run_impl<float>(p, n0, n1, wavelet, traces, (hipStream_t)stream, vhist)
               : run_impl<double>(p, n0, n1, wavelet, traces, (hipStream_t)stream, vhist);
}

int pe_run_store_strains(const pe_shot *p, int32_t n0, int32_t n1, const double *wavelet,
                         void *traces, void *ghist, void *stream)
{
    int rc = pe_check_shot(p);
    if (rc) return rc;
    if (n0 < 0 || n1 < n0) return fail(PE_EINVAL, "bad step range [%d, %d)", n0, n1);
    if (!wavelet && p->n_src > 0) return fail(PE_EINVAL, "wavelet is NULL");
    if (!ghist) return fail(PE_EINVAL, "ghist is NULL");
    return p->precision == PE_F32
               ? run_impl<float>(p, n0, n1, wavelet, traces, (hipStream_t)stream, ghist, 1)
               : run_impl<double>(p, n0, n1, wavelet, traces, (hipStream_t)stream, ghist, 1);
}

int pe_tile_info(int32_t precision, int32_t half, int32_t variant, int32_t *info3)
{
    if (precision != PE_F32 && precision != PE_F64)
        return fail(PE_EINVAL, "precision %d is neither PE_F32 nor PE_F64", precision);
    if (half != 0 && half != 1) return fail(PE_EINVAL, "half must be 0 (stress) or 1 (velocity)");
    int tx = 0, ty = 0, pad = 0, fr = 0;
    const int n = pe::tile_variants(precision, half, variant, &tx, &ty, &pad, &fr);
    if (variant < 0 || variant >= n) return fail(PE_ERANGE, "variant %d >= %d", variant, n);
    if (info3) {
        info3[0] = tx;
        info3[1] = ty;
        info3[2] = pad | (fr << 8);
    }
    return n;
}

int pe_run_timed(const pe_shot *p, int32_t n0, int32_t n1, const double *wavelet,
                 void *traces, void *stream, double *ms_kernels)
{
    int rc = pe_check_shot(p);
    if (rc) return rc;
    if (n0 < 0 || n1 <= n0) return fail(PE_EINVAL, "bad step range [%d, %d)", n0, n1);
    if (!wavelet && p->n_src > 0) return fail(PE_EINVAL, "wavelet is NULL");
    if (!ms_kernels) return fail(PE_EINVAL, "ms_kernels is NULL");
    return p->precision == PE_F32
               ? run_timed_impl<float>(p, n0, n1, wavelet, traces, (hipStream_t)stream, ms_kernels)
               : run_timed_impl<double>(p, n0, n1, wavelet, traces, (hipStream_t)stream, ms_kernels);
}

int pe_check_adjoint(const pe_shot *p, const pe_adjoint *a)
{
    int rc = pe_check_shot(p);
    if (rc) return rc;
    if (!a) return fail(PE_EINVAL, "adjoint descriptor is NULL");
    if (p->cpml_width > 0) {
        if (a->run_acc && a->hist_kind != PE_HIST_STRAIN)
            return fail(PE_EINVAL, "the gradient of a CPML shot needs the strain history "
                                   "(PE_HIST_STRAIN): its strains depend on the memory variables");
        for (int k = 0; k < 3; ++k)
            if (!a->cpml_psib[k] || !a->cpml_work[k])
                return fail(PE_EINVAL, "CPML adjoint memory variables / scratch of axis %d are "
                                       "NULL", k);
    }
    for (int k = 0; k < 3; ++k)
        if (!a->lv[k]) return fail(PE_EINVAL, "velocity adjoint %d is NULL", k);
    for (int k = 0; k < 6; ++k)
        if (!a->ls[k]) return fail(PE_EINVAL, "stress adjoint %d is NULL", k + 1);
    if (!a->work) return fail(PE_EINVAL, "the 6*ncell adjoint scratch is NULL");
    if (a->n_proj < 0 || a->n_proj > 4)
        return fail(PE_EINVAL, "n_proj = %d outside [0, 4]", a->n_proj);
    if (a->n_proj > 0 && !a->proj) return fail(PE_EINVAL, "n_proj > 0 needs proj");
    if (a->zbreak < 0) return fail(PE_EINVAL, "zbreak must be >= 0");
    if (a->kernel != PE_KERNEL_DIRECT && a->kernel != PE_KERNEL_TILED)
        return fail(PE_EINVAL, "unknown adjoint kernel family %d", a->kernel);
    if (a->kernel == PE_KERNEL_TILED && !a->wface)
        return fail(PE_EINVAL, "the tiled adjoint needs the 3*ncell face-weight scratch");
    if (a->hist_kind != PE_HIST_VELOCITY && a->hist_kind != PE_HIST_STRAIN)
        return fail(PE_EINVAL, "unknown history kind %d", a->hist_kind);
    if (a->run_acc) {
        if (!a->run_ptr || !a->mat_ptr || !a->mat_runs)
            return fail(PE_EINVAL, "run_acc needs run_ptr, mat_ptr and mat_runs");
        const int64_t nchunk = a->zbreak > 0 ? (p->nz + a->zbreak - 1) / a->zbreak : 1;
        if ((int64_t)a->n_runs < (int64_t)p->ny * p->nx * nchunk)
            return fail(PE_EINVAL, "n_runs = %d is below the %lld column chunks", a->n_runs,
                        (long long)p->ny * p->nx * nchunk);
    }
    return PE_OK;
}

int pe_adjoint_reset(const pe_shot *p, const pe_adjoint *a, void *stream)
{
    int rc = pe_check_adjoint(p, a);
    if (rc) return rc;
    hipError_t e = p->precision == PE_F32
                       ? pe::adjoint_reset<float>(p, a, (hipStream_t)stream)
                       : pe::adjoint_reset<double>(p, a, (hipStream_t)stream);
    return e == hipSuccess ? PE_OK : hip_fail(e, "adjoint reset");
}

int pe_adjoint_run(const pe_shot *p, const pe_adjoint *a, int32_t n_hi, int32_t n_lo,
                   const void *vhist, const void *residual, double *grad_wavelet, void *stream)
{
    int rc = pe_check_adjoint(p, a);
    if (rc) return rc;
    if (n_lo < 0 || n_hi < n_lo) return fail(PE_EINVAL, "bad step range [%d, %d)", n_lo, n_hi);
    if (a->run_acc && !vhist && n_hi > n_lo)
        return fail(PE_EINVAL, "the stiffness gradient needs the velocity history");
    hipError_t e = p->precision == PE_F32
                       ? pe::adjoint_run<float>(p, a, n_hi, n_lo, vhist, residual, grad_wavelet,
                                                (hipStream_t)stream)
                       : pe::adjoint_run<double>(p, a, n_hi, n_lo, vhist, residual, grad_wavelet,
                                                 (hipStream_t)stream);
    return e == hipSuccess ? PE_OK : hip_fail(e, "adjoint step launch");
}

int pe_adjoint_gradient(const pe_shot *p, const pe_adjoint *a, double *grad, void *stream)
{
    int rc = pe_check_adjoint(p, a);
    if (rc) return rc;
    if (!a->run_acc) return fail(PE_EINVAL, "no run accumulators (run_acc is NULL)");
    if (!grad) return fail(PE_EINVAL, "grad is NULL");
    hipError_t e = pe::adjoint_finalize(p, a, grad, (hipStream_t)stream);
    return e == hipSuccess ? PE_OK : hip_fail(e, "adjoint gradient launch");
}

int pe_prep_rotate(const int32_t *src, int32_t nx_s, int32_t ny_s, int32_t nz_s, double h_spec,
                   double h, double ca, double sa, int32_t nd, int32_t ndz, int32_t m,
                   int32_t *lab, void *stream)
{
    if (!src || !lab) return fail(PE_EINVAL, "src and lab are required");
    if (nx_s < 1 || ny_s < 1 || nz_s < 1 || nd < 1 || ndz < 1 || m < 0 || h <= 0 || h_spec <= 0)
        return fail(PE_EINVAL, "bad re-grid geometry");
    const int64_t n = (int64_t)(ndz + 2 * m) * (nd + 2 * m) * (nd + 2 * m);
    if (n >= ((int64_t)1 << 31)) return fail(PE_ERANGE, "re-grid exceeds 2^31-1 cells");
    hipError_t e = pe::prep_rotate(src, nx_s, ny_s, nz_s, h_spec, h, ca, sa, nd, ndz, m, lab,
                                   (hipStream_t)stream);
    return e == hipSuccess ? PE_OK : hip_fail(e, "re-grid launch");
}

int pe_prep_label_power(const double *seeds, const double *w, int32_t n_seed, double h,
                        double D, int32_t nd, int32_t ndz, int32_t m, int32_t *lab, void *stream)
{
    if (!seeds || !w || !lab || n_seed < 1) return fail(PE_EINVAL, "seeds, weights and lab are required");
    if (nd < 1 || ndz < 1 || m < 0 || h <= 0 || D <= 0) return fail(PE_EINVAL, "bad grid");
    const int64_t n = (int64_t)(ndz + 2 * m) * (nd + 2 * m) * (nd + 2 * m);
    if (n >= ((int64_t)1 << 31)) return fail(PE_ERANGE, "grid exceeds 2^31-1 cells");
    hipError_t e = pe::prep_label_power(seeds, w, n_seed, h, D, nd, ndz, m, lab,
                                        (hipStream_t)stream);
    return e == hipSuccess ? PE_OK : hip_fail(e, "labeller launch");
}

int pe_prep_edt_damp(const int32_t *lab, int32_t nz, int32_t ny, int32_t nx, double fluid_damp,
                     const float *pz, const float *py, const float *px, int32_t *s0, int32_t *s1,
                     float *damp, void *stream)
{
    if (!lab || !pz || !py || !px || !s0 || !s1 || !damp)
        return fail(PE_EINVAL, "every array is required");
    if (nz < 1 || ny < 1 || nx < 1 || nz > 1024 || ny > 1024 || nx > 1024)
        return fail(PE_ERANGE, "grid sides must lie in [1, 1024]");
    hipError_t e = pe::prep_edt_damp(lab, nz, ny, nx, fluid_damp, pz, py, px, s0, s1, damp,
                                     (hipStream_t)stream);
    return e == hipSuccess ? PE_OK : hip_fail(e, "distance-transform launch");
}

int pe_prep_vmax(const int32_t *mats, int32_t nz, int32_t ny, int32_t nx, const void *Ct,
                 const void *rho_t, int32_t precision, double *partial, int32_t n_partial,
                 void *stream)
{
    if (!mats || !Ct || !rho_t || !partial || n_partial < 1)
        return fail(PE_EINVAL, "every array is required");
    if (nz < 1 || ny < 1 || nx < 1) return fail(PE_EINVAL, "bad grid");
    hipError_t e = precision == PE_F32
                       ? pe::prep_vmax<float>(mats, nz, ny, nx, (const float *)Ct,
                                              (const float *)rho_t, partial, n_partial,
                                              (hipStream_t)stream)
                       : pe::prep_vmax<double>(mats, nz, ny, nx, (const double *)Ct,
                                               (const double *)rho_t, partial, n_partial,
                                               (hipStream_t)stream);
    return e == hipSuccess ? PE_OK : hip_fail(e, "vmax launch");
}

int pe_dense_run(const pe_shot *p, const pe_dense *d, int32_t n0, int32_t n1,
                 const double *wavelet, void *traces, void *stream)
{
    int rc = pe_check_shot(p);
    if (rc) return rc;
    if (!d || !d->Cflat || !d->rv[0] || !d->rv[1] || !d->rv[2])
        return fail(PE_EINVAL, "the dense twin needs Cflat and three face-density volumes");
    if (!p->g) return fail(PE_EINVAL, "the dense twin needs the 6 x ncell strain scratch g");
    if (p->cpml_width > 0) return fail(PE_EINVAL, "the dense twin runs the sponge boundary only");
    if (n1 < n0 || n0 < 0) return fail(PE_EINVAL, "bad step range");
    if (!wavelet && p->n_src > 0) return fail(PE_EINVAL, "wavelet is NULL");
    hipStream_t st = (hipStream_t)stream;
    auto go = [&](auto tag) -> int {
        using R = decltype(tag);
        const pe::Shot<R> s = pe::make_shot<R>(p);
        for (int32_t n = n0; n < n1; ++n) {
            hipError_t e = pe::launch_dense_half<R>(s, 0, d->Cflat, d->act_mask, d->rv, st);
            if (e == hipSuccess)
                e = pe::launch_source_record<R>(s, p, wavelet, n, true, (R *)traces, false, st);
            if (e == hipSuccess) e = pe::launch_dense_half<R>(s, 1, d->Cflat, d->act_mask, d->rv, st);
            if (e == hipSuccess) e = pe::launch_dense_half<R>(s, 2, d->Cflat, d->act_mask, d->rv, st);
            if (e == hipSuccess && traces)
                e = pe::launch_source_record<R>(s, p, wavelet, n, false, (R *)traces, true, st);
            if (e != hipSuccess) return hip_fail(e, "dense step launch");
        }
        return PE_OK;
    };
    return p->precision == PE_F32 ? go(float(0)) : go(double(0));
}

int64_t pe_batch_table_bytes(const pe_shot *shots, int32_t n_shot)
{
    BatchLayout L;
    if (batch_layout(shots, n_shot, &L)) return 0;
    return (int64_t)L.total;
}

int pe_batch_setup(pe_batch *b, void *stream)
{
    (void)stream;
    BatchLayout L;
    int rc = batch_check(b, &L);
    if (rc) return rc;
    return b->shots[0].precision == PE_F32 ? batch_setup_impl<float>(b, L)
                                           : batch_setup_impl<double>(b, L);
}

int pe_batch_setup_noalloc(pe_batch *b, void *stream)
{
    (void)stream;
    if (!b) return fail(PE_EINVAL, "batch is NULL");
    BatchLayout L;
    int rc = batch_layout(b->shots, b->n_shot, &L);
    if (rc) return rc;
    // step counters first, then the table at the next 256-byte boundary
    const size_t step_bytes = ((size_t)b->n_shot * sizeof(int32_t) + 255) / 256 * 256;
    if (L.total > (size_t)PE_BATCH_SCRATCH_BYTES - step_bytes)
        return fail(PE_ERANGE,
                    "batch table of %zu bytes (+%zu bytes of step counters) exceeds the "
                    "%d-byte library scratch of pe_batch_setup_noalloc: use pe_batch_setup "
                    "with a caller-allocated table",
                    L.total, step_bytes, (int)PE_BATCH_SCRATCH_BYTES);
    void *base = nullptr;
    rc = scratch_address(&base);
    if (rc) return rc;
    b->step = reinterpret_cast<int32_t *>(base);
    b->table = reinterpret_cast<unsigned char *>(base) + step_bytes;
    rc = batch_check(b, &L);
    if (rc) return rc;
    return b->shots[0].precision == PE_F32 ? batch_setup_impl<float>(b, L)
                                           : batch_setup_impl<double>(b, L);
}

int pe_batch_step(const pe_batch *b, int32_t steps, void *stream)
{
    BatchLayout L;
    int rc = batch_check(b, &L);
    if (rc) return rc;
    if (b->nb[0] != L.nb[0] || b->nb[1] != L.nb[1] || b->nb[2] != L.nb[2])
        return fail(PE_EINVAL, "batch changed since pe_batch_setup");
    if (steps < 0) return fail(PE_EINVAL, "negative step count");
    return b->shots[0].precision == PE_F32
               ? batch_step_impl<float>(b, L, steps, (hipStream_t)stream)
               : batch_step_impl<double>(b, L, steps, (hipStream_t)stream);
}

int pe_batch_graph(const pe_batch *b, int32_t steps, void *stream, void **graph_exec)
{
    if (!graph_exec) return fail(PE_EINVAL, "graph_exec is NULL");
    hipStream_t st = (hipStream_t)stream;
    if (!st) return fail(PE_EINVAL, "graph capture needs a non-default stream");
    hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return hip_fail(e, "hipStreamBeginCapture");
    int rc = pe_batch_step(b, steps, stream);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(st, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
    hipGraphExec_t ex = nullptr;
    e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return hip_fail(e, "hipGraphInstantiate");
    *graph_exec = (void *)ex;
    return PE_OK;
}

int pe_graph_launch(void *graph_exec, void *stream)
{
    if (!graph_exec) return fail(PE_EINVAL, "graph_exec is NULL");
    hipError_t e = hipGraphLaunch((hipGraphExec_t)graph_exec, (hipStream_t)stream);
    return e == hipSuccess ? PE_OK : hip_fail(e, "hipGraphLaunch");
}

int pe_graph_destroy(void *graph_exec)
{
    if (!graph_exec) return PE_OK;
    hipError_t e = hipGraphExecDestroy((hipGraphExec_t)graph_exec);
    return e == hipSuccess ? PE_OK : hip_fail(e, "hipGraphExecDestroy");
}

}  // extern "C"
