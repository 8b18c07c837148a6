/*
 * pefdtd -- C ABI of the MI355X (gfx950) label-indexed anisotropic elastic
 * staggered-grid FDTD solver.
 *
 * The upstream solver has no C ABI: its hot path is the Python function
 * sim/core/fdtd.py::forward_fused_labels (lines 1618-1816) launching
 * NVRTC-compiled CUDA kernels through cupy.RawKernel.  Each entry point below
 * names the upstream code it replaces.  The Python package
 * pulse_echo_azimuthal_simulation_amd binds them with ctypes and keeps
 * forward_fused_labels' signature on top.
 *
 * Conventions
 *  - Every pointer inside pe_shot is a DEVICE pointer (HBM), owned by the
 *    caller.  The library allocates nothing.
 *  - Grids are C order (nz, ny, nx), x fastest; flat index (iz*ny+iy)*nx+ix.
 *  - Real arrays are float (precision PE_F32) or double (PE_F64).
 *  - Functions return PE_OK (0) or a negative PE_E* code; pe_last_error()
 *    then holds a message (thread-local).
 *  - `stream` is a hipStream_t (NULL = the default stream).  All calls are
 *    asynchronous on that stream.
 */
#ifndef PEFDTD_H
#define PEFDTD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PE_ABI_VERSION 10

enum { PE_OK = 0, PE_EINVAL = -1, PE_EHIP = -2, PE_ERANGE = -3 };
enum { PE_F32 = 0, PE_F64 = 1 };
enum { PE_RECORD_PRESSURE = 0, PE_RECORD_VX = 1, PE_RECORD_VY = 2, PE_RECORD_VZ = 3 };
/* Kernel family for the two half-steps.
 *  PE_KERNEL_DIRECT: one thread per cell, strains materialised in `g`
 *                    (the upstream two-kernel shape; needs g).
 *  PE_KERNEL_TILED:  z-marching LDS-tiled kernels, strains computed on the
 *                    fly inside the stress kernel (no g). */
enum { PE_KERNEL_DIRECT = 0, PE_KERNEL_TILED = 1 };
/* Per-step history the adjoint's stiffness gradient reads. */
enum { PE_HIST_VELOCITY = 0, PE_HIST_STRAIN = 1 };

/* One shot: medium, fields, source and receivers.  Replaces the device state
 * forward_fused_labels builds at sim/core/fdtd.py:1678-1784. */
typedef struct pe_shot {
    int32_t nz, ny, nx;          /* grid; nz*ny*nx < 2^31 */
    int32_t nc;                  /* staggered stencil half-length, 1..4 */
    double coeffs[4];            /* c_0..c_{nc-1}, cast to the real type */
    double dt;                   /* time step (s) */
    double inv_h;                /* 1 / h */
    double dt_h;                 /* dt / h */
    int32_t n_mat;               /* rows of Ct / rho_t */
    int32_t precision;           /* PE_F32 or PE_F64 */
    uint64_t act_mask;           /* bit 6I+J: some material has C_IJ != 0 */
    const int32_t *mats;         /* material index per cell = label + 1 */
    const void *Ct;              /* n_mat x 36 stiffness rows (C_IJ at 6I+J) */
    const void *rho_t;           /* n_mat densities */
    const void *damp;            /* per-cell damping factor, or NULL (= 1) */
    void *v[3];                  /* vx, vy, vz */
    void *s[6];                  /* s1..s6 (Voigt order xx, yy, zz, yz, xz, xy) */
    void *g;                     /* 6*ncell strain scratch (PE_KERNEL_DIRECT) */
    int32_t n_src;               /* source footprint cells */
    const int64_t *src_idx;      /* flat indices, unique */
    const void *src_w;           /* weights */
    int32_t n_rec;               /* receiver groups */
    const int32_t *rec_ptr;      /* n_rec+1 offsets into rec_idx / rec_w */
    const int64_t *rec_idx;      /* flat indices */
    const void *rec_w;           /* weights */
    int32_t record_mode;         /* PE_RECORD_* */
    int32_t kernel;              /* PE_KERNEL_* */
    int32_t tile[2];             /* PE_KERNEL_TILED tile variant: [0] stress, [1] velocity */
    int32_t zchunk[2];           /* z planes per workgroup, 0 = automatic */
    /* CPML absorbing boundary (no upstream counterpart: the upstream absorber
     * is the sponge folded into damp).  cpml_width L = 0: none.  Otherwise
     * axis a (0 z, 1 y, 2 x, length n >= 2L+2) has a band of L+1 cells at
     * each end (i <= L or i >= n-1-L, band index i or i-(n-2L-2)).  A cell
     * in any band is the plain cell plus corrections: for each band axis (z,
     * then y, then x) each of its three derivatives dd, taken on its own, is
     * passed through psi' = fma(b, psi, a*dd) and contributes
     * fma(dd, 1/kappa - 1, psi') to its strain (or acceleration), so that
     * the stretched derivative is dd/kappa + psi'.
     *   cpml_prof[a]: real[2][3][n]: [stagger: 0 integer points, 1 half
     *                 points][b, a, 1/kappa][index]
     *   cpml_psi[a]:  real[6][band volume], zero-initialised, advanced in
     *                 place; band volume (2L+2)*ny*nx (z), nz*(2L+2)*nx (y),
     *                 nz*ny*(2L+2) (x); components 0..2 belong to the stress
     *                 half-step, 3..5 to the velocity half-step
     *   cpml_corr:    real[6][ncell] (PE_KERNEL_TILED): the corrections a
     *                 band pre-pass writes for the tiled kernels to add (no
     *                 initial value needed; only band cells are used)
     * Forward runs and batches only (the adjoint and the strain history
     * refuse it). */
    int32_t cpml_width;
    int32_t cpml_reserved;       /* set to 0 */
    const void *cpml_prof[3];
    void *cpml_psi[3];
    void *cpml_corr;
} pe_shot;

/* ABI version (PE_ABI_VERSION) and the last error message of this thread. */
int pe_abi_version(void);
const char *pe_last_error(void);

/* Validate a shot descriptor without touching device memory. */
int pe_check_shot(const pe_shot *shot);

/* Stress half-step: strains from v, moved strains, label-indexed stiffness,
 * deferred damping s <- s*damp + dt*C:q.
 * Replaces strains_zp + stress_lab_u and their slab loop,
 * sim/core/fdtd.py:1102-1177, 594-689, 1787-1797. */
int pe_stress(const pe_shot *shot, void *stream);

/* Explosive source s1..s3[k] += (real)wavelet[n] * w_k.
 * Replaces sim/core/fdtd.py:1798-1800.  wavelet: device double array. */
int pe_source(const pe_shot *shot, const double *wavelet, int32_t n, void *stream);

/* Velocity half-step with face densities from labels and damping.
 * Replaces vel_zp, sim/core/fdtd.py:1179-1257, 1801-1802. */
int pe_velocity(const pe_shot *shot, void *stream);

/* Receiver groups into row n of traces (nt x n_rec, real, device).
 * Replaces sim/core/fdtd.py:1803-1813. */
int pe_record(const pe_shot *shot, void *traces, int32_t n, void *stream);

/* Steps n0..n1-1 of the time loop: stress, source, velocity, record.
 * Replaces the hot loop sim/core/fdtd.py:1786-1815 (without the host
 * progress callback, which the caller runs between pe_run calls). */
int pe_run(const pe_shot *shot, int32_t n0, int32_t n1, const double *wavelet,
           void *traces, void *stream);

/* pe_run that also stores v at the start of every step n into
 * vhist + (n-n0)*3*ncell (vx, vy, vz planes), the history the adjoint's
 * stiffness gradient reads.  Replaces the history store of
 * vendor/ringfwi/elastic3d.py:301-305 (_grad_forward). */
int pe_run_store(const pe_shot *shot, int32_t n0, int32_t n1, const double *wavelet,
                 void *traces, void *vhist, void *stream);

/* pe_run that also stores the strains g1..g6 of v at the start of every
 * step n (the values its stress half-step computes, 6 x ncell per step) into
 * ghist + (n-n0)*6*ncell: the history of PE_HIST_STRAIN adjoints, which then
 * need not recompute them; with a CPML the strains with their band
 * corrections.  No upstream counterpart (the upstream keeps
 * velocities, vendor/ringfwi/elastic3d.py:301-305). */
int pe_run_store_strains(const pe_shot *shot, int32_t n0, int32_t n1, const double *wavelet,
                         void *traces, void *ghist, void *stream);

/* Dense-table twin of the step: the upstream forward_fused
 * (sim/core/fdtd.py:831-921, kernels stress_acc / vel_update at
 * sim/core/fdtd.py:392-583), the partner of the upstream's bit-equivalence
 * gate (sim/core/validate_labels.py:66-100).  Stiffness as 36 per-cell
 * planes Cflat[(6I+J)*ncell + i] (real; act_mask bit 6I+J: plane nonzero
 * somewhere), face densities as three per-cell volumes rv[a]; per step:
 * strains into shot->g and stress (no damping), source, velocity (no
 * damping), then v and s times damp, then the record (post-damp).  Where
 * damp == 1 the traces equal pe_run's on the matching label volume bit for
 * bit.  The shot's mats / Ct / rho_t are not read; sponge boundary only. */
typedef struct pe_dense {
    const void *Cflat;           /* 36 x ncell real */
    uint64_t act_mask;
    const void *rv[3];           /* face densities x, y, z: ncell real each */
} pe_dense;
int pe_dense_run(const pe_shot *shot, const pe_dense *dense, int32_t n0, int32_t n1,
                 const double *wavelet, void *traces, void *stream);

/* Tile variants of the tiled kernels, for tuning.  half: 0 stress, 1
 * velocity.  Returns the number of variants at that precision (>= 1) and,
 * when info3 is not NULL, writes (tile x, tile y, flags) of `variant`:
 * flags bits 0-7 = LDS row padding, bits 8-9 = stress-kernel mode: 1 the
 * frame ring (each v plane's frame fetched together with the z rings' plane
 * and kept in LDS until used), 2 the one-barrier pipeline (v tiles double
 * buffered, strain ring 4 planes deep); PE_ERANGE when `variant` does not
 * exist.  No upstream
 * counterpart (the upstream kernels have fixed launch shapes). */
int pe_tile_info(int32_t precision, int32_t half, int32_t variant, int32_t *info3);

/* pe_run with a hipEvent pair on `stream` around every half-step kernel:
 * ms_kernels[0] = summed stress-kernel time, [1] = velocity-kernel time,
 * [2] = source/record time, in milliseconds.  Blocks until the run ends.
 * No upstream counterpart (the upstream solver has no kernel timing); it
 * serves benchmarks that report per-kernel HBM roofline fractions. */
int pe_run_timed(const pe_shot *shot, int32_t n0, int32_t n1, const double *wavelet,
                 void *traces, void *stream, double *ms_kernels);

/* ------------------------------------------------------------------------
 * Adjoint (reverse-mode) step.  The upstream repository's only exact
 * adjoint is the vendored CPU array-path one,
 * vendor/ringfwi/elastic3d.py:348-463 (misfit_and_gradient: reverse loop
 * :386-461, stiffness gradient :412-419); these entry points give the same
 * reverse-mode derivative for the label-indexed step of pe_run, on the GPU,
 * with per-material gradients in place of its 21 per-voxel maps.
 *
 * Deterministic gradient reduction: the caller cuts every (y, x) column
 * into runs of constant material, also cut at z % zbreak == 0 when
 * zbreak > 0, and numbers the runs column by column (flat y*nx + x), in z
 * order inside a column.  With nchunk = zbreak > 0 ? ceil(nz / zbreak) : 1,
 * run_ptr[c*nchunk + k] is the first run of column c at or after
 * z = k*zbreak.  Each run owns one row of run_acc (K doubles, K = n_proj ?
 * n_proj : 36), updated by one thread once per run per step, in step
 * order; mat_ptr / mat_runs list each material's runs in ascending order.
 * The tiled gradient kernel (n_proj > 0) walks z-chunks of zbreak planes.
 * ------------------------------------------------------------------------ */
typedef struct pe_adjoint {
    void *lv[3];                 /* adjoint of vx, vy, vz (ncell real each) */
    void *ls[6];                 /* adjoint of s1..s6 */
    void *work;                  /* 6*ncell real scratch */
    void *wface;                 /* 3*ncell real scratch (PE_KERNEL_TILED) */
    int32_t kernel;              /* PE_KERNEL_DIRECT or PE_KERNEL_TILED */
    int32_t hist_kind;           /* PE_HIST_VELOCITY (3 x ncell per step) or PE_HIST_STRAIN (6) */
    int32_t n_proj;              /* 0: dJ/dCt, all 36 entries; 1..4: projections */
    const void *proj;            /* n_mat x n_proj x 36 real (n_proj > 0) */
    int32_t zbreak;              /* runs also cut at z % zbreak == 0 (0: never) */
    int32_t n_runs;
    const int32_t *run_ptr;      /* ny*nx*nchunk + 1: first run of each column chunk */
    const int32_t *mat_ptr;      /* n_mat + 1 offsets into mat_runs */
    const int32_t *mat_runs;     /* n_runs run indices grouped by material */
    double *run_acc;             /* n_runs x K accumulators, or NULL: no gradient */
    /* CPML shots (shot->cpml_width > 0; with a gradient PE_HIST_STRAIN,
     * since a step's strains depend on the forward memory variables):
     * cpml_psib[a] the adjoints of the memory variables,
     * laid out like shot->cpml_psi[a] (6 x band volume; zeroed by
     * pe_adjoint_reset, carried from step to step), cpml_work[a] a scratch
     * of 3 x band volume of axis a.  No upstream counterpart. */
    void *cpml_psib[3];
    void *cpml_work[3];
} pe_adjoint;

/* Validate an adjoint descriptor against its shot. */
int pe_check_adjoint(const pe_shot *shot, const pe_adjoint *adj);

/* Zero lv, ls, run_acc and (CPML) cpml_psib (fill kernels, no memset
 * nodes); for PE_KERNEL_TILED also fill wface with d * dt_h / rho_face. */
int pe_adjoint_reset(const pe_shot *shot, const pe_adjoint *adj, void *stream);

/* Reverse steps n = n_hi-1 down to n_lo.  On entry (lv, ls) are the
 * adjoints of the state after step n_hi-1; on return, of the state before
 * step n_lo.  vhist: (n_hi-n_lo) slots, slot n-n_lo = v at the start of step n
 * (3 x ncell real, PE_HIST_VELOCITY) or its strains (6 x ncell,
 * PE_HIST_STRAIN, as written by pe_run_store_strains); read only when
 * run_acc is set.  residual: (nt x n_rec)
 * real trace adjoint dJ/dy indexed by absolute step, or NULL.
 * grad_wavelet: double per absolute step, receives dJ/dwavelet[n], or NULL.
 * Receiver cells must be unique within each group.  Replaces the reverse
 * loop of vendor/ringfwi/elastic3d.py:386-461. */
int pe_adjoint_run(const pe_shot *shot, const pe_adjoint *adj, int32_t n_hi, int32_t n_lo,
                   const void *vhist, const void *residual, double *grad_wavelet, void *stream);

/* grad[m*K + k] = sum over the runs of material m of run_acc (K as above):
 * dJ/dCt[m][k] (n_proj = 0) or the projections.  Replaces the per-voxel
 * gradient maps g[C_IJ] of vendor/ringfwi/elastic3d.py:412-419 summed over
 * each grain. */
int pe_adjoint_gradient(const pe_shot *shot, const pe_adjoint *adj, double *grad, void *stream);

/* ------------------------------------------------------------------------
 * Multi-shot batches.  The upstream sweep engine solves azimuths one after
 * another, each a separate forward_fused_labels call
 * (sim/pipeline/sweep_runner.py:322-388).  A batch advances many shots (of
 * any grid sizes, dt and step counts) together: every batched step is one
 * launch per half-step over the workgroups of all shots, plus two small
 * per-shot launches (source / pressure record; velocity record and step
 * advance).  Each shot keeps its own step counter on the device, so the
 * launches take no per-step host arguments and a captured graph of k steps
 * can be replayed until every shot is done (finished shots exit at once).
 * All shots share one precision and the tile variants of shots[0]; every
 * shot must be PE_KERNEL_TILED.  Shots may have a CPML (cpml_width > 0, with
 * zeroed memory variables): each half-step of such a batch is preceded by
 * one band pre-pass launch over the band cells of every CPML shot.  Results
 * are bit-identical to pe_run of each shot alone.
 * ------------------------------------------------------------------------ */
typedef struct pe_batch {
    int32_t n_shot;
    const pe_shot *shots;            /* host array of n_shot descriptors */
    const double *const *wavelet;    /* host array: device wavelet of each shot (>= nt) */
    void *const *traces;             /* host array: device (nt x n_rec) real traces, or NULL */
    const int32_t *nt;               /* host array: step count of each shot */
    void *table;                     /* device scratch of pe_batch_table_bytes() bytes */
    int32_t *step;                   /* device, n_shot counters: next step of each shot */
    int64_t nb[3];                   /* set by pe_batch_setup: workgroups per half-step, and
                                        of the CPML band pre-passes */
} pe_batch;

/* Device scratch needed by a batch of these shots (0 on invalid input). */
int64_t pe_batch_table_bytes(const pe_shot *shots, int32_t n_shot);

/* Validate, write the device table (blocking copy) and zero the step
 * counters.  Must precede pe_batch_step / pe_batch_graph. */
int pe_batch_setup(pe_batch *batch, void *stream);

/* pe_batch_setup for callers that cannot allocate device memory (code under
 * graph capture, serving loops, harnesses that forbid hipMalloc).  No
 * upstream counterpart.  Ignores batch->table and batch->step on entry and
 * puts both in device memory the library owns: a static array of
 * PE_BATCH_SCRATCH_BYTES per device (held until the process ends), the step
 * counters in the first 256-byte-rounded n_shot*4 bytes, then the table; on
 * return batch->table / batch->step hold those device addresses, and
 * pe_batch_step / pe_batch_graph use them as usual.  Allocates nothing (the
 * table is uploaded with a blocking copy and the counters zeroed, as by
 * pe_batch_setup).  Returns PE_ERANGE, before touching the device, when the
 * table and counters exceed PE_BATCH_SCRATCH_BYTES (pe_batch_table_bytes
 * gives the table size).  The array is shared by every batch set up this
 * way on one device and keeps the last setup's contents: a later setup
 * replaces the table an earlier batch (or a graph captured from it) reads,
 * so such batches must not overlap on one device. */
#define PE_BATCH_SCRATCH_BYTES (8 * 1024 * 1024)
int pe_batch_setup_noalloc(pe_batch *batch, void *stream);

/* Enqueue `steps` batched steps (asynchronous, graph-capturable). */
int pe_batch_step(const pe_batch *batch, int32_t steps, void *stream);

/* Capture `steps` batched steps on `stream` into an executable graph
 * (*graph_exec, a hipGraphExec_t); replay with pe_graph_launch. */
int pe_batch_graph(const pe_batch *batch, int32_t steps, void *stream, void **graph_exec);
int pe_graph_launch(void *graph_exec, void *stream);
int pe_graph_destroy(void *graph_exec);

/* ------------------------------------------------------------------------
 * Per-azimuth shot preparation on the device (the upstream does these on
 * the host for every azimuth, sim/pipeline/sweep_runner.py:331-341).  All
 * pointers are device pointers; results equal the host computation.
 * ------------------------------------------------------------------------ */

/* Re-grid the disk labels src (nx_s x ny_s x nz_s, x-major, -1 outside)
 * rotated by the angle whose cosine / sine are ca / sa, into lab
 * (nz x n x n int32, nz = ndz + 2m, n = nd + 2m, couplant -1 in the
 * m-cell pad).  Replaces the nearest-neighbour path of
 * sim/core/rotation_test.py:77-91, 125-132. */
int pe_prep_rotate(const int32_t *src, int32_t nx_s, int32_t ny_s, int32_t nz_s, double h_spec,
                   double h, double ca, double sa, int32_t nd, int32_t ndz, int32_t m,
                   int32_t *lab, void *stream);

/* Single rasterisation of the rotated disk: every solver point inside the
 * disk of diameter D takes the seed minimising |x - s|^2 - w_s (float64,
 * all n_seed seeds, lowest index on ties); seeds (n_seed x 3, metres,
 * already rotated) and weights w are device arrays.  Same grid and pad as
 * pe_prep_rotate.  Replaces the single_raster path of
 * sim/core/rotation_test.py:93-124 and the labeller of
 * sim/core/specimen.py:298-335. */
int pe_prep_label_power(const double *seeds, const double *w, int32_t n_seed, double h,
                        double D, int32_t nd, int32_t ndz, int32_t m, int32_t *lab,
                        void *stream);

/* damp = (pz[z] * py[y]) * px[x] * (float)exp(-fluid_damp * d), d the
 * Euclidean distance from each couplant cell (lab < 0) to the nearest
 * specimen cell (exact, three separable passes); s0, s1: ncell int32
 * scratch.  Replaces sweep_runner.py:340-341 (scipy EDT) composed with the
 * sponge of sim/core/fdtd.py:1693-1700.  Grid sides up to 1024. */
int pe_prep_edt_damp(const int32_t *lab, int32_t nz, int32_t ny, int32_t nx, double fluid_damp,
                     const float *pz, const float *py, const float *px, int32_t *s0, int32_t *s1,
                     float *damp, void *stream);

/* partial[b] = block maxima of max(C11, C22, C33)[m] / min(rho, face
 * densities) over the cells (float64); the stable dt is
 * safety * cfl * h / sqrt(max partial).  Replaces the volume pass of
 * safe_dt_labels, sim/core/fdtd.py:1606-1615. */
int pe_prep_vmax(const int32_t *mats, int32_t nz, int32_t ny, int32_t nx, const void *Ct,
                 const void *rho_t, int32_t precision, double *partial, int32_t n_partial,
                 void *stream);

#ifdef __cplusplus
}
#endif

#endif /* PEFDTD_H */
