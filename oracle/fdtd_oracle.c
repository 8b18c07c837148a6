/*
 * CPU ORACLE -- test infrastructure only.  Nothing in the product links,
 * loads or calls this file; only tests/, the smoke check and the benchmark's
 * CPU-baseline leg do, and only as the checker.
 *
 * A plain C restatement of one time step of the label-indexed, deferred-damped
 * velocity-stress staggered-grid FDTD of the upstream solver
 * (sim/core/fdtd.py::forward_fused_labels, lines 1618-1816), in float32 and
 * float64.  Each piece cites the upstream lines it follows:
 *
 *   strains     sim/core/fdtd.py:1102-1177 (strains_zp; stencils :387-390)
 *   moved       sim/core/fdtd.py:612-642   (tap order z, y, x; 0.5 / 0.25)
 *   stress      sim/core/fdtd.py:643-657   (deferred damping, act mask)
 *   source      sim/core/fdtd.py:1798-1800 (s1..s3 += f32(w[n]) * w_k)
 *   velocity    sim/core/fdtd.py:1179-1257 (vel_zp; face density :1237-1240)
 *   record      sim/core/fdtd.py:1803-1813 (-(s1+s2+s3)/3 weighted sum)
 *
 * Floating-point convention.  The upstream kernels are compiled by NVRTC with
 * its default multiply-add contraction, so every "acc += c * x" statement is a
 * fused multiply-add.  This file spells those as fma()/fmaf() and is compiled
 * with -ffp-contract=off so that nothing else is fused.  The one statement
 * whose contraction is ambiguous upstream, "sv = s*dmp; sv += dt*acc", is
 * taken as fma(dt, acc, s*dmp).  The float instantiation runs with x86
 * flush-to-zero and denormals-are-zero set: with both conventions this file
 * reproduces the upstream float32 solver archives bit for bit (checked in
 * tests/test_oracle_golden.py), i.e. the upstream runs flushed subnormals.
 * The double instantiation keeps IEEE subnormals.
 *
 * Layout: C order (nz, ny, nx), x fastest, flat idx = (iz*ny + iy)*nx + ix.
 * Staggering (z, y, x half-cell offsets): vx (0,0,+), vy (0,+,0), vz (+,0,0),
 * s1..s3 (0,0,0), s4 (+,+,0), s5 (+,0,+), s6 (0,+,+).
 */
#include <math.h>
#include <xmmintrin.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    int32_t nz, ny, nx;
    int32_t nc;                 /* stencil half-length, 1..4 */
    double coeffs[4];
    double dt, inv_h, dt_h;
    int32_t n_mat;
    uint64_t act_mask;          /* bit 6I+J: some material has C_IJ != 0 */
    const int32_t *mats;        /* material index = label + 1, ncell */
    const void *Ct;             /* n_mat * 36, real */
    const void *rho_t;          /* n_mat, real */
    const void *damp;           /* ncell real or NULL */
    void *v[3];                 /* vx, vy, vz */
    void *s[6];                 /* s1..s6 */
    void *g;                    /* scratch 6 * ncell real */
    int32_t n_src;
    const int64_t *src_idx;
    const void *src_w;          /* n_src real */
    int32_t n_rec;
    const int32_t *rec_ptr;     /* n_rec + 1 */
    const int64_t *rec_idx;
    const void *rec_w;          /* real */
    int32_t record_mode;        /* 0 pressure, 1 vx, 2 vy, 3 vz */
    /* CPML (no upstream counterpart; cpml_width 0 = none).  Axis a (0 z,
     * 1 y, 2 x) of length n: band cells i <= L or i >= n-1-L, band index
     * i or i - (n - 2L - 2); prof[a] = real[2 staggers][b, a, 1/kappa][n];
     * psi[a] = real[6][band volume], advanced in place (components 0..2 by
     * the stress half-step, 3..5 by the velocity half-step). */
    int32_t cpml_width;
    int32_t cpml_unused;
    const void *cpml_prof[3];
    void *cpml_psi[3];
} orc_shot;

/* CPML band geometry of axis a: is index i in the band, its band index, and
 * the flat index of cell (iz, iy, ix) within one psi component. */
static inline int pml_in(int i, int n, int L) { return i <= L || i >= n - 1 - L; }
static inline int pml_bi(int i, int n, int L) { return i <= L ? i : i - (n - 2 * L - 2); }
static inline int64_t pml_vol(const orc_shot *S, int a)
{
    const int64_t W = 2 * (int64_t)S->cpml_width + 2;
    if (a == 0) return W * S->ny * S->nx;
    if (a == 1) return (int64_t)S->nz * W * S->nx;
    return (int64_t)S->nz * S->ny * W;
}
static inline int64_t pml_flat(const orc_shot *S, int a, int iz, int iy, int ix)
{
    const int L = S->cpml_width;
    const int64_t W = 2 * (int64_t)L + 2;
    if (a == 0) return ((int64_t)pml_bi(iz, S->nz, L) * S->ny + iy) * S->nx + ix;
    if (a == 1) return ((int64_t)iz * W + pml_bi(iy, S->ny, L)) * S->nx + ix;
    return ((int64_t)iz * S->ny + iy) * W + pml_bi(ix, S->nx, L);
}
static inline int pml_cell(const orc_shot *S, int iz, int iy, int ix)
{
    const int L = S->cpml_width;
    return L > 0 && (pml_in(iz, S->nz, L) || pml_in(iy, S->ny, L) || pml_in(ix, S->nx, L));
}

/* Position of each strain component J: 0 = centre, 1 = yz, 2 = xz, 3 = xy. */
static const int VP[6] = {0, 0, 0, 1, 2, 3};
static const int POS[4][3] = {{0, 0, 0}, {1, 1, 0}, {1, 0, 1}, {0, 1, 1}};

/* Tap offsets (flat) and per-axis guards of strain J moved to position p. */
typedef struct {
    int ntap;
    int64_t off[4];
    int guard[3];           /* +1: i <= n-2, -1: i >= 1, 0: none (z, y, x) */
} move_t;

static void move_table(move_t mv[4][6], int nz, int ny, int nx)
{
    const int64_t strides[3] = {(int64_t)nx * ny, nx, 1};
    for (int p = 0; p < 4; ++p)
        for (int J = 0; J < 6; ++J) {
            move_t *m = &mv[p][J];
            const int src = VP[J];
            m->ntap = 1;
            m->off[0] = 0;
            for (int a = 0; a < 3; ++a) {
                const int d = src == p ? 0 : POS[p][a] - POS[src][a];
                m->guard[a] = d > 0 ? 1 : (d < 0 ? -1 : 0);
                if (d == 0) continue;
                const int o0 = d > 0 ? 0 : -1, o1 = d > 0 ? 1 : 0;
                int64_t nxt[4];
                int k = 0;
                for (int t = 0; t < m->ntap; ++t) {
                    nxt[k++] = m->off[t] + o0 * strides[a];
                    nxt[k++] = m->off[t] + o1 * strides[a];
                }
                for (int t = 0; t < k; ++t) m->off[t] = nxt[t];
                m->ntap = k;
            }
        }
}

/* Planted-mistake switch for the adjoint tests (0 = exact adjoint):
 * 1 drops the face-density average, 2 weights 4-tap moves by 1/2,
 * 3 forgets the damping of the stress adjoint, 4 drops the memory variable's
 * share (a P) of a CPML derivative's adjoint. */
int orc_adj_plant = 0;

#define REAL float
#define SUF _f32
#define FMA fmaf
#define FP_ENTER const unsigned int fp_saved_ = _mm_getcsr(); _mm_setcsr(fp_saved_ | 0x8040u);
#define FP_EXIT _mm_setcsr(fp_saved_);
#include "fdtd_oracle_body.h"
#include "fdtd_adjoint_body.h"
#undef REAL
#undef SUF
#undef FMA
#undef FP_ENTER
#undef FP_EXIT

#define REAL double
#define SUF _f64
#define FMA fma
#define FP_ENTER
#define FP_EXIT
#include "fdtd_oracle_body.h"
#include "fdtd_adjoint_body.h"
#undef REAL
#undef SUF
#undef FMA
#undef FP_ENTER
#undef FP_EXIT
