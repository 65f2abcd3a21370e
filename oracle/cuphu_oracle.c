/*
 * cuphu_oracle.c -- CPU restatement of the reference's Laplace PCG/IRLS
 * unwrapping path.  TEST INFRASTRUCTURE ONLY: it is the checker used by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, and is
 * never linked into or called by the product library.
 *
 * Each function cites the reference file:line it restates.  Floating-point
 * operation order is spelled out (explicit fmaf, -ffp-contract=off) so that
 * the FP32 restatements are bit-comparable with the HIP kernels, which are
 * compiled with the same contraction rules.
 *
 * Parity status per function (how each one is pinned, see tests/):
 *   o_wrapped_phase   closed forms (ramp, quadrant cases) + glibc atan2f ulp bound
 *   o_phase_diffs     closed forms; wrap range; reference ramp fixtures
 *   o_boxcar2d        scipy.ndimage.uniform_filter(mode='mirror')
 *   o_costs           hand-evaluated special cases (rho=0, masked sentinels,
 *                     short clamp, sigsqshortmin floor, defo rounding) and
 *                     hand-evaluated sigsq with each rhopow term alone
 *                     (2 cstd1, cstd2 ln n at n = e^k, cstd3 n) and with the
 *                     documented defaults at n = 10; SNAPHU's own values are
 *                     NOT available (parity vs SNAPHU unpinned; DEF_*
 *                     constants unverifiable here)
 *   o_residues        vortex (+-1), conservative field (0), flow-curl identity
 *   o_assemble/matvec dense numpy Laplacian built from the arc list
 *   o_reweight        numpy closed form
 *   o_pcg_irls_*      dense float64 IRLS (every pass a dense solve, the delta
 *                     schedule with a binding floor) on scenes with residues,
 *                     both drivers, on the continuous solution and the
 *                     ambiguities; CG iteration bound (<= npix) for the FP32
 *                     driver; 256^2 ramp closed form through both drivers
 *   o_conncomp        scipy.sparse.csgraph.connected_components; reference
 *                     test_conncomp.py known answers
 *   o_integrate_phase closed form (zero flows, integer flows)
 *   o_tile_axis/stitch reference test_auto_tiling.py / geometry formula KATs;
 *                     hand-built two-tile cases
 *   o_wls_*           dense numpy solve + central finite differences
 *   o_fix_cycle_spikes reference test_fix_cycle_spikes.py cases (isolated
 *                     row / column, smooth ramp, sustained band, mask);
 *                     brute-force sorted medians
 *   o_neighbor_feedback reference test_laplace_neighbor_feedback.py cases
 *                     (constant offset, local gradient, sparse boundary,
 *                     far-from-data fade, single tile); hand-evaluated
 *                     profile (moving median, gap fill, feather)
 *   o_pcg_step_f32 / o_pcg_fused_step_f32  textbook dense float64 Jacobi-PCG
 *   o_incrcost_smooth hand-evaluated cases, exact integer arithmetic
 *   o_conncomp_gated  scipy.sparse.csgraph on the gated arc set
 *   o_bridge*         reference test_bridge.py cases; scipy.ndimage label
 *                     (8-connected), binary_erosion, maximum/minimum_filter;
 *                     numpy brute force for the nearest boundary pairs
 * The outputs of one small seeded scene are also frozen in
 * tests/golden/scene64x80.npz, which the oracle must keep reproducing.
 */
#include "cuphu_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <limits.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define O_TWOPI_F     6.28318530717958647692f
#define O_INV_TWOPI_F 0.15915494309189534561f
#define O_EPS_F       1e-12f
#define O_PI_F        3.14159265358979323846f
#define O_TWOPI_D     6.28318530717958647692
#define O_EPS_D       1e-12
#define O_LARGESHORT  32000

void o_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* ── a1: wrapped phase ───────────────────────────────────────────────────────
 * cuphu_solver.cpp:1707-1712 (host atan2, +2pi when negative) followed by
 * WrapPhaseKernel cuphu_util.cu:33-46 (p - 2pi*floor(p/2pi)).
 * correctly_rounded=1 evaluates atan2 in double and rounds once (the form
 * the HIP ingest kernel uses); 0 uses the C library's atan2f like the
 * reference host loop. */
static inline float wrap_to_2pi_f(float p) {
    float t = floorf(p / O_TWOPI_F);
    return p - O_TWOPI_F * t;
}
void o_wrapped_phase(const float *re, const float *im, size_t n, int correctly_rounded, float *phi) {
    for (size_t k = 0; k < n; ++k) {
        float a = correctly_rounded ? (float)atan2((double)im[k], (double)re[k])
                                    : atan2f(im[k], re[k]);
        if (a < 0.0f) a += 6.28318530717958648f;
        phi[k] = wrap_to_2pi_f(a);
    }
}

/* ── a2: wrapped phase differences in cycles ─────────────────────────────────
 * RangeDiffsKernel cuphu_util.cu:54-66 (right - left) and AzimuthDiffsKernel
 * cuphu_util.cu:80-92 (top - bottom); wrap_cycles cuphu_util.cu:27-30. */
static inline float wrap_cycles_f(float d) { return d - floorf(d + 0.5f); }
void o_phase_diffs(const float *phi, int nrow, int ncol, float *rng, float *az) {
    for (int r = 0; r < nrow; ++r)
        for (int c = 0; c < ncol - 1; ++c) {
            float d = (phi[(size_t)r * ncol + c + 1] - phi[(size_t)r * ncol + c]) / O_TWOPI_F;
            rng[(size_t)r * (ncol - 1) + c] = wrap_cycles_f(d);
        }
    for (int r = 0; r < nrow - 1; ++r)
        for (int c = 0; c < ncol; ++c) {
            float d = (phi[(size_t)r * ncol + c] - phi[(size_t)(r + 1) * ncol + c]) / O_TWOPI_F;
            az[(size_t)r * ncol + c] = wrap_cycles_f(d);
        }
}

/* ── a4: separable boxcar with mirror padding ────────────────────────────────
 * cuphu_boxcar2d cuphu_util.cu:237-287: MirrorPadKernel (:156-181, reflect
 * without edge repeat), BoxcarRowKernel (:104-122: taps summed k=0..2kc in
 * FP32, then / (2kc+1)) over all padded rows, BoxcarColKernel (:131-148).
 * Reflection is repeated until in range, which equals the reference
 * wherever the reference stays in bounds (pad < extent). */
static inline int mirror_idx(int i, int n) {
    if (n == 1) return 0;
    for (;;) {
        if (i < 0) i = -i;
        else if (i > n - 1) i = 2 * (n - 1) - i;
        else return i;
    }
}
int o_boxcar2d(const float *src, int nrow, int ncol, int kr, int kc, float *dst) {
    int prow = nrow + 2 * kr, wr = 2 * kr + 1, wc = 2 * kc + 1;
    float *tmp = (float *)malloc((size_t)prow * ncol * sizeof(float));
    if (!tmp) return -1;
    for (int pr = 0; pr < prow; ++pr) {
        int sr = mirror_idx(pr - kr, nrow);
        for (int c = 0; c < ncol; ++c) {
            float sum = 0.0f;
            for (int k = 0; k < wc; ++k) {
                int sc = mirror_idx(c + k - kc, ncol);
                sum += src[(size_t)sr * ncol + sc];
            }
            tmp[(size_t)pr * ncol + c] = sum / (float)wc;
        }
    }
    for (int r = 0; r < nrow; ++r)
        for (int c = 0; c < ncol; ++c) {
            float sum = 0.0f;
            for (int k = 0; k < wr; ++k) sum += tmp[(size_t)(r + k) * ncol + c];
            dst[(size_t)r * ncol + c] = sum / (float)wr;
        }
    free(tmp);
    return 0;
}

/* ── a5: SNAPHU statistical costs ────────────────────────────────────────────
 * Constants: upload_smooth_constants cuphu_cost.cu:41-59.
 * Smooth: BuildSmoothCol/RowCostsKernel cuphu_cost.cu:77-183 (offset
 * truncated toward zero, :126).  Defo: BuildDefoCol/RowCostsKernel
 * cuphu_cost.cu:208-323 (offset rounded to nearest even, :257; dzmax and
 * laycost :249-255).  Masked arcs (weight 0): :98-103, :225-232.
 * Windows: range diffs boxcar (kperp/2 rows, kpar/2 cols), azimuth diffs
 * transposed (cuphu_cost.cu:366-371).  Output: flat row arcs then column
 * arcs, 2 shorts (offset, sigsq) per arc for smooth, 4 (offset, sigsq,
 * dzmax, laycost) for defo.  Arithmetic is uncontracted FP64. */
typedef struct { double rho0, thresh, rhopow, sigsqrhoconst, sigsqcorr, nsc, nsc2, costscale, defomax; int sqmin; } OCC;
static OCC o_cost_consts(const OCostParams *p) {
    OCC k;
    k.rho0 = p->rhosconst1 / p->ncorrlooks + p->rhosconst2;
    k.thresh = p->defothreshfactor * k.rho0;
    k.rhopow = 2.0 * p->cstd1 + p->cstd2 * log(p->ncorrlooks) + p->cstd3 * p->ncorrlooks;
    k.sigsqrhoconst = 2.0 / 12.0;
    k.sigsqcorr = p->sigsqcorr;
    k.nsc = p->nshortcycle;
    k.nsc2 = k.nsc * k.nsc;
    k.costscale = p->costscale;
    k.defomax = p->defomax;
    k.sqmin = (int)p->sigsqshortmin;
    return k;
}
static void o_one_cost(const OCC *k, float ca, float cb, float dpsi, float avg, short w, int defo, short *out) {
    if (w == 0) {
        out[0] = O_LARGESHORT / 2; out[1] = O_LARGESHORT;
        if (defo) { out[2] = O_LARGESHORT; out[3] = 0; }
        return;
    }
    /* rho < thresh -> 0 (:107); written as !(rho >= thresh) so a NaN
     * coherence also maps to 0 instead of reaching an undefined int cast. */
    double rho = 0.5 * ((double)ca + (double)cb);
    if (!(rho >= k->thresh)) rho = 0.0;
    double t = k->sigsqrhoconst * pow(1.0 - rho, k->rhopow);
    double sigsqrho = (t + k->sigsqcorr) * k->nsc2;
    double dp = (double)dpsi, av = (double)avg;
    double offset = k->nsc * (dp - (rho > 0 ? av : 0.5 * av));
    double sigsq_f = sigsqrho / (k->costscale * (double)w);
    if (sigsq_f > 32767.0) sigsq_f = 32767.0;   /* same result as the later short clamp, no int overflow */
    int sigsq_i = (int)sigsq_f;
    if (sigsq_i < k->sqmin) sigsq_i = k->sqmin;
    if (offset > 32767.0) offset = 32767.0;
    if (offset < -32768.0) offset = -32768.0;
    int off_i = defo ? (int)rint(offset) : (int)offset;
    if (off_i < SHRT_MIN) off_i = SHRT_MIN;
    if (off_i > SHRT_MAX) off_i = SHRT_MAX;
    if (sigsq_i > SHRT_MAX) sigsq_i = SHRT_MAX;
    out[0] = (short)off_i; out[1] = (short)sigsq_i;
    if (defo) {
        double dz = rint(k->nsc * k->defomax);
        int dzmax = dz > 32767.0 ? 32767 : (dz < 1.0 ? 1 : (int)dz);
        int laycost = (int)sqrt((double)(dzmax * sigsq_i));
        if (laycost > SHRT_MAX) laycost = SHRT_MAX;
        out[2] = (short)dzmax; out[3] = (short)laycost;
    }
}
void o_costs(const float *phi, const float *corr, const short *weights, int nrow, int ncol,
             const OCostParams *p, int kperp, int kpar, int defo, short *out) {
    OCC k = o_cost_consts(p);
    size_t nrowarc = (size_t)(nrow - 1) * ncol, ncolarc = (size_t)nrow * (ncol - 1);
    float *rng = (float *)malloc(ncolarc * sizeof(float) + 1);
    float *az = (float *)malloc(nrowarc * sizeof(float) + 1);
    float *arng = (float *)malloc(ncolarc * sizeof(float) + 1);
    float *aaz = (float *)malloc(nrowarc * sizeof(float) + 1);
    o_phase_diffs(phi, nrow, ncol, rng, az);
    if (ncol > 1) o_boxcar2d(rng, nrow, ncol - 1, kperp / 2, kpar / 2, arng);
    if (nrow > 1) o_boxcar2d(az, nrow - 1, ncol, kpar / 2, kperp / 2, aaz);
    int ne = defo ? 4 : 2;
    for (int r = 0; r < nrow - 1; ++r)
        for (int c = 0; c < ncol; ++c) {
            size_t a = (size_t)r * ncol + c;
            short w = weights ? weights[a] : 1;
            o_one_cost(&k, corr[(size_t)r * ncol + c], corr[(size_t)(r + 1) * ncol + c], az[a], aaz[a], w, defo,
                       out + ne * a);
        }
    for (int r = 0; r < nrow; ++r)
        for (int c = 0; c < ncol - 1; ++c) {
            size_t a = (size_t)r * (ncol - 1) + c;
            short w = weights ? weights[nrowarc + a] : 1;
            o_one_cost(&k, corr[(size_t)r * ncol + c], corr[(size_t)r * ncol + c + 1], rng[a], arng[a], w, defo,
                       out + ne * (nrowarc + a));
        }
    free(rng); free(az); free(arng); free(aaz);
}

/* ── a3: residues (plaquette charges) ────────────────────────────────────────
 * No reference kernel (SNAPHU computes them internally; consumed at
 * snaphu_cs2parse.c:205-226 as residue[row][col] over (nrow-1)x(ncol-1)).
 * Charge = rint of the loop sum A(r,c)->B(r,c+1)->C(r+1,c+1)->D(r+1,c)->A of
 * the wrapped differences in cycles, summed left to right in FP32:
 * ((rng[r,c] - az[r,c+1]) - rng[r+1,c]) + az[r,c]. */
void o_residues(const float *phi, int nrow, int ncol, signed char *res) {
    size_t nrowarc = (size_t)(nrow - 1) * ncol, ncolarc = (size_t)nrow * (ncol - 1);
    float *rng = (float *)malloc(ncolarc * sizeof(float) + 1);
    float *az = (float *)malloc(nrowarc * sizeof(float) + 1);
    o_phase_diffs(phi, nrow, ncol, rng, az);
    for (int r = 0; r < nrow - 1; ++r)
        for (int c = 0; c < ncol - 1; ++c) {
            float s = rng[(size_t)r * (ncol - 1) + c] - az[(size_t)r * ncol + c + 1];
            s = s - rng[(size_t)(r + 1) * (ncol - 1) + c];
            s = s + az[(size_t)r * ncol + c];
            res[(size_t)r * (ncol - 1) + c] = (signed char)lrintf(s);
        }
    free(rng); free(az);
}

/* ── a6: arc weights (laplace_arc_weight cuphu_laplace.cu:65-69,
 * laplace_init_weights :198-205) ──────────────────────────────────────────── */
static inline float arc_weight_f(short s) {
    if (s <= 0 || s >= O_LARGESHORT) return 0.0f;
    return 1.0f / (float)s;
}
void o_arc_weights(const short *sigsq, size_t narcs, size_t stride, float *w) {
    for (size_t i = 0; i < narcs; ++i) w[i] = arc_weight_f(sigsq[i * stride]);
}

/* laplace_wrap_pi cuphu_laplace.cu:71-73, written with one explicit fma. */
static inline float wrap_pi_f(float x) { return fmaf(-O_TWOPI_F, rintf(x * O_INV_TWOPI_F), x); }
static inline double wrap_pi_d(double x) { return fma(-O_TWOPI_D, rint(x / O_TWOPI_D), x); }

/* ── a7: assemble (laplace_assemble cuphu_laplace.cu:88-148) ─────────────────
 * Arc order per pixel: down (outgoing), up (incoming), right (outgoing),
 * left (incoming); arcs with w <= 0 skipped; b = fma(-+w, g, b);
 * diag = sum(w) + 1e-12. */
void o_assemble(const float *w, const float *phi, int nrow, int ncol, float *b, float *diag) {
    const float *row_w = w, *col_w = w + (size_t)(nrow - 1) * ncol;
    for (int r = 0; r < nrow; ++r)
        for (int c = 0; c < ncol; ++c) {
            size_t i = (size_t)r * ncol + c;
            float pa = phi[i], bv = 0.0f, ws = 0.0f;
            if (r < nrow - 1) { float a = row_w[i]; if (a > 0.0f) { bv = fmaf(-a, wrap_pi_f(phi[i + ncol] - pa), bv); ws += a; } }
            if (r > 0) { float a = row_w[i - ncol]; if (a > 0.0f) { bv = fmaf(a, wrap_pi_f(pa - phi[i - ncol]), bv); ws += a; } }
            if (c < ncol - 1) { float a = col_w[(size_t)r * (ncol - 1) + c]; if (a > 0.0f) { bv = fmaf(-a, wrap_pi_f(phi[i + 1] - pa), bv); ws += a; } }
            if (c > 0) { float a = col_w[(size_t)r * (ncol - 1) + c - 1]; if (a > 0.0f) { bv = fmaf(a, wrap_pi_f(pa - phi[i - 1]), bv); ws += a; } }
            b[i] = bv;
            diag[i] = ws + O_EPS_F;
        }
}

/* ── a8: matvec (laplace_matvec cuphu_laplace.cu:156-190) ────────────────────
 * y = 1e-12*x_a, then fma(w, x_a - x_nb, y) over down, up, right, left.  A
 * missing neighbour contributes fma(0, x_a - 0, y), which equals y for
 * finite x (the HIP kernels see zero-weight padding there). */
void o_matvec(const float *w, const float *x, int nrow, int ncol, float *y) {
    const float *row_w = w, *col_w = w + (size_t)(nrow - 1) * ncol;
#pragma omp parallel for schedule(static)
    for (int r = 0; r < nrow; ++r)
        for (int c = 0; c < ncol; ++c) {
            size_t i = (size_t)r * ncol + c;
            float xa = x[i], yv = O_EPS_F * xa;
            float wd = (r < nrow - 1) ? row_w[i] : 0.0f, xd = (r < nrow - 1) ? x[i + ncol] : 0.0f;
            float wu = (r > 0) ? row_w[i - ncol] : 0.0f, xu = (r > 0) ? x[i - ncol] : 0.0f;
            float wrt = (c < ncol - 1) ? col_w[(size_t)r * (ncol - 1) + c] : 0.0f, xr = (c < ncol - 1) ? x[i + 1] : 0.0f;
            float wl = (c > 0) ? col_w[(size_t)r * (ncol - 1) + c - 1] : 0.0f, xl = (c > 0) ? x[i - 1] : 0.0f;
            yv = fmaf(wd, xa - xd, yv);
            yv = fmaf(wu, xa - xu, yv);
            yv = fmaf(wrt, xa - xr, yv);
            yv = fmaf(wl, xa - xl, yv);
            y[i] = yv;
        }
}

/* ── a11: IRLS reweight (laplace_reweight cuphu_laplace.cu:214-247) ──────────
 * w = w0 / sqrt(fma(e, e, delta*delta)), e = (u_h - u_t) - wrap(phi_h - phi_t);
 * w0 recomputed from sigsq (never compounds).  Correctly rounded division
 * and sqrt replace the reference's rsqrtf so the value is reproducible. */
static inline float reweight_one_f(short s, float ph, float pt, float uh, float ut, float d2) {
    float w0 = arc_weight_f(s);
    if (w0 <= 0.0f) return 0.0f;
    float g = wrap_pi_f(ph - pt);
    float e = (uh - ut) - g;
    return w0 / sqrtf(fmaf(e, e, d2));
}
void o_reweight(const short *sigsq, size_t stride, const float *phi, const float *u, float delta,
                int nrow, int ncol, float *w) {
    size_t nrowarc = (size_t)(nrow - 1) * ncol, narcs = nrowarc + (size_t)nrow * (ncol - 1);
    float d2 = delta * delta;
#pragma omp parallel for schedule(static)
    for (long long ii = 0; ii < (long long)narcs; ++ii) {
        size_t i = (size_t)ii, t, h;
        if (i < nrowarc) { t = i; h = i + ncol; }
        else { size_t j = i - nrowarc; size_t r = j / (ncol - 1), c = j % (ncol - 1); t = r * ncol + c; h = t + 1; }
        w[i] = reweight_one_f(sigsq[i * stride], phi[h], phi[t], u[h], u[t], d2);
    }
}

/* ── a12: congruence (laplace_congruence cuphu_laplace.cu:256-266) ───────────
 * u = fma(2pi, rint((u - phi) * 1/(2pi)), phi). */
void o_congruence(const float *phi, float *u, size_t n) {
    for (size_t i = 0; i < n; ++i) {
        float d = u[i] - phi[i];
        u[i] = fmaf(O_TWOPI_F, rintf(d * O_INV_TWOPI_F), phi[i]);
    }
}

/* ── a10: PCG/IRLS driver, FP32 vectors with FP64 reductions ─────────────────
 * cuphu_laplace_unwrap_gpu cuphu_laplace.cu:351-552: weights from sigsq
 * (:410), (N_IRLS+1) solves (:413) each assembling b/diag (:416), u=0 before
 * the first (:427), r = b - A u (:431-434), Jacobi z (:437), CG loop with
 * pq <= 0 break (:460), rz_new < tol^2 rz0 stop (:478), reweight with
 * delta = max(dmin, dmax*0.5^outer) (:516-522); anchor to phi[0] (:534-539)
 * and congruence (:549).  Scalars are FP64 (the reference's float atomics
 * are nondeterministic), vector updates use fmaf and z = r * (1/diag). */
static double dot_f32(const float *a, const float *b, size_t n) {
    double s = 0.0;
    for (size_t i = 0; i < n; ++i) s += (double)a[i] * (double)b[i];
    return s;
}
int o_pcg_irls_f32(const short *sigsq, size_t stride, const float *phi, int nrow, int ncol,
                   const OLapOptions *o, const float *u0, float *u, OLapStats *st) {
    size_t n = (size_t)nrow * ncol, narcs = (size_t)(nrow - 1) * ncol + (size_t)nrow * (ncol - 1);
    float *w = (float *)malloc((narcs + 1) * sizeof(float));
    float *b = (float *)malloc(n * sizeof(float)), *dinv = (float *)malloc(n * sizeof(float));
    float *r = (float *)malloc(n * sizeof(float)), *z = (float *)malloc(n * sizeof(float));
    float *p = (float *)malloc(n * sizeof(float)), *q = (float *)malloc(n * sizeof(float));
    if (!w || !b || !dinv || !r || !z || !p || !q) return -1;
    o_arc_weights(sigsq, narcs, stride, w);
    if (st) memset(st, 0, sizeof(*st));
    for (int outer = 0; outer <= o->n_irls; ++outer) {
        o_assemble(w, phi, nrow, ncol, b, dinv);
        for (size_t i = 0; i < n; ++i) dinv[i] = 1.0f / dinv[i];
        if (outer == 0) {
            if (u0) memcpy(u, u0, n * sizeof(float)); else memset(u, 0, n * sizeof(float));
        }
        o_matvec(w, u, nrow, ncol, r);
        for (size_t i = 0; i < n; ++i) { r[i] = b[i] - r[i]; z[i] = r[i] * dinv[i]; p[i] = z[i]; }
        double rz = dot_f32(r, z, n), rz0 = rz, rzl = rz;
        int iter = 0, status = 0;
        if (!(rz0 > 0.0)) status = (rz0 == 0.0) ? 3 : 4;
        else {
            status = 1;
            for (; iter < o->max_iter; ++iter) {
                o_matvec(w, p, nrow, ncol, q);
                double pq = dot_f32(p, q, n);
                if (!(pq > 0.0)) { status = isfinite(pq) ? 2 : 4; break; }
                float af = (float)(rz / pq);
                for (size_t i = 0; i < n; ++i) {
                    u[i] = fmaf(af, p[i], u[i]);
                    r[i] = fmaf(-af, q[i], r[i]);
                    z[i] = r[i] * dinv[i];
                }
                double rzn = dot_f32(r, z, n);
                rzl = rzn;
                if (rzn < o->tol * o->tol * rz0) { iter++; status = 0; break; }
                float bf = (float)(rzn / rz);
                rz = rzn;
                for (size_t i = 0; i < n; ++i) p[i] = fmaf(bf, p[i], z[i]);
            }
        }
        if (st) {
            st->iters[outer] = iter; st->status[outer] = status;
            st->relres[outer] = (rz0 > 0.0) ? sqrt(rzl / rz0) : 0.0;
            st->npass = outer + 1;
        }
        if (outer < o->n_irls) {
            float delta = o->delta_max * powf(0.5f, (float)outer);
            if (delta < o->delta_min) delta = o->delta_min;
            o_reweight(sigsq, stride, phi, u, delta, nrow, ncol, w);
        }
    }
    float off = phi[0] - u[0];
    if (off != 0.0f) for (size_t i = 0; i < n; ++i) u[i] = u[i] + off;
    if (!o->continuous) o_congruence(phi, u, n);
    free(w); free(b); free(dinv); free(r); free(z); free(p); free(q);
    return 0;
}

/* FP64 restatement of the same algorithm (the accuracy oracle): weights
 * 1/sigsq, wrapped gradients, stencil, reductions and updates all in
 * double, OpenMP-parallel with a deterministic (static, per-thread ordered)
 * reduction.  Output is the anchored, congruent solution in double. */
static void assemble_d(const double *w, const float *phi, int nrow, int ncol, double *b, double *dinv) {
    const double *row_w = w, *col_w = w + (size_t)(nrow - 1) * ncol;
#pragma omp parallel for schedule(static)
    for (int r = 0; r < nrow; ++r)
        for (int c = 0; c < ncol; ++c) {
            size_t i = (size_t)r * ncol + c;
            double pa = phi[i], bv = 0.0, ws = 0.0;
            if (r < nrow - 1) { double a = row_w[i]; if (a > 0.0) { bv = fma(-a, wrap_pi_d((double)phi[i + ncol] - pa), bv); ws += a; } }
            if (r > 0) { double a = row_w[i - ncol]; if (a > 0.0) { bv = fma(a, wrap_pi_d(pa - (double)phi[i - ncol]), bv); ws += a; } }
            if (c < ncol - 1) { double a = col_w[(size_t)r * (ncol - 1) + c]; if (a > 0.0) { bv = fma(-a, wrap_pi_d((double)phi[i + 1] - pa), bv); ws += a; } }
            if (c > 0) { double a = col_w[(size_t)r * (ncol - 1) + c - 1]; if (a > 0.0) { bv = fma(a, wrap_pi_d(pa - (double)phi[i - 1]), bv); ws += a; } }
            b[i] = bv;
            dinv[i] = 1.0 / (ws + O_EPS_D);
        }
}
static void matvec_d(const double *w, const double *x, int nrow, int ncol, double *y) {
    const double *row_w = w, *col_w = w + (size_t)(nrow - 1) * ncol;
#pragma omp parallel for schedule(static)
    for (int r = 0; r < nrow; ++r)
        for (int c = 0; c < ncol; ++c) {
            size_t i = (size_t)r * ncol + c;
            double xa = x[i], yv = O_EPS_D * xa;
            if (r < nrow - 1) yv = fma(row_w[i], xa - x[i + ncol], yv);
            if (r > 0) yv = fma(row_w[i - ncol], xa - x[i - ncol], yv);
            if (c < ncol - 1) yv = fma(col_w[(size_t)r * (ncol - 1) + c], xa - x[i + 1], yv);
            if (c > 0) yv = fma(col_w[(size_t)r * (ncol - 1) + c - 1], xa - x[i - 1], yv);
            y[i] = yv;
        }
}
static double dot_d(const double *a, const double *b, size_t n) {
    int nt = 1;
#ifdef _OPENMP
    nt = omp_get_max_threads();
#endif
    double part[256] = {0};
    if (nt > 256) nt = 256;
#pragma omp parallel num_threads(nt)
    {
        int t = 0, T = 1;
#ifdef _OPENMP
        t = omp_get_thread_num(); T = omp_get_num_threads();
#endif
        size_t lo = n * (size_t)t / (size_t)T, hi = n * (size_t)(t + 1) / (size_t)T;
        double s = 0.0;
        for (size_t i = lo; i < hi; ++i) s += a[i] * b[i];
        part[t] = s;
    }
    double s = 0.0;
    for (int t = 0; t < nt; ++t) s += part[t];
    return s;
}
/* The two sweeps of an FP64 iteration with their inner products fused in
 * (q = L p with p.q; x, r, z updates with r.z), so each vector is streamed
 * once per iteration.  Partial sums run over a static partition per thread
 * and are added in thread order: deterministic for a given thread count. */
static int omp_threads_capped(void) {
    int nt = 1;
#ifdef _OPENMP
    nt = omp_get_max_threads();
#endif
    return nt > 256 ? 256 : nt;
}
/* One output row of the FP64 stencil: arcs in the order down, up, right,
 * left (as matvec_d); the interior columns run branch-free so the compiler
 * vectorizes them.  wd/xd (wu/xu) are NULL on the last (first) row. */
#define O_MV_BODY(c, DN, UP)                                         \
    {                                                                \
        double xa = x[c], yv = O_EPS_D * xa;                         \
        if (DN) yv = fma(wd[c], xa - xd[c], yv);                     \
        if (UP) yv = fma(wu[c], xa - xu[c], yv);                     \
        yv = fma(wc[c], xa - x[(c) + 1], yv);                        \
        yv = fma(wc[(c) - 1], xa - x[(c) - 1], yv);                  \
        y[c] = yv;                                                   \
    }
static void matvec_row_d(const double *restrict wd, const double *restrict wu, const double *restrict wc,
                         const double *restrict x, const double *restrict xd, const double *restrict xu,
                         int ncol, double *restrict y) {
    int dn = wd != NULL, up = wu != NULL;
    for (int c = 0; c < ncol; c += (ncol > 1 ? ncol - 1 : 1)) {   /* the two edge columns */
        double xa = x[c], yv = O_EPS_D * xa;
        if (dn) yv = fma(wd[c], xa - xd[c], yv);
        if (up) yv = fma(wu[c], xa - xu[c], yv);
        if (c < ncol - 1) yv = fma(wc[c], xa - x[c + 1], yv);
        if (c > 0) yv = fma(wc[c - 1], xa - x[c - 1], yv);
        y[c] = yv;
    }
    if (dn && up) { for (int c = 1; c < ncol - 1; ++c) O_MV_BODY(c, 1, 1) }
    else if (dn)  { for (int c = 1; c < ncol - 1; ++c) O_MV_BODY(c, 1, 0) }
    else if (up)  { for (int c = 1; c < ncol - 1; ++c) O_MV_BODY(c, 0, 1) }
    else          { for (int c = 1; c < ncol - 1; ++c) O_MV_BODY(c, 0, 0) }
}
#undef O_MV_BODY
/* a.b over n elements with four interleaved partial sums (a fixed order). */
static double dot4_d(const double *restrict a, const double *restrict b, size_t n) {
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    size_t i = 0;
    for (; i + 4 <= n; i += 4) {
        s0 += a[i] * b[i]; s1 += a[i + 1] * b[i + 1]; s2 += a[i + 2] * b[i + 2]; s3 += a[i + 3] * b[i + 3];
    }
    for (; i < n; ++i) s0 += a[i] * b[i];
    return (s0 + s1) + (s2 + s3);
}
static double matvec_dot_d(const double *w, const double *x, int nrow, int ncol, double *y) {
    const double *row_w = w, *col_w = w + (size_t)(nrow - 1) * ncol;
    int nt = omp_threads_capped();
    double part[256] = {0};
#pragma omp parallel num_threads(nt)
    {
        int t = 0, T = 1;
#ifdef _OPENMP
        t = omp_get_thread_num(); T = omp_get_num_threads();
#endif
        int rlo = (int)((long long)nrow * t / T), rhi = (int)((long long)nrow * (t + 1) / T);
        double s = 0.0;
        for (int r = rlo; r < rhi; ++r) {
            size_t i = (size_t)r * ncol;
            matvec_row_d(r < nrow - 1 ? row_w + i : NULL, r > 0 ? row_w + i - ncol : NULL,
                         col_w + (size_t)r * (ncol - 1), x + i, r < nrow - 1 ? x + i + ncol : NULL,
                         r > 0 ? x + i - ncol : NULL, ncol, y + i);
            s += dot4_d(x + i, y + i, (size_t)ncol);
        }
        part[t] = s;
    }
    double s = 0.0;
    for (int t = 0; t < nt; ++t) s += part[t];
    return s;
}
static double update_xrz_d(double alpha, const double *restrict p, const double *restrict q,
                           const double *restrict dinv, double *restrict u, double *restrict r,
                           double *restrict z, size_t n) {
    int nt = omp_threads_capped();
    double part[256] = {0};
#pragma omp parallel num_threads(nt)
    {
        int t = 0, T = 1;
#ifdef _OPENMP
        t = omp_get_thread_num(); T = omp_get_num_threads();
#endif
        size_t lo = n * (size_t)t / (size_t)T, hi = n * (size_t)(t + 1) / (size_t)T;
        double s = 0.0;
        for (size_t a = lo; a < hi; a += 4096) {          /* blocks stay in L1 for the dot */
            size_t e = a + 4096 < hi ? a + 4096 : hi;
            for (size_t i = a; i < e; ++i) {
                u[i] = fma(alpha, p[i], u[i]);
                r[i] = fma(-alpha, q[i], r[i]);
                z[i] = r[i] * dinv[i];
            }
            s += dot4_d(r + a, z + a, e - a);
        }
        part[t] = s;
    }
    double s = 0.0;
    for (int t = 0; t < nt; ++t) s += part[t];
    return s;
}
int o_pcg_irls_f64(const short *sigsq, size_t stride, const float *phi, int nrow, int ncol,
                   const OLapOptions *o, const float *u0, double *u, OLapStats *st) {
    size_t n = (size_t)nrow * ncol, nrowarc = (size_t)(nrow - 1) * ncol,
           narcs = nrowarc + (size_t)nrow * (ncol - 1);
    double *w = (double *)malloc((narcs + 1) * sizeof(double));
    double *b = (double *)malloc(n * sizeof(double)), *dinv = (double *)malloc(n * sizeof(double));
    double *r = (double *)malloc(n * sizeof(double)), *z = (double *)malloc(n * sizeof(double));
    double *p = (double *)malloc(n * sizeof(double)), *q = (double *)malloc(n * sizeof(double));
    if (!w || !b || !dinv || !r || !z || !p || !q) return -1;
    for (size_t i = 0; i < narcs; ++i) {
        short s = sigsq[i * stride];
        w[i] = (s <= 0 || s >= O_LARGESHORT) ? 0.0 : 1.0 / (double)s;
    }
    if (st) memset(st, 0, sizeof(*st));
    for (int outer = 0; outer <= o->n_irls; ++outer) {
        assemble_d(w, phi, nrow, ncol, b, dinv);
        if (outer == 0) {
            if (u0) for (size_t i = 0; i < n; ++i) u[i] = u0[i];
            else memset(u, 0, n * sizeof(double));
        }
        matvec_d(w, u, nrow, ncol, r);
#pragma omp parallel for schedule(static)
        for (long long i = 0; i < (long long)n; ++i) { r[i] = b[i] - r[i]; z[i] = r[i] * dinv[i]; p[i] = z[i]; }
        double rz = dot_d(r, z, n), rz0 = rz, rzl = rz;
        int iter = 0, status = 0;
        if (!(rz0 > 0.0)) status = (rz0 == 0.0) ? 3 : 4;
        else {
            status = 1;
            for (; iter < o->max_iter; ++iter) {
                double pq = matvec_dot_d(w, p, nrow, ncol, q);
                if (!(pq > 0.0)) { status = isfinite(pq) ? 2 : 4; break; }
                double alpha = rz / pq;
                double rzn = update_xrz_d(alpha, p, q, dinv, u, r, z, n);
                rzl = rzn;
                if (rzn < o->tol * o->tol * rz0) { iter++; status = 0; break; }
                double beta = rzn / rz;
                rz = rzn;
#pragma omp parallel for schedule(static)
                for (long long i = 0; i < (long long)n; ++i) p[i] = fma(beta, p[i], z[i]);
            }
        }
        if (st) {
            st->iters[outer] = iter; st->status[outer] = status;
            st->relres[outer] = (rz0 > 0.0) ? sqrt(rzl / rz0) : 0.0;
            st->npass = outer + 1;
        }
        if (outer < o->n_irls) {
            float df = o->delta_max * powf(0.5f, (float)outer);
            if (df < o->delta_min) df = o->delta_min;
            double d2 = (double)df * (double)df;
#pragma omp parallel for schedule(static)
            for (long long ii = 0; ii < (long long)narcs; ++ii) {
                size_t i = (size_t)ii, t, h;
                if (i < nrowarc) { t = i; h = i + ncol; }
                else { size_t j = i - nrowarc; size_t rr = j / (ncol - 1), cc = j % (ncol - 1); t = rr * ncol + cc; h = t + 1; }
                short s = sigsq[i * stride];
                double w0 = (s <= 0 || s >= O_LARGESHORT) ? 0.0 : 1.0 / (double)s;
                if (w0 <= 0.0) { w[i] = 0.0; continue; }
                double e = (u[h] - u[t]) - wrap_pi_d((double)phi[h] - (double)phi[t]);
                w[i] = w0 / sqrt(fma(e, e, d2));
            }
        }
    }
    double off = (double)phi[0] - u[0];
    for (size_t i = 0; i < n; ++i) {
        double v = u[i] + off, d = v - (double)phi[i];
        u[i] = o->continuous ? v : fma(O_TWOPI_D, rint(d / O_TWOPI_D), (double)phi[i]);
    }
    free(w); free(b); free(dinv); free(r); free(z); free(p); free(q);
    return 0;
}

/* ── one Jacobi-PCG iteration, FP32 vectors (the loop body of
 * o_pcg_irls_f32 above, cuphu_laplace.cu:452-498, with the direction update
 * of :493-498 moved to the front, as the batched HIP solver runs it):
 * p = fmaf((float)beta, p, r*dinv), or p = r*dinv when beta == 0 (first
 * step); q = A p (o_matvec order); pq = sum p*q; alpha = gamma / pq, or
 * alpha_in when it is finite; x = fmaf((float)alpha, p, x);
 * r = fmaf(-(float)alpha, q, r); rz = sum r*(r*dinv).  FP64 sequential sums.
 * Returns 0, or 2 (no update of x, r) when pq <= 0. */
int o_pcg_step_f32(const float *w, const float *dinv, int nrow, int ncol, double gamma, double beta,
                   double alpha_in, float *x, float *r, float *p, float *q, double *pq, double *alpha,
                   double *rz) {
    size_t n = (size_t)nrow * ncol;
    float bf = (float)beta;
    for (size_t i = 0; i < n; ++i) {
        float z = r[i] * dinv[i];
        p[i] = (beta != 0.0) ? fmaf(bf, p[i], z) : z;
    }
    o_matvec(w, p, nrow, ncol, q);
    double s = dot_f32(p, q, n);
    *pq = s;
    if (!(s > 0.0)) return 2;
    double a = isfinite(alpha_in) ? alpha_in : gamma / s;
    *alpha = a;
    float af = (float)a;
    double g = 0.0;
    for (size_t i = 0; i < n; ++i) {
        x[i] = fmaf(af, p[i], x[i]);
        r[i] = fmaf(-af, q[i], r[i]);
        g += (double)r[i] * (double)(r[i] * dinv[i]);
    }
    *rz = g;
    return 0;
}

/* ── one fused PCG iteration, FP32 vectors (the batched solver's one-kernel
 * form of the loop body of o_pcg_irls_f32, cuphu_laplace.cu:452-498):
 * with (alpha, beta) of the previous iteration,
 *   x = fmaf(alpha, p, x); r = fmaf(-alpha, q, r)  (r = rt when rt != NULL);
 *   z = r*dinv; p = fmaf(beta, p, z) (p = z when beta == 0); q = A p;
 * start != 0 skips the x and r updates and sets p = z.  sums[0..3] =
 * p.q, r.z, q.(dinv r), q.(dinv q) in FP64, sequential. */
void o_pcg_fused_step_f32(const float *w, const float *dinv, int nrow, int ncol, double alpha, double beta,
                          int start, const float *rt, float *x, float *r, float *p, float *q, double *sums) {
    size_t n = (size_t)nrow * ncol;
    float af = (float)alpha, bf = (float)beta;
    for (size_t i = 0; i < n; ++i) {
        if (!start) {
            x[i] = fmaf(af, p[i], x[i]);
            r[i] = rt ? rt[i] : fmaf(-af, q[i], r[i]);
        }
        float z = r[i] * dinv[i];
        p[i] = (!start && beta != 0.0) ? fmaf(bf, p[i], z) : z;
    }
    o_matvec(w, p, nrow, ncol, q);
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    for (size_t i = 0; i < n; ++i) {
        double qd = (double)q[i], dd = (double)dinv[i], rd = (double)r[i];
        s0 += (double)p[i] * qd;
        s1 += rd * (double)(r[i] * dinv[i]);
        s2 += qd * dd * rd;
        s3 += qd * dd * qd;
    }
    sums[0] = s0; sums[1] = s1; sums[2] = s2; sums[3] = s3;
}

/* ── a14: connected components (cuphu_conncomp_gpu cuphu_conncomp.cu:289-428)
 * Valid: !(corr < 0.05f) and mask != 0 (MaskLabelsKernel :102-114, floor at
 * :324-325).  Edge between 4-neighbours when both valid and
 * fabsf(unw_a - unw_b) < pi (MergeLabelsKernel :178-218 with poscost NULL,
 * :331).  Component root = its minimum raster index (invariant :16-20).
 * Kept components: size >= max(1, (int)(minfrac*npix)), labelled 1.. in
 * increasing root index until maxncomps labels are used
 * (build_label_map :272-286, :412-417); everything else 0 (RelabelKernel). */
static uint32_t uf_find(uint32_t *par, uint32_t x) {
    while (par[x] != x) { par[x] = par[par[x]]; x = par[x]; }
    return x;
}
static void uf_union(uint32_t *par, uint32_t a, uint32_t b) {
    a = uf_find(par, a); b = uf_find(par, b);
    if (a == b) return;
    if (a < b) par[b] = a; else par[a] = b;
}
/* Incremental-cost gate of an arc (arc_confident, cuphu_conncomp.cu:84-93):
 * confident when min(poscost, negcost) >= thresh; no costs = confident. */
static int o_arc_ok(const short *pos, const short *neg, size_t a, long thresh) {
    if (!pos) return 1;
    short m = pos[a] < neg[a] ? pos[a] : neg[a];
    return (long)m >= thresh;
}
int o_conncomp_gated(const float *unw, const float *corr, const unsigned char *mask, int nrow, int ncol,
                     double minfrac, long maxncomps, const short *poscost, const short *negcost,
                     long conncompthresh, uint32_t *labels);
int o_conncomp(const float *unw, const float *corr, const unsigned char *mask, int nrow, int ncol,
               double minfrac, long maxncomps, uint32_t *labels) {
    return o_conncomp_gated(unw, corr, mask, nrow, ncol, minfrac, maxncomps, NULL, NULL, 0, labels);
}
/* The same with the MCF path's arc gate: row arc (r,c)-(r+1,c) at index
 * r*ncol + c, column arc (r,c)-(r,c+1) at (nrow-1)*ncol + r*(ncol-1) + c
 * (arc_index, cuphu_conncomp.cu:67-78). */
int o_conncomp_gated(const float *unw, const float *corr, const unsigned char *mask, int nrow, int ncol,
                     double minfrac, long maxncomps, const short *poscost, const short *negcost,
                     long conncompthresh, uint32_t *labels) {
    size_t n = (size_t)nrow * ncol;
    uint32_t *par = (uint32_t *)malloc(n * sizeof(uint32_t));
    uint32_t *cnt = (uint32_t *)calloc(n, sizeof(uint32_t));
    unsigned char *valid = (unsigned char *)malloc(n);
    if (!par || !cnt || !valid) return -1;
    for (size_t i = 0; i < n; ++i) {
        par[i] = (uint32_t)i;
        int bad = corr[i] < 0.05f;
        if (mask) bad = bad || (mask[i] == 0);
        valid[i] = !bad;
    }
    const float pth = O_PI_F;
    for (int r = 0; r < nrow; ++r)
        for (int c = 0; c < ncol; ++c) {
            size_t i = (size_t)r * ncol + c;
            if (!valid[i]) continue;
            size_t ra = (size_t)r * ncol + c, ca = (size_t)(nrow - 1) * ncol + (size_t)r * (ncol - 1) + c;
            if (c + 1 < ncol && valid[i + 1] && fabsf(unw[i] - unw[i + 1]) < pth &&
                o_arc_ok(poscost, negcost, ca, conncompthresh))
                uf_union(par, (uint32_t)i, (uint32_t)(i + 1));
            if (r + 1 < nrow && valid[i + ncol] && fabsf(unw[i] - unw[i + ncol]) < pth &&
                o_arc_ok(poscost, negcost, ra, conncompthresh))
                uf_union(par, (uint32_t)i, (uint32_t)(i + ncol));
        }
    for (size_t i = 0; i < n; ++i) if (valid[i]) cnt[uf_find(par, (uint32_t)i)]++;
    int min_size = (int)(minfrac * (double)(int)n);
    if (min_size < 1) min_size = 1;
    uint32_t *map = (uint32_t *)calloc(n, sizeof(uint32_t));
    uint32_t next = 1;
    for (size_t i = 0; i < n && (long)next - 1 < maxncomps; ++i)
        if (cnt[i] >= (uint32_t)min_size) map[i] = next++;
    for (size_t i = 0; i < n; ++i) labels[i] = valid[i] ? map[uf_find(par, (uint32_t)i)] : 0u;
    free(par); free(cnt); free(valid); free(map);
    return 0;
}

/* ── MCF path: incremental arc costs (ComputeIncrCostsKernel,
 * cuphu_incrcost.cu:55-98, SNAPHU CalcCostSmooth) ───────────────────────────
 * costs: smoothcostT pairs {offset, sigsq}.  sigsq == 32000 (LARGESHORT): 0.
 * idz1 = |f*nsc + off|, idz2pos/neg = |(f +- nflow)*nsc + off|;
 * cost1 = idz1^2 / sigsq (integer division), raw = idz2^2 / sigsq - cost1;
 * result = ceil(raw / nflow^2) if raw > 0 else floor(raw / nflow^2), clamped
 * to [-32000, 32000].  Returns 1 when min(poscost) >= 0 and min(negcost) >= 0
 * over all arcs (the early-exit test of cuphu_incrcost_early_exit, :146-191;
 * narcs == 0 gives 1). */
static long o_ceil_div(long a, long b) { return (a > 0L) ? (a + b - 1L) / b : a / b; }
static long o_floor_div(long a, long b) { return (a >= 0L) ? a / b : (a - b + 1L) / b; }
int o_incrcost_smooth(const short *costs, const short *flows, size_t narcs, long nshortcycle, int nflow,
                      short *poscost, short *negcost) {
    int minp = 32000, minn = 32000;
    for (size_t i = 0; i < narcs; ++i) {
        long off = costs[2 * i], sig = costs[2 * i + 1], f = flows[i];
        long pc = 0, nc = 0;
        if (sig != 32000) {
            long nfl = nflow;
            long idz1 = labs(f * nshortcycle + off);
            long idp = labs((f + nfl) * nshortcycle + off);
            long idn = labs((f - nfl) * nshortcycle + off);
            long c1 = (idz1 * idz1) / sig;
            long pr = (idp * idp) / sig - c1, nr = (idn * idn) / sig - c1;
            long nsq = nfl * nfl;
            pc = pr > 0L ? o_ceil_div(pr, nsq) : o_floor_div(pr, nsq);
            nc = nr > 0L ? o_ceil_div(nr, nsq) : o_floor_div(nr, nsq);
            if (pc > 32000L) pc = 32000L;
            if (pc < -32000L) pc = -32000L;
            if (nc > 32000L) nc = 32000L;
            if (nc < -32000L) nc = -32000L;
        }
        poscost[i] = (short)pc;
        negcost[i] = (short)nc;
        if ((int)pc < minp) minp = (int)pc;
        if ((int)nc < minn) minn = (int)nc;
    }
    return (minp >= 0 && minn >= 0) ? 1 : 0;
}

/* ── a13: integer flows -> phase (cuphu_integrate_phase_gpu cuphu_phase.cu) ──
 * unw[r,c] = phi[0,0] + rowoff[r] + coloff[r,c]; rowoff sums down column 0
 * the terms fma(-2pi, hflow, wrap(phi[i+1,0]-phi[i,0])) (:42-59), coloff
 * sums along row r fma(2pi, vflow, wrap(phi[r,j+1]-phi[r,j])) (:68-85);
 * wrap_angle x - 2pi*rint(x/2pi) (:27-29) as one fma.  The sums are carried
 * in FP64 and the three parts combined in FP64, rounded once to float. */
static inline float wrap_angle_f(float x) { return fmaf(-O_TWOPI_F, rintf(x / O_TWOPI_F), x); }
void o_integrate_phase(const float *phi, const short *hflows, const short *vflows, int nrow, int ncol,
                       float *unw) {
    double *rowoff = (double *)malloc((size_t)nrow * sizeof(double));
    rowoff[0] = 0.0;
    for (int r = 1; r < nrow; ++r) {
        float t = fmaf(-O_TWOPI_F, (float)hflows[(size_t)(r - 1) * ncol],
                       wrap_angle_f(phi[(size_t)r * ncol] - phi[(size_t)(r - 1) * ncol]));
        rowoff[r] = rowoff[r - 1] + (double)t;
    }
    double ref = (double)phi[0];
    for (int r = 0; r < nrow; ++r) {
        double co = 0.0;
        unw[(size_t)r * ncol] = (float)(ref + rowoff[r] + 0.0);
        for (int c = 1; c < ncol; ++c) {
            float t = fmaf(O_TWOPI_F, (float)vflows[(size_t)r * (ncol - 1) + c - 1],
                           wrap_angle_f(phi[(size_t)r * ncol + c] - phi[(size_t)r * ncol + c - 1]));
            co += (double)t;
            unw[(size_t)r * ncol + c] = (float)(ref + rowoff[r] + co);
        }
    }
    free(rowoff);
}

/* ── a15: tile geometry ──────────────────────────────────────────────────────
 * tile_size == 0: cuphu_solver.cpp:1759-1766 (first = t*(n-ovl)/nt,
 * last = (t+1)*(n-ovl)/nt + ovl, last tile ends at n).
 * tile_size > 0: fixed tiles of tile_size every tile_size-overlap pixels,
 * the last one truncated at n. */
int o_tile_axis(int n, int tile_size, int overlap, int ntiles, int *first, int *size, int max_tiles) {
    int k = 0;
    if (tile_size <= 0) {
        if (ntiles < 1) return -1;
        for (int t = 0; t < ntiles; ++t) {
            int f = (int)((long)t * (n - overlap) / ntiles);
            int l = (t == ntiles - 1) ? n : (int)((long)(t + 1) * (n - overlap) / ntiles) + overlap;
            if (k < max_tiles) { first[k] = f; size[k] = l - f; }
            k++;
        }
        return k;
    }
    int step = tile_size - overlap;
    if (step <= 0) return -1;
    for (int s = 0;; s += step) {
        int e = s + tile_size;
        if (e >= n) e = n;
        if (k < max_tiles) { first[k] = s; size[k] = e - s; }
        k++;
        if (e >= n) break;
    }
    return k;
}

/* ── a15: stitching (cuphu_solver.cpp:1846-2031) ─────────────────────────────
 * Registration: weighted median (sort by value, first element where the
 * running weight reaches half the total, :1867-1881) of L-R (or T-B)
 * overlap differences over pixels with corr >= 0.05 in both tiles and
 * valid conf mask, weight min(corr_L, corr_R) (:1899-1959); cycles = median
 * / 2pi, rounded unless laplace.  BFS from tile 0, right neighbour first,
 * then bottom (:1969-1999).  Merge: interior sides trimmed by ovl/2, unw +
 * (float)(tile_k*2pi), labels offset by the running per-tile maximum
 * (:2003-2031). */
typedef struct { float v, w; } OPair;
static int cmp_pair(const void *a, const void *b) {
    float x = ((const OPair *)a)->v, y = ((const OPair *)b)->v;
    return (x < y) ? -1 : (x > y) ? 1 : 0;
}
static double weighted_median(OPair *d, size_t n) {
    if (n == 0) return 0.0;
    qsort(d, n, sizeof(OPair), cmp_pair);
    double total = 0.0;
    for (size_t i = 0; i < n; ++i) total += d[i].w;
    double half = total / 2.0, acc = 0.0;
    for (size_t i = 0; i < n; ++i) { acc += d[i].w; if (acc >= half) return (double)d[i].v; }
    return (double)d[n - 1].v;
}
int o_stitch(int nrow, int ncol, int ntr, int ntc, const int *rf, const int *rs, const int *cf,
             const int *cs, int rovl, int covl, int laplace, const float *const *tunw,
             const uint32_t *const *tcc, const float *const *tcorr,
             const unsigned char *const *tconf, float *unw, uint32_t *cc, double *tile_k) {
    int nt = ntr * ntc;
    double *tk = (double *)calloc((size_t)nt, sizeof(double));
    memset(unw, 0, (size_t)nrow * ncol * sizeof(float));
    memset(cc, 0, (size_t)nrow * ncol * sizeof(uint32_t));
    if (nt > 1 && (rovl > 0 || covl > 0)) {
        char *vis = (char *)calloc((size_t)nt, 1);
        int *queue = (int *)malloc((size_t)nt * sizeof(int));
        int qh = 0, qt = 0;
        queue[qt++] = 0; vis[0] = 1;
        while (qh < qt) {
            int idx = queue[qh++], br = idx / ntc, bc = idx % ntc;
            for (int dir = 0; dir < 2; ++dir) {
                int nb;
                if (dir == 0) { if (bc + 1 >= ntc) continue; nb = br * ntc + bc + 1; }
                else { if (br + 1 >= ntr) continue; nb = (br + 1) * ntc + bc; }
                if (vis[nb]) continue;
                int A = idx, B = nb;
                int anr = rs[A / ntc], anc = cs[A % ntc], bnr = rs[B / ntc], bnc = cs[B % ntc];
                double k = 0.0;
                OPair *d = NULL; size_t nd = 0;
                if (dir == 0 && !(covl <= 0 || covl > anc || covl > bnc)) {
                    int nr = anr < bnr ? anr : bnr;
                    d = (OPair *)malloc((size_t)nr * covl * sizeof(OPair) + 1);
                    for (int r = 0; r < nr; ++r)
                        for (int c = 0; c < covl; ++c) {
                            size_t lp = (size_t)r * anc + (anc - covl + c), rp = (size_t)r * bnc + c;
                            if (tcorr[A][lp] < 0.05f || tcorr[B][rp] < 0.05f) continue;
                            if (tconf && tconf[A] && !tconf[A][lp]) continue;
                            if (tconf && tconf[B] && !tconf[B][rp]) continue;
                            float wv = tcorr[A][lp] < tcorr[B][rp] ? tcorr[A][lp] : tcorr[B][rp];
                            d[nd].v = tunw[A][lp] - tunw[B][rp]; d[nd].w = wv; nd++;
                        }
                } else if (dir == 1 && !(rovl <= 0 || rovl > anr || rovl > bnr)) {
                    int nc = anc < bnc ? anc : bnc;
                    d = (OPair *)malloc((size_t)rovl * nc * sizeof(OPair) + 1);
                    for (int r = 0; r < rovl; ++r)
                        for (int c = 0; c < nc; ++c) {
                            size_t tp = (size_t)(anr - rovl + r) * anc + c, bp = (size_t)r * bnc + c;
                            if (tcorr[A][tp] < 0.05f || tcorr[B][bp] < 0.05f) continue;
                            if (tconf && tconf[A] && !tconf[A][tp]) continue;
                            if (tconf && tconf[B] && !tconf[B][bp]) continue;
                            float wv = tcorr[A][tp] < tcorr[B][bp] ? tcorr[A][tp] : tcorr[B][bp];
                            d[nd].v = tunw[A][tp] - tunw[B][bp]; d[nd].w = wv; nd++;
                        }
                }
                if (d) {
                    double cyc = weighted_median(d, nd) / (2.0 * M_PI);
                    k = laplace ? cyc : round(cyc);
                    free(d);
                }
                tk[nb] = tk[idx] + k;
                vis[nb] = 1; queue[qt++] = nb;
            }
        }
        free(vis); free(queue);
    }
    uint32_t off = 0;
    for (int tr = 0; tr < ntr; ++tr)
        for (int tc = 0; tc < ntc; ++tc) {
            int t = tr * ntc + tc, tnr = rs[tr], tnc = cs[tc];
            uint32_t lmax = 0;
            for (size_t i = 0; i < (size_t)tnr * tnc; ++i) if (tcc[t][i] > lmax) lmax = tcc[t][i];
            float add = (float)(tk[t] * (2.0 * M_PI));
            int pr0 = (tr == 0) ? 0 : rovl / 2, pc0 = (tc == 0) ? 0 : covl / 2;
            int pr1 = (tr == ntr - 1) ? tnr : tnr - rovl / 2, pc1 = (tc == ntc - 1) ? tnc : tnc - covl / 2;
            for (int r = pr0; r < pr1; ++r)
                for (int c = pc0; c < pc1; ++c) {
                    size_t ti = (size_t)r * tnc + c, gi = (size_t)(rf[tr] + r) * ncol + (cf[tc] + c);
                    unw[gi] = tunw[t][ti] + add;
                    cc[gi] = (tcc[t][ti] == 0) ? 0 : tcc[t][ti] + off;
                }
            off += lmax;
        }
    if (tile_k) memcpy(tile_k, tk, (size_t)nt * sizeof(double));
    free(tk);
    return 0;
}

/* ── differentiable WLS: fixed weights, FP64 ─────────────────────────────────
 * L(w) u = b(w, phi) by Jacobi-PCG from 0 (same operator as a7/a8 with
 * double weights), mean removed.  Adjoint: L lambda = P gu, then
 * dLoss/dw_a = -(lambda_h - lambda_t) * (u_h - u_t - g_a). */
static int pcg_fixed_d(const double *w, const double *b, const double *dinv, int nrow, int ncol,
                       int max_iter, double tol, double *x, int *iters) {
    size_t n = (size_t)nrow * ncol;
    double *r = malloc(n * sizeof(double)), *z = malloc(n * sizeof(double)),
           *p = malloc(n * sizeof(double)), *q = malloc(n * sizeof(double));
    memset(x, 0, n * sizeof(double));
    for (size_t i = 0; i < n; ++i) { r[i] = b[i]; z[i] = r[i] * dinv[i]; p[i] = z[i]; }
    double rz = dot_d(r, z, n), rz0 = rz;
    int it = 0;
    if (rz0 > 0.0) {
        for (; it < max_iter; ++it) {
            matvec_d(w, p, nrow, ncol, q);
            double pq = dot_d(p, q, n);
            if (!(pq > 0.0)) break;
            double a = rz / pq;
            for (size_t i = 0; i < n; ++i) { x[i] += a * p[i]; r[i] -= a * q[i]; z[i] = r[i] * dinv[i]; }
            double rzn = dot_d(r, z, n);
            if (rzn < tol * tol * rz0) { it++; break; }
            double be = rzn / rz; rz = rzn;
            for (size_t i = 0; i < n; ++i) p[i] = z[i] + be * p[i];
        }
    }
    double m = 0.0;
    for (size_t i = 0; i < n; ++i) m += x[i];
    m /= (double)n;
    for (size_t i = 0; i < n; ++i) x[i] -= m;
    if (iters) *iters = it;
    free(r); free(z); free(p); free(q);
    return 0;
}
static void wls_rhs_d(const double *w, const double *phi, int nrow, int ncol, double *b, double *dinv) {
    const double *row_w = w, *col_w = w + (size_t)(nrow - 1) * ncol;
    for (int r = 0; r < nrow; ++r)
        for (int c = 0; c < ncol; ++c) {
            size_t i = (size_t)r * ncol + c;
            double pa = phi[i], bv = 0.0, ws = 0.0;
            if (r < nrow - 1) { double a = row_w[i]; if (a > 0.0) { bv -= a * wrap_pi_d(phi[i + ncol] - pa); ws += a; } }
            if (r > 0) { double a = row_w[i - ncol]; if (a > 0.0) { bv += a * wrap_pi_d(pa - phi[i - ncol]); ws += a; } }
            if (c < ncol - 1) { double a = col_w[(size_t)r * (ncol - 1) + c]; if (a > 0.0) { bv -= a * wrap_pi_d(phi[i + 1] - pa); ws += a; } }
            if (c > 0) { double a = col_w[(size_t)r * (ncol - 1) + c - 1]; if (a > 0.0) { bv += a * wrap_pi_d(pa - phi[i - 1]); ws += a; } }
            b[i] = bv; dinv[i] = 1.0 / (ws + O_EPS_D);
        }
}
int o_wls_solve_f64(const double *w, const double *phi, int nrow, int ncol, int max_iter, double tol,
                    double *u, int *iters) {
    size_t n = (size_t)nrow * ncol;
    double *b = malloc(n * sizeof(double)), *dinv = malloc(n * sizeof(double));
    wls_rhs_d(w, phi, nrow, ncol, b, dinv);
    pcg_fixed_d(w, b, dinv, nrow, ncol, max_iter, tol, u, iters);
    free(b); free(dinv);
    return 0;
}
int o_wls_adjoint_f64(const double *w, const double *phi, const double *u, const double *gu,
                      int nrow, int ncol, int max_iter, double tol, double *gw, double *lambda) {
    size_t n = (size_t)nrow * ncol, nrowarc = (size_t)(nrow - 1) * ncol,
           narcs = nrowarc + (size_t)nrow * (ncol - 1);
    double *b = malloc(n * sizeof(double)), *dinv = malloc(n * sizeof(double)),
           *lam = malloc(n * sizeof(double));
    wls_rhs_d(w, phi, nrow, ncol, b, dinv);
    double m = 0.0;
    for (size_t i = 0; i < n; ++i) m += gu[i];
    m /= (double)n;
    for (size_t i = 0; i < n; ++i) b[i] = gu[i] - m;
    pcg_fixed_d(w, b, dinv, nrow, ncol, max_iter, tol, lam, NULL);
    for (size_t i = 0; i < narcs; ++i) {
        size_t t, h;
        if (i < nrowarc) { t = i; h = i + ncol; }
        else { size_t j = i - nrowarc, rr = j / (ncol - 1), cc = j % (ncol - 1); t = rr * ncol + cc; h = t + 1; }
        double g = wrap_pi_d(phi[h] - phi[t]);
        gw[i] = -(lam[h] - lam[t]) * ((u[h] - u[t]) - g);
    }
    if (lambda) memcpy(lambda, lam, n * sizeof(double));
    free(b); free(dinv); free(lam);
    return 0;
}

/* ── post-passes (reference cuphu_solver.cpp:1385-1657) ─────────────────── */

static int cmp_float(const void *a, const void *b) {
    float x = *(const float *)a, y = *(const float *)b;
    return (x < y) ? -1 : (x > y) ? 1 : 0;
}
static int cmp_double(const void *a, const void *b) {
    double x = *(const double *)a, y = *(const double *)b;
    return (x < y) ? -1 : (x > y) ? 1 : 0;
}

/* Isolated whole-cycle row (by_row) or column spikes (:1385-1434): the
 * median of a line's valid pixels (the element of rank count/2, only for
 * count > 20) is compared with its two neighbours' medians, lines in order;
 * a line whose offsets from both round to the same nonzero cycle count,
 * each within 0.5 rad of it, is shifted back on its valid pixels and its
 * median updated before the next line is examined. */
static void fix_spikes_1d(float *unw, const unsigned char *mask, int nrow, int ncol, int by_row) {
    const double two_pi = 2.0 * M_PI, tol = 0.5;
    int nl = by_row ? nrow : ncol, len = by_row ? ncol : nrow;
    double *med = (double *)malloc(sizeof(double) * (size_t)(nl > 0 ? nl : 1));
    unsigned char *have = (unsigned char *)calloc((size_t)(nl > 0 ? nl : 1), 1);
    float *buf = (float *)malloc(sizeof(float) * (size_t)(len > 0 ? len : 1));
    for (int i = 0; i < nl; ++i) {
        int n = 0;
        for (int j = 0; j < len; ++j) {
            size_t g = by_row ? (size_t)i * ncol + j : (size_t)j * ncol + i;
            if (!mask || mask[g]) buf[n++] = unw[g];
        }
        med[i] = 0.0;
        if (n > 20) {
            qsort(buf, (size_t)n, sizeof(float), cmp_float);
            med[i] = (double)buf[n / 2];
            have[i] = 1;
        }
    }
    for (int i = 1; i < nl - 1; ++i) {
        if (!have[i] || !have[i - 1] || !have[i + 1]) continue;
        double up = med[i] - med[i - 1], down = med[i] - med[i + 1];
        double nu = round(up / two_pi), nd = round(down / two_pi);
        if (nu == 0.0 || nu != nd) continue;
        if (fabs(up - nu * two_pi) > tol || fabs(down - nd * two_pi) > tol) continue;
        float c = (float)(nu * two_pi);
        for (int j = 0; j < len; ++j) {
            size_t g = by_row ? (size_t)i * ncol + j : (size_t)j * ncol + i;
            if (!mask || mask[g]) unw[g] -= c;
        }
        med[i] -= c;
    }
    free(med); free(have); free(buf);
}

/* :1436-1441: rows, then columns */
void o_fix_cycle_spikes(float *unw, const unsigned char *mask, int nrow, int ncol) {
    fix_spikes_1d(unw, mask, nrow, ncol, 1);
    fix_spikes_1d(unw, mask, nrow, ncol, 0);
}

/* Moving median over a boundary profile (:1492-1507): for every measured
 * entry, the element of rank m/2 among the m measured entries within
 * +-w/2; unmeasured entries become 0. */
static void nf_moving_median(double *x, const unsigned char *have, int n, int w) {
    double *out = (double *)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
    double *seg = (double *)malloc(sizeof(double) * (size_t)(w + 1));
    int half = w / 2;
    for (int i = 0; i < n; ++i) {
        if (!have[i]) continue;
        int lo = i - half < 0 ? 0 : i - half, hi = i + half + 1 > n ? n : i + half + 1, m = 0;
        for (int j = lo; j < hi; ++j)
            if (have[j]) seg[m++] = x[j];
        qsort(seg, (size_t)m, sizeof(double), cmp_double);
        out[i] = seg[m / 2];
    }
    memcpy(x, out, sizeof(double) * (size_t)n);
    free(out); free(seg);
}

/* Gap fill with bounded reach (:1541-1583): linear interpolation between
 * measured entries at most 2*max_fill apart; otherwise, and before the first
 * / after the last measurement, a linear fade to 0 over max_fill entries. */
static void nf_fill_gaps(double *x, unsigned char *have, int n, int max_fill) {
    int *idx = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    int ni = 0;
    for (int i = 0; i < n; ++i)
        if (have[i]) idx[ni++] = i;
    if (ni == 0) { free(idx); return; }
    double *out = (double *)calloc((size_t)n, sizeof(double));
    unsigned char *oh = (unsigned char *)calloc((size_t)n, 1);
#define NF_FADE(src, i) do { int d_ = abs((i) - (src)); double w_ = 1.0 - (double)d_ / (double)max_fill; \
                             out[i] = x[src] * w_; oh[i] = 1; } while (0)
    int f = idx[0], l = idx[ni - 1];
    for (int i = (f - max_fill > 0 ? f - max_fill : 0); i < f; ++i) NF_FADE(f, i);
    for (int i = l + 1; i < (l + 1 + max_fill < n ? l + 1 + max_fill : n); ++i) NF_FADE(l, i);
    for (int k = 0; k + 1 < ni; ++k) {
        int a = idx[k], b = idx[k + 1], gap = b - a;
        double va = x[a], vb = x[b];
        if (gap <= 2 * max_fill) {
            for (int i = a + 1; i < b; ++i) {
                double t = (double)(i - a) / (double)gap;
                out[i] = va + t * (vb - va);
                oh[i] = 1;
            }
        } else {
            for (int i = a + 1; i <= a + max_fill; ++i) NF_FADE(a, i);
            for (int i = b - max_fill; i < b; ++i) NF_FADE(b, i);
        }
    }
#undef NF_FADE
    for (int k = 0; k < ni; ++k) { out[idx[k]] = x[idx[k]]; oh[idx[k]] = 1; }
    memcpy(x, out, sizeof(double) * (size_t)n);
    memcpy(have, oh, (size_t)n);
    free(idx); free(out); free(oh);
}

/* One internal boundary at column bc (vertical seam; by_col = 1) or row bc:
 * profile diff = unw[just before] - unw[at] where both are valid, moving
 * median (21), bounded gap fill (42), then unw += diff * (1 - d/feather) on
 * valid pixels d = 0..feather-1 past the seam (:1585-1649). */
static void nf_boundary(float *unw, const unsigned char *mask, int nrow, int ncol, int bc, int by_col,
                        int feather) {
    const int smooth = 21, max_fill = 2 * smooth;
    int n = by_col ? nrow : ncol, lim = by_col ? ncol : nrow;
    if (bc - 1 < 0 || bc + 1 >= lim) return;
    double *diff = (double *)calloc((size_t)n, sizeof(double));
    unsigned char *have = (unsigned char *)calloc((size_t)n, 1);
    for (int i = 0; i < n; ++i) {
        size_t lp = by_col ? (size_t)i * ncol + (bc - 1) : (size_t)(bc - 1) * ncol + i;
        size_t rp = by_col ? (size_t)i * ncol + bc : (size_t)bc * ncol + i;
        if ((mask && !mask[lp]) || (mask && !mask[rp])) continue;
        diff[i] = (double)unw[lp] - (double)unw[rp];
        have[i] = 1;
    }
    nf_moving_median(diff, have, n, smooth);
    nf_fill_gaps(diff, have, n, max_fill);
    int e1 = bc + feather < lim ? bc + feather : lim;
    for (int i = 0; i < n; ++i) {
        if (!have[i]) continue;
        for (int k = bc; k < e1; ++k) {
            size_t g = by_col ? (size_t)i * ncol + k : (size_t)k * ncol + i;
            if (mask && !mask[g]) continue;
            float w = 1.0f - (float)(k - bc) / (float)feather;
            unw[g] += (float)(diff[i] * w);
        }
    }
    free(diff); free(have);
}

/* Laplace neighbour-feedback refinement of internal tile seams (:1469-1657):
 * column seams first (tile c's first column + col_ovrlp/2, reference tile
 * formula :1475-1476), then row seams, each on the result of the previous. */
void o_neighbor_feedback(float *unw, const unsigned char *mask, int nrow, int ncol, int ntr, int ntc,
                         int row_ovrlp, int col_ovrlp, int feather) {
    if (feather <= 0) return;
    for (int tc = 1; tc < ntc; ++tc)
        nf_boundary(unw, mask, nrow, ncol, (int)((long)tc * (ncol - col_ovrlp) / ntc) + col_ovrlp / 2, 1, feather);
    for (int tr = 1; tr < ntr; ++tr)
        nf_boundary(unw, mask, nrow, ncol, (int)((long)tr * (nrow - row_ovrlp) / ntr) + row_ovrlp / 2, 0, feather);
}

/* ── phase bridging (reference cuphu_solver.cpp:1000-1366 and
 *    cuphu_bridge_gpu.cu) ──────────────────────────────────────────────── */

static uint32_t br_find(uint32_t *p, uint32_t x) {
    while (p[x] != x) x = p[x];
    return x;
}
static void br_union(uint32_t *p, uint32_t a, uint32_t b) {
    a = br_find(p, a);
    b = br_find(p, b);
    if (a == b) return;
    if (a < b) p[b] = a; else p[a] = b;   /* the smaller index stays root */
}

/* 8-connected regions of topo != 0 (cuphu_bridge_gpu.cu:1-20 and the
 * label entry point, ~:620-672): labels 1..N numbered by each region's
 * minimum raster index, 0 elsewhere.  Returns N. */
int o_bridge_label(const float *topo, int nrow, int ncol, uint32_t *labels) {
    size_t n = (size_t)nrow * ncol;
    uint32_t *p = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
    for (size_t i = 0; i < n; ++i) p[i] = (uint32_t)i;
    for (int r = 0; r < nrow; ++r)
        for (int c = 0; c < ncol; ++c) {
            size_t i = (size_t)r * ncol + c;
            if (topo[i] == 0.0f) continue;
            if (c + 1 < ncol && topo[i + 1] != 0.0f) br_union(p, (uint32_t)i, (uint32_t)(i + 1));
            if (r + 1 < nrow) {
                for (int dc = -1; dc <= 1; ++dc) {
                    int cc = c + dc;
                    if (cc < 0 || cc >= ncol) continue;
                    size_t j = (size_t)(r + 1) * ncol + cc;
                    if (topo[j] != 0.0f) br_union(p, (uint32_t)i, (uint32_t)j);
                }
            }
        }
    uint32_t *map = (uint32_t *)calloc(n ? n : 1, sizeof(uint32_t));
    uint32_t next = 1;
    for (size_t i = 0; i < n; ++i)
        if (topo[i] != 0.0f && br_find(p, (uint32_t)i) == i) map[i] = next++;
    for (size_t i = 0; i < n; ++i) labels[i] = (topo[i] != 0.0f) ? map[br_find(p, (uint32_t)i)] : 0u;
    free(p); free(map);
    return (int)next - 1;
}

static int br_remap(uint32_t *labels, size_t n, int N, const unsigned char *keep) {
    uint32_t *rm = (uint32_t *)calloc((size_t)N + 1, sizeof(uint32_t));
    uint32_t next = 1;
    for (int l = 1; l <= N; ++l)
        if (keep[l]) rm[l] = next++;
    for (size_t i = 0; i < n; ++i) labels[i] = rm[labels[i]];
    free(rm);
    return (int)next - 1;
}

/* regions smaller than min_num_pixel dropped, compact relabel (:1021-1032) */
int o_bridge_prune_small(uint32_t *labels, int nrow, int ncol, int N, int min_num_pixel) {
    size_t n = (size_t)nrow * ncol;
    long *cnt = (long *)calloc((size_t)N + 1, sizeof(long));
    unsigned char *keep = (unsigned char *)calloc((size_t)N + 1, 1);
    for (size_t i = 0; i < n; ++i) cnt[labels[i]]++;
    for (int l = 1; l <= N; ++l) keep[l] = cnt[l] >= min_num_pixel;
    int out = br_remap(labels, n, N, keep);
    free(cnt); free(keep);
    return out;
}

/* A region is dropped when none of its pixels survives a binary erosion by
 * the (2e+1)^2 square (circular: offsets with dr^2+dc^2 <= e^2) with the
 * outside counted as background; kept regions keep their full shape
 * (cuphu_bridge_gpu.cu:143-262). */
int o_bridge_erode_prune(uint32_t *labels, int nrow, int ncol, int N, int e, int circular) {
    if (e <= 0 || N == 0) return N;
    size_t n = (size_t)nrow * ncol;
    unsigned char *keep = (unsigned char *)calloc((size_t)N + 1, 1);
    for (int r = 0; r < nrow; ++r)
        for (int c = 0; c < ncol; ++c) {
            uint32_t l = labels[(size_t)r * ncol + c];
            if (!l || keep[l]) continue;
            int ok = 1;
            for (int dr = -e; dr <= e && ok; ++dr)
                for (int dc = -e; dc <= e; ++dc) {
                    if (circular && dr * dr + dc * dc > e * e) continue;
                    int rr = r + dr, cc = c + dc;
                    if (rr < 0 || rr >= nrow || cc < 0 || cc >= ncol || !labels[(size_t)rr * ncol + cc]) { ok = 0; break; }
                }
            if (ok) keep[l] = 1;
        }
    int out = br_remap(labels, n, N, keep);
    free(keep);
    return out;
}

/* boundary = labelled pixel whose 3x3 neighbourhood (outside = itself)
 * holds more than one label value (cuphu_bridge_gpu.cu:264-306) */
void o_bridge_boundary(const uint32_t *labels, int nrow, int ncol, unsigned char *b) {
    for (int r = 0; r < nrow; ++r)
        for (int c = 0; c < ncol; ++c) {
            uint32_t ctr = labels[(size_t)r * ncol + c], mx = ctr, mn = ctr;
            for (int dr = -1; dr <= 1; ++dr)
                for (int dc = -1; dc <= 1; ++dc) {
                    int rr = r + dr, cc = c + dc;
                    uint32_t v = (rr < 0 || rr >= nrow || cc < 0 || cc >= ncol) ? ctr : labels[(size_t)rr * ncol + cc];
                    if (v > mx) mx = v;
                    if (v < mn) mn = v;
                }
            b[(size_t)r * ncol + c] = (mx != mn) && ctr > 0;
        }
}

/* packed search key: distance quantised to 1/64 px above the two points'
 * per-region indices, 12 bits each (cuphu_bridge_gpu.cu:338-360) */
static uint64_t br_key(float d, int lp, int lq) {
    uint64_t qd = (uint64_t)(d * 64.0f);
    return (qd << 24) | ((uint64_t)lp << 12) | (uint64_t)lq;
}

/* Nearest boundary points between every two regions.  Each region's
 * boundary pixels in raster order, every max(1, m/max_samples)-th one up to
 * max_samples (<= 4096); for every ordered region pair the minimum packed
 * key over all point pairs (the reference's grid search covers the whole
 * image, so it is this minimum, :363-600).  distmat [(N+1)^2] gets the
 * quantised distance or -1, endpoints [(N+1)^2 * 4] y0, x0, y1, x1 or -1. */
void o_bridge_nn(const uint32_t *labels, const unsigned char *bnd, int nrow, int ncol, int N, int max_samples,
                 float *distmat, int *endpoints) {
    int n1 = N + 1;
    if (max_samples > 4096) max_samples = 4096;
    for (long i = 0; i < (long)n1 * n1; ++i) {
        distmat[i] = -1.0f;
        endpoints[4 * i] = endpoints[4 * i + 1] = endpoints[4 * i + 2] = endpoints[4 * i + 3] = -1;
    }
    if (N <= 1) return;
    size_t n = (size_t)nrow * ncol;
    long *cnt = (long *)calloc((size_t)n1, sizeof(long));
    for (size_t i = 0; i < n; ++i)
        if (bnd[i] && labels[i] && (int)labels[i] <= N) cnt[labels[i]]++;
    long *start = (long *)calloc((size_t)n1 + 1, sizeof(long));
    for (int l = 1; l <= N; ++l) start[l + 1] = start[l] + cnt[l];
    int *all = (int *)malloc(sizeof(int) * (size_t)(start[n1] ? start[n1] : 1));
    long *fill = (long *)calloc((size_t)n1, sizeof(long));
    for (size_t i = 0; i < n; ++i)
        if (bnd[i] && labels[i] && (int)labels[i] <= N) { uint32_t l = labels[i]; all[start[l] + fill[l]++] = (int)i; }
    /* subsampled points per region */
    int *px = (int *)malloc(sizeof(int) * (size_t)n1 * max_samples), *py = (int *)malloc(sizeof(int) * (size_t)n1 * max_samples);
    int *np = (int *)calloc((size_t)n1, sizeof(int));
    for (int l = 1; l <= N; ++l) {
        long m = cnt[l], stride = m / max_samples > 1 ? m / max_samples : 1;
        for (long k = 0; k < m && np[l] < max_samples; k += stride) {
            int idx = all[start[l] + k];
            px[(long)l * max_samples + np[l]] = idx % ncol;
            py[(long)l * max_samples + np[l]] = idx / ncol;
            np[l]++;
        }
    }
    for (int li = 1; li <= N; ++li)
        for (int lj = 1; lj <= N; ++lj) {
            if (li == lj) continue;
            uint64_t best = ~0ull;
            for (int a = 0; a < np[li]; ++a)
                for (int b = 0; b < np[lj]; ++b) {
                    float dx = (float)(px[(long)lj * max_samples + b] - px[(long)li * max_samples + a]);
                    float dy = (float)(py[(long)lj * max_samples + b] - py[(long)li * max_samples + a]);
                    float dxx = dx * dx, dyy = dy * dy;
                    float d = sqrtf(dxx + dyy);
                    uint64_t k = br_key(d, a, b);
                    if (k < best) best = k;
                }
            if (best == ~0ull) continue;
            long t = (long)li * n1 + lj;
            int a = (int)((best >> 12) & 4095u), b = (int)(best & 4095u);
            distmat[t] = (float)(best >> 24) / 64.0f;
            endpoints[4 * t + 0] = py[(long)li * max_samples + a];
            endpoints[4 * t + 1] = px[(long)li * max_samples + a];
            endpoints[4 * t + 2] = py[(long)lj * max_samples + b];
            endpoints[4 * t + 3] = px[(long)lj * max_samples + b];
        }
    free(cnt); free(start); free(all); free(fill); free(px); free(py); free(np);
}

/* ramp basis (:1049-1067) */
static int br_basis(int rt, double row, double col, double g[6]) {
    switch (rt) {
    case 1: g[0] = row; g[1] = col; g[2] = 1; return 3;
    case 2: g[0] = row * row; g[1] = col * col; g[2] = row * col; g[3] = row; g[4] = col; g[5] = 1; return 6;
    case 3: g[0] = col; g[1] = 1; return 2;
    case 4: g[0] = row; g[1] = 1; return 2;
    case 5: g[0] = col * col; g[1] = col; g[2] = 1; return 3;
    case 6: g[0] = row * row; g[1] = row; g[2] = 1; return 3;
    default: return 0;
    }
}

/* normal equations over the reference region's pixels (every step-th row
 * and column at offset step/2 once the count exceeds max_samples, raster
 * order), Gaussian elimination with partial pivoting (:1069-1141) */
static int br_fit_ramp(const float *unw, const uint32_t *labels, int nrow, int ncol, uint32_t ref, int rt,
                       int max_samples, double *coef, int *nterm) {
    double dummy[6];
    *nterm = br_basis(rt, 0, 0, dummy);
    if (*nterm == 0) return 0;
    long count = 0;
    for (size_t i = 0; i < (size_t)nrow * ncol; ++i)
        if (labels[i] == ref && unw[i] != 0.0f) ++count;
    if (count == 0) return 0;
    int step = 1;
    if (max_samples > 0 && count > max_samples) step = (int)ceil(sqrt((double)count / max_samples));
    int half = step / 2, m = *nterm;
    double A[6][6] = {{0}}, b[6] = {0};
    for (int r = 0; r < nrow; ++r) {
        if (step > 1 && r % step != half) continue;
        for (int c = 0; c < ncol; ++c) {
            if (step > 1 && c % step != half) continue;
            size_t i = (size_t)r * ncol + c;
            if (labels[i] != ref || unw[i] == 0.0f) continue;
            double g[6];
            br_basis(rt, (double)r, (double)c, g);
            for (int a = 0; a < m; ++a) {
                b[a] += g[a] * (double)unw[i];
                for (int k = 0; k < m; ++k) A[a][k] += g[a] * g[k];
            }
        }
    }
    for (int p = 0; p < m; ++p) {
        int piv = p;
        for (int r = p + 1; r < m; ++r)
            if (fabs(A[r][p]) > fabs(A[piv][p])) piv = r;
        if (fabs(A[piv][p]) < 1e-12) return 0;
        if (piv != p) {
            for (int k = 0; k < 6; ++k) { double t = A[piv][k]; A[piv][k] = A[p][k]; A[p][k] = t; }
            double t = b[piv]; b[piv] = b[p]; b[p] = t;
        }
        for (int r = p + 1; r < m; ++r) {
            double f = A[r][p] / A[p][p];
            for (int k = p; k < m; ++k) A[r][k] -= f * A[p][k];
            b[r] -= f * b[p];
        }
    }
    for (int r = m - 1; r >= 0; --r) {
        double s = b[r];
        for (int k = r + 1; k < m; ++k) s -= A[r][k] * coef[k];
        coef[r] = s / A[r][r];
    }
    return 1;
}

static void br_apply_ramp(float *unw, int nrow, int ncol, int rt, const double *coef, int m, double sign,
                          const unsigned char *nz) {
    for (int r = 0; r < nrow; ++r)
        for (int c = 0; c < ncol; ++c) {
            size_t i = (size_t)r * ncol + c;
            if (!nz[i]) continue;
            double g[6], ramp = 0.0;
            br_basis(rt, (double)r, (double)c, g);
            for (int a = 0; a < m; ++a) ramp += coef[a] * g[a];
            unw[i] += (float)(sign * ramp);
        }
}

static int br_median(const float *unw, const uint32_t *labels, int nrow, int ncol, int cy, int cx, int rad,
                     uint32_t region, double *out) {
    int r0 = cy - rad > 0 ? cy - rad : 0, r1 = cy + rad + 1 < nrow ? cy + rad + 1 : nrow;
    int c0 = cx - rad > 0 ? cx - rad : 0, c1 = cx + rad + 1 < ncol ? cx + rad + 1 : ncol;
    if (r1 <= r0 || c1 <= c0) return 0;
    float *v = (float *)malloc(sizeof(float) * (size_t)(r1 - r0) * (c1 - c0));
    int m = 0;
    for (int r = r0; r < r1; ++r)
        for (int c = c0; c < c1; ++c)
            if (labels[(size_t)r * ncol + c] == region) v[m++] = unw[(size_t)r * ncol + c];
    if (m == 0) { free(v); return 0; }
    qsort(v, (size_t)m, sizeof(float), cmp_float);
    *out = (double)v[m / 2];
    free(v);
    return 1;
}

/* The whole pass (:1270-1353).  Regions are computed on unw with masked
 * pixels zeroed; the reference region is the largest (lowest label on a
 * tie); an optional ramp fitted on it is removed first and restored last;
 * regions are joined along a Prim minimum spanning tree of the boundary
 * distances, visited breadth-first from the reference region; each bridge
 * takes the whole-cycle jump between the two regions' window medians
 * around its endpoints.  cc receives the region labels.  Returns the
 * number of regions bridged (0 or 1: unw and cc untouched). */
int o_bridge(float *unw, uint32_t *cc, const unsigned char *mask, int nrow, int ncol, int radius,
             int min_num_pixel, int erosion_size, int max_samples, int ramp_type, int ramp_max_samples) {
    size_t n = (size_t)nrow * ncol;
    float *topo = (float *)malloc(sizeof(float) * (n ? n : 1));
    for (size_t i = 0; i < n; ++i) topo[i] = (mask && !mask[i]) ? 0.0f : unw[i];
    uint32_t *lab = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
    int N = o_bridge_label(topo, nrow, ncol, lab);
    free(topo);
    if (N > 1) N = o_bridge_prune_small(lab, nrow, ncol, N, min_num_pixel);
    if (N > 1) N = o_bridge_erode_prune(lab, nrow, ncol, N, erosion_size, 0);
    if (N > 1) N = o_bridge_erode_prune(lab, nrow, ncol, N, erosion_size, 1);
    if (N <= 1) { free(lab); return N; }
    int n1 = N + 1;
    long *cnt = (long *)calloc((size_t)n1, sizeof(long));
    for (size_t i = 0; i < n; ++i) cnt[lab[i]]++;
    int ref = 1;
    for (int l = 2; l <= N; ++l)
        if (cnt[l] > cnt[ref]) ref = l;
    double coef[6] = {0};
    int nterm = 0, ramp = 0;
    unsigned char *nz = NULL;
    if (ramp_type != 0 && br_fit_ramp(unw, lab, nrow, ncol, (uint32_t)ref, ramp_type, ramp_max_samples, coef, &nterm)) {
        ramp = 1;
        nz = (unsigned char *)malloc(n ? n : 1);
        for (size_t i = 0; i < n; ++i) nz[i] = unw[i] != 0.0f;
        br_apply_ramp(unw, nrow, ncol, ramp_type, coef, nterm, -1.0, nz);
    }
    unsigned char *bnd = (unsigned char *)malloc(n ? n : 1);
    o_bridge_boundary(lab, nrow, ncol, bnd);
    float *dm = (float *)malloc(sizeof(float) * (size_t)n1 * n1);
    int *ep = (int *)malloc(sizeof(int) * 4 * (size_t)n1 * n1);
    o_bridge_nn(lab, bnd, nrow, ncol, N, max_samples, dm, ep);
    /* Prim from ref (strict comparisons: the lowest label wins ties), then
     * BFS over the tree in edge-insertion order (:1169-1229) */
    unsigned char *in = (unsigned char *)calloc((size_t)n1, 1);
    float *bd = (float *)malloc(sizeof(float) * (size_t)n1);
    int *bp = (int *)malloc(sizeof(int) * (size_t)n1);
    int *ea = (int *)malloc(sizeof(int) * (size_t)n1), *eb = (int *)malloc(sizeof(int) * (size_t)n1);
    int ne = 0;
#define BR_D(i, j) (dm[(long)(i) * n1 + (j)] < 0 ? INFINITY : dm[(long)(i) * n1 + (j)])
    in[ref] = 1;
    for (int j = 1; j <= N; ++j) { bd[j] = (j == ref) ? INFINITY : BR_D(ref, j); bp[j] = ref; }
    for (int it = 1; it < N; ++it) {
        int u = -1;
        float ud = INFINITY;
        for (int j = 1; j <= N; ++j)
            if (!in[j] && bd[j] < ud) { ud = bd[j]; u = j; }
        if (u < 0) break;
        in[u] = 1;
        ea[ne] = bp[u]; eb[ne] = u; ne++;
        for (int j = 1; j <= N; ++j) {
            if (in[j]) continue;
            float d = BR_D(u, j);
            if (d < bd[j]) { bd[j] = d; bp[j] = u; }
        }
    }
#undef BR_D
    /* adjacency in edge order */
    int *deg = (int *)calloc((size_t)n1 + 1, sizeof(int)), *adj = (int *)malloc(sizeof(int) * (size_t)(2 * ne + 1));
    for (int k = 0; k < ne; ++k) { deg[ea[k] + 1]++; deg[eb[k] + 1]++; }
    for (int l = 0; l < n1; ++l) deg[l + 1] += deg[l];
    int *pos = (int *)calloc((size_t)n1, sizeof(int));
    for (int k = 0; k < ne; ++k) {
        adj[deg[ea[k]] + pos[ea[k]]++] = eb[k];
        adj[deg[eb[k]] + pos[eb[k]]++] = ea[k];
    }
    int rad = radius < (nrow < ncol ? nrow : ncol) / 2 ? radius : (nrow < ncol ? nrow : ncol) / 2;
    double *corr = (double *)calloc((size_t)n1, sizeof(double));
    unsigned char *vis = (unsigned char *)calloc((size_t)n1, 1);
    int *q = (int *)malloc(sizeof(int) * (size_t)n1);
    int qh = 0, qt = 0;
    q[qt++] = ref;
    vis[ref] = 1;
    while (qh < qt) {
        int u = q[qh++];
        for (int k = deg[u]; k < deg[u + 1]; ++k) {
            int v = adj[k];
            if (vis[v]) continue;
            vis[v] = 1;
            q[qt++] = v;
            long t = (long)u * n1 + v;
            double m0, m1;
            int ok0 = br_median(unw, lab, nrow, ncol, ep[4 * t], ep[4 * t + 1], rad, (uint32_t)u, &m0);
            int ok1 = br_median(unw, lab, nrow, ncol, ep[4 * t + 2], ep[4 * t + 3], rad, (uint32_t)v, &m1);
            if (!ok0 || !ok1) continue;
            double diff = m1 - (m0 + corr[u]);
            double nj = floor((fabs(diff) + M_PI) / (2.0 * M_PI));
            if (diff > 0) nj *= -1.0;
            corr[v] = nj * (2.0 * M_PI);
        }
    }
    for (size_t i = 0; i < n; ++i) {
        uint32_t l = lab[i];
        cc[i] = l;
        if (l) unw[i] += (float)corr[l];
    }
    if (ramp) br_apply_ramp(unw, nrow, ncol, ramp_type, coef, nterm, 1.0, nz);
    free(lab); free(cnt); free(nz); free(bnd); free(dm); free(ep); free(in); free(bd); free(bp); free(ea); free(eb);
    free(deg); free(adj); free(pos); free(corr); free(vis); free(q);
    return N;
}
