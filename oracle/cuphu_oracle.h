/*
 * cuphu_oracle.h -- CPU restatement of the reference's Laplace unwrapping
 * path, used ONLY as the checker by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py.  Nothing in the product links or calls it.
 *
 * Array conventions follow the reference: row-major pixels; flat arc arrays
 * with row arcs first ((nrow-1)*ncol, arc [r,c] = (r,c)->(r+1,c)) then
 * column arcs (nrow*(ncol-1), arc [r,c] = (r,c)->(r,c+1)).
 */
#ifndef CUPHU_ORACLE_H_
#define CUPHU_ORACLE_H_
#include <stddef.h>
#include <stdint.h>

typedef struct OCostParams {
    double ncorrlooks, rhosconst1, rhosconst2, cstd1, cstd2, cstd3;
    double defothreshfactor, sigsqcorr, costscale, nshortcycle, defomax;
    long sigsqshortmin;
} OCostParams;

/* continuous = 1 returns the anchored solution without the final
 * congruence step, so tests can compare the continuous WLS solution. */
typedef struct OLapOptions {
    int n_irls; float delta_max, delta_min; int max_iter; double tol; int continuous;
} OLapOptions;

typedef struct OLapStats {
    int npass; int iters[33]; double relres[33]; int status[33];
} OLapStats;

/* a1 */ void o_wrapped_phase(const float *re, const float *im, size_t n, int correctly_rounded, float *phi);
/* a2 */ void o_phase_diffs(const float *phi, int nrow, int ncol, float *rng, float *az);
/* a4 */ int  o_boxcar2d(const float *src, int nrow, int ncol, int kr, int kc, float *dst);
/* a5 */ void o_costs(const float *phi, const float *corr, const short *weights, int nrow, int ncol,
                      const OCostParams *p, int kperp, int kpar, int defo, short *out);
/* a3 */ void o_residues(const float *phi, int nrow, int ncol, signed char *res);
/* a6 */ void o_arc_weights(const short *sigsq, size_t narcs, size_t stride, float *w);
/* a7 */ void o_assemble(const float *w, const float *phi, int nrow, int ncol, float *b, float *diag);
/* a8 */ void o_matvec(const float *w, const float *x, int nrow, int ncol, float *y);
/* a11 */ void o_reweight(const short *sigsq, size_t stride, const float *phi, const float *u,
                          float delta, int nrow, int ncol, float *w);
/* a12 */ void o_congruence(const float *phi, float *u, size_t n);
/* a10 */ int o_pcg_irls_f32(const short *sigsq, size_t stride, const float *phi, int nrow, int ncol,
                             const OLapOptions *o, const float *u0, float *unw, OLapStats *st);
/* a10 */ int o_pcg_irls_f64(const short *sigsq, size_t stride, const float *phi, int nrow, int ncol,
                             const OLapOptions *o, const float *u0, double *unw, OLapStats *st);
/* one Jacobi-PCG iteration in the batched solver's order, FP32 vectors */
int o_pcg_step_f32(const float *w, const float *dinv, int nrow, int ncol, double gamma, double beta,
                   double alpha_in, float *x, float *r, float *p, float *q, double *pq, double *alpha,
                   double *rz);
/* one fused PCG iteration (the batched solver's one-kernel form), FP32 vectors */
void o_pcg_fused_step_f32(const float *w, const float *dinv, int nrow, int ncol, double alpha, double beta,
                          int start, const float *rt, float *x, float *r, float *p, float *q, double *sums);
/* MCF path: incremental arc costs at the given flows; returns the early-exit flag */
int o_incrcost_smooth(const short *costs, const short *flows, size_t narcs, long nshortcycle, int nflow,
                      short *poscost, short *negcost);
/* a14 with the MCF path's incremental-cost arc gate (costs NULL = no gate) */
int o_conncomp_gated(const float *unw, const float *corr, const unsigned char *mask, int nrow, int ncol,
                     double minfrac, long maxncomps, const short *poscost, const short *negcost,
                     long conncompthresh, uint32_t *labels);
/* a14 */ int o_conncomp(const float *unw, const float *corr, const unsigned char *mask, int nrow,
                         int ncol, double minfrac, long maxncomps, uint32_t *labels);
/* a13 */ void o_integrate_phase(const float *phi, const short *hflows, const short *vflows,
                                 int nrow, int ncol, float *unw);
/* a15 */ int o_tile_axis(int n, int tile_size, int overlap, int ntiles, int *first, int *size, int max_tiles);
/* a15 */ int o_stitch(int nrow, int ncol, int ntr, int ntc, const int *rf, const int *rs, const int *cf,
                       const int *cs, int rovl, int covl, int laplace, const float *const *tunw,
                       const uint32_t *const *tcc, const float *const *tcorr,
                       const unsigned char *const *tconf, float *unw, uint32_t *cc, double *tile_k);
/* differentiable WLS (fixed weights, FP64) */
int o_wls_solve_f64(const double *w, const double *phi, int nrow, int ncol, int max_iter, double tol,
                    double *u, int *iters);
int o_wls_adjoint_f64(const double *w, const double *phi, const double *u, const double *gu,
                      int nrow, int ncol, int max_iter, double tol, double *gw, double *lambda);
/* post-passes: isolated whole-cycle row/column spikes; tile-seam neighbour feedback */
void o_fix_cycle_spikes(float *unw, const unsigned char *mask, int nrow, int ncol);
void o_neighbor_feedback(float *unw, const unsigned char *mask, int nrow, int ncol, int ntr, int ntc,
                         int row_ovrlp, int col_ovrlp, int feather);
/* phase bridging and its stages */
int o_bridge_label(const float *topo, int nrow, int ncol, uint32_t *labels);
int o_bridge_prune_small(uint32_t *labels, int nrow, int ncol, int N, int min_num_pixel);
int o_bridge_erode_prune(uint32_t *labels, int nrow, int ncol, int N, int e, int circular);
void o_bridge_boundary(const uint32_t *labels, int nrow, int ncol, unsigned char *b);
void o_bridge_nn(const uint32_t *labels, const unsigned char *bnd, int nrow, int ncol, int N, int max_samples,
                 float *distmat, int *endpoints);
int o_bridge(float *unw, uint32_t *cc, const unsigned char *mask, int nrow, int ncol, int radius,
             int min_num_pixel, int erosion_size, int max_samples, int ramp_type, int ramp_max_samples);
void o_set_threads(int n);
#endif
