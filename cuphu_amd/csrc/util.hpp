// util.hpp -- launchers of util_kernels.hip and conncomp_kernels.hip.
#pragma once
#include "common.hpp"

namespace cuphu {

void util_wrap_phase(float *d, size_t n, hipStream_t st);
void phase_from_igram(const float2 *ig, float *ph, size_t n, hipStream_t st);
void util_phase_diffs(const float *ph, float *rng, float *az, int nrow, int ncol, hipStream_t st);
// tmp must hold (nrow + 2*kr) * ncol floats.
void util_boxcar(const float *src, float *dst, float *tmp, int nrow, int ncol, int kr, int kc, hipStream_t st);
// out: flat row-then-col arcs, 2 shorts per arc (smooth) or 4 (defo).
void util_costs(const float *ph, const float *corr, const short *weights, int nrow, int ncol, const CostConsts &k,
                int kperp, int kpar, int defo, short *out, hipStream_t st);
void util_residues(const float *ph, int nrow, int ncol, signed char *res, hipStream_t st);
void util_integrate(const float *ph, const short *hflows, const short *vflows, int nrow, int ncol, float *unw,
                    hipStream_t st);

// Connected components of one image (stride = elements per row of the unw,
// corr, mask and label arrays).  Returns the number of labels assigned
// (labels are 1..K).
// poscost/negcost (flat arc layout, both or neither): the MCF path's arc gate
// min(poscost, negcost) >= conncompthresh.
// scr (nullable): device scratch of a caller that keeps it between calls --
// cnt, flags, rank of n >= nrow * stride entries each, tmp of tmp_bytes >=
// conncomp_scan_bytes(n) and one int; without it the call allocates its own.
struct CcScratch {
    uint32_t *cnt, *flags, *rank;
    unsigned char *tmp;
    size_t tmp_bytes;
    int *count;
    size_t n;
};
size_t conncomp_scan_bytes(size_t n);
int conncomp_run(const float *unw, const float *corr, const unsigned char *mask, int nrow, int ncol, int stride,
                 double minfrac, long maxncomps, uint32_t *labels, hipStream_t st, float corr_thresh = 0.05f,
                 const short *poscost = nullptr, const short *negcost = nullptr, long conncompthresh = 0,
                 const CcScratch *scr = nullptr);

// Incremental costs of smooth-cost arcs at the given flows (SNAPHU
// CalcCostSmooth for a +-nflow change; costs are {offset, sigsq} pairs).
// mins: 2 device ints of scratch.  Returns 1 when min(poscost) >= 0 and
// min(negcost) >= 0 (no improving unit augmentation), else 0.  Synchronous.
int util_incrcost(const short *costs, const short *flows, size_t narcs, long nshortcycle, int nflow, short *poscost,
                  short *negcost, int *mins, hipStream_t st);

// Tile-pair registration (stitch_kernels.hip).  A pair's overlap window is
// nr x nc samples; sample (r, c) compares pool element a_off + r*a_pitch + c
// (tile A) with b_off + r*b_pitch + c (tile B) and, when a confidence mask
// is given, its bytes at ca_off + r*ca_stride + c and cb_off + r*cb_stride + c.
struct PairDesc {
    long long a_off, b_off, ca_off, cb_off;
    int a_pitch, b_pitch, ca_stride, cb_stride;
    int nr, nc;
};
// Selection state per pair; value is the weighted median of A - B (0 for a
// pair without samples) once done.
struct PairSel {
    double half, below;
    uint32_t prefix;
    int done;
    float value;
    int pad_;
};
// Post-passes on a stitched scene (postpass.hip), in place on the device.
// Isolated whole-cycle row/column spike correction.
void fix_cycle_spikes_dev(float *d_unw, const unsigned char *d_mask, int nrow, int ncol, hipStream_t st);
// Tile-seam neighbour feedback: column seams (first column of the right
// tile's kept half) in order, then row seams; feather <= 0 does nothing.
void neighbor_feedback_dev(float *d_unw, const unsigned char *d_mask, int nrow, int ncol, const std::vector<int> &col_seams,
                           const std::vector<int> &row_seams, int feather, hipStream_t st);

// Phase bridging of disconnected regions (bridge.hip): unw corrected in
// place by whole cycles per region, cc set to the region labels; both left
// untouched when fewer than two regions remain.  d_mask may be NULL.
void bridge_dev(float *d_unw, uint32_t *d_cc, const unsigned char *d_mask, int nrow, int ncol,
                const CuPhuBridgeParams &bp, hipStream_t st);

// Samples per histogram block = 256 * 16; chunks covering nsamples.
int pair_chunks(long long nsamples);
// Weighted medians of all pairs (4 passes of 8 bits).  d_hist holds
// npairs * 256 doubles; max_chunks = max over pairs of pair_chunks().
void pair_medians(const PairDesc *d_pairs, int npairs, int max_chunks, const float *unw, const float *corr,
                  const unsigned char *conf, PairSel *d_sel, double *d_hist, hipStream_t st);

}  // namespace cuphu
