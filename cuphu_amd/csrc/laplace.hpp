// laplace.hpp -- host-side declarations of the Laplace PCG/IRLS machinery.
#pragma once
#include "common.hpp"

namespace cuphu {

// Stencil tile geometry: a workgroup of 4 waves covers TW columns (64 lanes x
// 4 pixels) by 4*RPW rows; each wave owns RPW consecutive rows.  The fused
// solver's kernels also run a narrow shape (TileShape below).
constexpr int kTileW = 256;
constexpr int kLdsStride = 264;   // LDS row stride (elements); tile col j -> lds col 4+j
constexpr int kLdsC0 = 4;
constexpr int kBlock = 256;

// One instance (a tile or an interferogram) inside the pooled planes.
// pitch is the row stride; wcols (a multiple of 64, >= ncol) the columns
// the iteration kernel covers.  The fused solver rounds the stride up to a
// multiple of 256 elements; the columns from wcols to pitch then hold zeros
// that only the once-per-pass kernels touch.
struct InstDesc {
    int nrow, ncol, pitch, tiles_x;
    int tiles_y, block_begin, row_begin, wcols;
    int band;        // tile columns per band of the launch order (0: raster order)
    int pad_;
    long long off;   // element offset of pixel (0,0) in every pixel plane
};

// Per-instance solver scalars, device resident.
struct SolveScal {
    double gamma, delta, gamma_prev, alpha, alpha_prev, beta, gamma0, relres;
    int iters, active, status, replace;   // replace: the last residual replacement was taken
    // x update of the next fused q-free launch: 0 = defer alpha_j p_j,
    // 1 = apply the deferred alpha_{j-1} p_{j-1} (alpha_prev, p from the
    // output plane set) then alpha_j p_j, 2 = apply alpha_j p_j (no deferral)
    int xmode, pad2;
};

enum RedMode { RED_START = 0, RED_PQ = 1, RED_RZ = 2, RED_RR = 3, RED_IT = 4 };

// Instance source description for ingest (pointers already offset to the
// instance's (0,0); src_stride = elements per source row).
struct IngestSrc {
    const float2 *igram;     // complex64 or NULL
    const float *phase;      // used when igram is NULL
    const float *corr;
    const unsigned char *mask;
    const float *u0;         // warm start or NULL
    int src_stride, pad;
    CostConsts cc;
};

// Pooled per-group planes.  T = float (FP32 mode) or double (FP64 mode).
template <class T> struct Planes {
    float *phi = nullptr;           // wrapped phase (float in both modes)
    float *corr = nullptr;          // pitched coherence copy
    unsigned char *mask = nullptr;  // pitched mask copy (1 = valid)
    short2 *sigsq = nullptr;        // {down, right} arc sigsq
    T *w2 = nullptr;                // {down, right} arc weights interleaved (2 T per pixel)
    // s holds q = L p; wv is the plain solver's z.
    T *b = nullptr, *dinv = nullptr, *x = nullptr, *r = nullptr, *p = nullptr, *s = nullptr, *wv = nullptr;
    // Second plane set: a stencil kernel that reads a vector at halo pixels
    // reads it from one set and writes its update to the other, so no
    // workgroup overwrites a value a neighbouring workgroup still reads.
    // Two-kernel solver: direction p/p2.  Fused solver: (r, p, s) and
    // (r2, p2, q2); wv then holds the true residual of a replacement.
    T *p2 = nullptr, *r2 = nullptr, *q2 = nullptr;
};

// Workgroup tile of the fused solver's kernels (weights, assemble, the
// iteration, the residual monitor, the x finalize): lpr lanes per tile row,
// rpw rows per lane group.  lpr 64: 256 columns x 4*rpw rows (rpw 2, 4, 8);
// lpr 32: 128 columns x 8*rpw rows (rpw 2, 4).  Every other kernel uses
// lpr 64.
struct TileShape {
    int rpw = 4, lpr = 64;
    int th() const { return 4 * rpw * (64 / lpr); }
    int tw() const { return 4 * lpr; }
    bool operator==(const TileShape &o) const { return rpw == o.rpw && lpr == o.lpr; }
};

// ── launchers (laplace_kernels.hip) ─────────────────────────────────────────
template <class T>
void launch_ingest(const IngestSrc *d_src, const InstDesc *d_desc, const int *d_rowinst, int nrows_total,
                   const Planes<T> &pl, int has_u0, int max_pitch, hipStream_t st);
template <class T>
void launch_weights(const InstDesc *d_desc, const int *d_blk, int nblk, const Planes<T> &pl, int reweight,
                    T delta2, TileShape sh, hipStream_t st);
template <class T>
void launch_assemble(const InstDesc *d_desc, const int *d_blk, int nblk, const Planes<T> &pl, TileShape sh,
                     hipStream_t st);
// start: partial r.z (force: also for inactive instances)
template <class T>
void launch_rz(const InstDesc *d_desc, const int *d_blk, int nblk, const Planes<T> &pl, const SolveScal *scal,
               int force, double2 *partials, int rpw, hipStream_t st);
// parity 0 reads direction p and writes p2; parity 1 the reverse.
template <class T>
void launch_pq(const InstDesc *d_desc, const int *d_blk, int nblk, const Planes<T> &pl, const SolveScal *scal,
               double2 *partials, int rpw, int parity, hipStream_t st);
template <class T>
void launch_xr(const InstDesc *d_desc, const int *d_blk, int nblk, const Planes<T> &pl, const SolveScal *scal,
               double2 *partials, int rpw, int parity, hipStream_t st);
// true residual into p2 (+ partial r_t.z_t), then r := p2 where replaced
template <class T>
void launch_rr(const InstDesc *d_desc, const int *d_blk, int nblk, const Planes<T> &pl, const SolveScal *scal,
               double2 *partials, int rpw, hipStream_t st);
template <class T>
void launch_rr_commit(const InstDesc *d_desc, const int *d_blk, int nblk, const Planes<T> &pl, const SolveScal *scal,
                      int rpw, hipStream_t st);
void launch_reduce(const InstDesc *d_desc, int ninst, const double2 *partials, SolveScal *scal, int mode,
                   int max_iter, double tol2, hipStream_t st);
// Fused solver: one stencil kernel per PCG iteration (parity 0 reads plane
// set 0 and writes set 1; start = 1 forms the first direction from set 0),
// the true residual of a replacement, and the reduction (modes RED_START,
// RED_IT, RED_RR).
template <class T>
void launch_it(const InstDesc *d_desc, const int *d_blk, int nblk, const Planes<T> &pl, const SolveScal *scal,
               double4 *partials, int rpw, int parity, int start, hipStream_t st);
template <class T>
void launch_rr_it(const InstDesc *d_desc, const int *d_blk, int nblk, const Planes<T> &pl, const SolveScal *scal,
                  double4 *partials, int rpw, hipStream_t st);
// scal_in (default: scal) is the state the reduction starts from; the new
// state goes to scal.
void launch_reduce_it(const InstDesc *d_desc, int ninst, const double4 *partials, SolveScal *scal, int mode,
                      int max_iter, double tol2, hipStream_t st, int defer_x = 0,
                      const SolveScal *scal_in = nullptr);
// The same iteration without a stored q (q_j = A p_j recomputed from p_j):
// plane sets (r, p) and (r2, p2); and its true-residual kernel.
template <class T>
void launch_iq(const InstDesc *d_desc, const int *d_blk, int nblk, const Planes<T> &pl, const SolveScal *scal,
               double4 *partials, TileShape sh, int parity, int start, hipStream_t st);
// The same iteration in the merged tiling (k_iq_mix): d_tiles lists the
// launch's ntiles workgroups as {instance, tile column, tile row, kind}.
template <class T>
void launch_iq_mix(const InstDesc *d_desc, const int4 *d_tiles, int ntiles, const Planes<T> &pl,
                   const SolveScal *scal, double4 *partials, int rpw, int parity, int start, hipStream_t st);
// The same iteration with the previous launch's reduction folded in (for
// grids of a few resident rounds, where the reduction's own launch is a
// large part of an iteration): reduces part_prev in the reduction's order,
// advances scal_prev to the next state exactly as the reduction would
// (written to scal_next by the instance's first tile) and runs on it.
template <class T>
void launch_iq_fold(const InstDesc *d_desc, const int *d_blk, int nblk, const Planes<T> &pl,
                    const SolveScal *scal_prev, SolveScal *scal_next, const double4 *part_prev, double4 *partials,
                    TileShape sh, int parity, int max_iter, double tol2, hipStream_t st);
// The folded iteration in the merged tiling (k_iq_fold_mix).
template <class T>
void launch_iq_fold_mix(const InstDesc *d_desc, const int4 *d_tiles, int ntiles, const Planes<T> &pl,
                        const SolveScal *scal_prev, SolveScal *scal_next, const double4 *part_prev,
                        double4 *partials, int rpw, int parity, int max_iter, double tol2, hipStream_t st);
template <class T>
void launch_rr_iq(const InstDesc *d_desc, const int *d_blk, int nblk, const Planes<T> &pl, const SolveScal *scal,
                  double4 *partials, TileShape sh, hipStream_t st);
// After a solve: apply an x update still deferred (xmode 1) when it stopped.
template <class T>
void launch_x_finalize(const InstDesc *d_desc, const int *d_blk, int nblk, const Planes<T> &pl, const SolveScal *scal,
                       TileShape sh, hipStream_t st);
template <class T>
void launch_final(const InstDesc *d_desc, const int *d_rowinst, int nrows_total, const Planes<T> &pl,
                  float *out_pool, int max_pitch, hipStream_t st);

// Plain (reference launch sequence) kernels on one instance.
template <class T>
void launch_plain_matvec(const InstDesc &d, const T *w2, const T *x, T *y, hipStream_t st);
template <class T> void launch_plain_axpy(T *y, const T *x, T alpha, size_t n, hipStream_t st);
template <class T> void launch_plain_vmul(T *z, const T *r, const T *dinv, size_t n, hipStream_t st);
template <class T> void launch_plain_pz(T *p, T *z, const T *r, const T *dinv, size_t n, hipStream_t st);
template <class T> void launch_plain_axpby(T *p, const T *z, T beta, size_t n, hipStream_t st);
template <class T> void launch_plain_dot(const T *a, const T *b, size_t n, double *partials, double *out,
                                         hipStream_t st);
constexpr int kDotParts = 1024;

}  // namespace cuphu
