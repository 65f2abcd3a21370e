/*
 * cuphu.h -- C ABI of the MI355X (gfx950) Laplace PCG/IRLS phase-unwrapping
 * library (libcuphu_hip.so).
 *
 * The structs, enums and the reference-compatible entry points keep the
 * field order, types, meanings and ownership rules of the reference header
 * (src/cxx/cuphu/cuphu.h), so a caller written against it links unchanged.
 * Differences, all deliberate:
 *   - every entry point returns an int status (0 = success) and never lets
 *     a C++ exception cross the ABI; cuphu_last_error() returns the
 *     thread-local message of the last failure;
 *   - streams are passed as an opaque cuphu_stream_t (a hipStream_t), so the
 *     header needs no HIP include;
 *   - the network-flow initialisers (MCF/MST, SNAPHU's CPU TreeSolve) are
 *     not part of this library: init 'mcf'/'mst' runs the reference's GPU
 *     stages around a host solver that the caller registers with
 *     cuphu_set_flow_solver, and returns CUPHU_ERR_UNSUPPORTED without one.
 *
 * Array conventions (same as the reference): pixels are row-major
 * [r*ncol + c]; arc arrays are flat, row arcs first ((nrow-1)*ncol entries,
 * arc [r,c] joins (r,c)->(r+1,c)), then column arcs (nrow*(ncol-1) entries,
 * arc [r,c] joins (r,c)->(r,c+1)).  "d_" pointers are device pointers.
 */
#ifndef CUPHU_H_
#define CUPHU_H_

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Opaque HIP stream handle (hipStream_t); NULL selects the null stream. */
typedef void *cuphu_stream_t;

/* ── status codes ─────────────────────────────────────────────────────────── */
enum {
    CUPHU_OK              = 0,
    CUPHU_ERR_INVALID     = -2,   /* bad argument (shape, pointer, option)   */
    CUPHU_ERR_UNSUPPORTED = -3,   /* feature not in this library (MCF/MST)   */
    CUPHU_ERR_HIP         = -4,   /* HIP runtime error                       */
    CUPHU_ERR_NOMEM       = -1,   /* host allocation failure (ref: :1704)    */
    CUPHU_ERR_INTERNAL    = -5
};

/* ── cost mode (ref cuphu.h:19-23) ────────────────────────────────────────── */
typedef enum {
    CUPHU_COST_SMOOTH = 0,
    CUPHU_COST_DEFO   = 1,
    CUPHU_COST_TOPO   = 2
} CuPhuCostMode;

/* ── init method (ref cuphu.h:26-30) ──────────────────────────────────────── */
typedef enum {
    CUPHU_INIT_MST     = 0,
    CUPHU_INIT_MCF     = 1,
    CUPHU_INIT_LAPLACE = 2
} CuPhuInitMethod;

/* ── run-time parameters (ref cuphu.h:33-92; identical layout) ────────────── */
typedef struct CuPhuParams {
    double nlooks;
    double ncorrlooks;
    double defothreshfactor;
    double rhosconst1;
    double rhosconst2;
    double cstd1, cstd2, cstd3;
    double sigsqcorr;
    double costscale;
    double nshortcycle;
    long   sigsqshortmin;

    double orbitradius;
    double altitude;
    double earthradius;
    double baseline;
    double baselineangle;
    double bperp;
    int    transmitmode;
    long   nlooksrange;
    long   nlooksaz;
    long   nlooksother;
    long   ncorrlooksrange;
    long   ncorrlooksaz;
    double nearrange;
    double dr, da;
    double rangeres, azres;
    double lambda;

    double kds, specularexp, dzrcritfactor;
    int    shadow;
    double dzeimin;
    long   laywidth;
    double layminei, sloperatiofactor, sigsqei;

    double dzlaypeak, azdzfactor, dzeifactor, dzeiweight;
    double dzlayfactor, layconst, layfalloffconst, sigsqlayfactor;

    double defoazdzfactor, defomax, defolayconst;

    int    kperpdpsi, kpardpsi;
    double p;
    double maxcost;
    double costscaleambight;
    double dnomincangle;
    double initdzr, initdzstep, threshold;
    long   initmaxflow, arcmaxflowconst, maxflow, nshortcycleL;
    long   cs2scalefactor;

    double minconncompfrac;
    long   conncompthresh;
    long   maxncomps;
} CuPhuParams;

/* ── tile geometry (ref cuphu.h:95-138; identical layout) ─────────────────── */
typedef struct CuPhuTileParams {
    int ntilerow, ntilecol;
    int rowovrlp, colovrlp;
    int tilecostthresh;
    int minregionsize;
    int nproc;
    int ngpustreams;
    int single_tile_reoptimize;
    int laplace_neighbor_feedback;
    int laplace_neighbor_feedback_feather;
    int fix_cycle_spikes;
} CuPhuTileParams;

/* ── phase-bridging parameters (ref cuphu.h:142-162) ─────────────────────── */
typedef enum {
    CUPHU_RAMP_NONE = 0,
    CUPHU_RAMP_LINEAR,
    CUPHU_RAMP_QUADRATIC,
    CUPHU_RAMP_LINEAR_RANGE,
    CUPHU_RAMP_LINEAR_AZIMUTH,
    CUPHU_RAMP_QUADRATIC_RANGE,
    CUPHU_RAMP_QUADRATIC_AZIMUTH
} CuPhuRampType;

typedef struct CuPhuBridgeParams {
    int enabled;
    int radius;
    int min_num_pixel;
    int erosion_size;
    int max_boundary_samples;
    CuPhuRampType ramp_type;
    int ramp_max_num_sample;
} CuPhuBridgeParams;

/* ── result handle (ref cuphu.h:165-169) ──────────────────────────────────── */
typedef struct CuPhuResult {
    float    *unw;        /* nrow*ncol, malloc'd; caller frees with free() */
    uint32_t *conncomp;   /* nrow*ncol, malloc'd; caller frees with free() */
    int       nrow, ncol;
} CuPhuResult;

/* ── SNAPHU cost element types (same layout as SNAPHU's smoothcostT/costT) ── */
typedef struct CuPhuSmoothCost { short offset; short sigsq; } CuPhuSmoothCost;
typedef struct CuPhuCost { short offset; short sigsq; short dzmax; short laycost; } CuPhuCost;
#ifndef CUPHU_NO_SNAPHU_TYPE_ALIASES
typedef CuPhuSmoothCost smoothcostT;
typedef CuPhuCost       costT;
#endif

/* Sentinel for a masked arc's sigsq (SNAPHU LARGESHORT; ref cuphu_laplace.cu:62). */
#define CUPHU_LARGESHORT 32000

/* ════════════════════════════════════════════════════════════════════════════
 * Reference-compatible entry points
 * ════════════════════════════════════════════════════════════════════════════ */

/* Defaults (ref cuphu.h:177-179; values ref cuphu_solver.cpp:148-249). */
void cuphu_default_params(CuPhuParams *p);
void cuphu_default_tile_params(CuPhuTileParams *tp);
void cuphu_default_bridge_params(CuPhuBridgeParams *bp);

/* Full pipeline on host arrays (ref cuphu.h:218-234, cuphu_solver.cpp:1674).
 * init_meth CUPHU_INIT_LAPLACE: cost modes smooth, defo and topo (topo =
 * defo kernels, ref cuphu_solver.cpp:441-450) give identical sigsq.
 * Multi-tile requests use the reference tile formula
 * (cuphu_solver.cpp:1759-1766) and stitching (:1846-2031); all tiles are
 * solved in one batched launch sequence.
 * init_meth CUPHU_INIT_MCF / CUPHU_INIT_MST: the reference's GPU stages
 * (costs, incremental costs, flow integration, gated conncomp) around the
 * host solver registered with cuphu_set_flow_solver, per tile, the tiles
 * stitched by whole 2 pi cycles (cuphu_solver.cpp:1928) with their labels
 * concatenated; CUPHU_ERR_UNSUPPORTED without a solver.
 * Post-passes, as in the reference (:2037-2098): laplace_neighbor_feedback
 * (Laplace) and fix_cycle_spikes run on multi-tile results (the single-tile
 * path skips them, :1714-1723); bridge->enabled runs on every result.
 * single_tile_reoptimize is a full-scene CPU TreeSolve: with more than one
 * tile it returns CUPHU_ERR_UNSUPPORTED, on one tile it has nothing to do. */
int cuphu_unwrap(const float *igram_r, const float *igram_i, const float *corr,
                 const float *mag, const unsigned char *mask,
                 int nrow, int ncol,
                 CuPhuCostMode cost_mode, CuPhuInitMethod init_meth,
                 const CuPhuParams *params, const CuPhuTileParams *tile,
                 const CuPhuBridgeParams *bridge,
                 const unsigned char *orig_mask,
                 int gpu_id, CuPhuResult *result);

/* Cost arrays on host arrays (ref cuphu.h:243-256, cuphu_solver.cpp:2103).
 * *costs_out is malloc'd (caller frees with free()); flat row-then-col
 * CuPhuSmoothCost (smooth) or CuPhuCost (defo/topo).  Returns the element
 * size in bytes (4 or 8) on success, a negative status on failure. */
int cuphu_build_costs_gpu(const float *igram_r, const float *igram_i,
                          const float *corr, const float *mag,
                          const short *weights,
                          int nrow, int ncol, CuPhuCostMode cost_mode,
                          const CuPhuParams *params, int gpu_id,
                          void **costs_out, size_t *cost_elem_sz);

/* Integer flows -> unwrapped phase (ref cuphu.h:261-270, cuphu_phase.cu:109).
 * Device pointers: d_phase [nrow*ncol], d_hflows row-arc flows
 * [(nrow-1)*ncol], d_vflows column-arc flows [nrow*(ncol-1)], d_unw out. */
int cuphu_integrate_phase_gpu(const float *d_phase, const short *d_hflows,
                              const short *d_vflows, int nrow, int ncol,
                              int gpu_id, float *d_unw);

/* Connected components (ref cuphu.h:275-289, cuphu_conncomp.cu:289).
 * Device pointers.  d_poscost/d_negcost: NULL for the Laplace criterion
 * (every arc confident), or both set (flat arc layout, from
 * cuphu_incrcost_gpu at the solved flows) to join only arcs with
 * min(poscost, negcost) >= params->conncompthresh. */
int cuphu_conncomp_gpu(const float *d_unw, const float *d_corr,
                       const unsigned char *d_mask, int nrow, int ncol,
                       const short *d_poscost, const short *d_negcost,
                       const CuPhuParams *params, int gpu_id,
                       uint32_t *d_labels_out);

/* Incremental arc costs at the given flows (ref cuphu_incrcost.cu:146-191,
 * cuphu_incrcost_early_exit): SNAPHU CalcCostSmooth of every arc for a
 * +-nflow flow change, into caller-allocated d_poscost/d_negcost [narcs]
 * (flat row-arc-then-column-arc layout, like d_costs and d_flows).
 * *no_improvement (nullable) = 1 when min(poscost) >= 0 and
 * min(negcost) >= 0: no unit augmentation lowers the cost, so the caller may
 * skip its network-flow solve.  Smooth costs only, as in the reference. */
int cuphu_incrcost_gpu(const CuPhuSmoothCost *d_costs, const short *d_flows,
                       int narcs, long nshortcycle, int nflow,
                       short *d_poscost, short *d_negcost, int *no_improvement,
                       cuphu_stream_t stream);

/* Host network-flow solver for init 'mcf'/'mst'.  The reference links
 * SNAPHU's CPU MCFInitFlows/MSTInitFlows and TreeSolve here
 * (cuphu_solver.cpp:512-880); that code is not part of this library, so a
 * caller registers a solver.  costs: host array of narcs CuPhuSmoothCost
 * (cost_mode smooth) or CuPhuCost (defo/topo), flat row-arc-then-column-arc;
 * phase: wrapped phase [nrow*ncol]; init_method: CUPHU_INIT_MCF or
 * CUPHU_INIT_MST; flows_out: narcs shorts, same arc layout.  Return 0 on
 * success.  With a solver registered, cuphu_unwrap/cuphu_unwrap_ex with
 * init 'mcf'/'mst' run the reference's GPU sequence around it on each tile:
 * costs -> solver -> incremental costs at the solved flows (smooth) ->
 * flow integration -> connected components gated by the incremental costs
 * (min(poscost, negcost) >= params->conncompthresh).  Without one they
 * return CUPHU_ERR_UNSUPPORTED. */
typedef int (*cuphu_flow_solver_fn)(const void *costs, int cost_mode, const float *phase,
                                   int nrow, int ncol, int init_method, short *flows_out,
                                   void *user);
/* Register (fn != NULL) or clear (fn == NULL) the process-wide solver. */
int cuphu_set_flow_solver(cuphu_flow_solver_fn fn, void *user);

/* Device-pointer building blocks (ref cuphu_util.cu:187-287). */
int cuphu_wrap_phase(float *d_phase, int n, cuphu_stream_t stream);
int cuphu_phase_diffs(const float *d_phase, float *d_dpsi_rng, float *d_dpsi_az,
                      int nrow, int ncol, cuphu_stream_t stream);
int cuphu_boxcar2d(const float *d_src, float *d_dst, int nrow, int ncol,
                   int kr, int kc, cuphu_stream_t stream);

/* Device cost builders (ref cuphu_cost.cu:338-351, :417-430).  The callee
 * allocates *d_costs_out with hipMalloc; the caller frees it with hipFree. */
int cuphu_build_smooth_costs_gpu(const float *d_phase, const float *d_corr,
                                 const short *d_roww, const short *d_colw,
                                 int nrow, int ncol, const CuPhuParams *params,
                                 int kperpdpsi, int kpardpsi,
                                 CuPhuSmoothCost **d_costs_out,
                                 cuphu_stream_t stream);
int cuphu_build_defo_costs_gpu(const float *d_phase, const float *d_corr,
                               const short *d_roww, const short *d_colw,
                               int nrow, int ncol, const CuPhuParams *params,
                               int kperpdpsi, int kpardpsi,
                               CuPhuCost **d_costs_out, cuphu_stream_t stream);

/* Laplace PCG/IRLS on device arrays (ref cuphu_laplace.cu:351-363).
 * Same argument meaning; IRLS schedule from CUPHU_N_IRLS /
 * CUPHU_IRLS_DELTA_MAX / CUPHU_IRLS_DELTA_MIN like the reference. */
int cuphu_laplace_unwrap_gpu(const CuPhuSmoothCost *d_smooth_costs,
                             const float *d_phase, int nrow, int ncol,
                             double nshortcycle, int max_iter, float tol,
                             int verbose, float *d_unw, cuphu_stream_t stream);

/* ════════════════════════════════════════════════════════════════════════════
 * Additions (no counterpart in the reference)
 * ════════════════════════════════════════════════════════════════════════════ */

/* cuphu_laplace_unwrap_gpu without any device allocation.  No counterpart
 * in the reference: it is for callers that cannot allocate (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, kernels and results as cuphu_laplace_unwrap_gpu; the
 * solver's workspace is a fixed region inside the library (one per GPU,
 * held until the process ends, CUPHU_NOALLOC_MAX_ELEMENTS * 69 bytes plus
 * 4 MiB), and the launches are eager (no graph objects).  An instance may
 * have at most CUPHU_NOALLOC_MAX_ELEMENTS row-pitched elements (nrow x
 * ncol rounded up to 256 columns, FP32; e.g. 1024 x 1024); larger ones
 * return CUPHU_ERR_INVALID before anything runs.  Calls on one GPU share
 * the region, which keeps the last call's contents (the solver initialises
 * everything it reads); calls that overlapped on one GPU would race on it.
 * Within a process the library's lock serialises its calls and each call
 * returns only after its GPU work has finished. */
#define CUPHU_NOALLOC_MAX_ELEMENTS 1048576
int cuphu_laplace_unwrap_gpu_noalloc(const CuPhuSmoothCost *d_smooth_costs,
                                     const float *d_phase, int nrow, int ncol,
                                     double nshortcycle, int max_iter, float tol,
                                     int verbose, float *d_unw,
                                     cuphu_stream_t stream);

/* Thread-local message of the last failed call ("" if none). */
const char *cuphu_last_error(void);
int cuphu_gpu_count(void);
/* Writes the device name into buf (NUL-terminated). */
int cuphu_gpu_name(int device_id, char *buf, int buflen);
/* total_mem_bytes / cu_count / arch name ("gfx950") of a device. */
int cuphu_gpu_info(int device_id, long long *total_mem_bytes, int *cu_count,
                   char *arch_buf, int arch_buflen);

/* Residue (plaquette charge) extraction.  d_res is int8
 * [(nrow-1)*(ncol-1)]: res[r,c] = rint(rng[r,c] - az[r,c+1] - rng[r+1,c]
 * + az[r,c]) with the wrapped differences (in cycles) of
 * cuphu_phase_diffs, summed left to right in FP32. */
int cuphu_residues_gpu(const float *d_phase, int nrow, int ncol,
                       signed char *d_res, cuphu_stream_t stream);

/* PCG/IRLS options.  cuphu_default_laplace_options() fills the reference
 * defaults (N_IRLS 16, delta 0.5 -> 0.01, max_iter 1000, tol 1e-3) and then
 * applies the reference's environment overrides (CUPHU_N_IRLS,
 * CUPHU_IRLS_DELTA_MAX, CUPHU_IRLS_DELTA_MIN, CUPHU_LAPLACE_MAX_ITER,
 * CUPHU_LAPLACE_TOL, CUPHU_LAPLACE_VERBOSE). */
enum {
    CUPHU_SOLVER_PLAIN   = 0, /* reference launch sequence: 7 launches and 2
                                 blocking scalar reads per PCG iteration     */
    CUPHU_SOLVER_BATCHED = 1, /* all instances at once, device-resident
                                 scalars, HIP graph chunks; the reference's
                                 recurrence in two fused stencil/update
                                 kernels and two reductions per iteration    */
    CUPHU_SOLVER_FUSED   = 2  /* as BATCHED with one stencil kernel and one
                                 reduction per iteration: beta from the
                                 expansion of r.z (equal to the reference's
                                 sum up to FP rounding); default             */
};
#define CUPHU_MAX_IRLS 32
typedef struct CuPhuLaplaceOptions {
    int    n_irls;        /* IRLS re-solves after the first (ref N_IRLS)     */
    float  delta_max;     /* delta_k = max(delta_min, delta_max * 0.5^k)     */
    float  delta_min;
    int    max_iter;      /* PCG iteration cap per solve                     */
    double tol;           /* stop when r.z < tol^2 * (r.z)_0                 */
    int    solver;        /* CUPHU_SOLVER_PLAIN / _BATCHED / _FUSED          */
    int    precision;     /* 32: FP32 vectors, FP64 reductions; 64: all FP64 */
    int    rr_interval;   /* batched: true-residual replacement period (50)  */
    int    use_graph;     /* batched: capture iteration chunks in hipGraphs  */
    int    group_bytes_mb;/* batched: solve instances in groups whose working
                             set fits this many MB (0 = all at once)         */
    int    verbose;
    int    autotune;      /* batched/fused: time the stencil tile heights (8,
                             16, 32 rows) on the first solve of a shape class
                             and use the fastest (0 = fixed 16 rows).  The
                             tiling orders the FP64 partial sums, so results
                             may then differ in the last bits between
                             processes; the default keeps them reproducible */
    int    stitch_integer;/* tiled scenes: round each tile's stitch offset to
                             whole cycles, so the scene stays congruent with
                             the wrapped phase (0 = the reference's
                             fractional Laplace offsets, :1928)             */
    int    merge_labels;  /* tiled scenes: label connected components on the
                             stitched scene, so a component crossing tile
                             seams gets one label and the size threshold is
                             minconncompfrac of the scene (0 = the
                             reference's per-tile labels offset by running
                             count, cuphu_solver.cpp:2025-2029)             */
} CuPhuLaplaceOptions;
void cuphu_default_laplace_options(CuPhuLaplaceOptions *o);

/* Per-instance solve record. */
enum {
    CUPHU_SOLVE_CONVERGED = 0, /* every solve met the tolerance             */
    CUPHU_SOLVE_MAXITER   = 1, /* at least one solve stopped at max_iter    */
    CUPHU_SOLVE_BREAKDOWN = 2, /* p.Ap <= 0 stopped a solve (ref :460)      */
    CUPHU_SOLVE_EMPTY     = 3, /* r.z == 0 at a solve start (ref :450)      */
    CUPHU_SOLVE_NONFINITE = 4  /* a reduction was not finite                */
};
typedef struct CuPhuSolveStats {
    int    status;                       /* worst status over the passes     */
    int    npass;                        /* solves run (n_irls + 1)          */
    int    iters[CUPHU_MAX_IRLS + 1];    /* PCG iterations per solve         */
    double relres[CUPHU_MAX_IRLS + 1];   /* sqrt(rz/rz0) at each solve's end */
    int    pass_status[CUPHU_MAX_IRLS + 1];
} CuPhuSolveStats;

/* One unwrapping instance (an interferogram or a tile) on the device. */
typedef struct CuPhuInstance {
    int nrow, ncol;
    const float *d_igram;          /* complex64 interleaved [nrow*ncol*2], or NULL */
    const float *d_phase;          /* wrapped phase [0,2pi) if d_igram is NULL      */
    const float *d_corr;           /* [nrow*ncol]                                   */
    const unsigned char *d_mask;   /* [nrow*ncol] or NULL                           */
    const float *d_u0;             /* optional warm start for the first solve       */
    double nlooks;                 /* per-instance looks (cost model)               */
    float *d_unw;                  /* out [nrow*ncol] (anchored, congruent)         */
    uint32_t *d_conncomp;          /* out [nrow*ncol] or NULL to skip conncomp      */
} CuPhuInstance;

/* Unwrap ninst independent device-resident instances in one launch sequence
 * (ingest -> SNAPHU sigsq -> PCG/IRLS -> anchor + congruence -> conncomp).
 * base supplies every cost/conncomp parameter except nlooks/ncorrlooks,
 * which come from each instance.  stats may be NULL, else [ninst]. */
int cuphu_laplace_batch_gpu(const CuPhuInstance *inst, int ninst,
                            CuPhuCostMode cost_mode, const CuPhuParams *base,
                            const CuPhuLaplaceOptions *opt,
                            CuPhuSolveStats *stats, cuphu_stream_t stream);

/* cuphu_unwrap with explicit options and statistics.  igram_c64 (nullable)
 * is an interleaved complex64 host array used instead of igram_r/igram_i
 * (which may then be NULL).  opt NULL = cuphu_default_laplace_options();
 * stats (nullable) receives ntilerow*ntilecol records. */
int cuphu_unwrap_ex(const float *igram_r, const float *igram_i, const float *igram_c64,
                    const float *corr, const unsigned char *mask, int nrow, int ncol,
                    CuPhuCostMode cost_mode, CuPhuInitMethod init_meth,
                    const CuPhuParams *params, const CuPhuTileParams *tile,
                    const CuPhuBridgeParams *bridge, const unsigned char *orig_mask,
                    int gpu_id, const CuPhuLaplaceOptions *opt, CuPhuSolveStats *stats,
                    CuPhuResult *result);

/* Connected components of an already-unwrapped phase on host arrays, with
 * an explicit coherence floor (the reference's regrow_conncomp,
 * src/cuphu/_conncomp.py:12-66, run on the GPU).  labels is caller-owned
 * [nrow*ncol]. */
int cuphu_regrow_conncomp(const float *unw, const float *corr, const unsigned char *mask,
                          int nrow, int ncol, double corr_thresh, double minfrac,
                          long maxncomps, int gpu_id, uint32_t *labels);

/* Whole-scene device-resident unwrap with tiling and stitching.
 * Tiling: tile_size > 0 selects fixed-size tiles of tile_size with overlap
 * `overlap` (the last tile on each axis is ragged); tile_size == 0 uses the
 * reference formula with ntilerow x ntilecol tiles.  All tiles are solved
 * by cuphu_laplace_batch_gpu-style batching, stitched by the reference's
 * coherence-weighted overlap median and BFS (fractional cycles), and merged
 * into d_unw / d_conncomp.  d_orig_mask (nullable) gates the stitch
 * samples like the reference's orig_mask.  stats may be NULL else [ntiles];
 * *ntiles_out (nullable) receives the tile count. */
int cuphu_unwrap_scene_gpu(const float *d_igram, const float *d_corr,
                           const unsigned char *d_mask,
                           const unsigned char *d_orig_mask,
                           int nrow, int ncol, double nlooks,
                           CuPhuCostMode cost_mode, const CuPhuParams *params,
                           int tile_size, int overlap,
                           int ntilerow, int ntilecol,
                           const CuPhuLaplaceOptions *opt,
                           float *d_unw, uint32_t *d_conncomp,
                           CuPhuSolveStats *stats, int *ntiles_out,
                           cuphu_stream_t stream);

/* Tile geometry on one axis.  tile_size > 0: fixed ragged tiles, starts
 * 0, step, 2*step, ... (step = tile_size - overlap) until a tile reaches n;
 * tile_size == 0: the reference formula with ntiles tiles.  Writes up to
 * max_tiles (first, size) pairs; returns the tile count or < 0. */
int cuphu_tile_axis(int n, int tile_size, int overlap, int ntiles,
                    int *first, int *size, int max_tiles);

/* Host-only stitch of solved tiles (the reference's registration, BFS and
 * merge, cuphu_solver.cpp:1846-2031).  tile_unw/tile_cc/tile_corr/
 * tile_conf are arrays of ntr*ntc host pointers to tile arrays
 * (tile_conf entries nullable).  laplace != 0 keeps fractional offsets.
 * tile_k_out (nullable) receives the per-tile cycle offsets. */
int cuphu_stitch_tiles_host(int nrow, int ncol, int ntr, int ntc,
                            const int *row_first, const int *row_size,
                            const int *col_first, const int *col_size,
                            int row_ovrlp, int col_ovrlp, int laplace,
                            const float *const *tile_unw,
                            const uint32_t *const *tile_cc,
                            const float *const *tile_corr,
                            const unsigned char *const *tile_conf,
                            float *unw_out, uint32_t *cc_out,
                            double *tile_k_out);

/* Weighted least-squares solve with fixed per-arc weights (the
 * differentiable path; no IRLS, no anchor, no congruence):
 * L(w) u = b(w, phi), u returned with its mean removed.  d_w is the flat
 * arc layout [narcs] (row arcs then column arcs); precision 32 takes float
 * arrays, 64 double arrays (d_w, d_phase, d_u all of that type). */
int cuphu_wls_solve_gpu(const void *d_w, const void *d_phase, int nrow,
                        int ncol, int precision, int max_iter, double tol,
                        void *d_u, int *iters_out, double *relres_out,
                        cuphu_stream_t stream);

/* Adjoint of cuphu_wls_solve_gpu: given u and dLoss/du (d_gu), solves
 * L(w) lambda = P d_gu (P removes the mean) with the same operator and
 * writes dLoss/dw[a] = -(lambda_h - lambda_t) * (u_h - u_t - g_a) into
 * d_gw [narcs], g_a = wrap(phi_h - phi_t).  d_lambda (nullable) receives
 * lambda. */
int cuphu_wls_adjoint_gpu(const void *d_w, const void *d_phase, const void *d_u,
                          const void *d_gu, int nrow, int ncol, int precision,
                          int max_iter, double tol, void *d_gw, void *d_lambda,
                          int *iters_out, double *relres_out,
                          cuphu_stream_t stream);

/* Live timing of the PCG kernels.  While enabled, the first two PCG
 * iterations of every iteration chunk are launched outside the chunk's HIP
 * graph, the second with HIP events around its kernels on the solver's
 * stream (it follows the first back to back, as the graph's launches do);
 * the durations and algorithmic bytes are accumulated per kernel:
 *   pq: (BATCHED) p = z + beta p, q = L p, p.q -- reads r, 1/diag, p_old and
 *       two arc weights, writes p_new and q (28 B per active pixel in FP32,
 *       56 B in FP64);
 *   xr: (BATCHED) x += alpha p, r -= alpha q, r.z -- reads x, p, r, q,
 *       1/diag, writes x and r (28 B / 56 B per active pixel);
 *   it: (FUSED) one whole iteration -- q = L p and 1/diag are recomputed;
 *       every other launch defers its x update to the next, so a pair of
 *       launches reads r, p and two arc weights twice, x and the previous p
 *       once, and writes r, p twice and x once (60 B / 120 B per active pixel
 *       per pair, 30 B / 60 B per launch on average); both launches of the
 *       pair are timed.  With CUPHU_FUSED_STORED_Q=1 the kernel that also
 *       reads and writes q, without deferral (40 B / 80 B per launch).
 * cuphu_set_profiling resets the counters. */
typedef struct CuPhuKernelProfile {
    long long samples;        /* timed iterations                          */
    double    px_sum;         /* their summed active pixels                */
    double    pq_ms_sum;      /* summed duration of the pq kernel (ms)     */
    double    pq_bytes_sum;   /* its summed algorithmic bytes              */
    double    xr_ms_sum;      /* summed duration of the xr kernel (ms)     */
    double    xr_bytes_sum;   /* its summed algorithmic bytes              */
    double    it_ms_sum;      /* summed duration of the fused kernel (ms)  */
    double    it_bytes_sum;   /* its summed algorithmic bytes              */
    long long iterations;     /* all PCG iterations launched while enabled */
} CuPhuKernelProfile;
int cuphu_set_profiling(int enabled);
int cuphu_get_profile(CuPhuKernelProfile *out);

/* ── test hooks (like the reference's private _*_test bindings,
 *    cuphu_py.cu:608-633): run single kernels on host arrays, FP32. ─────── */

/* Ingest kernel: wrapped phase (FP64 atan2 rounded once) and the SNAPHU
 * sigsq of every arc (flat row-then-col layout, mask-derived weights as
 * cuphu_solver.cpp:410-426).  igram_c64 interleaved; mask nullable. */
int cuphu_test_ingest(const float *igram_c64, const float *corr, const unsigned char *mask,
                      int nrow, int ncol, const CuPhuParams *params,
                      float *phase_out, short *sigsq_out);

/* Laplace operator kernels.  Inputs: phase, flat sigsq [narcs], x, delta.
 * Outputs (each nullable): w0 flat [narcs] (1/sigsq), b, dinv, r = b - L x
 * (assemble kernel, LDS stencil), lx = L x (plain global-gather stencil),
 * wrw flat [narcs] (reweight with u = x), cong = congruence(x anchored). */
int cuphu_test_laplace_ops(const float *phase, const short *sigsq, const float *x,
                           int nrow, int ncol, float delta,
                           float *w0, float *b, float *dinv, float *r, float *lx,
                           float *wrw, float *cong);

/* One batched-path PCG iteration (FP32) through its two kernels and
 * reductions: p = r*dinv + beta*p (p = r*dinv when beta == 0), q = L p,
 * pq = p.q, alpha = gamma / pq, x += alpha p, r -= alpha q, rz = r.(r*dinv).
 * Updates x, r, p, q in place; returns pq, alpha and rz.  w is the flat arc
 * layout [narcs]; rpw = rows per wave (2, 4, 8). */
int cuphu_test_pcg_step(const float *w, const float *dinv, int nrow, int ncol,
                        double gamma, double beta, int rpw,
                        float *x, float *r, float *p, float *q,
                        double *pq, double *alpha, double *rz);

/* One fused-solver PCG iteration (FP32) through its stencil kernel:
 * x += alpha p; r -= alpha q (r = rt when rt != NULL, the residual
 * replacement); z = r*dinv; p = z + beta p (p = z when beta == 0); q = L p;
 * start != 0 skips the x, r updates and sets p = z.  Updates x, r, p, q in
 * place; sums[4] = p.q, r.z, q.(dinv r), q.(dinv q).  stored_q != 0 runs
 * the kernel that reads q from memory; 0 runs the default kernel, which
 * recomputes q = L p from p (the q passed in must then be L p) and leaves
 * q unchanged.  rpw: rows per lane group, 2, 4 or 8 on 256-column tiles, or
 * -2, -4 on the 128-column tiles of the default kernel.  dinv must be the
 * assemble kernel's 1/diag of w. */
int cuphu_test_pcg_fused_step(const float *w, const float *dinv, int nrow, int ncol,
                              double alpha, double beta, int start, const float *rt, int rpw,
                              int stored_q, float *x, float *r, float *p, float *q, double *sums);


/* ── Optional post-passes on a stitched scene ─────────────────────────────
 * Isolated whole-cycle row/column spike correction (reference
 * cuphu_solver.cpp:1385-1441): the median of each row (then each column) of
 * valid pixels is compared with its neighbours' medians, and a line offset
 * from both by the same nonzero whole number of cycles (within 0.5 rad) is
 * shifted back on its valid pixels.  d_mask may be NULL.  In place. */
int cuphu_fix_cycle_spikes_gpu(float *d_unw, const unsigned char *d_mask, int nrow, int ncol,
                               cuphu_stream_t stream);
/* Laplace tile-seam neighbour feedback (reference cuphu_solver.cpp:1469-1657):
 * seams of the reference tile formula for ntr x ntc tiles with the given
 * overlaps; each seam's per-row (per-column) mismatch profile, smoothed and
 * gap-filled, is feathered into the next feather_px pixels.  In place. */
int cuphu_neighbor_feedback_gpu(float *d_unw, const unsigned char *d_mask, int nrow, int ncol, int ntr, int ntc,
                                int row_ovrlp, int col_ovrlp, int feather_px, cuphu_stream_t stream);
/* The same two passes on host arrays, with the reference's test-hook names
 * and arguments (cuphu.h:397-419 there, which return void; these return a
 * status). */
int cuphu_laplace_neighbor_feedback_test(float *unw, const unsigned char *mask, int nrow, int ncol, int ntr,
                                         int ntc, int row_ovrlp, int col_ovrlp, int feather_px);
int cuphu_fix_cycle_spikes_test(float *unw, const unsigned char *mask, int nrow, int ncol);


/* Phase bridging (reference cuphu_solver.cpp:1000-1366, cuphu_bridge_gpu.cu;
 * isce3's bridge_unwrapped_phase()): regions of nonzero, unmasked phase
 * (8-connected, pruned by size and erosion) are joined along a minimum
 * spanning tree of their boundary distances and each shifted by the whole
 * cycles that match its neighbour's median across the bridge; d_conncomp
 * receives the region labels.  With fewer than two regions both arrays are
 * left as they are.  At most 4096 regions after pruning. */
int cuphu_bridge_gpu(float *d_unw, uint32_t *d_conncomp, const unsigned char *d_mask, int nrow, int ncol,
                     const CuPhuBridgeParams *bp, cuphu_stream_t stream);
/* The same on host arrays, with the reference's test-hook name and
 * arguments (there it returns void). */
int cuphu_bridge_apply_test(float *unw, uint32_t *conncomp, int nrow, int ncol, const CuPhuBridgeParams *bp,
                            const unsigned char *mask, int gpu_id);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* CUPHU_H_ */
