// solver.hip -- host driver of the Laplace PCG/IRLS path
// (reference: src/cxx/cuphu/cuphu_laplace.cu:351-552 driver,
// cuphu_solver.cpp:370-510 solve_tile, :1674-2101 cuphu_unwrap).
//
// Batched solver: all instances (tiles of a scene, or independent
// interferograms) share pooled pitched planes and are solved together, in
// groups whose working set is bounded by group_bytes_mb.  Per IRLS pass:
// weights -> assemble (+ r = b - L x) -> start (u, L u, r.u, Lu.u) ->
// reduce, then chunks of `rr_interval` fused PCG steps (each a tile kernel
// and a per-instance reduction kernel) closed by a true-residual
// replacement.  Each chunk is captured once in a HIP graph and replayed;
// the host reads the per-instance scalars (a few hundred bytes) once per
// chunk, the fused solver with the next chunk already issued.  Converged
// instances are masked inside the kernels.
//
// Plain solver: the reference's launch sequence for one instance at a
// time (matvec, dot, axpy, axpy, vmul, dot, axpby with two blocking scalar
// reads per iteration), on the same planes and kernels' arithmetic.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <type_traits>

#include "laplace.hpp"
#include "stitch.hpp"
#include "util.hpp"

namespace cuphu {

static thread_local std::string g_last_error;
void set_last_error(const std::string &msg) { g_last_error = msg; }

static std::recursive_mutex g_mu;   // serialises calls (shared workspace and stream)

// Host network-flow solver for init 'mcf'/'mst' (cuphu_set_flow_solver).
static cuphu_flow_solver_fn g_flow_solver = nullptr;
static void *g_flow_solver_user = nullptr;

// ── per-device cached resources ─────────────────────────────────────────────
struct DevRes {
    hipStream_t stream = nullptr;
    hipEvent_t ev_in = nullptr, ev_out = nullptr;
    hipEvent_t stage_ev[2] = {nullptr, nullptr};   // staged host copies
    std::map<std::string, DevBuf<unsigned char>> bufs;
    std::map<std::string, PinBuf<unsigned char>> pins;
    // A fixed device region (the no-allocation entry point): while `arena` is
    // set, get() hands out 256-byte aligned pieces of it in request order
    // and never allocates.
    unsigned char *arena = nullptr;
    size_t arena_cap = 0, arena_used = 0;
    std::map<std::string, std::pair<void *, size_t>> arena_bufs;
    void *get(const std::string &name, size_t bytes) {
        if (arena) {
            auto it = arena_bufs.find(name);
            if (it != arena_bufs.end() && it->second.second >= bytes) return it->second.first;
            size_t off = (arena_used + 255) & ~(size_t)255;
            if (off > arena_cap || bytes > arena_cap - off)
                throw Error(CUPHU_ERR_INVALID, "scratch region exhausted (" + std::to_string(off + bytes) + " of " +
                                                   std::to_string(arena_cap) + " bytes)");
            arena_used = off + bytes;
            arena_bufs[name] = {arena + off, bytes};
            return arena + off;
        }
        DevBuf<unsigned char> &b = bufs[name];
        if (b.n < bytes) b.alloc(bytes);
        return b.p;
    }
    void *pin(const std::string &name, size_t bytes) {
        if (arena) throw Error(CUPHU_ERR_INTERNAL, "pinned staging requested inside the no-allocation path");
        PinBuf<unsigned char> &b = pins[name];
        if (b.n < bytes) b.alloc(bytes);
        return b.p;
    }
};
// Routes every workspace request of one call to a fixed region; restores
// the allocating workspace when the call ends (normally or by exception).
struct ArenaScope {
    DevRes &r;
    ArenaScope(DevRes &res, unsigned char *base, size_t cap) : r(res) {
        r.arena = base;
        r.arena_cap = cap;
        r.arena_used = 0;
        r.arena_bufs.clear();
    }
    ~ArenaScope() {
        r.arena = nullptr;
        r.arena_cap = r.arena_used = 0;
        r.arena_bufs.clear();
    }
};
static DevRes &devres() {
    static std::map<int, DevRes> m;
    int dev = 0;
    CUPHU_HIP_CHECK(hipGetDevice(&dev));
    DevRes &r = m[dev];
    if (!r.stream) {
        CUPHU_HIP_CHECK(hipStreamCreateWithFlags(&r.stream, hipStreamNonBlocking));
        CUPHU_HIP_CHECK(hipEventCreateWithFlags(&r.ev_in, hipEventDisableTiming));
        CUPHU_HIP_CHECK(hipEventCreateWithFlags(&r.ev_out, hipEventDisableTiming));
    }
    return r;
}
template <class T> static T *ws(DevRes &r, const char *name, size_t count) {
    return reinterpret_cast<T *>(r.get(name, std::max<size_t>(count, 1) * sizeof(T) + 256));
}

// Connected-component scratch for images of up to n = nrow * stride
// entries, kept in the workspace between calls.
static CcScratch cc_scratch(DevRes &res, size_t n) {
    CcScratch scr;
    scr.n = n;
    scr.cnt = ws<uint32_t>(res, "cc_cnt", n);
    scr.flags = ws<uint32_t>(res, "cc_flags", n);
    scr.rank = ws<uint32_t>(res, "cc_rank", n);
    scr.tmp_bytes = conncomp_scan_bytes(n);
    scr.tmp = ws<unsigned char>(res, "cc_tmp", scr.tmp_bytes);
    scr.count = ws<int>(res, "cc_count", 1);
    return scr;
}

// ── host <-> device copies of caller buffers through pinned staging ────────
// Pageable device-to-host copies run at a few GB/s; staging through two
// pinned chunks runs at the link rate.  The DMA of one chunk overlaps the
// CPU copy (split over threads) of the previous one.
static void par_memcpy(void *dst, const void *src, size_t n) {
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const size_t nth = n >= ((size_t)8 << 20) ? std::min<size_t>(8, hw) : 1;
    if (nth <= 1) { std::memcpy(dst, src, n); return; }
    const size_t part = ((n + nth - 1) / nth + 4095) & ~(size_t)4095;
    std::vector<std::thread> th;
    for (size_t t = 0; t < nth && t * part < n; ++t) {
        const size_t off = t * part, len = std::min(part, n - off);
        th.emplace_back([=]() { std::memcpy((char *)dst + off, (const char *)src + off, len); });
    }
    for (auto &x : th) x.join();
}
static constexpr size_t kStageChunk = (size_t)64 << 20;
static void stage_events(DevRes &r) {
    for (auto &e : r.stage_ev)
        if (!e) CUPHU_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
}
static void copy_d2h_staged(void *dst, const void *src, size_t n, DevRes &r, hipStream_t st) {
    if (n < ((size_t)8 << 20)) {
        CUPHU_HIP_CHECK(hipMemcpyAsync(dst, src, n, hipMemcpyDeviceToHost, st));
        CUPHU_HIP_CHECK(hipStreamSynchronize(st));
        return;
    }
    stage_events(r);
    unsigned char *pb[2] = {(unsigned char *)r.pin("stage0", kStageChunk), (unsigned char *)r.pin("stage1", kStageChunk)};
    const size_t nch = (n + kStageChunk - 1) / kStageChunk;
    auto len = [&](size_t k) { return std::min(kStageChunk, n - k * kStageChunk); };
    for (size_t k = 0; k < nch; ++k) {
        // buffer k&1: the CPU copy of chunk k-2 out of it is done, and an
        // earlier host-to-device DMA from it precedes this one on the stream
        CUPHU_HIP_CHECK(hipMemcpyAsync(pb[k & 1], (const char *)src + k * kStageChunk, len(k), hipMemcpyDeviceToHost, st));
        CUPHU_HIP_CHECK(hipEventRecord(r.stage_ev[k & 1], st));
        if (k >= 1) {
            CUPHU_HIP_CHECK(hipEventSynchronize(r.stage_ev[(k - 1) & 1]));
            par_memcpy((char *)dst + (k - 1) * kStageChunk, pb[(k - 1) & 1], len(k - 1));
        }
    }
    CUPHU_HIP_CHECK(hipEventSynchronize(r.stage_ev[(nch - 1) & 1]));
    par_memcpy((char *)dst + (nch - 1) * kStageChunk, pb[(nch - 1) & 1], len(nch - 1));
}
static void copy_h2d_staged(void *dst, const void *src, size_t n, DevRes &r, hipStream_t st) {
    if (n < ((size_t)8 << 20)) {
        CUPHU_HIP_CHECK(hipMemcpyAsync(dst, src, n, hipMemcpyHostToDevice, st));
        CUPHU_HIP_CHECK(hipStreamSynchronize(st));   // the caller's buffer may change after return
        return;
    }
    stage_events(r);
    unsigned char *pb[2] = {(unsigned char *)r.pin("stage0", kStageChunk), (unsigned char *)r.pin("stage1", kStageChunk)};
    const size_t nch = (n + kStageChunk - 1) / kStageChunk;
    for (size_t k = 0; k < nch; ++k) {
        const size_t len = std::min(kStageChunk, n - k * kStageChunk);
        // buffer k&1 is free once the last DMA that used it (chunk k-2, or an
        // earlier copy's) has finished; an event never recorded is complete
        CUPHU_HIP_CHECK(hipEventSynchronize(r.stage_ev[k & 1]));
        par_memcpy(pb[k & 1], (const char *)src + k * kStageChunk, len);
        CUPHU_HIP_CHECK(hipMemcpyAsync((char *)dst + k * kStageChunk, pb[k & 1], len, hipMemcpyHostToDevice, st));
        CUPHU_HIP_CHECK(hipEventRecord(r.stage_ev[k & 1], st));
    }
    // the staging buffers are reused by the next copy, which waits on these events
}

// Orders the library's work stream after the caller's stream on entry and
// the caller's stream after the work stream on exit.
struct StreamJoin {
    hipStream_t user, work;
    DevRes &r;
    StreamJoin(DevRes &res, hipStream_t u) : user(u), work(res.stream), r(res) {
        CUPHU_HIP_CHECK(hipEventRecord(r.ev_in, user));
        CUPHU_HIP_CHECK(hipStreamWaitEvent(work, r.ev_in, 0));
    }
    ~StreamJoin() {
        if (hipEventRecord(r.ev_out, work) == hipSuccess) (void)hipStreamWaitEvent(user, r.ev_out, 0);
    }
};

// ── options ─────────────────────────────────────────────────────────────────
static void env_int(const char *k, int *v) { if (const char *e = std::getenv(k)) *v = std::atoi(e); }
static void env_float(const char *k, float *v) { if (const char *e = std::getenv(k)) *v = (float)std::atof(e); }

void default_options(CuPhuLaplaceOptions *o) {
    std::memset(o, 0, sizeof(*o));
    // cuphu_laplace.cu:385-390, cuphu_solver.cpp:468-472
    o->n_irls = 16;
    o->delta_max = 0.5f;
    o->delta_min = 0.01f;
    o->max_iter = 1000;
    o->tol = 1e-3;
    o->solver = CUPHU_SOLVER_FUSED;
    o->precision = 32;
    o->rr_interval = 50;
    o->use_graph = 1;
    o->group_bytes_mb = 0;
    env_int("CUPHU_N_IRLS", &o->n_irls);
    env_float("CUPHU_IRLS_DELTA_MAX", &o->delta_max);
    env_float("CUPHU_IRLS_DELTA_MIN", &o->delta_min);
    env_int("CUPHU_LAPLACE_MAX_ITER", &o->max_iter);
    if (const char *e = std::getenv("CUPHU_LAPLACE_TOL")) o->tol = std::atof(e);
    if (const char *e = std::getenv("CUPHU_LAPLACE_VERBOSE")) o->verbose = std::max(1, std::atoi(e));
    if (const char *e = std::getenv("CUPHU_LAPLACE_SOLVER"))
        o->solver = (std::strcmp(e, "plain") == 0)     ? CUPHU_SOLVER_PLAIN
                    : (std::strcmp(e, "batched") == 0) ? CUPHU_SOLVER_BATCHED
                                                       : CUPHU_SOLVER_FUSED;
    env_int("CUPHU_LAPLACE_PRECISION", &o->precision);
    env_int("CUPHU_LAPLACE_GROUP_MB", &o->group_bytes_mb);
    env_int("CUPHU_LAPLACE_GRAPH", &o->use_graph);
    env_int("CUPHU_LAPLACE_AUTOTUNE", &o->autotune);
}

static void check_options(const CuPhuLaplaceOptions &o) {
    CUPHU_REQUIRE(o.n_irls >= 0 && o.n_irls <= CUPHU_MAX_IRLS, "n_irls must be in [0, 32]");
    CUPHU_REQUIRE(o.precision == 32 || o.precision == 64, "precision must be 32 or 64");
    CUPHU_REQUIRE(o.solver == CUPHU_SOLVER_PLAIN || o.solver == CUPHU_SOLVER_BATCHED || o.solver == CUPHU_SOLVER_FUSED,
                  "unknown solver");
    CUPHU_REQUIRE(o.tol >= 0.0 && o.max_iter >= 0, "tol and max_iter must be >= 0");
}

// Rows per lane group of the stencil tiles: CUPHU_LAPLACE_RPW (2, 4 or 8),
// else 4 in FP32 and 2 in FP64.  The FP64 iteration kernel holds twice the
// registers (292 VGPRs at 4 rows: one workgroup per CU; about 198 at 2):
// the 8192^2 scene's FP64 solve took 6.23 s in 128 x 32 tiles and 5.80 s
// in 128 x 16 ones (256 x 8: 5.85 s).
static int rows_per_wave(int precision) {
    int rpw = 0;
    env_int("CUPHU_LAPLACE_RPW", &rpw);
    if (rpw == 2 || rpw == 4 || rpw == 8) return rpw;
    return precision == 64 ? 2 : 4;
}

// ── layout ──────────────────────────────────────────────────────────────────
struct Layout {
    std::vector<InstDesc> desc;   // global: off, row_begin (block_begin unused)
    long long total_el = 0;
    int total_rows = 0, max_pitch = 64;
};
// stride256: round each row stride up to a multiple of 256 elements (the
// fused solver; the iteration kernel still covers only wcols columns).
// Measured on k_iq, same pixels per launch, traced averages over 401
// launches: 1843-column tiles (the reference tile formula at 8192^2) 491.7 ->
// 460.0 us; 1856 476.4 -> 458.0; 1920 470.1 -> 447.0; 2600 428.1 -> 419.5;
// 1080 462.7 -> 457.9; 700 447.2 -> 443.6.  Strides that are already
// multiples of 256 are unchanged (moving them to a larger multiple was
// slower).
// Band width (tile columns) of an instance's launch order.  Instances
// wider than 4096 columns are walked in bands of equal width of at most
// 2048 columns, like 2048-wide instances, whose halo rows meet in L2 (a
// single 8192-wide instance read 5.6 % more DRAM bytes in raster order).
// Measured on k_iq, same box, alternating: one 8192-wide instance 0.390 ->
// 0.371 ms per launch; 4096-wide tiles 0.436 -> 0.438 (two bands); 2560-wide
// tiles 0.464 -> 0.478 with bands of 16 and 4 tiles, hence no bands up to
// 4096 columns.  CUPHU_LAPLACE_BANDS=0 keeps raster order.
static int band_of(int tiles_x, TileShape sh) {
    int on = 1;
    env_int("CUPHU_LAPLACE_BANDS", &on);
    const int b = 2048 / sh.tw();
    if (!on || tiles_x <= 2 * b) return 0;
    const int nbands = (tiles_x + b - 1) / b;
    return (tiles_x + nbands - 1) / nbands;
}

static Layout make_layout(const std::vector<std::pair<int, int>> &dims, TileShape sh, bool stride256 = false) {
    Layout L;
    for (auto &dm : dims) {
        InstDesc d{};
        d.nrow = dm.first;
        d.ncol = dm.second;
        d.wcols = pitch_of(d.ncol);
        d.pitch = stride256 ? (d.wcols + 255) / 256 * 256 : d.wcols;
        d.tiles_x = (int)ceil_div((size_t)d.pitch, (size_t)sh.tw());
        d.tiles_y = (int)ceil_div((size_t)d.nrow, (size_t)sh.th());
        d.band = band_of(d.tiles_x, sh);
        d.row_begin = L.total_rows;
        d.off = L.total_el;
        L.total_el += (long long)d.nrow * d.pitch;
        L.total_rows += d.nrow;
        L.max_pitch = std::max(L.max_pitch, d.pitch);
        L.desc.push_back(d);
    }
    return L;
}

template <class T> struct PoolSet {
    Planes<T> pl;
    float *out = nullptr;
    uint32_t *lab = nullptr;
};
template <class T> static PoolSet<T> alloc_pools(DevRes &r, long long n) {
    PoolSet<T> P;
    const char *sfx = sizeof(T) == 4 ? "32" : "64";
    auto nm = [&](const char *b) { return std::string(b) + sfx; };
    P.pl.phi = ws<float>(r, "phi", n);
    P.pl.corr = ws<float>(r, "corr", n);
    P.pl.mask = ws<unsigned char>(r, "mask", n);
    P.pl.sigsq = ws<short2>(r, "sigsq", n);
    P.pl.w2 = ws<T>(r, nm("w2").c_str(), 2 * n);
    P.pl.b = ws<T>(r, nm("b").c_str(), n);
    P.pl.dinv = ws<T>(r, nm("dinv").c_str(), n);
    P.pl.x = ws<T>(r, nm("x").c_str(), n);
    P.pl.r = ws<T>(r, nm("r").c_str(), n);
    P.pl.p = ws<T>(r, nm("p").c_str(), n);
    P.pl.s = ws<T>(r, nm("s").c_str(), n);
    P.pl.wv = ws<T>(r, nm("wv").c_str(), n);
    P.pl.p2 = ws<T>(r, nm("p2").c_str(), n);
    P.pl.r2 = ws<T>(r, nm("r2").c_str(), n);
    P.pl.q2 = ws<T>(r, nm("q2").c_str(), n);
    P.out = ws<float>(r, "out", n);
    P.lab = ws<uint32_t>(r, "lab", n);
    return P;
}

static void record_pass(CuPhuSolveStats *st, int pass, const SolveScal &s) {
    if (!st) return;
    st->iters[pass] = s.iters;
    st->relres[pass] = s.relres;
    st->pass_status[pass] = s.status;
    st->npass = pass + 1;
    // worst status: nonfinite > breakdown > maxiter > empty > converged
    auto rank = [](int v) {
        switch (v) {
        case CUPHU_SOLVE_NONFINITE: return 4;
        case CUPHU_SOLVE_BREAKDOWN: return 3;
        case CUPHU_SOLVE_MAXITER: return 2;
        case CUPHU_SOLVE_EMPTY: return 1;
        default: return 0;
        }
    };
    if (pass == 0 || rank(s.status) > rank(st->status)) st->status = s.status;
}

static float pass_delta(const CuPhuLaplaceOptions &o, int pass) {
    // delta used to reweight before solve `pass` (after solve pass-1),
    // cuphu_laplace.cu:516-522.
    float d = o.delta_max * powf(0.5f, (float)(pass - 1));
    return std::max(o.delta_min, d);
}

// ── batched solve of one group ──────────────────────────────────────────────
struct GroupDev {
    InstDesc *desc;
    int *blk;
    SolveScal *scal;
    double2 *part;
    double4 *part4;   // fused solver's four partial sums per workgroup
    SolveScal *scal2;    // the folded iteration's second state buffer
    double4 *part4b;     // and its second partial-sum buffer
    SolveScal *hstate[2] = {nullptr, nullptr};   // pinned host copies of the state (chunk pipelining), or null
    int4 *mix = nullptr;   // the merged iteration launch's tile list (k_iq_mix), or null
    int ninst, nblk;
    int nmix = 0;
    std::vector<long long> px;   // host: logical pixels per instance
};

// Live timing of the two PCG kernels (cuphu_set_profiling).
struct KernelProf {
    bool on = false;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    hipEvent_t ev2[4] = {nullptr, nullptr, nullptr, nullptr};   // the second chunk in flight (fused solver)
    CuPhuKernelProfile acc{};
};
static KernelProf g_prof;

// Algorithmic HBM bytes per active pixel of the two PCG kernels.
//   direction + operator: reads r, dinv, p_old and the {down, right} arc
//   pair, writes p_new and q: 7 values.
//   update: reads x, p, r, q, dinv, writes x and r: 7 values.
constexpr double kPqValuesPerPx = 7.0, kXrValuesPerPx = 7.0;

template <class T>
static void solve_group_batched(const Planes<T> &pl, const GroupDev &g, const CuPhuLaplaceOptions &o, int rpw,
                                bool fixed_weights, hipStream_t st, const std::vector<CuPhuSolveStats *> &stats,
                                std::vector<SolveScal> &h) {
    // Chunk length: even (the direction kernel alternates its two planes, so
    // an even chunk leaves the current direction in p and p2 free for the
    // true residual), at least 2.
    int ch = o.rr_interval > 0 ? o.rr_interval : 50;
    ch = std::max(2, std::min(ch, std::max(o.max_iter, 2)));
    const int CH = ch + (ch & 1);
    const double tol2 = o.tol * o.tol;
    const bool prof = g_prof.on;
    if (prof && !g_prof.ev[0])
        for (auto &e : g_prof.ev) CUPHU_HIP_CHECK(hipEventCreate(&e));
    auto step = [&](int parity) {
        launch_pq<T>(g.desc, g.blk, g.nblk, pl, g.scal, g.part, rpw, parity, st);
        launch_reduce(g.desc, g.ninst, g.part, g.scal, RED_PQ, o.max_iter, tol2, st);
        launch_xr<T>(g.desc, g.blk, g.nblk, pl, g.scal, g.part, rpw, parity, st);
        launch_reduce(g.desc, g.ninst, g.part, g.scal, RED_RZ, o.max_iter, tol2, st);
    };
    // With profiling on, the first two steps of a chunk run outside the
    // graph: step 0 keeps the GPU busy so that step 1, with events around
    // each kernel, runs back to back like the graph's launches; the graph
    // holds CH-2.
    auto chunk = [&]() {
        for (int i = prof ? 2 : 0; i < CH; ++i) step(i & 1);
        if (o.rr_interval > 0) {
            launch_rr<T>(g.desc, g.blk, g.nblk, pl, g.scal, g.part, rpw, st);
            launch_reduce(g.desc, g.ninst, g.part, g.scal, RED_RR, o.max_iter, tol2, st);
            launch_rr_commit<T>(g.desc, g.blk, g.nblk, pl, g.scal, rpw, st);
        }
    };
    hipGraphExec_t gexec = nullptr;
    if (o.use_graph) {
        hipGraph_t graph = nullptr;
        CUPHU_HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
        try {
            chunk();
        } catch (...) {
            hipGraph_t dummy;
            (void)hipStreamEndCapture(st, &dummy);
            throw;
        }
        CUPHU_HIP_CHECK(hipStreamEndCapture(st, &graph));
        CUPHU_HIP_CHECK(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
        CUPHU_HIP_CHECK(hipGraphDestroy(graph));
    }
    h.resize((size_t)g.ninst);
    const int npass = fixed_weights ? 1 : o.n_irls + 1;
    try {
        for (int pass = 0; pass < npass; ++pass) {
            if (!fixed_weights) {
                T d2 = T(0);
                if (pass > 0) {
                    float df = pass_delta(o, pass);
                    d2 = (sizeof(T) == 4) ? (T)(df * df) : (T)((double)df * (double)df);
                }
                launch_weights<T>(g.desc, g.blk, g.nblk, pl, pass > 0 ? 1 : 0, d2, TileShape{rpw, 64}, st);
                launch_assemble<T>(g.desc, g.blk, g.nblk, pl, TileShape{rpw, 64}, st);
            }
            launch_rz<T>(g.desc, g.blk, g.nblk, pl, g.scal, 1, g.part, rpw, st);
            launch_reduce(g.desc, g.ninst, g.part, g.scal, RED_START, o.max_iter, tol2, st);
            CUPHU_HIP_CHECK(hipMemcpyAsync(h.data(), g.scal, h.size() * sizeof(SolveScal), hipMemcpyDeviceToHost, st));
            CUPHU_HIP_CHECK(hipStreamSynchronize(st));
            int nch = 0;
            auto any_active = [&]() {
                for (auto &s : h) if (s.active) return true;
                return false;
            };
            while (any_active()) {
                double active_px = 0.0;
                if (prof) {
                    for (int i = 0; i < g.ninst; ++i)
                        if (h[(size_t)i].active) active_px += (double)g.px[(size_t)i];
                    step(0);
                    CUPHU_HIP_CHECK(hipEventRecord(g_prof.ev[0], st));
                    launch_pq<T>(g.desc, g.blk, g.nblk, pl, g.scal, g.part, rpw, 1, st);
                    CUPHU_HIP_CHECK(hipEventRecord(g_prof.ev[1], st));
                    launch_reduce(g.desc, g.ninst, g.part, g.scal, RED_PQ, o.max_iter, tol2, st);
                    CUPHU_HIP_CHECK(hipEventRecord(g_prof.ev[2], st));
                    launch_xr<T>(g.desc, g.blk, g.nblk, pl, g.scal, g.part, rpw, 1, st);
                    CUPHU_HIP_CHECK(hipEventRecord(g_prof.ev[3], st));
                    launch_reduce(g.desc, g.ninst, g.part, g.scal, RED_RZ, o.max_iter, tol2, st);
                }
                if (gexec) CUPHU_HIP_CHECK(hipGraphLaunch(gexec, st));
                else chunk();
                CUPHU_HIP_CHECK(hipMemcpyAsync(h.data(), g.scal, h.size() * sizeof(SolveScal), hipMemcpyDeviceToHost, st));
                CUPHU_HIP_CHECK(hipStreamSynchronize(st));
                if (prof) {
                    float ms_pq = 0.0f, ms_xr = 0.0f;
                    CUPHU_HIP_CHECK(hipEventElapsedTime(&ms_pq, g_prof.ev[0], g_prof.ev[1]));
                    CUPHU_HIP_CHECK(hipEventElapsedTime(&ms_xr, g_prof.ev[2], g_prof.ev[3]));
                    CuPhuKernelProfile &a = g_prof.acc;
                    a.samples += 1;
                    a.px_sum += active_px;
                    a.pq_ms_sum += ms_pq;
                    a.pq_bytes_sum += active_px * kPqValuesPerPx * sizeof(T);
                    a.xr_ms_sum += ms_xr;
                    a.xr_bytes_sum += active_px * kXrValuesPerPx * sizeof(T);
                    a.iterations += CH;
                }
                ++nch;
                if ((long long)nch * CH > (long long)o.max_iter + CH) break;   // cannot happen: reduce caps iterations
            }
            if (o.verbose)
                for (int i = 0; i < g.ninst; ++i)
                    fprintf(stderr, "[cuphu-pcg] inst %d pass %d iters %d relres %.3e status %d\n", i, pass, h[i].iters,
                            h[i].relres, h[i].status);
            for (int i = 0; i < g.ninst; ++i) record_pass(stats[(size_t)i], pass, h[(size_t)i]);
        }
    } catch (...) {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        throw;
    }
    if (gexec) CUPHU_HIP_CHECK(hipGraphExecDestroy(gexec));
}

// Algorithmic HBM bytes per active pixel of the fused iteration kernels
// (1/diag is recomputed from the arcs in both):
//   k_iq (default): reads x, r, p and the {down, right} arc pair, writes
//   x, r, p -- q = A p is recomputed, never stored (8 values);
//   k_it (CUPHU_FUSED_STORED_Q=1): reads x, r, p, q and the arc pair,
//   writes x, r, p, q (10 values).
static bool fused_stored_q() {
    static const bool v = [] {
        const char *e = std::getenv("CUPHU_FUSED_STORED_Q");
        return e && std::atoi(e) != 0;
    }();
    return v;
}

// Row strides rounded up to 256 elements (make_layout) on the fused q-free
// solver, whose iteration kernel skips the extra columns (the other solvers
// would process them every iteration); CUPHU_LAPLACE_STRIDE=64 turns the
// rounding off.
static bool wide_stride(const CuPhuLaplaceOptions &o) {
    int v = 0;
    env_int("CUPHU_LAPLACE_STRIDE", &v);
    return v != 64 && o.solver == CUPHU_SOLVER_FUSED && !fused_stored_q();
}

// Folded reductions (launch_iq_fold) for grids of at most this many
// workgroups, where the reduction launch it replaces is a large part of an
// iteration and every workgroup's extra read of its instance's partial sums
// is short.  Measured (same GPU, alternating runs, fused solver with and
// without folding): 1024^2 (256 workgroups; k_iq 9.0 us + k_reduce_it
// 4.5 us per iteration unfolded) 13.07 -> 15.17 Mpixel/s; 2048^2 (1024
// workgroups; 29.0 + 4.7 us) 20.13 -> 20.59.  CUPHU_FUSED_FOLD=0 turns it
// off, =1 forces it on.
constexpr int kFoldMaxBlocks = 1024;
static bool fold_reductions(int nblk) {
    int v = -1;
    env_int("CUPHU_FUSED_FOLD", &v);
    if (v == 0 || v == 1) return v == 1;
    return nblk <= kFoldMaxBlocks;
}

// Fused solver: one stencil kernel and one reduction per PCG iteration.
template <class T>
static void solve_group_fused(const Planes<T> &pl, const GroupDev &g, const CuPhuLaplaceOptions &o, TileShape sh,
                              bool fixed_weights, hipStream_t st, const std::vector<CuPhuSolveStats *> &stats,
                              std::vector<SolveScal> &h) {
    // Chunk length: even, so a chunk (parities 1, 0, 1, ..., 0) leaves the
    // state in plane set 1 where the start kernel put it.
    int ch = o.rr_interval > 0 ? o.rr_interval : 50;
    ch = std::max(2, std::min(ch, std::max(o.max_iter, 2)));
    const int CH = ch + (ch & 1);
    const double tol2 = o.tol * o.tol;
    const bool prof = g_prof.on;
    if (prof && !g_prof.ev[0])
        for (auto &e : g_prof.ev) CUPHU_HIP_CHECK(hipEventCreate(&e));
    if (prof && !g_prof.ev2[0])
        for (auto &e : g_prof.ev2) CUPHU_HIP_CHECK(hipEventCreate(&e));
    const bool sq = fused_stored_q();
    CUPHU_REQUIRE(!sq || sh.lpr == 64, "the stored-q kernels run 256-column tiles only");
    const int rpw = sh.rpw;
    // values moved per pixel by two consecutive launches: k_it 10 + 10;
    // k_iq 6 (x deferred: reads r, p, arc pair; writes r, p) + 9 (also reads
    // x and the previous p, writes x)
    const double pair_values_per_px = sq ? 20.0 : 15.0;
    auto it = [&](int parity, int start) {
        if (sq) launch_it<T>(g.desc, g.blk, g.nblk, pl, g.scal, g.part4, rpw, parity, start, st);
        else if (g.nmix) launch_iq_mix<T>(g.desc, g.mix, g.nmix, pl, g.scal, g.part4, rpw, parity, start, st);
        else launch_iq<T>(g.desc, g.blk, g.nblk, pl, g.scal, g.part4, sh, parity, start, st);
    };
    auto step = [&](int i) {
        it(1 - (i & 1), 0);
        launch_reduce_it(g.desc, g.ninst, g.part4, g.scal, RED_IT, o.max_iter, tol2, st);
    };
    // Steps first..last-1 of a chunk with the reductions folded: step `first`
    // runs on the state in g.scal and leaves its sums in g.part4; each later
    // step k (counted from first) reduces step k-1's sums, advancing the
    // state from one buffer to the other (state after step k-2 in g.scal when
    // k is odd, g.scal2 when even; it writes the other), while writing its own
    // sums to the other partial buffer; one reduction after the last step
    // brings the state back to g.scal.  The same states, bit for bit, as one
    // reduction launch per step.
    CUPHU_REQUIRE(g.scal2 && g.part4b, "fused solver group without its second state / partial buffers");
    const bool fold = !sq && fold_reductions(g.nblk);
    auto folded_steps = [&](int first, int last) {
        if (first >= last) return;
        double4 *pb[2] = {g.part4, g.part4b};
        it(1 - (first & 1), 0);   // into pb[0] = g.part4
        for (int i = first + 1; i < last; ++i) {
            const int k = i - first;
            const SolveScal *sin = (k & 1) ? g.scal : g.scal2;
            SolveScal *sout = (k & 1) ? g.scal2 : g.scal;
            if (g.nmix)
                launch_iq_fold_mix<T>(g.desc, g.mix, g.nmix, pl, sin, sout, pb[(k - 1) & 1], pb[k & 1], rpw,
                                      1 - (i & 1), o.max_iter, tol2, st);
            else
                launch_iq_fold<T>(g.desc, g.blk, g.nblk, pl, sin, sout, pb[(k - 1) & 1], pb[k & 1], sh,
                                  1 - (i & 1), o.max_iter, tol2, st);
        }
        const int k = last - 1 - first;   // the last step: state after step k-1 is in (k odd ? scal2 : scal)
        launch_reduce_it(g.desc, g.ninst, pb[k & 1], g.scal, RED_IT, o.max_iter, tol2, st, 0,
                         k == 0 ? g.scal : ((k & 1) ? g.scal2 : g.scal));
    };
    // With profiling on, the first two steps of a chunk run outside the
    // graph, each between two events (a deferral pair: together they move
    // the average bytes of the steady state); the graph holds CH-2.
    auto chunk = [&]() {
        if (fold) folded_steps(prof ? 2 : 0, CH);
        else
            for (int i = prof ? 2 : 0; i < CH; ++i) step(i);
        if (o.rr_interval > 0) {
            if (sq) launch_rr_it<T>(g.desc, g.blk, g.nblk, pl, g.scal, g.part4, rpw, st);
            else launch_rr_iq<T>(g.desc, g.blk, g.nblk, pl, g.scal, g.part4, sh, st);
            launch_reduce_it(g.desc, g.ninst, g.part4, g.scal, RED_RR, o.max_iter, tol2, st);
        }
    };
    hipGraphExec_t gexec = nullptr;
    if (o.use_graph) {
        hipGraph_t graph = nullptr;
        CUPHU_HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
        try {
            chunk();
        } catch (...) {
            hipGraph_t dummy;
            (void)hipStreamEndCapture(st, &dummy);
            throw;
        }
        CUPHU_HIP_CHECK(hipStreamEndCapture(st, &graph));
        CUPHU_HIP_CHECK(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
        CUPHU_HIP_CHECK(hipGraphDestroy(graph));
    }
    // Chunk pipelining: with pinned state buffers, the next chunk is issued
    // before the host has read the state after the current one, so the GPU
    // does not wait for the host between chunks (1024^2: 14.7 -> 15.2
    // Mpixel/s; 8192^2: the 120 chunk boundaries of an unwrap left the GPU
    // idle about 130 us each).  It is issued early only while an instance
    // active in the last state read can still be active after the chunk in
    // flight (iterations + chunk < max_iter), so a solve that runs to the
    // cap ends without an extra chunk.  One that converges inside a chunk
    // leaves the chunk after it a no-op (its kernels skip stopped instances
    // and the reductions keep their states).  Whenever instances are still
    // active after a read with nothing in flight, the next chunk is issued
    // then, as in the serial loop, so the results are those of the
    // one-chunk-at-a-time loop whatever the prediction.
    const bool pipe = g.hstate[0] && g.hstate[1];
    struct Done {   // waits for a chunk still in flight before its event and state buffer go
        hipEvent_t e[2] = {nullptr, nullptr};
        ~Done() {
            for (auto &x : e)
                if (x) {
                    (void)hipEventSynchronize(x);
                    (void)hipEventDestroy(x);
                }
        }
    } done;
    if (pipe)
        for (auto &x : done.e) CUPHU_HIP_CHECK(hipEventCreateWithFlags(&x, hipEventDisableTiming));
    h.resize((size_t)g.ninst);
    // Issues one chunk into slot `slot` (profiled pair first when profiling
    // is on) and the copy of the state after it.
    auto issue = [&](int slot) {
        if (prof) {
            hipEvent_t *ev = slot ? g_prof.ev2 : g_prof.ev;
            CUPHU_HIP_CHECK(hipEventRecord(ev[0], st));
            it(1, 0);
            CUPHU_HIP_CHECK(hipEventRecord(ev[1], st));
            launch_reduce_it(g.desc, g.ninst, g.part4, g.scal, RED_IT, o.max_iter, tol2, st);
            CUPHU_HIP_CHECK(hipEventRecord(ev[2], st));
            it(0, 0);
            CUPHU_HIP_CHECK(hipEventRecord(ev[3], st));
            launch_reduce_it(g.desc, g.ninst, g.part4, g.scal, RED_IT, o.max_iter, tol2, st);
        }
        if (gexec) CUPHU_HIP_CHECK(hipGraphLaunch(gexec, st));
        else chunk();
        if (pipe) {
            CUPHU_HIP_CHECK(hipMemcpyAsync(g.hstate[slot], g.scal, (size_t)g.ninst * sizeof(SolveScal),
                                           hipMemcpyDeviceToHost, st));
            CUPHU_HIP_CHECK(hipEventRecord(done.e[slot], st));
        } else {
            CUPHU_HIP_CHECK(hipMemcpyAsync(h.data(), g.scal, h.size() * sizeof(SolveScal), hipMemcpyDeviceToHost, st));
        }
    };
    // Waits for the chunk in `slot` and reads the state after it into h;
    // `before` is the state it started from (the profiled pair's bytes).
    auto collect = [&](int slot, const std::vector<SolveScal> &before, int pass) {
        if (pipe) {
            CUPHU_HIP_CHECK(hipEventSynchronize(done.e[slot]));
            std::memcpy(h.data(), g.hstate[slot], h.size() * sizeof(SolveScal));
        } else {
            CUPHU_HIP_CHECK(hipStreamSynchronize(st));
        }
        if (o.verbose > 1 && o.rr_interval > 0)
            for (int i = 0; i < g.ninst; ++i)
                if (h[(size_t)i].active)
                    fprintf(stderr, "[cuphu-rr] inst %d pass %d iters %d relres %.3e true/recursive r.z %.4g %s\n", i,
                            pass, h[i].iters, h[i].relres, h[i].delta, h[i].replace ? "replaced" : "kept");
        if (prof) {
            // both launches of a deferral pair (the first defers the x
            // update, the second applies two), on the instances active
            // when the chunk started
            hipEvent_t *ev = slot ? g_prof.ev2 : g_prof.ev;
            double active_px = 0.0;
            for (int i = 0; i < g.ninst; ++i)
                if (before[(size_t)i].active) active_px += (double)g.px[(size_t)i];
            float ms0 = 0.0f, ms1 = 0.0f;
            CUPHU_HIP_CHECK(hipEventElapsedTime(&ms0, ev[0], ev[1]));
            CUPHU_HIP_CHECK(hipEventElapsedTime(&ms1, ev[2], ev[3]));
            CuPhuKernelProfile &a = g_prof.acc;
            a.samples += 2;
            a.px_sum += 2.0 * active_px;
            a.it_ms_sum += (double)ms0 + (double)ms1;
            a.it_bytes_sum += active_px * pair_values_per_px * sizeof(T);
            a.iterations += CH;
        }
    };
    const int npass = fixed_weights ? 1 : o.n_irls + 1;
    try {
        for (int pass = 0; pass < npass; ++pass) {
            TraceRange tr("cuphu:pcg_solve");
            if (!fixed_weights) {
                T d2 = T(0);
                if (pass > 0) {
                    float df = pass_delta(o, pass);
                    d2 = (sizeof(T) == 4) ? (T)(df * df) : (T)((double)df * (double)df);
                }
                launch_weights<T>(g.desc, g.blk, g.nblk, pl, pass > 0 ? 1 : 0, d2, sh, st);
                launch_assemble<T>(g.desc, g.blk, g.nblk, pl, sh, st);
            }
            it(0, 1);
            launch_reduce_it(g.desc, g.ninst, g.part4, g.scal, RED_START, o.max_iter, tol2, st, sq ? 0 : 1);
            CUPHU_HIP_CHECK(hipMemcpyAsync(h.data(), g.scal, h.size() * sizeof(SolveScal), hipMemcpyDeviceToHost, st));
            CUPHU_HIP_CHECK(hipStreamSynchronize(st));
            auto any_active = [&]() {
                for (auto &s : h) if (s.active) return true;
                return false;
            };
            // chunks issued and collected; `before` is the state the chunk
            // being collected started from
            int issued = 0, collected = 0;
            std::vector<SolveScal> before = h;
            auto can_issue = [&]() { return (long long)issued * CH <= (long long)o.max_iter + CH; };
            // h is the state after the last chunk read; one chunk is in flight
            auto may_run_after = [&]() {
                for (auto &s : h)
                    if (s.active && (long long)s.iters + CH < (long long)o.max_iter) return true;
                return false;
            };
            if (any_active()) {
                issue(issued & 1);
                ++issued;
            }
            while (collected < issued) {
                if (pipe && can_issue() && may_run_after()) {   // the next chunk, before this one's state is known
                    issue(issued & 1);
                    ++issued;
                }
                collect(collected & 1, before, pass);
                ++collected;
                before = h;
                if (!any_active()) break;   // a chunk still in flight finds every instance stopped: a no-op
                if ((!pipe || collected == issued) && can_issue()) {
                    issue(issued & 1);
                    ++issued;
                }
            }
            if (!sq) launch_x_finalize<T>(g.desc, g.blk, g.nblk, pl, g.scal, sh, st);
            if (o.verbose)
                for (int i = 0; i < g.ninst; ++i)
                    fprintf(stderr, "[cuphu-pcg] inst %d pass %d iters %d relres %.3e status %d\n", i, pass, h[i].iters,
                            h[i].relres, h[i].status);
            for (int i = 0; i < g.ninst; ++i) record_pass(stats[(size_t)i], pass, h[(size_t)i]);
        }
    } catch (...) {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        throw;
    }
    if (gexec) CUPHU_HIP_CHECK(hipGraphExecDestroy(gexec));
}

template <class T>
static void solve_group(const Planes<T> &pl, const GroupDev &g, const CuPhuLaplaceOptions &o, TileShape sh,
                        bool fixed_weights, hipStream_t st, const std::vector<CuPhuSolveStats *> &stats,
                        std::vector<SolveScal> &h) {
    if (o.solver == CUPHU_SOLVER_FUSED) {
        solve_group_fused<T>(pl, g, o, sh, fixed_weights, st, stats, h);
    } else {
        CUPHU_REQUIRE(sh.lpr == 64, "the batched solver runs 256-column tiles only");
        solve_group_batched<T>(pl, g, o, sh.rpw, fixed_weights, st, stats, h);
    }
}

// ── plain solve of one instance (reference launch sequence) ─────────────────
template <class T>
static void solve_instance_plain(const Planes<T> &pl, const GroupDev &g, const InstDesc &d, const CuPhuLaplaceOptions &o,
                                 int rpw, DevRes &res, hipStream_t st, CuPhuSolveStats *stats) {
    size_t n = (size_t)d.nrow * d.pitch;
    T *x = pl.x + d.off, *r = pl.r + d.off, *z = pl.wv + d.off, *p = pl.p + d.off, *q = pl.s + d.off;
    const T *dinv = pl.dinv + d.off;
    double *parts = ws<double>(res, "dotparts", kDotParts);
    double *dres = ws<double>(res, "dotres", 1);
    auto dot = [&](const T *a, const T *b) {
        launch_plain_dot<T>(a, b, n, parts, dres, st);
        double v = 0.0;
        CUPHU_HIP_CHECK(hipMemcpyAsync(&v, dres, sizeof(double), hipMemcpyDeviceToHost, st));
        CUPHU_HIP_CHECK(hipStreamSynchronize(st));
        return v;
    };
    const double tol2 = o.tol * o.tol;
    for (int pass = 0; pass <= o.n_irls; ++pass) {
        T d2 = T(0);
        if (pass > 0) {
            float df = pass_delta(o, pass);
            d2 = (sizeof(T) == 4) ? (T)(df * df) : (T)((double)df * (double)df);
        }
        launch_weights<T>(g.desc, g.blk, g.nblk, pl, pass > 0 ? 1 : 0, d2, TileShape{rpw, 64}, st);
        launch_assemble<T>(g.desc, g.blk, g.nblk, pl, TileShape{rpw, 64}, st);
        launch_plain_pz<T>(p, z, r, dinv, n, st);
        double rz = dot(r, z), rz0 = rz, rzl = rz;
        int iter = 0, status;
        if (!(rz0 > 0.0)) status = (rz0 == 0.0) ? CUPHU_SOLVE_EMPTY : CUPHU_SOLVE_NONFINITE;
        else {
            status = CUPHU_SOLVE_MAXITER;
            for (; iter < o.max_iter; ++iter) {
                launch_plain_matvec<T>(d, pl.w2, pl.p, pl.s, st);
                double pq = dot(p, q);
                if (!(pq > 0.0)) { status = std::isfinite(pq) ? CUPHU_SOLVE_BREAKDOWN : CUPHU_SOLVE_NONFINITE; break; }
                T a = (T)(rz / pq);
                launch_plain_axpy<T>(x, p, a, n, st);
                launch_plain_axpy<T>(r, q, -a, n, st);
                launch_plain_vmul<T>(z, r, dinv, n, st);
                double rzn = dot(r, z);
                rzl = rzn;
                if (o.verbose && (iter % 20 == 0))
                    fprintf(stderr, "[cuphu-plain] pass %d iter %d rel_res %.3e\n", pass, iter, std::sqrt(rzn / rz0));
                if (rzn < tol2 * rz0) { iter++; status = CUPHU_SOLVE_CONVERGED; break; }
                T b = (T)(rzn / rz);
                rz = rzn;
                launch_plain_axpby<T>(p, z, b, n, st);
            }
        }
        SolveScal s{};
        s.iters = iter;
        s.status = status;
        s.relres = (rz0 > 0.0) ? std::sqrt(std::max(rzl, 0.0) / rz0) : 0.0;
        record_pass(stats, pass, s);
    }
}

// ── batch driver ────────────────────────────────────────────────────────────
struct BatchOut {
    Layout lay;
    std::vector<int> ncomp;
    float *unw = nullptr;       // pitched pool
    uint32_t *lab = nullptr;    // pitched pool
    float *corr = nullptr;      // pitched pool
    unsigned char *mask = nullptr;
};

// ── tile shape ──────────────────────────────────────────────────────────────
// Every tile is processed whole, so an instance occupies ceil(wcols / TW) * TW
// x ceil(nrow / TH) * TH lanes' pixels of the iteration kernel whatever its
// size: a width 64 past a multiple of 256 leaves 3/4 of its last 256-column
// tile idle.
static double tiled_px(const Layout &L, TileShape sh) {
    double n = 0.0;
    for (auto &d : L.desc)
        n += (double)(ceil_div((size_t)d.wcols, (size_t)sh.tw()) * (size_t)sh.tw()) *
             (double)(ceil_div((size_t)d.nrow, (size_t)sh.th()) * (size_t)sh.th());
    return n;
}

// Blocks of a launch that spans 4 resident rounds (3 workgroups x 256 CUs
// each).
constexpr size_t kNarrowMinBlocks = 4 * 768;

static size_t tiled_blocks(const Layout &L, TileShape sh) {
    size_t n = 0;
    for (auto &d : L.desc)
        n += ceil_div((size_t)d.pitch, (size_t)sh.tw()) * ceil_div((size_t)d.nrow, (size_t)sh.th());
    return n;
}

// Default shape: rows_per_wave() (CUPHU_LAPLACE_RPW; 4 in FP32, 2 in FP64) rows per lane group and
// 256-column tiles, or, on the fused solver, 128-column tiles when every
// instance's width is a multiple of 256, both shapes leave exactly as many
// lanes idle and the launch spans at least 4 resident rounds.  Measured on one GPU, alternating pairs:
//   - 8192^2 scene in 2048-column tiles (no idle lanes either way): the
//     128 x 32 tile reads 5 % fewer halo pixels; k_iq 0.4343 -> 0.4324 ms
//     per launch, 24.86 -> 25.00 Mpixel/s;
//   - one 2048^2 image (1024 workgroups, 1.3 per slot): 19.85 -> 19.35;
//   - 64 interferograms of mixed shapes, where 128-column tiles leave fewer
//     lanes idle: 27.87 -> 27.76 (idle lanes cost little in this HBM-bound
//     kernel: their loads are never issued).
// CUPHU_LAPLACE_LPR=64 or 32 forces the width (64 or 32 lanes per tile row).
static TileShape default_shape(const Layout &L, const CuPhuLaplaceOptions &o) {
    const int rpw = rows_per_wave(o.precision);
    const TileShape wide{rpw, 64}, narrow{rpw, 32};
    if (o.solver != CUPHU_SOLVER_FUSED || fused_stored_q() || rpw > 4) return wide;
    int lpr = 0;
    env_int("CUPHU_LAPLACE_LPR", &lpr);
    if (lpr == 64) return wide;
    if (lpr == 32) return narrow;
    // widths that are multiples of 256 (so the row stride has no extra
    // columns either: a 128-column tile that fits a 128-column instance
    // exactly ran 355 us on a 128-element stride but 523 us on a 256-element
    // one)
    for (auto &d : L.desc)
        if (d.wcols % 256 != 0) return wide;
    return tiled_px(L, narrow) == tiled_px(L, wide) && tiled_blocks(L, narrow) >= kNarrowMinBlocks ? narrow : wide;
}

static void retile(Layout &L, TileShape sh) {
    for (auto &d : L.desc) {
        d.tiles_x = (int)ceil_div((size_t)d.pitch, (size_t)sh.tw());
        d.tiles_y = (int)ceil_div((size_t)d.nrow, (size_t)sh.th());
        d.band = band_of(d.tiles_x, sh);
    }
}

// ── merged tiling of the iteration kernel (k_iq_mix) ────────────────────────
// On 256-column tiles, an instance whose last tile column holds 64 work
// columns (wcols = 256 k + 64, as the reference tile formula's 1843-column
// tiles have) runs that column in 64 x 4TH tall tiles: one workgroup instead
// of four quarter-filled ones; with 128 work columns, in 128 x 2TH tiles, one
// instead of two half-filled ones.  CUPHU_LAPLACE_MERGE=0 turns it off, =2
// takes it on grids of any size (default 1: beyond one resident round).
// Returns the last tile column's kind: 0 (256 x TH), 1 (64 x 4TH), 2
// (128 x 2TH).
static int tall_column(const InstDesc &d, TileShape sh) {
    if (sh.lpr != 64) return 0;
    const int rem = d.wcols - (d.tiles_x - 1) * sh.tw();
    return rem == 64 ? 1 : (rem == 128 ? 2 : 0);
}
// Grids that fit in one resident round (3 workgroups x 256 CUs) keep the
// plain tiling: every workgroup is resident at once there, so fewer of them
// save nothing and the launch waits for the slower tall tiles.  Measured on
// single images, folded reductions, merged against plain: 1100 x 1300 (414
// workgroups) 0.7 % slower, 1500 x 1400 (528) 0.8 % faster; 1843 x 1843
// (928) 3.5 % and 2048 x 1600 (896) 6.5 % faster.
constexpr size_t kMergeMinBlocks = 768;
static int merge_mode() {
    int v = 1;
    env_int("CUPHU_LAPLACE_MERGE", &v);
    return v;
}
static bool merge_allowed(TileShape sh) {
    return merge_mode() != 0 && sh.lpr == 64 && (sh.rpw == 2 || sh.rpw == 4) && !fused_stored_q();
}
// Appends instance i's workgroups (descriptor d, tile coordinates in
// 256 x TH units) in the launch order of get_tile (raster, or band by band),
// with the tall column's groups of four (or two) tiles at the place of their
// first tile.
static void mix_tiles(const InstDesc &d, int i, TileShape sh, std::vector<int4> &out) {
    const int kind = tall_column(d, sh), cover = kind == 1 ? 4 : 2;
    const int n = d.tiles_x * d.tiles_y;
    for (int local = 0; local < n; ++local) {
        int tx, ty;
        if (d.band > 0) {
            const int bsz = d.band * d.tiles_y;
            const int k = local / bsz;
            const int bw = std::min(d.band, d.tiles_x - k * d.band);
            const int rem = local - k * bsz;
            ty = rem / bw;
            tx = k * d.band + rem % bw;
        } else {
            tx = local % d.tiles_x;
            ty = local / d.tiles_x;
        }
        if (kind && tx == d.tiles_x - 1) {
            if (ty % cover == 0) out.push_back(make_int4(i, tx, ty, kind));
        } else {
            out.push_back(make_int4(i, tx, ty, 0));
        }
    }
}

// ── tile-shape autotuner ────────────────────────────────────────────────────
// The stencil kernels' tile shape trades halo re-reads and idle lanes against
// registers and occupancy, and the best value depends on the instance
// shapes.  On the first batched/fused solve of a shape class (precision,
// widest pitch, stride rounding, instance count, stored and work pixel
// count classes) each candidate
// (256-column tiles of 8, 16 or 32 rows; on the fused solver also
// 128-column tiles of 16 or 32 rows) is timed on the real pooled planes with
// the fused kernel's steady-state step form before the solve starts (scratch
// scalars; reads plane set 1, writes set 0, which the first assemble resets;
// x is not touched): one warm-up launch each, then four rounds of one timed
// launch per candidate, minimum per candidate; the fastest is taken unless
// the default shape is within 3 % of it, and cached for the process.  Opt-in (options.autotune or
// CUPHU_LAPLACE_AUTOTUNE=1): the tiling fixes the order of the FP64 partial
// sums, so a timing-dependent choice can change results in the last bits
// between processes.  CUPHU_LAPLACE_RPW forces the default shape.
static int ilog2_ceil(long long v) {
    int k = 0;
    while ((1LL << k) < v && k < 62) ++k;
    return k;
}

template <class T>
static TileShape autotune_shape(Layout L, const Planes<T> &pl, const CuPhuLaplaceOptions &o, DevRes &res,
                                hipStream_t st) {
    if (std::getenv("CUPHU_LAPLACE_RPW")) return default_shape(L, o);
    static std::map<std::string, TileShape> cache;
    const int ninst = (int)L.desc.size();
    // 128-column tiles only where default_shape would consider them: every
    // work width a multiple of 256 (a 128-column tile on a wider row stride
    // ran 523 us against 355 us on its own stride)
    bool widths256 = true;
    long long work_px = 0;
    for (auto &d : L.desc) {
        widths256 = widths256 && d.wcols % 256 == 0;
        work_px += (long long)d.nrow * d.wcols;
    }
    const bool narrow_ok = o.solver == CUPHU_SOLVER_FUSED && !fused_stored_q() && widths256;
    // shape class: precision, widest stride, stride rounding, instance count,
    // stored and work pixel counts (log2 classes) and the 128-column choice
    const std::string key = std::to_string(sizeof(T)) + "/" + std::to_string(L.max_pitch) + "/" +
                            std::to_string((int)wide_stride(o)) + "/" + std::to_string(ilog2_ceil(ninst)) + "/" +
                            std::to_string(ilog2_ceil(L.total_el)) + "/" + std::to_string(ilog2_ceil(work_px)) + "/" +
                            std::to_string((int)narrow_ok);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    std::vector<TileShape> cand = {{2, 64}, {4, 64}, {8, 64}};
    if (narrow_ok) {
        cand.push_back({2, 32});
        cand.push_back({4, 32});
    }
    // one descriptor set per candidate, so that the candidates can be timed
    // in interleaved rounds
    const int ncand = (int)cand.size();
    std::vector<InstDesc *> d_desc((size_t)ncand);
    std::vector<int *> d_blk((size_t)ncand);
    std::vector<int> nblk((size_t)ncand);
    size_t maxblk = 0;
    for (int ci = 0; ci < ncand; ++ci) {
        retile(L, cand[(size_t)ci]);
        std::vector<InstDesc> gd = L.desc;
        std::vector<int> blk;
        for (int i = 0; i < ninst; ++i) {
            gd[(size_t)i].block_begin = (int)blk.size();
            for (int k = 0; k < gd[(size_t)i].tiles_x * gd[(size_t)i].tiles_y; ++k) blk.push_back(i);
        }
        const std::string sfx = std::to_string(ci);
        d_desc[(size_t)ci] = ws<InstDesc>(res, ("tunedesc" + sfx).c_str(), (size_t)ninst);
        d_blk[(size_t)ci] = ws<int>(res, ("tuneblk" + sfx).c_str(), blk.size());
        nblk[(size_t)ci] = (int)blk.size();
        maxblk = std::max(maxblk, blk.size());
        CUPHU_HIP_CHECK(hipMemcpyAsync(d_desc[(size_t)ci], gd.data(), gd.size() * sizeof(InstDesc),
                                       hipMemcpyHostToDevice, st));
        CUPHU_HIP_CHECK(hipMemcpyAsync(d_blk[(size_t)ci], blk.data(), blk.size() * sizeof(int), hipMemcpyHostToDevice,
                                       st));
        // the host vectors go out of scope: finish the copies first
        CUPHU_HIP_CHECK(hipStreamSynchronize(st));
    }
    double4 *d_part = ws<double4>(res, "tunepart", maxblk);
    SolveScal *d_scal = ws<SolveScal>(res, "tunescal", (size_t)ninst);
    // scratch scalars for the steady-state step form: q recomputed, r and p
    // updated, x untouched (xmode 0)
    std::vector<SolveScal> zs((size_t)ninst);
    std::memset(zs.data(), 0, zs.size() * sizeof(SolveScal));
    for (auto &z : zs) { z.active = 1; z.alpha = 0.1; z.beta = 0.5; z.xmode = 0; }
    CUPHU_HIP_CHECK(hipMemcpyAsync(d_scal, zs.data(), zs.size() * sizeof(SolveScal), hipMemcpyHostToDevice, st));
    hipEvent_t e0, e1;
    CUPHU_HIP_CHECK(hipEventCreate(&e0));
    CUPHU_HIP_CHECK(hipEventCreate(&e1));
    // 1 warm-up launch per candidate, then kRounds rounds of one timed launch
    // per candidate, minimum per candidate: a clock or power transient during
    // the timing then reaches every candidate alike instead of deciding
    // against the one it happened to meet.
    constexpr int kRounds = 4;
    std::vector<float> ms_min((size_t)ncand, 1e30f);
    try {
        for (int ci = 0; ci < ncand; ++ci)
            launch_iq<T>(d_desc[(size_t)ci], d_blk[(size_t)ci], nblk[(size_t)ci], pl, d_scal, d_part,
                         cand[(size_t)ci], 1, 0, st);
        for (int rep = 0; rep < kRounds; ++rep)
            for (int ci = 0; ci < ncand; ++ci) {
                CUPHU_HIP_CHECK(hipEventRecord(e0, st));
                launch_iq<T>(d_desc[(size_t)ci], d_blk[(size_t)ci], nblk[(size_t)ci], pl, d_scal, d_part,
                             cand[(size_t)ci], 1, 0, st);
                CUPHU_HIP_CHECK(hipEventRecord(e1, st));
                CUPHU_HIP_CHECK(hipEventSynchronize(e1));
                float ms = 0.0f;
                CUPHU_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
                ms_min[(size_t)ci] = std::min(ms_min[(size_t)ci], ms);
            }
    } catch (...) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        throw;
    }
    CUPHU_HIP_CHECK(hipEventDestroy(e0));
    CUPHU_HIP_CHECK(hipEventDestroy(e1));
    // The default shape is kept unless another candidate is at least
    // kMargin faster: the candidates often differ by less than the timing's
    // own spread, and a shape that only wins on noise can lose several per
    // cent over the solve.
    constexpr float kMargin = 0.03f;
    const TileShape def = default_shape(L, o);
    int ibest = -1, idef = -1;
    for (int ci = 0; ci < ncand; ++ci) {
        if (cand[(size_t)ci] == def) idef = ci;
        if (ibest < 0 || ms_min[(size_t)ci] < ms_min[(size_t)ibest]) ibest = ci;
    }
    if (idef >= 0 && ms_min[(size_t)ibest] > (1.0f - kMargin) * ms_min[(size_t)idef]) ibest = idef;
    const TileShape best = cand[(size_t)ibest];
    if (o.verbose)
        for (int ci = 0; ci < ncand; ++ci)
            fprintf(stderr, "[cuphu-autotune] %s rows/lane group %d, lanes/row %d: %.4f ms%s\n", key.c_str(),
                    cand[(size_t)ci].rpw, cand[(size_t)ci].lpr, ms_min[(size_t)ci],
                    ci == ibest ? " (chosen)" : (ci == idef ? " (default)" : ""));
    cache[key] = best;
    return best;
}

// Zeroes columns wcols..pitch-1 of two pixel planes: one thread per (row,
// gap column) of the pooled rows, max_gap = the widest gap.
template <class T>
__global__ void k_zero_gaps(const InstDesc *descs, const int *rowinst, int nrows_total, int max_gap, T *a, T *b) {
    long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= (long long)nrows_total * max_gap) return;
    const int row = (int)(i / max_gap), j = (int)(i % max_gap);
    const InstDesc d = descs[rowinst[row]];
    const int c = d.wcols + j;
    if (c >= d.pitch) return;
    const long long e = d.off + (long long)(row - d.row_begin) * d.pitch + c;
    a[e] = T(0);
    b[e] = T(0);
}

// ingest(planes, layout, d_desc, d_rowinst) fills phi/corr/mask/sigsq and x.
template <class T, class Ingest>
static void run_batch(const std::vector<std::pair<int, int>> &dims, const CuPhuLaplaceOptions &o, double minfrac,
                      long maxncomps, bool do_conncomp, DevRes &res, hipStream_t st, CuPhuSolveStats *stats,
                      BatchOut &out, Ingest &&ingest) {
    const bool stride256 = wide_stride(o);
    Layout L = make_layout(dims, TileShape{rows_per_wave(o.precision), 64}, stride256);
    int ninst = (int)dims.size();
    PoolSet<T> P = alloc_pools<T>(res, L.total_el);
    InstDesc *d_gdesc = ws<InstDesc>(res, "gdesc", (size_t)ninst);
    int *d_rowinst = ws<int>(res, "rowinst", (size_t)L.total_rows);
    std::vector<int> rowinst((size_t)L.total_rows);
    for (int i = 0; i < ninst; ++i)
        for (int rr = 0; rr < L.desc[(size_t)i].nrow; ++rr) rowinst[(size_t)(L.desc[(size_t)i].row_begin + rr)] = i;
    CUPHU_HIP_CHECK(hipMemcpyAsync(d_gdesc, L.desc.data(), (size_t)ninst * sizeof(InstDesc), hipMemcpyHostToDevice, st));
    CUPHU_HIP_CHECK(hipMemcpyAsync(d_rowinst, rowinst.data(), rowinst.size() * sizeof(int), hipMemcpyHostToDevice, st));
    if (stride256) {
        // The q-free iteration kernel covers columns below wcols only, so it
        // never writes the stride gap (wcols..pitch-1) of the second plane
        // set (r2, p2), which the residual monitor and the x finalize read
        // over the whole pitch: the gap must hold zeros, whatever a previous
        // call left there.  (Plane set 0 is rewritten over the whole pitch by
        // the assemble kernel every pass; q2 is not used by this solver.)
        int max_gap = 0;
        for (auto &d : L.desc) max_gap = std::max(max_gap, d.pitch - d.wcols);
        if (max_gap > 0) {
            long long n = (long long)L.total_rows * max_gap;
            hipLaunchKernelGGL(k_zero_gaps<T>, dim3((unsigned)ceil_div((size_t)n, 256)), dim3(256), 0, st, d_gdesc,
                               d_rowinst, L.total_rows, max_gap, P.pl.r2, P.pl.p2);
            CUPHU_HIP_CHECK(hipGetLastError());
        }
    }
    {
        TraceRange tr("cuphu:ingest");
        ingest(P.pl, L, d_gdesc, d_rowinst);
    }
    // tile shape (the element layout does not depend on it, only the
    // workgroup tiling)
    const TileShape sh = (o.solver == CUPHU_SOLVER_PLAIN || !o.autotune) ? default_shape(L, o)
                                                                          : autotune_shape<T>(L, P.pl, o, res, st);
    retile(L, sh);

    // groups of consecutive instances bounded by the working-set budget
    // bytes per pixel one PCG iteration touches: w2 pair + x, r, p, p2, q, dinv
    const double bpe = 8.0 * sizeof(T);
    long long cap = o.group_bytes_mb > 0 ? (long long)((double)o.group_bytes_mb * 1048576.0 / bpe) : (1LL << 62);
    std::vector<std::pair<int, int>> groups;
    for (int i = 0; i < ninst;) {
        int j = i;
        long long el = 0;
        while (j < ninst && (j == i || el + (long long)L.desc[(size_t)j].nrow * L.desc[(size_t)j].pitch <= cap)) {
            el += (long long)L.desc[(size_t)j].nrow * L.desc[(size_t)j].pitch;
            ++j;
        }
        groups.push_back({i, j});
        i = j;
    }
    size_t maxblk = 0, maxinst = 0;
    for (auto &gr : groups) {
        size_t nb = 0;
        for (int i = gr.first; i < gr.second; ++i) nb += (size_t)L.desc[(size_t)i].tiles_x * L.desc[(size_t)i].tiles_y;
        if (o.solver == CUPHU_SOLVER_PLAIN)
            for (int i = gr.first; i < gr.second; ++i)
                maxblk = std::max(maxblk, (size_t)L.desc[(size_t)i].tiles_x * L.desc[(size_t)i].tiles_y);
        maxblk = std::max(maxblk, nb);
        maxinst = std::max(maxinst, (size_t)(gr.second - gr.first));
    }
    GroupDev g{};
    g.desc = ws<InstDesc>(res, "grpdesc", maxinst);
    g.blk = ws<int>(res, "grpblk", maxblk);
    g.scal = ws<SolveScal>(res, "grpscal", maxinst);
    g.part = ws<double2>(res, "grppart", maxblk);
    g.part4 = ws<double4>(res, "grppart4", maxblk);
    g.scal2 = ws<SolveScal>(res, "grpscal2", maxinst);
    g.part4b = ws<double4>(res, "grppart4b", maxblk);
    if (!res.arena) {   // pinned state copies for the pipelined chunk loop (none in the no-allocation path)
        g.hstate[0] = (SolveScal *)res.pin("grpstate0", maxinst * sizeof(SolveScal));
        g.hstate[1] = (SolveScal *)res.pin("grpstate1", maxinst * sizeof(SolveScal));
    }
    std::vector<SolveScal> hs;
    std::vector<CuPhuSolveStats> dummy((size_t)ninst);
    auto sp = [&](int i) { return stats ? stats + i : &dummy[(size_t)i]; };
    for (int i = 0; i < ninst; ++i) std::memset(sp(i), 0, sizeof(CuPhuSolveStats));

    // the merged iteration launch, where some instance has a tall column of
    // at least two tall tiles' worth of rows
    const bool merge = o.solver == CUPHU_SOLVER_FUSED && merge_allowed(sh);
    if (merge) g.mix = ws<int4>(res, "grpmix", maxblk);
    auto upload_group = [&](int a, int b) {
        std::vector<InstDesc> gd;
        std::vector<int> blk;
        std::vector<int4> mix;
        bool merges = false;
        for (int i = a; i < b; ++i) {
            InstDesc d = L.desc[(size_t)i];
            d.block_begin = (int)blk.size();
            for (int k = 0; k < d.tiles_x * d.tiles_y; ++k) blk.push_back(i - a);
            if (merge) {
                mix_tiles(d, i - a, sh, mix);
                merges = merges || (tall_column(d, sh) != 0 && d.tiles_y > 1);
            }
            gd.push_back(d);
        }
        CUPHU_HIP_CHECK(hipMemcpyAsync(g.desc, gd.data(), gd.size() * sizeof(InstDesc), hipMemcpyHostToDevice, st));
        CUPHU_HIP_CHECK(hipMemcpyAsync(g.blk, blk.data(), blk.size() * sizeof(int), hipMemcpyHostToDevice, st));
        g.nmix = merges && (blk.size() > kMergeMinBlocks || merge_mode() == 2) ? (int)mix.size() : 0;
        if (g.nmix)
            CUPHU_HIP_CHECK(hipMemcpyAsync(g.mix, mix.data(), mix.size() * sizeof(int4), hipMemcpyHostToDevice, st));
        g.ninst = b - a;
        g.nblk = (int)blk.size();
        g.px.clear();
        for (int i = a; i < b; ++i) g.px.push_back((long long)L.desc[(size_t)i].nrow * L.desc[(size_t)i].ncol);
        CUPHU_HIP_CHECK(hipStreamSynchronize(st));   // host vectors die at scope end
    };
    TraceRange tr_solve("cuphu:pcg_irls");
    if (o.solver == CUPHU_SOLVER_PLAIN) {
        for (int i = 0; i < ninst; ++i) {
            upload_group(i, i + 1);
            solve_instance_plain<T>(P.pl, g, L.desc[(size_t)i], o, sh.rpw, res, st, sp(i));
        }
    } else {
        for (auto &gr : groups) {
            upload_group(gr.first, gr.second);
            std::vector<CuPhuSolveStats *> gs;
            for (int i = gr.first; i < gr.second; ++i) gs.push_back(sp(i));
            solve_group<T>(P.pl, g, o, sh, false, st, gs, hs);
        }
    }
    launch_final<T>(d_gdesc, d_rowinst, L.total_rows, P.pl, P.out, L.max_pitch, st);
    out.ncomp.assign((size_t)ninst, 0);
    TraceRange tr_cc("cuphu:conncomp");
    if (do_conncomp) {
        // one scratch set in the workspace, sized for the largest instance, serves every call
        size_t maxn = 1;
        for (int i = 0; i < ninst; ++i)
            maxn = std::max(maxn, (size_t)L.desc[(size_t)i].nrow * (size_t)L.desc[(size_t)i].pitch);
        const CcScratch scr = cc_scratch(res, maxn);
        for (int i = 0; i < ninst; ++i) {
            const InstDesc &d = L.desc[(size_t)i];
            out.ncomp[(size_t)i] = conncomp_run(P.out + d.off, P.pl.corr + d.off, P.pl.mask + d.off, d.nrow, d.ncol,
                                                d.pitch, minfrac, maxncomps, P.lab + d.off, st, 0.05f, nullptr,
                                                nullptr, 0, &scr);
        }
    }
    out.lay = L;
    out.unw = P.out;
    out.lab = P.lab;
    out.corr = P.pl.corr;
    out.mask = P.pl.mask;
}

// Standard ingest from IngestSrc descriptors.
template <class T>
static auto src_ingest(const std::vector<IngestSrc> &srcs, bool has_u0, DevRes &res, hipStream_t st) {
    return [&srcs, has_u0, &res, st](const Planes<T> &pl, const Layout &L, const InstDesc *d_desc, const int *d_rowinst) {
        IngestSrc *d_src = ws<IngestSrc>(res, "srcs", srcs.size());
        CUPHU_HIP_CHECK(hipMemcpyAsync(d_src, srcs.data(), srcs.size() * sizeof(IngestSrc), hipMemcpyHostToDevice, st));
        launch_ingest<T>(d_src, d_desc, d_rowinst, L.total_rows, pl, has_u0 ? 1 : 0, L.max_pitch, st);
        CUPHU_HIP_CHECK(hipStreamSynchronize(st));
    };
}

// ── scene: tiling, batched solve, stitch, merge ─────────────────────────────
__global__ void k_merge(const float *unw_pool, const uint32_t *lab_pool, const long long *toff, const int *tpitch,
                        const int *rtab, const int *rloc, const int *ctab, const int *cloc, int ntc, const float *add,
                        const uint32_t *coff, int nrow, int ncol, float *out, uint32_t *cc) {
    long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= (long long)nrow * ncol) return;
    int r = (int)(i / ncol), c = (int)(i % ncol);
    int t = rtab[r] * ntc + ctab[c];
    long long e = toff[t] + (long long)rloc[r] * tpitch[t] + cloc[c];
    out[i] = unw_pool[e] + add[t];
    if (cc) {
        uint32_t l = lab_pool[e];
        cc[i] = l ? l + coff[t] : 0u;
    }
}

__global__ void k_copy_rows(const float *src, int spitch, float *dst, int nrow, int ncol) {
    long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= (long long)nrow * ncol) return;
    dst[i] = src[(i / ncol) * spitch + (i % ncol)];
}

// Registration, offsets and merge of solved tiles held in pooled planes
// (bo: pitched unw / labels / coherence / solve-mask pools and the tile
// layout, tiles in row-major order): each pair's coherence-weighted median
// overlap difference on the GPU (cuphu_solver.cpp:1867-1959), the BFS over
// the tile grid on the host, whole-cycle offsets when `integer`, overlap
// halves trimmed and labels offset by the running component count
// (:1969-2031).  d_conf: scene-level confidence mask (or NULL: the pooled
// solve mask when pooled_mask, else none).  d_cc NULL skips the labels.
static void stitch_tiles(const BatchOut &bo, int nrow, int ncol, int ntr, int ntc, const std::vector<int> &rf,
                         const std::vector<int> &rs, const std::vector<int> &cf, const std::vector<int> &cs, int rovl,
                         int covl, const unsigned char *d_conf, bool pooled_mask, bool integer, float *d_unw,
                         uint32_t *d_cc, DevRes &res, hipStream_t st) {
    TraceRange tr("cuphu:stitch");
    // pair registrations on the GPU (weighted medians of the overlap
    // differences), tile offsets by the BFS over the tile grid on the host
    const int nt = ntr * ntc;
    const unsigned char *conf = d_conf ? d_conf : (pooled_mask ? bo.mask : nullptr);
    std::vector<PairDesc> pairs;
    std::vector<std::pair<int, bool>> pair_of;
    int max_chunks = 0;
    for (int t = 0; t < nt; ++t) {
        const int br = t / ntc, bc = t % ntc;
        const InstDesc &A = bo.lay.desc[(size_t)t];
        for (int h = 1; h >= 0; --h) {
            if (h ? (bc + 1 >= ntc) : (br + 1 >= ntr)) continue;
            const int u = h ? t + 1 : t + ntc;
            const InstDesc &B = bo.lay.desc[(size_t)u];
            PairDesc pd{};
            pd.a_pitch = A.pitch;
            pd.b_pitch = B.pitch;
            // window origin in A (B's is its (0, 0)); empty windows as the
            // host registration (:1884-1890)
            int ar = 0, ac = 0;
            if (h) {
                if (covl > 0 && covl <= A.ncol && covl <= B.ncol) {
                    pd.nr = std::min(A.nrow, B.nrow);
                    pd.nc = covl;
                    ac = A.ncol - covl;
                }
            } else if (rovl > 0 && rovl <= A.nrow && rovl <= B.nrow) {
                pd.nr = rovl;
                pd.nc = std::min(A.ncol, B.ncol);
                ar = A.nrow - rovl;
            }
            pd.a_off = A.off + (long long)ar * A.pitch + ac;
            pd.b_off = B.off;
            if (d_conf) {   // scene-level mask: tile origins in the scene
                pd.ca_off = (long long)(rf[(size_t)br] + ar) * ncol + cf[(size_t)bc] + ac;
                pd.cb_off = (long long)rf[(size_t)(u / ntc)] * ncol + cf[(size_t)(u % ntc)];
                pd.ca_stride = pd.cb_stride = ncol;
            } else {        // pooled solve mask: same elements as unw
                pd.ca_off = pd.a_off;
                pd.cb_off = pd.b_off;
                pd.ca_stride = A.pitch;
                pd.cb_stride = B.pitch;
            }
            max_chunks = std::max(max_chunks, pair_chunks((long long)pd.nr * pd.nc));
            pairs.push_back(pd);
            pair_of.push_back({t, h != 0});
        }
    }
    const int npairs = (int)pairs.size();
    CUPHU_REQUIRE(npairs <= 65535, "too many tile pairs for the registration grid");
    std::vector<double> hk((size_t)nt, 0.0), vk((size_t)nt, 0.0);
    if (npairs > 0) {
        PairDesc *d_pairs = ws<PairDesc>(res, "s_pairs", (size_t)npairs);
        PairSel *d_sel = ws<PairSel>(res, "s_sel", (size_t)npairs);
        double *d_hist = ws<double>(res, "s_hist", (size_t)npairs * 256);
        CUPHU_HIP_CHECK(hipMemcpyAsync(d_pairs, pairs.data(), (size_t)npairs * sizeof(PairDesc),
                                       hipMemcpyHostToDevice, st));
        pair_medians(d_pairs, npairs, max_chunks, bo.unw, bo.corr, conf, d_sel, d_hist, st);
        std::vector<PairSel> sel((size_t)npairs);
        CUPHU_HIP_CHECK(hipMemcpyAsync(sel.data(), d_sel, (size_t)npairs * sizeof(PairSel), hipMemcpyDeviceToHost, st));
        CUPHU_HIP_CHECK(hipStreamSynchronize(st));
        for (int i = 0; i < npairs; ++i) {
            // cycles: fractional for the Laplace path (:1928), whole cycles
            // when the congruent stitch is asked for
            double cyc = (double)sel[(size_t)i].value / (2.0 * M_PI);
            if (integer) cyc = std::round(cyc);
            (pair_of[(size_t)i].second ? hk : vk)[(size_t)pair_of[(size_t)i].first] = cyc;
        }
    }
    std::vector<double> tk = bfs_offsets(ntr, ntc, hk, vk);
    // merge tables (later tiles win where odd overlaps make trims meet)
    std::vector<int> rtab((size_t)nrow, 0), rloc((size_t)nrow, 0), ctab((size_t)ncol, 0), cloc((size_t)ncol, 0);
    for (int tr = 0; tr < ntr; ++tr) {
        int a, b;
        trim_range(tr, ntr, rs[(size_t)tr], rovl, &a, &b);
        for (int r = a; r < b; ++r) { rtab[(size_t)(rf[(size_t)tr] + r)] = tr; rloc[(size_t)(rf[(size_t)tr] + r)] = r; }
    }
    for (int tc = 0; tc < ntc; ++tc) {
        int a, b;
        trim_range(tc, ntc, cs[(size_t)tc], covl, &a, &b);
        for (int c = a; c < b; ++c) { ctab[(size_t)(cf[(size_t)tc] + c)] = tc; cloc[(size_t)(cf[(size_t)tc] + c)] = c; }
    }
    std::vector<long long> toff((size_t)nt);
    std::vector<int> tpitch((size_t)nt);
    std::vector<float> add((size_t)nt);
    std::vector<uint32_t> coff((size_t)nt);
    uint32_t off = 0;
    for (int t = 0; t < nt; ++t) {
        toff[(size_t)t] = bo.lay.desc[(size_t)t].off;
        tpitch[(size_t)t] = bo.lay.desc[(size_t)t].pitch;
        add[(size_t)t] = (float)(tk[(size_t)t] * (2.0 * M_PI));
        coff[(size_t)t] = off;
        off += (uint32_t)bo.ncomp[(size_t)t];
    }
    long long *d_toff = ws<long long>(res, "m_toff", (size_t)nt);
    int *d_tpitch = ws<int>(res, "m_tpitch", (size_t)nt);
    int *d_rtab = ws<int>(res, "m_rtab", (size_t)nrow), *d_rloc = ws<int>(res, "m_rloc", (size_t)nrow);
    int *d_ctab = ws<int>(res, "m_ctab", (size_t)ncol), *d_cloc = ws<int>(res, "m_cloc", (size_t)ncol);
    float *d_add = ws<float>(res, "m_add", (size_t)nt);
    uint32_t *d_coff = ws<uint32_t>(res, "m_coff", (size_t)nt);
    CUPHU_HIP_CHECK(hipMemcpyAsync(d_toff, toff.data(), (size_t)nt * 8, hipMemcpyHostToDevice, st));
    CUPHU_HIP_CHECK(hipMemcpyAsync(d_tpitch, tpitch.data(), (size_t)nt * 4, hipMemcpyHostToDevice, st));
    CUPHU_HIP_CHECK(hipMemcpyAsync(d_rtab, rtab.data(), (size_t)nrow * 4, hipMemcpyHostToDevice, st));
    CUPHU_HIP_CHECK(hipMemcpyAsync(d_rloc, rloc.data(), (size_t)nrow * 4, hipMemcpyHostToDevice, st));
    CUPHU_HIP_CHECK(hipMemcpyAsync(d_ctab, ctab.data(), (size_t)ncol * 4, hipMemcpyHostToDevice, st));
    CUPHU_HIP_CHECK(hipMemcpyAsync(d_cloc, cloc.data(), (size_t)ncol * 4, hipMemcpyHostToDevice, st));
    CUPHU_HIP_CHECK(hipMemcpyAsync(d_add, add.data(), (size_t)nt * 4, hipMemcpyHostToDevice, st));
    CUPHU_HIP_CHECK(hipMemcpyAsync(d_coff, coff.data(), (size_t)nt * 4, hipMemcpyHostToDevice, st));
    long long npx = (long long)nrow * ncol;
    hipLaunchKernelGGL(k_merge, dim3((unsigned)ceil_div((size_t)npx, 256)), dim3(256), 0, st, bo.unw, bo.lab, d_toff,
                       d_tpitch, d_rtab, d_rloc, d_ctab, d_cloc, ntc, d_add, d_coff, nrow, ncol, d_unw, d_cc);
    CUPHU_HIP_CHECK(hipGetLastError());
}


template <class T>
static void scene_impl(const float2 *d_igram, const float *d_corr, const unsigned char *d_mask,
                       const unsigned char *d_conf, int nrow, int ncol, double nlooks, const CuPhuParams &prm,
                       int tile_size, int rovl, int covl, int ntr_req, int ntc_req, const CuPhuLaplaceOptions &o,
                       float *d_unw, uint32_t *d_cc, CuPhuSolveStats *stats, int *ntiles_out, DevRes &res,
                       hipStream_t st) {
    std::vector<int> rf(4096), rs(4096), cf(4096), cs(4096);
    int ntr = tile_axis(nrow, tile_size, rovl, ntr_req, rf.data(), rs.data(), 4096);
    int ntc = tile_axis(ncol, tile_size, covl, ntc_req, cf.data(), cs.data(), 4096);
    CUPHU_REQUIRE(ntr > 0 && ntc > 0 && ntr <= 4096 && ntc <= 4096, "invalid tiling");
    for (int t = 0; t < ntr; ++t) CUPHU_REQUIRE(rs[(size_t)t] >= 1, "tiling produced an empty tile (rows)");
    for (int t = 0; t < ntc; ++t) CUPHU_REQUIRE(cs[(size_t)t] >= 1, "tiling produced an empty tile (cols)");
    int nt = ntr * ntc;
    if (ntiles_out) *ntiles_out = nt;
    CostConsts cc = make_cost_consts(prm, nlooks);
    std::vector<IngestSrc> srcs((size_t)nt);
    std::vector<std::pair<int, int>> dims((size_t)nt);
    for (int tr = 0; tr < ntr; ++tr)
        for (int tc = 0; tc < ntc; ++tc) {
            IngestSrc &s = srcs[(size_t)(tr * ntc + tc)];
            size_t o0 = (size_t)rf[(size_t)tr] * ncol + cf[(size_t)tc];
            std::memset(&s, 0, sizeof(s));
            s.igram = d_igram + o0;
            s.corr = d_corr + o0;
            s.mask = d_mask ? d_mask + o0 : nullptr;
            s.src_stride = ncol;
            s.cc = cc;
            dims[(size_t)(tr * ntc + tc)] = {rs[(size_t)tr], cs[(size_t)tc]};
        }
    BatchOut bo;
    // merge_labels: components are labelled once on the stitched scene
    const bool scene_cc = d_cc != nullptr && o.merge_labels != 0 && nt > 1;
    run_batch<T>(dims, o, prm.minconncompfrac, prm.maxncomps, d_cc != nullptr && !scene_cc, res, st, stats, bo,
                 src_ingest<T>(srcs, false, res, st));
    if (nt == 1) {
        const InstDesc &d = bo.lay.desc[0];
        CUPHU_HIP_CHECK(hipMemcpy2DAsync(d_unw, (size_t)ncol * 4, bo.unw + d.off, (size_t)d.pitch * 4, (size_t)ncol * 4,
                                         (size_t)nrow, hipMemcpyDeviceToDevice, st));
        if (d_cc)
            CUPHU_HIP_CHECK(hipMemcpy2DAsync(d_cc, (size_t)ncol * 4, bo.lab + d.off, (size_t)d.pitch * 4,
                                             (size_t)ncol * 4, (size_t)nrow, hipMemcpyDeviceToDevice, st));
        CUPHU_HIP_CHECK(hipStreamSynchronize(st));
        return;
    }
    stitch_tiles(bo, nrow, ncol, ntr, ntc, rf, rs, cf, cs, rovl, covl, d_conf, d_mask != nullptr,
                 o.stitch_integer != 0, d_unw, scene_cc ? nullptr : d_cc, res, st);
    if (scene_cc) {
        const CcScratch scr = cc_scratch(res, (size_t)nrow * ncol);
        conncomp_run(d_unw, d_corr, d_mask, nrow, ncol, ncol, prm.minconncompfrac, prm.maxncomps, d_cc, st, 0.05f,
                     nullptr, nullptr, 0, &scr);
    }
    CUPHU_HIP_CHECK(hipStreamSynchronize(st));
}

// ── fixed-weight WLS (differentiable path) ──────────────────────────────────
template <class T>
__global__ void k_wls_ingest(const T *w, const float *phase, const T *rhs, InstDesc d, Planes<T> pl) {
    long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= (long long)d.nrow * d.pitch) return;
    int r = (int)(i / d.pitch), c = (int)(i % d.pitch);
    long long e = d.off + i;
    size_t nra = (size_t)(d.nrow - 1) * d.ncol;
    T wd = T(0), wr = T(0);
    float ph = 0.0f;
    if (c < d.ncol) {
        ph = phase[(size_t)r * d.ncol + c];
        if (r < d.nrow - 1) wd = w[(size_t)r * d.ncol + c];
        if (c < d.ncol - 1) wr = w[nra + (size_t)r * (d.ncol - 1) + c];
    }
    pl.w2[2 * e] = wd;
    pl.w2[2 * e + 1] = wr;
    pl.phi[e] = ph;
    pl.x[e] = T(0);
    (void)rhs;
}
// r = b = rhs (mean removed) for the adjoint solve; x = p = s = 0.
template <class T>
__global__ void k_wls_rhs(const T *rhs, double mean, InstDesc d, Planes<T> pl) {
    long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= (long long)d.nrow * d.pitch) return;
    int r = (int)(i / d.pitch), c = (int)(i % d.pitch);
    long long e = d.off + i;
    T v = (c < d.ncol) ? (T)((double)rhs[(size_t)r * d.ncol + c] - mean) : T(0);
    pl.b[e] = v;
    pl.r[e] = v;
}
template <class T>
__global__ void k_sum_part(const T *v, int nrow, int ncol, int stride, double *parts) {
    double s = 0.0;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < (long long)nrow * ncol; i += (long long)gridDim.x * 256)
        s += (double)v[(i / ncol) * stride + (i % ncol)];
    __shared__ double sh[256];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) parts[blockIdx.x] = sh[0];
}
template <class T>
__global__ void k_out_centered(const T *x, InstDesc d, double mean, T *out) {
    long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= (long long)d.nrow * d.ncol) return;
    out[i] = (T)((double)x[d.off + (i / d.ncol) * d.pitch + (i % d.ncol)] - mean);
}
template <class T>
__global__ void k_wls_grad(const T *lam, InstDesc d, const T *u, const float *phase, T *gw) {
    long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    long long nra = (long long)(d.nrow - 1) * d.ncol, na = nra + (long long)d.nrow * (d.ncol - 1);
    if (i >= na) return;
    long long t, h;
    if (i < nra) { t = i; h = i + d.ncol; }
    else { long long j = i - nra, r = j / (d.ncol - 1), c = j % (d.ncol - 1); t = r * d.ncol + c; h = t + 1; }
    auto lp = [&](long long k) { return lam[d.off + (k / d.ncol) * d.pitch + (k % d.ncol)]; };
    T g = Prec<T>::wrap((T)phase[h] - (T)phase[t]);
    gw[i] = -(lp(h) - lp(t)) * ((u[h] - u[t]) - g);
}

template <class T>
static double device_mean(const T *v, int nrow, int ncol, int stride, DevRes &res, hipStream_t st) {
    const int nb = 256;
    double *parts = ws<double>(res, "meanparts", nb);
    hipLaunchKernelGGL(k_sum_part<T>, dim3(nb), dim3(256), 0, st, v, nrow, ncol, stride, parts);
    CUPHU_HIP_CHECK(hipGetLastError());
    std::vector<double> h(nb);
    CUPHU_HIP_CHECK(hipMemcpyAsync(h.data(), parts, nb * sizeof(double), hipMemcpyDeviceToHost, st));
    CUPHU_HIP_CHECK(hipStreamSynchronize(st));
    double s = 0.0;
    for (double x : h) s += x;
    return s / ((double)nrow * ncol);
}

template <class T>
static void wls_impl(const T *d_w, const float *d_phase, const T *d_u, const T *d_gu, int nrow, int ncol, int max_iter,
                     double tol, T *d_out_u, T *d_gw, T *d_lambda, int *iters_out, double *relres_out, DevRes &res,
                     hipStream_t st) {
    CuPhuLaplaceOptions o;
    default_options(&o);
    o.n_irls = 0;
    o.max_iter = max_iter;
    o.tol = tol;
    o.solver = CUPHU_SOLVER_FUSED;
    o.precision = sizeof(T) == 8 ? 64 : 32;   // picks the default tile shape
    Layout L0 = make_layout({{nrow, ncol}}, TileShape{rows_per_wave(o.precision), 64});
    const TileShape sh = default_shape(L0, o);
    Layout L = make_layout({{nrow, ncol}}, sh);
    PoolSet<T> P = alloc_pools<T>(res, L.total_el);
    InstDesc d = L.desc[0];
    unsigned gsz = (unsigned)ceil_div((size_t)d.nrow * d.pitch, 256);
    hipLaunchKernelGGL(k_wls_ingest<T>, dim3(gsz), dim3(256), 0, st, d_w, d_phase, (const T *)nullptr, d, P.pl);
    CUPHU_HIP_CHECK(hipGetLastError());
    GroupDev g{};
    g.desc = ws<InstDesc>(res, "grpdesc", 1);
    int nb = d.tiles_x * d.tiles_y;
    g.blk = ws<int>(res, "grpblk", (size_t)nb);
    g.scal = ws<SolveScal>(res, "grpscal", 1);
    g.part = ws<double2>(res, "grppart", (size_t)nb);
    g.part4 = ws<double4>(res, "grppart4", (size_t)nb);
    g.scal2 = ws<SolveScal>(res, "grpscal2", 1);
    g.part4b = ws<double4>(res, "grppart4b", (size_t)nb);
    g.hstate[0] = (SolveScal *)res.pin("grpstate0", sizeof(SolveScal));
    g.hstate[1] = (SolveScal *)res.pin("grpstate1", sizeof(SolveScal));
    g.ninst = 1;
    g.nblk = nb;
    g.px = {(long long)nrow * ncol};
    InstDesc gd = d;
    gd.block_begin = 0;
    std::vector<int> blk((size_t)nb, 0);
    CUPHU_HIP_CHECK(hipMemcpyAsync(g.desc, &gd, sizeof(InstDesc), hipMemcpyHostToDevice, st));
    CUPHU_HIP_CHECK(hipMemcpyAsync(g.blk, blk.data(), (size_t)nb * sizeof(int), hipMemcpyHostToDevice, st));
    launch_assemble<T>(g.desc, g.blk, g.nblk, P.pl, sh, st);
    if (d_gu) {
        double m = device_mean<T>(d_gu, nrow, ncol, ncol, res, st);
        hipLaunchKernelGGL(k_wls_rhs<T>, dim3(gsz), dim3(256), 0, st, d_gu, m, gd, P.pl);
        CUPHU_HIP_CHECK(hipGetLastError());
    }
    std::vector<SolveScal> h;
    CuPhuSolveStats stt{};
    std::vector<CuPhuSolveStats *> sp = {&stt};
    solve_group<T>(P.pl, g, o, sh, true, st, sp, h);
    if (iters_out) *iters_out = stt.iters[0];
    if (relres_out) *relres_out = stt.relres[0];
    double mx = device_mean<T>(P.pl.x + d.off, nrow, ncol, d.pitch, res, st);
    unsigned gp = (unsigned)ceil_div((size_t)nrow * ncol, 256);
    if (d_out_u) {
        hipLaunchKernelGGL(k_out_centered<T>, dim3(gp), dim3(256), 0, st, P.pl.x, d, mx, d_out_u);
        CUPHU_HIP_CHECK(hipGetLastError());
    }
    if (d_lambda) {
        hipLaunchKernelGGL(k_out_centered<T>, dim3(gp), dim3(256), 0, st, P.pl.x, d, mx, d_lambda);
        CUPHU_HIP_CHECK(hipGetLastError());
    }
    if (d_gw) {
        size_t na = (size_t)(nrow - 1) * ncol + (size_t)nrow * (ncol - 1);
        hipLaunchKernelGGL(k_wls_grad<T>, dim3((unsigned)ceil_div(na, 256)), dim3(256), 0, st, P.pl.x, d, d_u,
                           d_phase, d_gw);
        CUPHU_HIP_CHECK(hipGetLastError());
    }
    CUPHU_HIP_CHECK(hipStreamSynchronize(st));
}

// ── flat arc arrays <-> pitched planes ──────────────────────────────────────
// sig points at the sigsq field of element 0; stride in shorts per element
// (2 for CuPhuSmoothCost arrays, 4 for CuPhuCost, 1 for a plain short array).
__global__ void k_costs_to_planes(const short *sig, int stride_shorts, const float *phase, InstDesc d, float *phi,
                                  float *corr, unsigned char *mask, short2 *sigsq) {
    long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= (long long)d.nrow * d.pitch) return;
    int r = (int)(i / d.pitch), c = (int)(i % d.pitch);
    long long e = d.off + i;
    size_t nra = (size_t)(d.nrow - 1) * d.ncol;
    short sd = 0, sr = 0;
    float ph = 0.0f;
    if (c < d.ncol) {
        ph = phase[(size_t)r * d.ncol + c];
        if (r < d.nrow - 1) sd = sig[((size_t)r * d.ncol + c) * stride_shorts];
        if (c < d.ncol - 1) sr = sig[(nra + (size_t)r * (d.ncol - 1) + c) * stride_shorts];
    }
    phi[e] = ph;
    corr[e] = (c < d.ncol) ? 1.0f : 0.0f;
    mask[e] = (c < d.ncol) ? 1 : 0;
    sigsq[e] = make_short2(sd, sr);
}
template <class T> __global__ void k_zero(T *p, long long n) {
    long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = T(0);
}
// w2 planes -> flat arcs (row arcs then column arcs)
template <class T> __global__ void k_planes_to_flat_w(const T *w2, InstDesc d, T *w) {
    long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    long long nra = (long long)(d.nrow - 1) * d.ncol, na = nra + (long long)d.nrow * (d.ncol - 1);
    if (i >= na) return;
    if (i < nra) w[i] = w2[2 * (d.off + (i / d.ncol) * d.pitch + (i % d.ncol))];
    else {
        long long j = i - nra;
        w[i] = w2[2 * (d.off + (j / (d.ncol - 1)) * d.pitch + (j % (d.ncol - 1))) + 1];
    }
}
// flat arcs -> w2 planes (zero where no arc exists)
template <class T> __global__ void k_flat_w_to_planes(const T *w, InstDesc d, T *w2) {
    long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= (long long)d.nrow * d.pitch) return;
    int r = (int)(i / d.pitch), c = (int)(i % d.pitch);
    size_t nra = (size_t)(d.nrow - 1) * d.ncol;
    T wd = T(0), wr = T(0);
    if (c < d.ncol) {
        if (r < d.nrow - 1) wd = w[(size_t)r * d.ncol + c];
        if (c < d.ncol - 1) wr = w[nra + (size_t)r * (d.ncol - 1) + c];
    }
    w2[2 * (d.off + i)] = wd;
    w2[2 * (d.off + i) + 1] = wr;
}

// ── complex split planes -> interleaved ─────────────────────────────────────
__global__ void k_interleave(const float *re, const float *im, float2 *out, size_t n) {
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) out[i] = make_float2(re[i], im[i]);
}

// Non-finite inputs as the reference's Python layer scrubs them before the
// solve (_unwrap.py:515-523): an interferogram element with a non-finite
// part becomes 0 + 0i, a non-finite coherence 0.
__global__ void k_scrub_inputs(float2 *ig, float *corr, size_t n) {
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float2 v = ig[i];
    if (!(isfinite(v.x) && isfinite(v.y))) ig[i] = make_float2(0.0f, 0.0f);
    if (!isfinite(corr[i])) corr[i] = 0.0f;
}

}  // namespace cuphu

using namespace cuphu;

// ════════════════════════════════════════════════════════════════════════════
// C ABI
// ════════════════════════════════════════════════════════════════════════════
extern "C" {

const char *cuphu_last_error(void) { return g_last_error.c_str(); }

void cuphu_default_laplace_options(CuPhuLaplaceOptions *o) {
    if (o) default_options(o);
}

// SNAPHU defaults (cuphu_solver.cpp:148-217 fills these from SNAPHU's DEF_*
// macros; SNAPHU's header is not available, the values below are SNAPHU's
// documented defaults).  Only nlooks/ncorrlooks, costscale, nshortcycle,
// sigsqshortmin, rhosconst1/2, cstd1-3, defothreshfactor, sigsqcorr,
// defomax, kperpdpsi/kpardpsi and the conncomp fields act on this path.
void cuphu_default_params(CuPhuParams *p) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->rhosconst1 = 1.3;
    p->rhosconst2 = 0.14;
    p->cstd1 = 0.4;
    p->cstd2 = 0.35;
    p->cstd3 = 0.06;
    p->defothreshfactor = 1.2;
    p->sigsqcorr = 0.05;
    p->nlooks = 23.8;
    p->ncorrlooks = 23.8;
    p->costscale = 100.0;
    p->nshortcycle = 200.0;
    p->sigsqshortmin = 1;
    p->kperpdpsi = 7;
    p->kpardpsi = 7;
    p->defomax = 1.2;
    p->minconncompfrac = 0.01;
    p->conncompthresh = 300;
    p->maxncomps = 32;
    // network-flow / topography fields (unused by the Laplace path)
    p->p = -1.0;
    p->maxcost = 1000.0;
    p->costscaleambight = 80.0;
    p->dnomincangle = 0.01;
    p->initmaxflow = 9999;
    p->arcmaxflowconst = 3;
    p->maxflow = 4;
    p->cs2scalefactor = 8;
    p->orbitradius = 7153000.0;
    p->earthradius = 6378000.0;
    p->baseline = 150.0;
    p->baselineangle = 0.25 * M_PI;
    p->transmitmode = 1;
    p->nlooksrange = 1;
    p->nlooksaz = 5;
    p->nlooksother = 1;
    p->ncorrlooksrange = 3;
    p->ncorrlooksaz = 15;
    p->nearrange = 831000.0;
    p->dr = 8.0;
    p->da = 20.0;
    p->rangeres = 10.0;
    p->azres = 6.0;
    p->lambda = 0.0565647;
    p->kds = 0.02;
    p->specularexp = 8.0;
    p->dzrcritfactor = 2.0;
    p->dzeimin = -4.0;
    p->laywidth = 16;
    p->layminei = 1.25;
    p->sloperatiofactor = 1.18;
    p->sigsqei = 100.0;
    p->dzlaypeak = -2.0;
    p->azdzfactor = 0.99;
    p->dzeifactor = 4.0;
    p->dzeiweight = 0.5;
    p->dzlayfactor = 1.0;
    p->layconst = 0.9;
    p->layfalloffconst = 2.0;
    p->sigsqlayfactor = 0.1;
    p->defoazdzfactor = 1.0;
    p->defolayconst = 0.9;
    p->initdzr = 2048.0;
    p->initdzstep = 100.0;
    p->threshold = 0.001;
}

// cuphu_solver.cpp:219-238
void cuphu_default_tile_params(CuPhuTileParams *tp) {
    if (!tp) return;
    std::memset(tp, 0, sizeof(*tp));
    tp->ntilerow = 1;
    tp->ntilecol = 1;
    tp->rowovrlp = 0;
    tp->colovrlp = 0;
    tp->tilecostthresh = 500;
    tp->minregionsize = 100;
    tp->nproc = 1;
    tp->ngpustreams = 2;
    tp->laplace_neighbor_feedback_feather = 200;
}

// cuphu_solver.cpp:240-249
void cuphu_default_bridge_params(CuPhuBridgeParams *bp) {
    if (!bp) return;
    std::memset(bp, 0, sizeof(*bp));
    bp->radius = 500;
    bp->min_num_pixel = 14;
    bp->erosion_size = 2;
    bp->max_boundary_samples = 4096;
    bp->ramp_type = CUPHU_RAMP_NONE;
    bp->ramp_max_num_sample = 1000000;
}

int cuphu_set_flow_solver(cuphu_flow_solver_fn fn, void *user) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    g_flow_solver = fn;
    g_flow_solver_user = fn ? user : nullptr;
    return 0;
}

int cuphu_set_profiling(int enabled) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    g_prof.on = enabled != 0;
    g_prof.acc = CuPhuKernelProfile{};
    return 0;
}

int cuphu_get_profile(CuPhuKernelProfile *out) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    if (!out) return CUPHU_ERR_INVALID;
    *out = g_prof.acc;
    return 0;
}

int cuphu_gpu_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int cuphu_gpu_name(int device_id, char *buf, int buflen) {
    return guarded([&]() {
        CUPHU_REQUIRE(buf && buflen > 0, "buffer required");
        hipDeviceProp_t prop;
        CUPHU_HIP_CHECK(hipGetDeviceProperties(&prop, device_id));
        std::snprintf(buf, (size_t)buflen, "%s", prop.name);
        return 0;
    });
}

int cuphu_gpu_info(int device_id, long long *total_mem, int *cu_count, char *arch_buf, int arch_buflen) {
    return guarded([&]() {
        hipDeviceProp_t prop;
        CUPHU_HIP_CHECK(hipGetDeviceProperties(&prop, device_id));
        if (total_mem) *total_mem = (long long)prop.totalGlobalMem;
        if (cu_count) *cu_count = prop.multiProcessorCount;
        if (arch_buf && arch_buflen > 0) std::snprintf(arch_buf, (size_t)arch_buflen, "%s", prop.gcnArchName);
        return 0;
    });
}

static CuPhuLaplaceOptions opts_or_default(const CuPhuLaplaceOptions *o) {
    CuPhuLaplaceOptions r;
    if (o) r = *o;
    else default_options(&r);
    check_options(r);
    return r;
}

int cuphu_laplace_batch_gpu(const CuPhuInstance *inst, int ninst, CuPhuCostMode cost_mode, const CuPhuParams *base,
                            const CuPhuLaplaceOptions *opt, CuPhuSolveStats *stats, cuphu_stream_t stream) {
    return guarded([&]() {
        std::lock_guard<std::recursive_mutex> lk(g_mu);
        TraceRange tr_entry("cuphu:laplace_batch");
        CUPHU_REQUIRE(inst && ninst > 0 && base, "instances and params required");
        CUPHU_REQUIRE(cost_mode >= CUPHU_COST_SMOOTH && cost_mode <= CUPHU_COST_TOPO, "invalid cost mode");
        CuPhuLaplaceOptions o = opts_or_default(opt);
        DevRes &res = devres();
        StreamJoin join(res, (hipStream_t)stream);
        hipStream_t st = res.stream;
        std::vector<IngestSrc> srcs((size_t)ninst);
        std::vector<std::pair<int, int>> dims((size_t)ninst);
        bool has_u0 = false, want_cc = false;
        for (int i = 0; i < ninst; ++i) {
            const CuPhuInstance &I = inst[i];
            CUPHU_REQUIRE(I.nrow >= 1 && I.ncol >= 1, "instance dimensions must be >= 1");
            CUPHU_REQUIRE(I.d_igram || I.d_phase, "instance needs d_igram or d_phase");
            CUPHU_REQUIRE(I.d_corr && I.d_unw, "instance needs d_corr and d_unw");
            CUPHU_REQUIRE(I.nlooks >= 1.0, "nlooks must be >= 1");
            IngestSrc &s = srcs[(size_t)i];
            std::memset(&s, 0, sizeof(s));
            s.igram = reinterpret_cast<const float2 *>(I.d_igram);
            s.phase = I.d_phase;
            s.corr = I.d_corr;
            s.mask = I.d_mask;
            s.u0 = I.d_u0;
            s.src_stride = I.ncol;
            s.cc = make_cost_consts(*base, I.nlooks);
            has_u0 = has_u0 || I.d_u0;
            want_cc = want_cc || I.d_conncomp;
            dims[(size_t)i] = {I.nrow, I.ncol};
        }
        BatchOut bo;
        if (o.precision == 64)
            run_batch<double>(dims, o, base->minconncompfrac, base->maxncomps, want_cc, res, st, stats, bo,
                              src_ingest<double>(srcs, has_u0, res, st));
        else
            run_batch<float>(dims, o, base->minconncompfrac, base->maxncomps, want_cc, res, st, stats, bo,
                             src_ingest<float>(srcs, has_u0, res, st));
        for (int i = 0; i < ninst; ++i) {
            const InstDesc &d = bo.lay.desc[(size_t)i];
            const CuPhuInstance &I = inst[i];
            CUPHU_HIP_CHECK(hipMemcpy2DAsync(I.d_unw, (size_t)I.ncol * 4, bo.unw + d.off, (size_t)d.pitch * 4,
                                             (size_t)I.ncol * 4, (size_t)I.nrow, hipMemcpyDeviceToDevice, st));
            if (I.d_conncomp)
                CUPHU_HIP_CHECK(hipMemcpy2DAsync(I.d_conncomp, (size_t)I.ncol * 4, bo.lab + d.off, (size_t)d.pitch * 4,
                                                 (size_t)I.ncol * 4, (size_t)I.nrow, hipMemcpyDeviceToDevice, st));
        }
        CUPHU_HIP_CHECK(hipStreamSynchronize(st));
        return 0;
    });
}

int cuphu_unwrap_scene_gpu(const float *d_igram, const float *d_corr, const unsigned char *d_mask,
                           const unsigned char *d_orig_mask, int nrow, int ncol, double nlooks,
                           CuPhuCostMode cost_mode, const CuPhuParams *params, int tile_size, int overlap,
                           int ntilerow, int ntilecol, const CuPhuLaplaceOptions *opt, float *d_unw,
                           uint32_t *d_conncomp, CuPhuSolveStats *stats, int *ntiles_out, cuphu_stream_t stream) {
    return guarded([&]() {
        std::lock_guard<std::recursive_mutex> lk(g_mu);
        TraceRange tr_entry("cuphu:unwrap_scene");
        CUPHU_REQUIRE(d_igram && d_corr && d_unw && params, "igram, corr, unw and params required");
        CUPHU_REQUIRE(nrow >= 1 && ncol >= 1, "dimensions must be >= 1");
        CUPHU_REQUIRE(nlooks >= 1.0, "nlooks must be >= 1");
        CUPHU_REQUIRE(cost_mode >= CUPHU_COST_SMOOTH && cost_mode <= CUPHU_COST_TOPO, "invalid cost mode");
        CUPHU_REQUIRE(overlap >= 0 && tile_size >= 0, "tile_size and overlap must be >= 0");
        CUPHU_REQUIRE(tile_size == 0 || tile_size > overlap, "tile_size must exceed overlap");
        CuPhuLaplaceOptions o = opts_or_default(opt);
        DevRes &res = devres();
        StreamJoin join(res, (hipStream_t)stream);
        const unsigned char *conf = d_orig_mask;
        if (o.precision == 64)
            scene_impl<double>(reinterpret_cast<const float2 *>(d_igram), d_corr, d_mask, conf, nrow, ncol, nlooks,
                               *params, tile_size, overlap, overlap, ntilerow, ntilecol, o, d_unw, d_conncomp, stats,
                               ntiles_out, res, res.stream);
        else
            scene_impl<float>(reinterpret_cast<const float2 *>(d_igram), d_corr, d_mask, conf, nrow, ncol, nlooks,
                              *params, tile_size, overlap, overlap, ntilerow, ntilecol, o, d_unw, d_conncomp, stats,
                              ntiles_out, res, res.stream);
        return 0;
    });
}

int cuphu_incrcost_gpu(const CuPhuSmoothCost *d_costs, const short *d_flows, int narcs, long nshortcycle, int nflow,
                       short *d_poscost, short *d_negcost, int *no_improvement, cuphu_stream_t stream) {
    return guarded([&]() {
        std::lock_guard<std::recursive_mutex> lk(g_mu);
        CUPHU_REQUIRE(narcs >= 0 && nflow >= 1 && nshortcycle >= 1, "invalid argument");
        CUPHU_REQUIRE(narcs == 0 || (d_costs && d_flows && d_poscost && d_negcost), "null argument");
        DevRes &res = devres();
        StreamJoin join(res, (hipStream_t)stream);
        int *mins = ws<int>(res, "incrmins", 2);
        int flag = util_incrcost(reinterpret_cast<const short *>(d_costs), d_flows, (size_t)narcs, nshortcycle, nflow,
                                 d_poscost, d_negcost, mins, res.stream);
        if (no_improvement) *no_improvement = flag;
        return 0;
    });
}

// ── MCF / MST initialisation through a registered host solver ──────────────

__global__ void k_mask_arc_weights(const unsigned char *mask, int nrow, int ncol, short *w) {
    long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    long long nra = (long long)(nrow - 1) * ncol, na = nra + (long long)nrow * (ncol - 1);
    if (i >= na) return;
    long long t, h;
    if (i < nra) { t = i; h = i + ncol; }
    else { long long j = i - nra, r = j / (ncol - 1), c = j % (ncol - 1); t = r * ncol + c; h = t + 1; }
    w[i] = (mask[t] != 0 && mask[h] != 0) ? 1 : 0;
}

// One tile on the MCF path (cuphu_solver.cpp:370-510 costs, :512-880 the
// CPU network-flow solve, :900-993 the GPU tail): wrapped phase and SNAPHU
// costs on the GPU, the registered host solver for the integer flows, then
// on the GPU the incremental costs at the solved flows (smooth costs), the
// flow integration and the connected components gated by the incremental
// costs.  Inputs and outputs are device arrays.
static int mcf_tile(const float2 *d_ig, const float *d_corr, const unsigned char *d_mask, int nrow, int ncol,
                    CuPhuCostMode cost_mode, CuPhuInitMethod init_meth, const CuPhuParams &prm, float *d_unw,
                    uint32_t *d_cc, DevRes &res, hipStream_t st) {
    size_t n = (size_t)nrow * ncol;
    size_t nra = (size_t)(nrow - 1) * ncol, na = nra + (size_t)nrow * (ncol - 1);
    int defo = cost_mode != CUPHU_COST_SMOOTH;
    size_t esz = defo ? 8 : 4;
    DevBuf<float> d_ph(n);
    DevBuf<short> d_w(d_mask ? na + 1 : 0), d_costs(na * esz / 2 + 2), d_flows(na + 1), d_pos(na + 1), d_neg(na + 1);
    phase_from_igram(d_ig, d_ph.get(), n, st);
    if (d_mask && na) {
        hipLaunchKernelGGL(k_mask_arc_weights, dim3((unsigned)ceil_div(na, 256)), dim3(256), 0, st, d_mask, nrow, ncol,
                           d_w.get());
        CUPHU_HIP_CHECK(hipGetLastError());
    }
    CostConsts k = make_cost_consts(prm, prm.ncorrlooks);
    util_costs(d_ph.get(), d_corr, d_mask ? d_w.get() : nullptr, nrow, ncol, k, prm.kperpdpsi, prm.kpardpsi, defo,
               d_costs.get(), st);
    std::vector<unsigned char> h_costs(na * esz + 8);
    std::vector<float> h_ph(n);
    std::vector<short> h_flows(na + 1, 0);
    if (na) CUPHU_HIP_CHECK(hipMemcpyAsync(h_costs.data(), d_costs.get(), na * esz, hipMemcpyDeviceToHost, st));
    CUPHU_HIP_CHECK(hipMemcpyAsync(h_ph.data(), d_ph.get(), n * 4, hipMemcpyDeviceToHost, st));
    CUPHU_HIP_CHECK(hipStreamSynchronize(st));
    int rc;
    {
        TraceRange tr("cuphu:host_flow_solver");
        rc = g_flow_solver(h_costs.data(), (int)cost_mode, h_ph.data(), nrow, ncol, (int)init_meth, h_flows.data(),
                           g_flow_solver_user);
    }
    if (rc != 0) throw Error(CUPHU_ERR_INTERNAL, "the registered network-flow solver failed (" + std::to_string(rc) + ")");
    if (na) CUPHU_HIP_CHECK(hipMemcpyAsync(d_flows.get(), h_flows.data(), na * 2, hipMemcpyHostToDevice, st));
    const short *pos = nullptr, *neg = nullptr;
    if (!defo) {
        int *mins = ws<int>(res, "incrmins", 2);
        util_incrcost(d_costs.get(), d_flows.get(), na, (long)prm.nshortcycle, 1, d_pos.get(), d_neg.get(), mins, st);
        pos = d_pos.get();
        neg = d_neg.get();
    }
    util_integrate(d_ph.get(), d_flows.get(), d_flows.get() + nra, nrow, ncol, d_unw, st);
    int ncomp = conncomp_run(d_unw, d_corr, d_mask, nrow, ncol, ncol, prm.minconncompfrac, prm.maxncomps, d_cc, st,
                             0.05f, pos, neg, prm.conncompthresh);
    CUPHU_HIP_CHECK(hipStreamSynchronize(st));
    return ncomp;
}

// Tiled 'mcf'/'mst' (cuphu_solver.cpp:1753-1808 with solve_tile's network-
// flow branch): the reference tile formula, each tile copied out of the
// scene and solved through mcf_tile (costs -> registered solver ->
// incremental costs -> integration -> gated conncomp, one tile after the
// other: the host solver is the serial part), then stitched by whole 2 pi
// cycles (:1928: network-flow tiles differ by exact integer cycles) with
// the labels concatenated.
static void mcf_scene(const float2 *d_ig, const float *d_corr, const unsigned char *d_mask,
                      const unsigned char *d_conf, int nrow, int ncol, int rovl, int covl, int ntr_req, int ntc_req,
                      CuPhuCostMode cost_mode, CuPhuInitMethod init_meth, const CuPhuParams &prm, float *d_unw,
                      uint32_t *d_cc, DevRes &res, hipStream_t st) {
    std::vector<int> rf(4096), rs(4096), cf(4096), cs(4096);
    int ntr = tile_axis(nrow, 0, rovl, ntr_req, rf.data(), rs.data(), 4096);
    int ntc = tile_axis(ncol, 0, covl, ntc_req, cf.data(), cs.data(), 4096);
    CUPHU_REQUIRE(ntr > 0 && ntc > 0 && ntr <= 4096 && ntc <= 4096, "invalid tiling");
    for (int t = 0; t < ntr; ++t) CUPHU_REQUIRE(rs[(size_t)t] >= 1, "tiling produced an empty tile (rows)");
    for (int t = 0; t < ntc; ++t) CUPHU_REQUIRE(cs[(size_t)t] >= 1, "tiling produced an empty tile (cols)");
    const int nt = ntr * ntc;
    BatchOut bo;
    for (int t = 0; t < nt; ++t) {
        InstDesc d{};
        d.nrow = rs[(size_t)(t / ntc)];
        d.ncol = cs[(size_t)(t % ntc)];
        d.pitch = d.wcols = d.ncol;
        d.off = bo.lay.total_el;
        bo.lay.total_el += (long long)d.nrow * d.ncol;
        bo.lay.desc.push_back(d);
    }
    const size_t tot = (size_t)bo.lay.total_el;
    float2 *ig = ws<float2>(res, "mcf_ig", tot);
    bo.unw = ws<float>(res, "mcf_unw", tot);
    bo.lab = ws<uint32_t>(res, "mcf_lab", tot);
    bo.corr = ws<float>(res, "mcf_corr", tot);
    bo.mask = d_mask ? ws<unsigned char>(res, "mcf_mask", tot) : nullptr;
    bo.ncomp.assign((size_t)nt, 0);
    for (int t = 0; t < nt; ++t) {
        const InstDesc &d = bo.lay.desc[(size_t)t];
        const size_t o0 = (size_t)rf[(size_t)(t / ntc)] * ncol + cf[(size_t)(t % ntc)];
        CUPHU_HIP_CHECK(hipMemcpy2DAsync(ig + d.off, (size_t)d.ncol * 8, d_ig + o0, (size_t)ncol * 8, (size_t)d.ncol * 8,
                                         (size_t)d.nrow, hipMemcpyDeviceToDevice, st));
        CUPHU_HIP_CHECK(hipMemcpy2DAsync(bo.corr + d.off, (size_t)d.ncol * 4, d_corr + o0, (size_t)ncol * 4,
                                         (size_t)d.ncol * 4, (size_t)d.nrow, hipMemcpyDeviceToDevice, st));
        if (d_mask)
            CUPHU_HIP_CHECK(hipMemcpy2DAsync(bo.mask + d.off, (size_t)d.ncol, d_mask + o0, (size_t)ncol, (size_t)d.ncol,
                                             (size_t)d.nrow, hipMemcpyDeviceToDevice, st));
        bo.ncomp[(size_t)t] = mcf_tile(ig + d.off, bo.corr + d.off, d_mask ? bo.mask + d.off : nullptr, d.nrow, d.ncol,
                                       cost_mode, init_meth, prm, bo.unw + d.off, bo.lab + d.off, res, st);
    }
    stitch_tiles(bo, nrow, ncol, ntr, ntc, rf, rs, cf, cs, rovl, covl, d_conf, d_mask != nullptr, true, d_unw, d_cc,
                 res, st);
    CUPHU_HIP_CHECK(hipStreamSynchronize(st));
}

// Internal tile seams of the reference tile formula (cuphu_solver.cpp:1475-
// 1476, :1586, :1623): each tile's first row/column past the first tile,
// plus half the overlap.
static std::vector<int> seams(int n, int ovl, int ntiles) {
    std::vector<int> out;
    if (ntiles < 2) return out;
    std::vector<int> f((size_t)ntiles), sz((size_t)ntiles);
    if (cuphu::tile_axis(n, 0, ovl, ntiles, f.data(), sz.data(), ntiles) != ntiles) return out;
    for (int t = 1; t < ntiles; ++t) out.push_back(f[(size_t)t] + ovl / 2);
    return out;
}

int cuphu_unwrap_ex(const float *igram_r, const float *igram_i, const float *igram_c64, const float *corr,
                    const unsigned char *mask, int nrow, int ncol, CuPhuCostMode cost_mode, CuPhuInitMethod init_meth,
                    const CuPhuParams *params, const CuPhuTileParams *tile, const CuPhuBridgeParams *bridge,
                    const unsigned char *orig_mask, int gpu_id, const CuPhuLaplaceOptions *opt, CuPhuSolveStats *stats,
                    CuPhuResult *result) {
    return guarded([&]() {
        std::lock_guard<std::recursive_mutex> lk(g_mu);
        TraceRange tr_entry("cuphu:unwrap");
        CUPHU_REQUIRE(result && corr && params && tile, "null argument");
        CUPHU_REQUIRE(igram_c64 || (igram_r && igram_i), "igram required");
        CUPHU_REQUIRE(nrow >= 1 && ncol >= 1, "dimensions must be >= 1");
        result->unw = nullptr;
        result->conncomp = nullptr;
        result->nrow = nrow;
        result->ncol = ncol;
        const bool mcf = init_meth != CUPHU_INIT_LAPLACE;
        if (mcf && !g_flow_solver)
            throw Error(CUPHU_ERR_UNSUPPORTED,
                        "init 'mcf'/'mst' needs a network-flow solver (SNAPHU's CPU solver is not part of this "
                        "library); register one with cuphu_set_flow_solver");
        CUPHU_REQUIRE(cost_mode >= CUPHU_COST_SMOOTH && cost_mode <= CUPHU_COST_TOPO, "invalid cost mode");
        if (tile->single_tile_reoptimize && tile->ntilerow * tile->ntilecol > 1)
            throw Error(CUPHU_ERR_UNSUPPORTED, "single_tile_reoptimize needs the CPU network-flow solver");
        CUPHU_REQUIRE(tile->ntilerow >= 1 && tile->ntilecol >= 1, "ntilerow/ntilecol must be >= 1");
        CUPHU_REQUIRE(tile->rowovrlp >= 0 && tile->colovrlp >= 0, "tile overlaps must be >= 0");
        CUPHU_REQUIRE(params->ncorrlooks >= 1.0, "nlooks must be >= 1");
        CuPhuLaplaceOptions o = opts_or_default(opt);
        DeviceGuard dg(gpu_id);
        DevRes &res = devres();
        hipStream_t st = res.stream;
        size_t n = (size_t)nrow * ncol;
        result->unw = (float *)std::malloc(n * sizeof(float));
        result->conncomp = (uint32_t *)std::malloc(n * sizeof(uint32_t));
        if (!result->unw || !result->conncomp) {
            std::free(result->unw); std::free(result->conncomp);
            result->unw = nullptr; result->conncomp = nullptr;
            return (int)CUPHU_ERR_NOMEM;
        }
        try {
            // the device copies of the caller's arrays live in the per-device
            // workspace, like the solver's planes, so repeated calls do not
            // allocate and free 20 B per pixel each time
            float *d_corr = ws<float>(res, "ex_corr", n), *d_unw = ws<float>(res, "ex_unw", n);
            float2 *d_ig = ws<float2>(res, "ex_ig", n);
            uint32_t *d_cc = ws<uint32_t>(res, "ex_cc", n);
            unsigned char *d_mask = mask ? ws<unsigned char>(res, "ex_mask", n) : nullptr;
            unsigned char *d_orig = orig_mask ? ws<unsigned char>(res, "ex_orig", n) : nullptr;
            if (igram_c64) {
                copy_h2d_staged(d_ig, igram_c64, n * 8, res, st);
            } else {
                DevBuf<float> d_re(n), d_im(n);
                copy_h2d_staged(d_re.get(), igram_r, n * 4, res, st);
                copy_h2d_staged(d_im.get(), igram_i, n * 4, res, st);
                hipLaunchKernelGGL(k_interleave, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, st, d_re.get(),
                                   d_im.get(), d_ig, n);
                CUPHU_HIP_CHECK(hipGetLastError());
                CUPHU_HIP_CHECK(hipStreamSynchronize(st));
            }
            copy_h2d_staged(d_corr, corr, n * 4, res, st);
            if (mask) copy_h2d_staged(d_mask, mask, n, res, st);
            if (orig_mask) copy_h2d_staged(d_orig, orig_mask, n, res, st);
            hipLaunchKernelGGL(k_scrub_inputs, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, st, d_ig,
                               d_corr, n);
            CUPHU_HIP_CHECK(hipGetLastError());
            // tile-stitching confidence mask: orig_mask, else mask (cuphu_solver.cpp:1696)
            const unsigned char *conf = orig_mask ? d_orig : nullptr;
            int single = (tile->ntilerow == 1 && tile->ntilecol == 1);
            int rov = single ? 0 : tile->rowovrlp, cov = single ? 0 : tile->colovrlp;
            if (mcf && single)
                mcf_tile(d_ig, d_corr, mask ? d_mask : nullptr, nrow, ncol, cost_mode, init_meth,
                         *params, d_unw, d_cc, res, st);
            else if (mcf)
                mcf_scene(d_ig, d_corr, mask ? d_mask : nullptr, conf, nrow, ncol, rov, cov,
                          tile->ntilerow, tile->ntilecol, cost_mode, init_meth, *params, d_unw, d_cc, res,
                          st);
            else if (o.precision == 64)
                scene_impl<double>(d_ig, d_corr, mask ? d_mask : nullptr, conf, nrow, ncol,
                                   params->ncorrlooks, *params, 0, rov, cov, tile->ntilerow, tile->ntilecol, o,
                                   d_unw, d_cc, stats, nullptr, res, st);
            else
                scene_impl<float>(d_ig, d_corr, mask ? d_mask : nullptr, conf, nrow, ncol,
                                  params->ncorrlooks, *params, 0, rov, cov, tile->ntilerow, tile->ntilecol, o,
                                  d_unw, d_cc, stats, nullptr, res, st);
            // optional post-passes in the reference's order (cuphu_solver.cpp:2037-2095); like the
            // reference, the single-tile path skips both (it returns before them, :1714-1723)
            if (!mcf && tile->laplace_neighbor_feedback && !single)
                neighbor_feedback_dev(d_unw, mask ? d_mask : nullptr, nrow, ncol,
                                      seams(ncol, cov, tile->ntilecol), seams(nrow, rov, tile->ntilerow),
                                      tile->laplace_neighbor_feedback_feather, st);
            if (tile->fix_cycle_spikes && !single)
                fix_cycle_spikes_dev(d_unw, mask ? d_mask : nullptr, nrow, ncol, st);
            // phase bridging, on single tiles too (:1721-1722, :2097-2098)
            if (bridge && bridge->enabled)
                bridge_dev(d_unw, d_cc, mask ? d_mask : nullptr, nrow, ncol, *bridge, st);
            copy_d2h_staged(result->unw, d_unw, n * 4, res, st);
            copy_d2h_staged(result->conncomp, d_cc, n * 4, res, st);
        } catch (...) {
            std::free(result->unw); std::free(result->conncomp);
            result->unw = nullptr; result->conncomp = nullptr;
            throw;
        }
        return 0;
    });
}

int cuphu_unwrap(const float *igram_r, const float *igram_i, const float *corr, const float *mag,
                 const unsigned char *mask, int nrow, int ncol, CuPhuCostMode cost_mode, CuPhuInitMethod init_meth,
                 const CuPhuParams *params, const CuPhuTileParams *tile, const CuPhuBridgeParams *bridge,
                 const unsigned char *orig_mask, int gpu_id, CuPhuResult *result) {
    (void)mag;   // the Laplace branch never reads the magnitude (cuphu_solver.cpp:458-510)
    if (!igram_r || !igram_i) {
        set_last_error("igram_r and igram_i required");
        return CUPHU_ERR_INVALID;
    }
    return cuphu_unwrap_ex(igram_r, igram_i, nullptr, corr, mask, nrow, ncol, cost_mode, init_meth, params, tile, bridge,
                           orig_mask, gpu_id, nullptr, nullptr, result);
}

int cuphu_regrow_conncomp(const float *unw, const float *corr, const unsigned char *mask, int nrow, int ncol,
                          double corr_thresh, double minfrac, long maxncomps, int gpu_id, uint32_t *labels) {
    return guarded([&]() {
        std::lock_guard<std::recursive_mutex> lk(g_mu);
        CUPHU_REQUIRE(unw && corr && labels && nrow >= 1 && ncol >= 1, "invalid argument");
        DeviceGuard dg(gpu_id);
        DevRes &res = devres();
        hipStream_t st = res.stream;
        size_t n = (size_t)nrow * ncol;
        DevBuf<float> du(n), dc(n);
        DevBuf<unsigned char> dm(mask ? n : 0);
        DevBuf<uint32_t> dl(n);
        CUPHU_HIP_CHECK(hipMemcpyAsync(du.get(), unw, n * 4, hipMemcpyHostToDevice, st));
        CUPHU_HIP_CHECK(hipMemcpyAsync(dc.get(), corr, n * 4, hipMemcpyHostToDevice, st));
        if (mask) CUPHU_HIP_CHECK(hipMemcpyAsync(dm.get(), mask, n, hipMemcpyHostToDevice, st));
        conncomp_run(du.get(), dc.get(), mask ? dm.get() : nullptr, nrow, ncol, ncol, minfrac, maxncomps, dl.get(), st,
                     (float)corr_thresh);
        CUPHU_HIP_CHECK(hipMemcpyAsync(labels, dl.get(), n * 4, hipMemcpyDeviceToHost, st));
        CUPHU_HIP_CHECK(hipStreamSynchronize(st));
        return 0;
    });
}

int cuphu_build_costs_gpu(const float *igram_r, const float *igram_i, const float *corr, const float *mag,
                          const short *weights, int nrow, int ncol, CuPhuCostMode cost_mode, const CuPhuParams *params,
                          int gpu_id, void **costs_out, size_t *cost_elem_sz) {
    (void)mag;
    return guarded([&]() {
        std::lock_guard<std::recursive_mutex> lk(g_mu);
        CUPHU_REQUIRE(igram_r && igram_i && corr && params && costs_out && cost_elem_sz, "null argument");
        CUPHU_REQUIRE(nrow >= 1 && ncol >= 1, "dimensions must be >= 1");
        CUPHU_REQUIRE(cost_mode >= CUPHU_COST_SMOOTH && cost_mode <= CUPHU_COST_TOPO, "invalid cost mode");
        *costs_out = nullptr;
        DeviceGuard dg(gpu_id);
        DevRes &res = devres();
        hipStream_t st = res.stream;
        size_t n = (size_t)nrow * ncol, na = (size_t)(nrow - 1) * ncol + (size_t)nrow * (ncol - 1);
        int defo = cost_mode != CUPHU_COST_SMOOTH;
        size_t esz = defo ? 8 : 4;
        DevBuf<float> d_re(n), d_im(n), d_corr(n), d_ph(n);
        DevBuf<float2> d_ig(n);
        DevBuf<short> d_w(weights ? na : 0), d_out(na * esz / 2 + 1);
        CUPHU_HIP_CHECK(hipMemcpyAsync(d_re.get(), igram_r, n * 4, hipMemcpyHostToDevice, st));
        CUPHU_HIP_CHECK(hipMemcpyAsync(d_im.get(), igram_i, n * 4, hipMemcpyHostToDevice, st));
        CUPHU_HIP_CHECK(hipMemcpyAsync(d_corr.get(), corr, n * 4, hipMemcpyHostToDevice, st));
        if (weights && na) CUPHU_HIP_CHECK(hipMemcpyAsync(d_w.get(), weights, na * 2, hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_interleave, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, st, d_re.get(), d_im.get(),
                           d_ig.get(), n);
        CUPHU_HIP_CHECK(hipGetLastError());
        phase_from_igram(d_ig.get(), d_ph.get(), n, st);
        CostConsts k = make_cost_consts(*params, params->ncorrlooks);
        util_costs(d_ph.get(), d_corr.get(), weights ? d_w.get() : nullptr, nrow, ncol, k, params->kperpdpsi,
                   params->kpardpsi, defo, d_out.get(), st);
        void *h = std::malloc(na * esz + 1);
        if (!h) return (int)CUPHU_ERR_NOMEM;
        CUPHU_HIP_CHECK(hipMemcpyAsync(h, d_out.get(), na * esz, hipMemcpyDeviceToHost, st));
        CUPHU_HIP_CHECK(hipStreamSynchronize(st));
        *costs_out = h;
        *cost_elem_sz = esz;
        return (int)esz;
    });
}

static int device_costs(const float *d_phase, const float *d_corr, const short *d_roww, const short *d_colw, int nrow,
                        int ncol, const CuPhuParams *params, int kperp, int kpar, int defo, void **out,
                        cuphu_stream_t stream) {
    return guarded([&]() {
        std::lock_guard<std::recursive_mutex> lk(g_mu);
        CUPHU_REQUIRE(d_phase && d_corr && params && out, "null argument");
        CUPHU_REQUIRE(nrow >= 1 && ncol >= 1, "dimensions must be >= 1");
        CUPHU_REQUIRE((d_roww == nullptr) == (d_colw == nullptr), "row and column weights go together");
        DevRes &res = devres();
        StreamJoin join(res, (hipStream_t)stream);
        hipStream_t st = res.stream;
        size_t nra = (size_t)(nrow - 1) * ncol, nca = (size_t)nrow * (ncol - 1), na = nra + nca;
        size_t esz = defo ? 8 : 4;
        DevBuf<short> w;
        if (d_roww) {
            w.alloc(na + 1);
            if (nra) CUPHU_HIP_CHECK(hipMemcpyAsync(w.get(), d_roww, nra * 2, hipMemcpyDeviceToDevice, st));
            if (nca) CUPHU_HIP_CHECK(hipMemcpyAsync(w.get() + nra, d_colw, nca * 2, hipMemcpyDeviceToDevice, st));
        }
        void *d = nullptr;
        CUPHU_HIP_CHECK(hipMalloc(&d, na * esz + 8));
        try {
            CostConsts k = make_cost_consts(*params, params->ncorrlooks);
            util_costs(d_phase, d_corr, d_roww ? w.get() : nullptr, nrow, ncol, k, kperp, kpar, defo, (short *)d, st);
        } catch (...) {
            (void)hipFree(d);
            throw;
        }
        *out = d;
        return 0;
    });
}

int cuphu_build_smooth_costs_gpu(const float *d_phase, const float *d_corr, const short *d_roww, const short *d_colw,
                                 int nrow, int ncol, const CuPhuParams *params, int kperpdpsi, int kpardpsi,
                                 CuPhuSmoothCost **d_costs_out, cuphu_stream_t stream) {
    return device_costs(d_phase, d_corr, d_roww, d_colw, nrow, ncol, params, kperpdpsi, kpardpsi, 0,
                        reinterpret_cast<void **>(d_costs_out), stream);
}
int cuphu_build_defo_costs_gpu(const float *d_phase, const float *d_corr, const short *d_roww, const short *d_colw,
                               int nrow, int ncol, const CuPhuParams *params, int kperpdpsi, int kpardpsi,
                               CuPhuCost **d_costs_out, cuphu_stream_t stream) {
    return device_costs(d_phase, d_corr, d_roww, d_colw, nrow, ncol, params, kperpdpsi, kpardpsi, 1,
                        reinterpret_cast<void **>(d_costs_out), stream);
}

int cuphu_wrap_phase(float *d_phase, int n, cuphu_stream_t stream) {
    return guarded([&]() {
        CUPHU_REQUIRE(d_phase && n >= 0, "invalid argument");
        util_wrap_phase(d_phase, (size_t)n, (hipStream_t)stream);
        return 0;
    });
}
int cuphu_phase_diffs(const float *d_phase, float *d_rng, float *d_az, int nrow, int ncol, cuphu_stream_t stream) {
    return guarded([&]() {
        CUPHU_REQUIRE(d_phase && d_rng && d_az && nrow >= 1 && ncol >= 1, "invalid argument");
        util_phase_diffs(d_phase, d_rng, d_az, nrow, ncol, (hipStream_t)stream);
        return 0;
    });
}
int cuphu_boxcar2d(const float *d_src, float *d_dst, int nrow, int ncol, int kr, int kc, cuphu_stream_t stream) {
    return guarded([&]() {
        CUPHU_REQUIRE(d_src && d_dst && nrow >= 1 && ncol >= 1 && kr >= 0 && kc >= 0, "invalid argument");
        DevBuf<float> tmp((size_t)(nrow + 2 * kr) * ncol);
        util_boxcar(d_src, d_dst, tmp.get(), nrow, ncol, kr, kc, (hipStream_t)stream);
        CUPHU_HIP_CHECK(hipStreamSynchronize((hipStream_t)stream));
        return 0;
    });
}
int cuphu_residues_gpu(const float *d_phase, int nrow, int ncol, signed char *d_res, cuphu_stream_t stream) {
    return guarded([&]() {
        CUPHU_REQUIRE(d_phase && d_res && nrow >= 1 && ncol >= 1, "invalid argument");
        util_residues(d_phase, nrow, ncol, d_res, (hipStream_t)stream);
        return 0;
    });
}
int cuphu_integrate_phase_gpu(const float *d_phase, const short *d_hflows, const short *d_vflows, int nrow, int ncol,
                              int gpu_id, float *d_unw) {
    return guarded([&]() {
        CUPHU_REQUIRE(d_phase && d_hflows && d_vflows && d_unw && nrow >= 1 && ncol >= 1, "invalid argument");
        DeviceGuard dg(gpu_id);
        util_integrate(d_phase, d_hflows, d_vflows, nrow, ncol, d_unw, nullptr);
        return 0;
    });
}
int cuphu_conncomp_gpu(const float *d_unw, const float *d_corr, const unsigned char *d_mask, int nrow, int ncol,
                       const short *d_poscost, const short *d_negcost, const CuPhuParams *params, int gpu_id,
                       uint32_t *d_labels_out) {
    return guarded([&]() {
        CUPHU_REQUIRE(d_unw && d_corr && params && d_labels_out && nrow >= 1 && ncol >= 1, "invalid argument");
        CUPHU_REQUIRE((d_poscost == nullptr) == (d_negcost == nullptr), "poscost and negcost go together");
        DeviceGuard dg(gpu_id);
        conncomp_run(d_unw, d_corr, d_mask, nrow, ncol, ncol, params->minconncompfrac, params->maxncomps, d_labels_out,
                     nullptr, 0.05f, d_poscost, d_negcost, params->conncompthresh);
        return 0;
    });
}

// ── the device Laplace unwrap, allocating or from the fixed scratch ──────────
// Scratch of the no-allocation entry point: the solver's pooled planes and
// its small per-group buffers for one instance of up to
// CUPHU_NOALLOC_MAX_ELEMENTS row-pitched elements in FP32.  Every GPU the
// library runs on holds its own copy until the process ends.
constexpr size_t kNoAllocSmallBytes = (size_t)4 << 20;
// Scratch bytes one instance needs (run_batch's requests): 17 pooled planes
// of row-pitched elements (phi, corr, sigsq, out, lab 4 B, mask 1 B; w2 two
// values and b, dinv, x, r, p, s, wv, p2, r2, q2 one value each: 21 + 12
// sizeof(T) bytes per element, 69 in FP32), each padded by up to 512
// bytes, plus the per-group buffers.  With at most
// CUPHU_NOALLOC_MAX_ELEMENTS elements an instance has at most 16384 rows
// and 11265 workgroup tiles (256 x 8 or 128 x 16 at the smallest), so the
// per-group buffers (row map 4 B per row; 100 B per tile: instance map,
// three partial-sum buffers and the merged launch's tile list; two scalar
// states per instance; the plain solver's 8 KB of dot partials) stay far
// below kNoAllocSmallBytes.
constexpr size_t noalloc_need(long long el, size_t tsize) {
    return (size_t)el * (21 + 12 * tsize) + 17 * 512 + kNoAllocSmallBytes;
}
constexpr size_t kNoAllocBytes = noalloc_need(CUPHU_NOALLOC_MAX_ELEMENTS, 4);
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];
static std::map<int, unsigned char *> g_scratch_dev;   // device -> address of its copy (under g_mu)

// Device address of the current device's scratch (the array's name is a
// host-side stand-in and must never reach a kernel or a copy).
static unsigned char *scratch_address() {
    int dev = 0;
    CUPHU_HIP_CHECK(hipGetDevice(&dev));
    auto it = g_scratch_dev.find(dev);
    if (it != g_scratch_dev.end()) return it->second;
    void *ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        throw Error(CUPHU_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_dev[dev] = (unsigned char *)ptr;
    return (unsigned char *)ptr;
}

static int laplace_unwrap_dev(const CuPhuSmoothCost *d_costs, const float *d_phase, int nrow, int ncol, int max_iter,
                              float tol, int verbose, float *d_unw, cuphu_stream_t stream, bool noalloc) {
    return guarded([&]() {
        std::lock_guard<std::recursive_mutex> lk(g_mu);
        CUPHU_REQUIRE(d_costs && d_phase && d_unw && nrow >= 1 && ncol >= 1, "invalid argument");
        CuPhuLaplaceOptions o;
        default_options(&o);
        o.max_iter = max_iter;
        o.tol = tol;
        o.verbose = verbose;
        check_options(o);
        DevRes &res = devres();
        unsigned char *base = nullptr;
        if (noalloc) {
            // eager launches (no graph objects to instantiate) and the default
            // tile shape (the tuner creates events); both give the same bits
            o.use_graph = 0;
            o.autotune = 0;
            int pitch = pitch_of(ncol);
            if (wide_stride(o)) pitch = (pitch + 255) / 256 * 256;
            long long el = (long long)nrow * pitch;
            size_t need = noalloc_need(el, o.precision == 64 ? 8 : 4);
            if (el > (long long)CUPHU_NOALLOC_MAX_ELEMENTS || need > kNoAllocBytes)
                throw Error(CUPHU_ERR_INVALID,
                            "instance of " + std::to_string(nrow) + " x " + std::to_string(ncol) + " needs " +
                                std::to_string(el) + " row-pitched elements (" + std::to_string(need) +
                                " scratch bytes); the no-allocation entry point holds " +
                                std::to_string(CUPHU_NOALLOC_MAX_ELEMENTS) + " elements (" +
                                std::to_string(kNoAllocBytes) + " bytes, FP32)");
            base = scratch_address();
        }
        StreamJoin join(res, (hipStream_t)stream);
        hipStream_t st = res.stream;
        std::unique_ptr<ArenaScope> arena;
        if (noalloc) arena.reset(new ArenaScope(res, base, kNoAllocBytes));
        auto ing = [&](auto &pl, const Layout &L, const InstDesc *, const int *) {
            const InstDesc &d = L.desc[0];
            unsigned gsz = (unsigned)ceil_div((size_t)d.nrow * d.pitch, 256);
            hipLaunchKernelGGL(k_costs_to_planes, dim3(gsz), dim3(256), 0, st, (const short *)d_costs + 1, 2, d_phase, d,
                               pl.phi, pl.corr, pl.mask, pl.sigsq);
            using TT = std::remove_pointer_t<decltype(pl.x)>;
            hipLaunchKernelGGL(k_zero<TT>, dim3(gsz), dim3(256), 0, st, pl.x + d.off, (long long)d.nrow * d.pitch);
            CUPHU_HIP_CHECK(hipGetLastError());
        };
        BatchOut bo;
        std::vector<std::pair<int, int>> dims = {{nrow, ncol}};
        if (o.precision == 64) run_batch<double>(dims, o, 0.0, 0, false, res, st, nullptr, bo, ing);
        else run_batch<float>(dims, o, 0.0, 0, false, res, st, nullptr, bo, ing);
        const InstDesc &d = bo.lay.desc[0];
        CUPHU_HIP_CHECK(hipMemcpy2DAsync(d_unw, (size_t)ncol * 4, bo.unw + d.off, (size_t)d.pitch * 4, (size_t)ncol * 4,
                                         (size_t)nrow, hipMemcpyDeviceToDevice, st));
        CUPHU_HIP_CHECK(hipStreamSynchronize(st));
        return 0;
    });
}

int cuphu_laplace_unwrap_gpu(const CuPhuSmoothCost *d_costs, const float *d_phase, int nrow, int ncol,
                             double nshortcycle, int max_iter, float tol, int verbose, float *d_unw,
                             cuphu_stream_t stream) {
    (void)nshortcycle;   // unused by the reference as well (cuphu_laplace.cu:365)
    return laplace_unwrap_dev(d_costs, d_phase, nrow, ncol, max_iter, tol, verbose, d_unw, stream, false);
}

int cuphu_laplace_unwrap_gpu_noalloc(const CuPhuSmoothCost *d_costs, const float *d_phase, int nrow, int ncol,
                                     double nshortcycle, int max_iter, float tol, int verbose, float *d_unw,
                                     cuphu_stream_t stream) {
    (void)nshortcycle;
    return laplace_unwrap_dev(d_costs, d_phase, nrow, ncol, max_iter, tol, verbose, d_unw, stream, true);
}

int cuphu_wls_solve_gpu(const void *d_w, const void *d_phase, int nrow, int ncol, int precision, int max_iter,
                        double tol, void *d_u, int *iters_out, double *relres_out, cuphu_stream_t stream) {
    return guarded([&]() {
        std::lock_guard<std::recursive_mutex> lk(g_mu);
        CUPHU_REQUIRE(d_w && d_phase && d_u && nrow >= 2 && ncol >= 2, "invalid argument");
        CUPHU_REQUIRE(precision == 32 || precision == 64, "precision must be 32 or 64");
        DevRes &res = devres();
        StreamJoin join(res, (hipStream_t)stream);
        if (precision == 64)
            wls_impl<double>((const double *)d_w, (const float *)d_phase, nullptr, nullptr, nrow, ncol, max_iter, tol,
                             (double *)d_u, nullptr, nullptr, iters_out, relres_out, res, res.stream);
        else
            wls_impl<float>((const float *)d_w, (const float *)d_phase, nullptr, nullptr, nrow, ncol, max_iter, tol,
                            (float *)d_u, nullptr, nullptr, iters_out, relres_out, res, res.stream);
        return 0;
    });
}

int cuphu_wls_adjoint_gpu(const void *d_w, const void *d_phase, const void *d_u, const void *d_gu, int nrow, int ncol,
                          int precision, int max_iter, double tol, void *d_gw, void *d_lambda, int *iters_out,
                          double *relres_out, cuphu_stream_t stream) {
    return guarded([&]() {
        std::lock_guard<std::recursive_mutex> lk(g_mu);
        CUPHU_REQUIRE(d_w && d_phase && d_u && d_gu && d_gw && nrow >= 2 && ncol >= 2, "invalid argument");
        CUPHU_REQUIRE(precision == 32 || precision == 64, "precision must be 32 or 64");
        DevRes &res = devres();
        StreamJoin join(res, (hipStream_t)stream);
        if (precision == 64)
            wls_impl<double>((const double *)d_w, (const float *)d_phase, (const double *)d_u, (const double *)d_gu, nrow,
                             ncol, max_iter, tol, nullptr, (double *)d_gw, (double *)d_lambda, iters_out, relres_out,
                             res, res.stream);
        else
            wls_impl<float>((const float *)d_w, (const float *)d_phase, (const float *)d_u, (const float *)d_gu, nrow,
                            ncol, max_iter, tol, nullptr, (float *)d_gw, (float *)d_lambda, iters_out, relres_out, res,
                            res.stream);
        return 0;
    });
}

// ── test hooks ──────────────────────────────────────────────────────────────
namespace {
struct HookInst {
    Layout L;
    PoolSet<float> P;
    GroupDev g{};
    InstDesc d{};
};
void hook_setup(HookInst &h, DevRes &res, int nrow, int ncol, TileShape sh, hipStream_t st) {
    h.L = make_layout({{nrow, ncol}}, sh);
    h.P = alloc_pools<float>(res, h.L.total_el);
    h.d = h.L.desc[0];
    int nb = h.d.tiles_x * h.d.tiles_y;
    h.g.desc = ws<InstDesc>(res, "grpdesc", 1);
    h.g.blk = ws<int>(res, "grpblk", (size_t)nb);
    h.g.scal = ws<SolveScal>(res, "grpscal", 1);
    h.g.part = ws<double2>(res, "grppart", (size_t)nb);
    h.g.part4 = ws<double4>(res, "grppart4", (size_t)nb);
    h.g.scal2 = ws<SolveScal>(res, "grpscal2", 1);
    h.g.part4b = ws<double4>(res, "grppart4b", (size_t)nb);
    h.g.ninst = 1;
    h.g.nblk = nb;
    std::vector<int> blk((size_t)nb, 0);
    CUPHU_HIP_CHECK(hipMemcpyAsync(h.g.desc, &h.d, sizeof(InstDesc), hipMemcpyHostToDevice, st));
    CUPHU_HIP_CHECK(hipMemcpyAsync(h.g.blk, blk.data(), (size_t)nb * sizeof(int), hipMemcpyHostToDevice, st));
    size_t nel = (size_t)h.L.total_el;
    unsigned g = (unsigned)ceil_div(2 * nel, 256);
    hipLaunchKernelGGL(k_zero<float>, dim3(g), dim3(256), 0, st, h.P.pl.w2, (long long)(2 * nel));
    for (float *p : {h.P.pl.b, h.P.pl.dinv, h.P.pl.x, h.P.pl.r, h.P.pl.p, h.P.pl.s, h.P.pl.wv, h.P.pl.phi,
                     h.P.pl.p2, h.P.pl.r2, h.P.pl.q2})
        hipLaunchKernelGGL(k_zero<float>, dim3(g), dim3(256), 0, st, p, (long long)nel);
    CUPHU_HIP_CHECK(hipGetLastError());
    CUPHU_HIP_CHECK(hipStreamSynchronize(st));
}
void to_plane(const float *host, float *dev, const InstDesc &d, hipStream_t st) {
    CUPHU_HIP_CHECK(hipMemcpy2DAsync(dev + d.off, (size_t)d.pitch * 4, host, (size_t)d.ncol * 4, (size_t)d.ncol * 4,
                                     (size_t)d.nrow, hipMemcpyHostToDevice, st));
}
void from_plane(const float *dev, float *host, const InstDesc &d, hipStream_t st) {
    CUPHU_HIP_CHECK(hipMemcpy2DAsync(host, (size_t)d.ncol * 4, dev + d.off, (size_t)d.pitch * 4, (size_t)d.ncol * 4,
                                     (size_t)d.nrow, hipMemcpyDeviceToHost, st));
}
}  // namespace

int cuphu_test_ingest(const float *igram_c64, const float *corr, const unsigned char *mask, int nrow, int ncol,
                      const CuPhuParams *params, float *phase_out, short *sigsq_out) {
    return guarded([&]() {
        std::lock_guard<std::recursive_mutex> lk(g_mu);
        CUPHU_REQUIRE(igram_c64 && corr && params && phase_out && sigsq_out && nrow >= 1 && ncol >= 1, "null argument");
        DevRes &res = devres();
        hipStream_t st = res.stream;
        size_t n = (size_t)nrow * ncol;
        DevBuf<float2> dig(n);
        DevBuf<float> dc(n);
        DevBuf<unsigned char> dm(mask ? n : 0);
        CUPHU_HIP_CHECK(hipMemcpyAsync(dig.get(), igram_c64, n * 8, hipMemcpyHostToDevice, st));
        CUPHU_HIP_CHECK(hipMemcpyAsync(dc.get(), corr, n * 4, hipMemcpyHostToDevice, st));
        if (mask) CUPHU_HIP_CHECK(hipMemcpyAsync(dm.get(), mask, n, hipMemcpyHostToDevice, st));
        std::vector<IngestSrc> srcs(1);
        std::memset(&srcs[0], 0, sizeof(IngestSrc));
        srcs[0].igram = dig.get();
        srcs[0].corr = dc.get();
        srcs[0].mask = mask ? dm.get() : nullptr;
        srcs[0].src_stride = ncol;
        srcs[0].cc = make_cost_consts(*params, params->ncorrlooks);
        Layout L = make_layout({{nrow, ncol}}, TileShape{4, 64});
        PoolSet<float> P = alloc_pools<float>(res, L.total_el);
        InstDesc *d_desc = ws<InstDesc>(res, "gdesc", 1);
        int *d_rowinst = ws<int>(res, "rowinst", (size_t)nrow);
        std::vector<int> ri((size_t)nrow, 0);
        CUPHU_HIP_CHECK(hipMemcpyAsync(d_desc, L.desc.data(), sizeof(InstDesc), hipMemcpyHostToDevice, st));
        CUPHU_HIP_CHECK(hipMemcpyAsync(d_rowinst, ri.data(), ri.size() * 4, hipMemcpyHostToDevice, st));
        auto ing = src_ingest<float>(srcs, false, res, st);
        ing(P.pl, L, d_desc, d_rowinst);
        const InstDesc &d = L.desc[0];
        from_plane(P.pl.phi, phase_out, d, st);
        std::vector<short2> sq((size_t)nrow * d.pitch);
        CUPHU_HIP_CHECK(hipMemcpyAsync(sq.data(), P.pl.sigsq + d.off, sq.size() * sizeof(short2), hipMemcpyDeviceToHost, st));
        CUPHU_HIP_CHECK(hipStreamSynchronize(st));
        size_t nra = (size_t)(nrow - 1) * ncol;
        for (int r = 0; r < nrow; ++r)
            for (int c = 0; c < ncol; ++c) {
                const short2 &v = sq[(size_t)r * d.pitch + c];
                if (r < nrow - 1) sigsq_out[(size_t)r * ncol + c] = v.x;
                if (c < ncol - 1) sigsq_out[nra + (size_t)r * (ncol - 1) + c] = v.y;
            }
        return 0;
    });
}

int cuphu_test_laplace_ops(const float *phase, const short *sigsq, const float *x, int nrow, int ncol, float delta,
                           float *w0, float *b, float *dinv, float *r, float *lx, float *wrw, float *cong) {
    return guarded([&]() {
        std::lock_guard<std::recursive_mutex> lk(g_mu);
        CUPHU_REQUIRE(phase && sigsq && x && nrow >= 1 && ncol >= 1, "null argument");
        DevRes &res = devres();
        hipStream_t st = res.stream;
        int lpr = 64;
        env_int("CUPHU_LAPLACE_LPR", &lpr);   // 32: the fused solver's 128-column tiles
        const TileShape sh{rows_per_wave(32), (lpr == 32 && rows_per_wave(32) <= 4) ? 32 : 64};
        HookInst h;
        hook_setup(h, res, nrow, ncol, sh, st);
        const InstDesc &d = h.d;
        size_t na = (size_t)(nrow - 1) * ncol + (size_t)nrow * (ncol - 1);
        DevBuf<short> dsig(na + 1);
        DevBuf<float> dph((size_t)nrow * ncol), dflat(na + 1);
        CUPHU_HIP_CHECK(hipMemcpyAsync(dsig.get(), sigsq, na * 2, hipMemcpyHostToDevice, st));
        CUPHU_HIP_CHECK(hipMemcpyAsync(dph.get(), phase, (size_t)nrow * ncol * 4, hipMemcpyHostToDevice, st));
        unsigned gp = (unsigned)ceil_div((size_t)nrow * d.pitch, 256), ga = (unsigned)ceil_div(na + 1, 256);
        hipLaunchKernelGGL(k_costs_to_planes, dim3(gp), dim3(256), 0, st, (const short *)dsig.get(), 1,
                           (const float *)dph.get(), d, h.P.pl.phi, h.P.pl.corr, h.P.pl.mask, h.P.pl.sigsq);
        CUPHU_HIP_CHECK(hipGetLastError());
        to_plane(x, h.P.pl.x, d, st);
        launch_weights<float>(h.g.desc, h.g.blk, h.g.nblk, h.P.pl, 0, 0.0f, sh, st);
        if (w0 && na) {
            hipLaunchKernelGGL(k_planes_to_flat_w<float>, dim3(ga), dim3(256), 0, st, h.P.pl.w2, d, dflat.get());
            CUPHU_HIP_CHECK(hipMemcpyAsync(w0, dflat.get(), na * 4, hipMemcpyDeviceToHost, st));
        }
        launch_assemble<float>(h.g.desc, h.g.blk, h.g.nblk, h.P.pl, sh, st);
        if (b) from_plane(h.P.pl.b, b, d, st);
        if (dinv) from_plane(h.P.pl.dinv, dinv, d, st);
        if (r) from_plane(h.P.pl.r, r, d, st);
        if (lx) {
            launch_plain_matvec<float>(d, h.P.pl.w2, h.P.pl.x, h.P.pl.wv, st);
            from_plane(h.P.pl.wv, lx, d, st);
        }
        if (wrw && na) {
            launch_weights<float>(h.g.desc, h.g.blk, h.g.nblk, h.P.pl, 1, delta * delta, sh, st);
            hipLaunchKernelGGL(k_planes_to_flat_w<float>, dim3(ga), dim3(256), 0, st, h.P.pl.w2, d, dflat.get());
            CUPHU_HIP_CHECK(hipMemcpyAsync(wrw, dflat.get(), na * 4, hipMemcpyDeviceToHost, st));
        }
        if (cong) {
            std::vector<int> ri((size_t)nrow, 0);
            int *d_rowinst = ws<int>(res, "rowinst", (size_t)nrow);
            InstDesc *d_gdesc = ws<InstDesc>(res, "gdesc", 1);
            CUPHU_HIP_CHECK(hipMemcpyAsync(d_rowinst, ri.data(), ri.size() * 4, hipMemcpyHostToDevice, st));
            CUPHU_HIP_CHECK(hipMemcpyAsync(d_gdesc, &d, sizeof(InstDesc), hipMemcpyHostToDevice, st));
            launch_final<float>(d_gdesc, d_rowinst, nrow, h.P.pl, h.P.out, d.pitch, st);
            from_plane(h.P.out, cong, d, st);
        }
        CUPHU_HIP_CHECK(hipStreamSynchronize(st));
        return 0;
    });
}

int cuphu_test_pcg_step(const float *w, const float *dinv, int nrow, int ncol, double gamma, double beta, int rpw,
                        float *x, float *r, float *p, float *q, double *pq, double *alpha, double *rz) {
    return guarded([&]() {
        std::lock_guard<std::recursive_mutex> lk(g_mu);
        CUPHU_REQUIRE(w && dinv && x && r && p && q && pq && alpha && rz && nrow >= 1 && ncol >= 1, "null argument");
        CUPHU_REQUIRE(rpw == 2 || rpw == 4 || rpw == 8, "rpw must be 2, 4 or 8");
        DevRes &res = devres();
        hipStream_t st = res.stream;
        HookInst h;
        hook_setup(h, res, nrow, ncol, TileShape{rpw, 64}, st);
        const InstDesc &d = h.d;
        size_t na = (size_t)(nrow - 1) * ncol + (size_t)nrow * (ncol - 1);
        DevBuf<float> dw(na + 1);
        if (na) CUPHU_HIP_CHECK(hipMemcpyAsync(dw.get(), w, na * 4, hipMemcpyHostToDevice, st));
        unsigned gp = (unsigned)ceil_div((size_t)nrow * d.pitch, 256);
        hipLaunchKernelGGL(k_flat_w_to_planes<float>, dim3(gp), dim3(256), 0, st, (const float *)dw.get(), d,
                           h.P.pl.w2);
        CUPHU_HIP_CHECK(hipGetLastError());
        to_plane(dinv, h.P.pl.dinv, d, st);
        to_plane(x, h.P.pl.x, d, st);
        to_plane(r, h.P.pl.r, d, st);
        to_plane(p, h.P.pl.p, d, st);
        SolveScal sc{};
        sc.gamma = gamma;
        sc.gamma0 = gamma;
        sc.beta = beta;
        sc.active = 1;
        sc.status = CUPHU_SOLVE_MAXITER;
        CUPHU_HIP_CHECK(hipMemcpyAsync(h.g.scal, &sc, sizeof(sc), hipMemcpyHostToDevice, st));
        // parity 0: direction p -> p2, q -> s
        launch_pq<float>(h.g.desc, h.g.blk, h.g.nblk, h.P.pl, h.g.scal, h.g.part, rpw, 0, st);
        launch_reduce(h.g.desc, 1, h.g.part, h.g.scal, RED_PQ, 1 << 30, 0.0, st);
        launch_xr<float>(h.g.desc, h.g.blk, h.g.nblk, h.P.pl, h.g.scal, h.g.part, rpw, 0, st);
        launch_reduce(h.g.desc, 1, h.g.part, h.g.scal, RED_RZ, 1 << 30, 0.0, st);
        from_plane(h.P.pl.x, x, d, st);
        from_plane(h.P.pl.r, r, d, st);
        from_plane(h.P.pl.p2, p, d, st);
        from_plane(h.P.pl.s, q, d, st);
        CUPHU_HIP_CHECK(hipMemcpyAsync(&sc, h.g.scal, sizeof(sc), hipMemcpyDeviceToHost, st));
        CUPHU_HIP_CHECK(hipStreamSynchronize(st));
        CUPHU_REQUIRE(sc.active == 1, "PCG step stopped (p.q not positive)");
        *pq = sc.delta;
        *alpha = sc.alpha;
        *rz = sc.gamma;
        return 0;
    });
}

int cuphu_test_pcg_fused_step(const float *w, const float *dinv, int nrow, int ncol, double alpha, double beta,
                              int start, const float *rt, int rpw, int stored_q, float *x, float *r, float *p,
                              float *q, double *sums) {
    return guarded([&]() {
        std::lock_guard<std::recursive_mutex> lk(g_mu);
        CUPHU_REQUIRE(w && dinv && x && r && p && q && sums && nrow >= 1 && ncol >= 1, "null argument");
        // rpw 2, 4, 8: rows per lane group on 256-column tiles; -2, -4: on
        // 128-column tiles (q-free kernel only)
        CUPHU_REQUIRE(rpw == 2 || rpw == 4 || rpw == 8 || (!stored_q && (rpw == -2 || rpw == -4)),
                      "rpw must be 2, 4 or 8 (or -2, -4 without stored q)");
        const TileShape sh = rpw > 0 ? TileShape{rpw, 64} : TileShape{-rpw, 32};
        DevRes &res = devres();
        hipStream_t st = res.stream;
        HookInst h;
        hook_setup(h, res, nrow, ncol, sh, st);
        const InstDesc &d = h.d;
        size_t na = (size_t)(nrow - 1) * ncol + (size_t)nrow * (ncol - 1);
        DevBuf<float> dw(na + 1);
        if (na) CUPHU_HIP_CHECK(hipMemcpyAsync(dw.get(), w, na * 4, hipMemcpyHostToDevice, st));
        unsigned gp = (unsigned)ceil_div((size_t)nrow * d.pitch, 256);
        hipLaunchKernelGGL(k_flat_w_to_planes<float>, dim3(gp), dim3(256), 0, st, (const float *)dw.get(), d,
                           h.P.pl.w2);
        CUPHU_HIP_CHECK(hipGetLastError());
        to_plane(dinv, h.P.pl.dinv, d, st);
        to_plane(x, h.P.pl.x, d, st);
        to_plane(r, h.P.pl.r, d, st);
        to_plane(p, h.P.pl.p, d, st);
        to_plane(q, h.P.pl.s, d, st);
        if (rt) to_plane(rt, h.P.pl.wv, d, st);
        SolveScal sc{};
        sc.alpha = alpha;
        sc.beta = beta;
        sc.replace = rt ? 1 : 0;
        sc.active = 1;
        sc.xmode = 2;   // apply this step's x update (no deferral)
        CUPHU_HIP_CHECK(hipMemcpyAsync(h.g.scal, &sc, sizeof(sc), hipMemcpyHostToDevice, st));
        // the merged tiling wherever the solver would take it (also on a
        // single tile row)
        std::vector<int4> mix;
        if (!stored_q && merge_allowed(sh) && tall_column(d, sh) != 0) {
            mix_tiles(d, 0, sh, mix);
            h.g.mix = ws<int4>(res, "grpmix", mix.size());
            h.g.nmix = (int)mix.size();
            CUPHU_HIP_CHECK(hipMemcpyAsync(h.g.mix, mix.data(), mix.size() * sizeof(int4), hipMemcpyHostToDevice, st));
        }
        // parity 0: reads set 0 (r, p, s), writes set 1 (r2, p2, q2)
        if (stored_q) launch_it<float>(h.g.desc, h.g.blk, h.g.nblk, h.P.pl, h.g.scal, h.g.part4, rpw, 0, start ? 1 : 0, st);
        else if (h.g.nmix)
            launch_iq_mix<float>(h.g.desc, h.g.mix, h.g.nmix, h.P.pl, h.g.scal, h.g.part4, sh.rpw, 0, start ? 1 : 0, st);
        else launch_iq<float>(h.g.desc, h.g.blk, h.g.nblk, h.P.pl, h.g.scal, h.g.part4, sh, 0, start ? 1 : 0, st);
        std::vector<double4> parts((size_t)h.g.nblk);
        CUPHU_HIP_CHECK(hipMemcpyAsync(parts.data(), h.g.part4, parts.size() * sizeof(double4), hipMemcpyDeviceToHost, st));
        from_plane(h.P.pl.x, x, d, st);
        from_plane(h.P.pl.r2, r, d, st);
        from_plane(h.P.pl.p2, p, d, st);
        if (stored_q) from_plane(h.P.pl.q2, q, d, st);
        CUPHU_HIP_CHECK(hipStreamSynchronize(st));
        double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
        for (auto &v : parts) { a0 += v.x; a1 += v.y; a2 += v.z; a3 += v.w; }
        sums[0] = a0; sums[1] = a1; sums[2] = a2; sums[3] = a3;
        return 0;
    });
}


int cuphu_fix_cycle_spikes_gpu(float *d_unw, const unsigned char *d_mask, int nrow, int ncol, cuphu_stream_t stream) {
    return guarded([&]() {
        std::lock_guard<std::recursive_mutex> lk(g_mu);
        CUPHU_REQUIRE(d_unw && nrow >= 1 && ncol >= 1, "invalid argument");
        DevRes &res = devres();
        StreamJoin join(res, (hipStream_t)stream);
        fix_cycle_spikes_dev(d_unw, d_mask, nrow, ncol, res.stream);
        return 0;
    });
}

int cuphu_neighbor_feedback_gpu(float *d_unw, const unsigned char *d_mask, int nrow, int ncol, int ntr, int ntc,
                                int row_ovrlp, int col_ovrlp, int feather_px, cuphu_stream_t stream) {
    return guarded([&]() {
        std::lock_guard<std::recursive_mutex> lk(g_mu);
        CUPHU_REQUIRE(d_unw && nrow >= 1 && ncol >= 1 && ntr >= 1 && ntc >= 1, "invalid argument");
        CUPHU_REQUIRE(row_ovrlp >= 0 && col_ovrlp >= 0, "overlaps must be >= 0");
        DevRes &res = devres();
        StreamJoin join(res, (hipStream_t)stream);
        neighbor_feedback_dev(d_unw, d_mask, nrow, ncol, seams(ncol, col_ovrlp, ntc), seams(nrow, row_ovrlp, ntr),
                              feather_px, res.stream);
        return 0;
    });
}

// Host-array forms with the reference's test-hook names and arguments
// (cuphu.h:397-419), run on the current GPU.
int cuphu_fix_cycle_spikes_test(float *unw, const unsigned char *mask, int nrow, int ncol) {
    return guarded([&]() {
        std::lock_guard<std::recursive_mutex> lk(g_mu);
        CUPHU_REQUIRE(unw && nrow >= 1 && ncol >= 1, "invalid argument");
        DevRes &res = devres();
        hipStream_t st = res.stream;
        size_t n = (size_t)nrow * ncol;
        DevBuf<float> du(n);
        DevBuf<unsigned char> dm(mask ? n : 0);
        CUPHU_HIP_CHECK(hipMemcpyAsync(du.get(), unw, n * 4, hipMemcpyHostToDevice, st));
        if (mask) CUPHU_HIP_CHECK(hipMemcpyAsync(dm.get(), mask, n, hipMemcpyHostToDevice, st));
        fix_cycle_spikes_dev(du.get(), mask ? dm.get() : nullptr, nrow, ncol, st);
        CUPHU_HIP_CHECK(hipMemcpyAsync(unw, du.get(), n * 4, hipMemcpyDeviceToHost, st));
        CUPHU_HIP_CHECK(hipStreamSynchronize(st));
        return 0;
    });
}

int cuphu_laplace_neighbor_feedback_test(float *unw, const unsigned char *mask, int nrow, int ncol, int ntr, int ntc,
                                         int row_ovrlp, int col_ovrlp, int feather_px) {
    return guarded([&]() {
        std::lock_guard<std::recursive_mutex> lk(g_mu);
        CUPHU_REQUIRE(unw && nrow >= 1 && ncol >= 1 && ntr >= 1 && ntc >= 1, "invalid argument");
        CUPHU_REQUIRE(row_ovrlp >= 0 && col_ovrlp >= 0, "overlaps must be >= 0");
        DevRes &res = devres();
        hipStream_t st = res.stream;
        size_t n = (size_t)nrow * ncol;
        DevBuf<float> du(n);
        DevBuf<unsigned char> dm(mask ? n : 0);
        CUPHU_HIP_CHECK(hipMemcpyAsync(du.get(), unw, n * 4, hipMemcpyHostToDevice, st));
        if (mask) CUPHU_HIP_CHECK(hipMemcpyAsync(dm.get(), mask, n, hipMemcpyHostToDevice, st));
        neighbor_feedback_dev(du.get(), mask ? dm.get() : nullptr, nrow, ncol, seams(ncol, col_ovrlp, ntc),
                              seams(nrow, row_ovrlp, ntr), feather_px, st);
        CUPHU_HIP_CHECK(hipMemcpyAsync(unw, du.get(), n * 4, hipMemcpyDeviceToHost, st));
        CUPHU_HIP_CHECK(hipStreamSynchronize(st));
        return 0;
    });
}


int cuphu_bridge_gpu(float *d_unw, uint32_t *d_conncomp, const unsigned char *d_mask, int nrow, int ncol,
                     const CuPhuBridgeParams *bp, cuphu_stream_t stream) {
    return guarded([&]() {
        std::lock_guard<std::recursive_mutex> lk(g_mu);
        CUPHU_REQUIRE(d_unw && d_conncomp && bp && nrow >= 1 && ncol >= 1, "invalid argument");
        DevRes &res = devres();
        StreamJoin join(res, (hipStream_t)stream);
        bridge_dev(d_unw, d_conncomp, d_mask, nrow, ncol, *bp, res.stream);
        return 0;
    });
}

// Host-array form with the reference's test-hook name and arguments
// (cuphu_solver.cpp:1360-1366; bp->enabled is not consulted, as there).
int cuphu_bridge_apply_test(float *unw, uint32_t *conncomp, int nrow, int ncol, const CuPhuBridgeParams *bp,
                            const unsigned char *mask, int gpu_id) {
    return guarded([&]() {
        std::lock_guard<std::recursive_mutex> lk(g_mu);
        CUPHU_REQUIRE(unw && conncomp && bp && nrow >= 1 && ncol >= 1, "invalid argument");
        DeviceGuard dg(gpu_id);
        DevRes &res = devres();
        hipStream_t st = res.stream;
        size_t n = (size_t)nrow * ncol;
        DevBuf<float> du(n);
        DevBuf<uint32_t> dc(n);
        DevBuf<unsigned char> dm(mask ? n : 0);
        CUPHU_HIP_CHECK(hipMemcpyAsync(du.get(), unw, n * 4, hipMemcpyHostToDevice, st));
        CUPHU_HIP_CHECK(hipMemcpyAsync(dc.get(), conncomp, n * 4, hipMemcpyHostToDevice, st));
        if (mask) CUPHU_HIP_CHECK(hipMemcpyAsync(dm.get(), mask, n, hipMemcpyHostToDevice, st));
        bridge_dev(du.get(), dc.get(), mask ? dm.get() : nullptr, nrow, ncol, *bp, st);
        CUPHU_HIP_CHECK(hipMemcpyAsync(unw, du.get(), n * 4, hipMemcpyDeviceToHost, st));
        CUPHU_HIP_CHECK(hipMemcpyAsync(conncomp, dc.get(), n * 4, hipMemcpyDeviceToHost, st));
        CUPHU_HIP_CHECK(hipStreamSynchronize(st));
        return 0;
    });
}

}  // extern "C"
