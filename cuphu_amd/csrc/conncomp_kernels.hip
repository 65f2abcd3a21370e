// conncomp_kernels.hip -- connected-component labelling of the unwrapped
// phase (reference: src/cxx/cuphu/cuphu_conncomp.cu, Laplace criterion).
//
// A pixel is valid when !(corr < 0.05) and mask != 0 (:102-114, :324-325).
// Two valid 4-neighbours join when fabsf(unw_a - unw_b) < pi (:178-218) and,
// after an MCF solve, their arc is confident: min(poscost, negcost) >=
// conncompthresh at the solved flows (:84-93; without costs every arc is).  The label of a component is
// decided by its minimum raster index (the invariant of :16-20), components
// smaller than max(1, (int)(minfrac*npix)) get 0, and the rest are numbered
// 1..maxncomps in increasing root order (build_label_map :272-286).
//
// GPU scheme: (1) union-find inside a 64x16 tile in LDS; every pixel points
// at its tile-local root and the local roots are flagged, (2) lock-free
// global union (atomicMin, parent <= self) across tile borders, (3) the
// local roots (a few per tile) find their final root with path halving,
// (4) every pixel takes its local root's final root (two loads) and the
// per-root sizes are counted with one atomic per distinct root per wave,
// (5) keep flags, device scan, relabel.  The output is a pure function of
// the partition, so it is bit-identical to a sequential union-find.
#include <hipcub/hipcub.hpp>

#include "common.hpp"
#include "util.hpp"

namespace cuphu {

constexpr int kCcTW = 64, kCcTH = 16, kCcN = kCcTW * kCcTH;

__device__ __forceinline__ bool cc_valid(const float *corr, const unsigned char *mask, long long i, float th) {
    bool bad = corr[i] < th;
    if (mask) bad = bad || (mask[i] == 0);
    return !bad;
}
__device__ __forceinline__ bool cc_join(const float *unw, long long a, long long b) {
    return fabsf(unw[a] - unw[b]) < kPiF;
}
// MCF path arc gate (arc_confident, cuphu_conncomp.cu:84-93): confident when
// min(poscost, negcost) >= thresh; no costs = confident.  Arc indices are the
// flat row-arc-then-column-arc layout (arc_index, :67-78).
struct CcGate {
    const short *pos, *neg;
    long thresh;
    long long nra;   // (nrow - 1) * ncol
    int ncol;
    __device__ __forceinline__ bool ok(long long a) const {
        if (!pos) return true;
        short m = pos[a] < neg[a] ? pos[a] : neg[a];
        return (long)m >= thresh;
    }
    __device__ __forceinline__ bool down(int r, int c) const { return ok((long long)r * ncol + c); }
    __device__ __forceinline__ bool right(int r, int c) const { return ok(nra + (long long)r * (ncol - 1) + c); }
};

__device__ __forceinline__ int lds_find(const int *par, int x) {
    int p = par[x];
    while (p != x) { x = p; p = par[x]; }
    return x;
}
__device__ void lds_union(int *par, int a, int b) {
    for (;;) {
        a = lds_find(par, a);
        b = lds_find(par, b);
        if (a == b) return;
        if (a < b) { int t = a; a = b; b = t; }   // a > b: link a under b
        int old = atomicMin(&par[a], b);
        if (old == a) return;
        a = old;
    }
}

__global__ __launch_bounds__(256) void k_cc_local(const float *unw, const float *corr, const unsigned char *mask,
                                                  int nrow, int ncol, int stride, float th, CcGate gate, uint32_t *lab,
                                                  uint32_t *rootflag) {
    __shared__ int par[kCcN];
    int tx = blockIdx.x, ty = blockIdx.y;
    int c0 = tx * kCcTW, r0 = ty * kCcTH;
    for (int l = threadIdx.x; l < kCcN; l += 256) {
        int r = r0 + l / kCcTW, c = c0 + l % kCcTW;
        bool v = r < nrow && c < ncol && cc_valid(corr, mask, (long long)r * stride + c, th);
        par[l] = v ? l : -1;
    }
    __syncthreads();
    for (int l = threadIdx.x; l < kCcN; l += 256) {
        if (par[l] < 0) continue;
        int lr = l / kCcTW, lc = l % kCcTW, r = r0 + lr, c = c0 + lc;
        long long g = (long long)r * stride + c;
        if (lc + 1 < kCcTW && par[l + 1] >= 0 && cc_join(unw, g, g + 1) && gate.right(r, c))
            lds_union(par, l, l + 1);
        if (lr + 1 < kCcTH && par[l + kCcTW] >= 0 && cc_join(unw, g, g + stride) && gate.down(r, c))
            lds_union(par, l, l + kCcTW);
    }
    __syncthreads();
    for (int l = threadIdx.x; l < kCcN; l += 256) {
        int r = r0 + l / kCcTW, c = c0 + l % kCcTW;
        if (r >= nrow || c >= ncol) continue;
        long long g = (long long)r * stride + c;
        uint32_t out = kInvalidLabel, isroot = 0u;
        if (par[l] >= 0) {
            int root = lds_find(par, l);
            out = (uint32_t)((long long)(r0 + root / kCcTW) * stride + (c0 + root % kCcTW));
            isroot = (root == l) ? 1u : 0u;
        }
        lab[g] = out;
        rootflag[g] = isroot;
    }
}

__device__ __forceinline__ uint32_t g_load(const uint32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t g_find(uint32_t *lab, uint32_t x) {
    uint32_t p = g_load(lab + x);
    while (p != x) { x = p; p = g_load(lab + x); }
    return x;
}
__device__ void g_union(uint32_t *lab, uint32_t a, uint32_t b) {
    for (;;) {
        a = g_find(lab, a);
        b = g_find(lab, b);
        if (a == b) return;
        if (a < b) { uint32_t t = a; a = b; b = t; }
        uint32_t old = atomicMin(lab + a, b);
        if (old == a) return;
        a = old;
    }
}

// Unions across tile borders: the right edge of a tile's last column and
// the down edge of its last row.
__global__ __launch_bounds__(256) void k_cc_border(const float *unw, int nrow, int ncol, int stride, CcGate gate,
                                                   uint32_t *lab) {
    long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= (long long)nrow * ncol) return;
    int r = (int)(i / ncol), c = (int)(i % ncol);
    const bool right = (c % kCcTW) == kCcTW - 1, bottom = (r % kCcTH) == kCcTH - 1;
    if (!right && !bottom) return;   // 1 pixel in 13: no label load for the others
    long long g = (long long)r * stride + c;
    if (g_load(lab + g) == kInvalidLabel) return;
    if (right && c + 1 < ncol && g_load(lab + g + 1) != kInvalidLabel && cc_join(unw, g, g + 1) &&
        gate.right(r, c))
        g_union(lab, (uint32_t)g, (uint32_t)(g + 1));
    if (bottom && r + 1 < nrow && g_load(lab + g + stride) != kInvalidLabel &&
        cc_join(unw, g, g + stride) && gate.down(r, c))
        g_union(lab, (uint32_t)g, (uint32_t)(g + stride));
}

// Final root of every tile-local root.  The forest no longer changes except
// for the halving stores, each of which points a node at one of its own
// ancestors, so every find still ends at the true root.  Along any path the
// indices decrease toward the root, which is the minimum of its component,
// so the stores are atomicMin: an entry only moves closer to the root, and a
// late halving store from another thread cannot overwrite a final root.
__global__ __launch_bounds__(256) void k_cc_flatten(int nrow, int ncol, int stride, const uint32_t *rootflag,
                                                    uint32_t *lab) {
    long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= (long long)nrow * ncol) return;
    long long g = (long long)(i / ncol) * stride + (i % ncol);
    if (!rootflag[g]) return;
    uint32_t x = (uint32_t)g;
    uint32_t p = g_load(lab + x);
    while (p != x) {
        uint32_t gp = g_load(lab + p);
        if (gp != p) atomicMin(lab + x, gp);
        x = p;
        p = gp;
    }
    atomicMin(lab + g, x);
}

// lab[g] := final root (the local root's entry after k_cc_flatten; only
// local-root entries are read, and a local root rewrites its entry with the
// value it already holds) and per-root pixel counts.  Each thread takes
// kCcRun pixels 256 apart (coalesced) and counts a run while the root stays
// the same, flushing when it changes; the last run is aggregated per wave.
// A component covering the image then costs one atomic per wave of
// kCcRun * 64 pixels instead of one per 64.
constexpr int kCcRun = 16;
__global__ __launch_bounds__(256) void k_cc_compress(int nrow, int ncol, int stride, uint32_t *lab, uint32_t *cnt) {
    const long long npx = (long long)nrow * ncol;
    const long long base = (long long)blockIdx.x * 256 * kCcRun + threadIdx.x;
    bool pending = false;
    uint32_t root = 0u, run = 0u;
    for (int k = 0; k < kCcRun; ++k) {
        const long long i = base + (long long)k * 256;
        if (i >= npx) break;
        const long long g = (long long)(i / ncol) * stride + (i % ncol);
        const uint32_t l = lab[g];
        if (l == kInvalidLabel) continue;
        const uint32_t r = lab[l];
        lab[g] = r;
        if (pending && r != root) atomicAdd(cnt + root, run);   // the run ended
        if (!pending || r != root) { root = r; run = 0u; }
        pending = true;
        ++run;
    }
    const int lane = threadIdx.x & 63;
    for (;;) {
        unsigned long long m = __ballot(pending);
        if (m == 0ull) break;
        int leader = __ffsll((long long)m) - 1;
        uint32_t lr = __shfl(root, leader, 64);
        bool mine = pending && root == lr;
        uint32_t tot = mine ? run : 0u;   // sum of the matching runs over the wave
        for (int o = 32; o > 0; o >>= 1) tot += __shfl_xor(tot, o, 64);
        if (lane == leader) atomicAdd(cnt + lr, tot);
        if (mine) pending = false;
    }
}

__global__ void k_cc_fill(uint32_t *p, size_t n, uint32_t v) {
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_cc_keep(const uint32_t *cnt, size_t n, uint32_t min_size, uint32_t *flags) {
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) flags[i] = (cnt[i] >= min_size) ? 1u : 0u;
}

__global__ __launch_bounds__(256) void k_cc_relabel(int nrow, int ncol, int stride, const uint32_t *rank,
                                                    const uint32_t *flags, long maxncomps, uint32_t *lab) {
    long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= (long long)nrow * ncol) return;
    long long g = (long long)(i / ncol) * stride + (i % ncol);
    uint32_t l = lab[g];
    uint32_t out = 0u;
    if (l != kInvalidLabel && flags[l] && (long)rank[l] < maxncomps) out = rank[l] + 1u;
    lab[g] = out;
}

__global__ void k_cc_count(const uint32_t *rank, const uint32_t *flags, size_t n, long maxncomps, int *out) {
    uint32_t total = rank[n - 1] + flags[n - 1];
    *out = (int)((long)total < maxncomps ? (long)total : (maxncomps > 0 ? maxncomps : 0));
}

size_t conncomp_scan_bytes(size_t n) {
    size_t tmp_bytes = 0;
    CUPHU_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, (const uint32_t *)nullptr,
                                                     (uint32_t *)nullptr, (int)n, (hipStream_t)0));
    return tmp_bytes + 16;
}

int conncomp_run(const float *unw, const float *corr, const unsigned char *mask, int nrow, int ncol, int stride,
                 double minfrac, long maxncomps, uint32_t *labels, hipStream_t st, float corr_thresh,
                 const short *poscost, const short *negcost, long conncompthresh, const CcScratch *scr) {
    if (nrow <= 0 || ncol <= 0) return 0;
    size_t n = (size_t)nrow * stride;
    size_t npx = (size_t)nrow * ncol;
    CUPHU_REQUIRE(n < 0xFFFFFFF0ull, "conncomp: image too large for 32-bit labels");
    CUPHU_REQUIRE(!scr || n <= scr->n, "conncomp: scratch smaller than the image");
    // scratch: the caller's, else allocated for this call
    DevBuf<uint32_t> own_cnt, own_flags, own_rank;
    DevBuf<int> own_count;
    auto mk = [](auto &b, size_t k) { b.alloc(k); return b.get(); };
    uint32_t *cnt = scr ? scr->cnt : mk(own_cnt, n);
    uint32_t *flags = scr ? scr->flags : mk(own_flags, n);
    uint32_t *rank = scr ? scr->rank : mk(own_rank, n);
    int *dcount = scr ? scr->count : mk(own_count, 1);
    unsigned g = (unsigned)ceil_div(n, 256), gp = (unsigned)ceil_div(npx, 256);
    dim3 tiles((unsigned)ceil_div((size_t)ncol, kCcTW), (unsigned)ceil_div((size_t)nrow, kCcTH));
    hipLaunchKernelGGL(k_cc_fill, dim3(g), dim3(256), 0, st, cnt, n, 0u);
    if (stride != ncol) hipLaunchKernelGGL(k_cc_fill, dim3(g), dim3(256), 0, st, labels, n, kInvalidLabel);
    CcGate gate{poscost, negcost, conncompthresh, (long long)(nrow - 1) * ncol, ncol};
    hipLaunchKernelGGL(k_cc_local, tiles, dim3(256), 0, st, unw, corr, mask, nrow, ncol, stride, corr_thresh, gate,
                       labels, flags);
    hipLaunchKernelGGL(k_cc_border, dim3(gp), dim3(256), 0, st, unw, nrow, ncol, stride, gate, labels);
    hipLaunchKernelGGL(k_cc_flatten, dim3(gp), dim3(256), 0, st, nrow, ncol, stride, flags, labels);
    hipLaunchKernelGGL(k_cc_compress, dim3((unsigned)ceil_div(npx, (size_t)256 * kCcRun)), dim3(256), 0, st, nrow, ncol,
                       stride, labels, cnt);
    int min_size = (int)(minfrac * (double)(int)npx);
    if (min_size < 1) min_size = 1;
    hipLaunchKernelGGL(k_cc_keep, dim3(g), dim3(256), 0, st, cnt, n, (uint32_t)min_size, flags);
    CUPHU_HIP_CHECK(hipGetLastError());
    size_t tmp_bytes = 0;
    CUPHU_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, flags, rank, (int)n, st));
    DevBuf<unsigned char> own_tmp;
    unsigned char *tmp = nullptr;
    if (scr) {
        CUPHU_REQUIRE(tmp_bytes <= scr->tmp_bytes, "conncomp: scan scratch too small");
        tmp = scr->tmp;
    } else {
        tmp = mk(own_tmp, tmp_bytes + 16);
    }
    CUPHU_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, flags, rank, (int)n, st));
    hipLaunchKernelGGL(k_cc_relabel, dim3(gp), dim3(256), 0, st, nrow, ncol, stride, rank, flags,
                       maxncomps, labels);
    hipLaunchKernelGGL(k_cc_count, dim3(1), dim3(1), 0, st, rank, flags, n, maxncomps, dcount);
    CUPHU_HIP_CHECK(hipGetLastError());
    int k = 0;
    CUPHU_HIP_CHECK(hipMemcpyAsync(&k, dcount, sizeof(int), hipMemcpyDeviceToHost, st));
    CUPHU_HIP_CHECK(hipStreamSynchronize(st));
    return k;
}

}  // namespace cuphu
