// bridge.hip -- phase bridging: whole-cycle offsets between disconnected
// regions of an unwrapped phase (reference: src/cxx/cuphu/cuphu_solver.cpp:
// 1000-1366 and cuphu_bridge_gpu.cu, a port of isce3's
// bridge_unwrapped_phase()).
//
// Pixel-scale stages run on the GPU:
//   regions   8-connected components of "unw != 0 and not masked":
//             union-find in 64 x 16 LDS tiles, then lock-free global unions
//             (atomicMin links, the smaller index stays root) across tile
//             borders, flatten, root flags + device scan -> labels 1..N
//             numbered by each region's minimum raster index
//   pruning   regions below min_num_pixel, then regions with no pixel left
//             after a binary erosion by a square and then a circular element
//   ramp      the optional ramp fit's samples are gathered in raster order
//             (the normal equations are summed on the host in that order,
//             so the coefficients are the reference's bits); removal and
//             restoration per pixel on the GPU
//   boundary  3x3 label-disagreement pixels, gathered in raster order
//   nearest   for every pair of regions the minimum packed key (1/64-px
//             distance, then the two points' per-region indices) over all
//             pairs of their subsampled boundary points: one workgroup per
//             region pair, no atomics
//   medians   exact rank selection (count/2) of a region's values in a
//             window around each bridge endpoint, 8 bits per pass
//   apply     whole-cycle correction per region, labels into conncomp
// Region-scale stages (Prim spanning tree over the region distances,
// breadth-first bridge order, the cumulative jump per bridge) are a few
// thousand numbers and run on the host, as in the reference.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <limits>
#include <queue>
#include <vector>

#include "common.hpp"
#include "util.hpp"

namespace cuphu {

namespace {

constexpr uint32_t kBrNone = 0xFFFFFFFFu;
constexpr int kBrMaxRegions = 4096;     // region-pair table (N+1)^2 x 8 B
constexpr int kBrMaxLocal = 4096;       // 12-bit per-region point index

__device__ __forceinline__ uint32_t br_load(const uint32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t br_find(uint32_t *par, uint32_t x) {
    uint32_t p = br_load(par + x);
    while (p != x) { x = p; p = br_load(par + x); }
    return x;
}
__device__ void br_union(uint32_t *par, uint32_t a, uint32_t b) {
    for (;;) {
        a = br_find(par, a);
        b = br_find(par, b);
        if (a == b) return;
        if (a < b) { uint32_t t = a; a = b; b = t; }   // link the larger root under the smaller
        uint32_t old = atomicMin(&par[a], b);
        if (old == a) return;
        a = old;
    }
}

// Regions are labelled in two levels, as the connected components are:
// union-find inside a 64 x 16 tile in LDS, then global unions over the
// forward 8-neighbour edges that cross a tile border.  Tile-local order is
// raster order, and every union links the larger root under the smaller, so
// each component's root is its smallest raster index at both levels.
constexpr int kBrTW = 64, kBrTH = 16, kBrTN = kBrTW * kBrTH;

__device__ __forceinline__ bool br_valid(const float *unw, const unsigned char *mask, size_t i) {
    return !(mask && !mask[i]) && unw[i] != 0.0f;
}
__device__ __forceinline__ int br_lds_find(const int *lp, int x) {
    int p = lp[x];
    while (p != x) { x = p; p = lp[x]; }
    return x;
}
__device__ void br_lds_union(int *lp, int a, int b) {
    for (;;) {
        a = br_lds_find(lp, a);
        b = br_lds_find(lp, b);
        if (a == b) return;
        if (a < b) { int t = a; a = b; b = t; }
        int old = atomicMin(&lp[a], b);
        if (old == a) return;
        a = old;
    }
}

// Every valid pixel's entry := the raster index of its tile-local root;
// invalid pixels kBrNone.  Forward edges inside the tile: E, SW, S, SE.
__global__ __launch_bounds__(256) void k_br_local(const float *unw, const unsigned char *mask, int nrow, int ncol,
                                                  uint32_t *par) {
    __shared__ int lp[kBrTN];
    const int c0 = (int)blockIdx.x * kBrTW, r0 = (int)blockIdx.y * kBrTH;
    for (int l = threadIdx.x; l < kBrTN; l += 256) {
        const int r = r0 + l / kBrTW, c = c0 + l % kBrTW;
        lp[l] = (r < nrow && c < ncol && br_valid(unw, mask, (size_t)r * ncol + c)) ? l : -1;
    }
    __syncthreads();
    // an entry that is -1 stays -1 (unions touch valid entries only)
    for (int l = threadIdx.x; l < kBrTN; l += 256) {
        if (lp[l] < 0) continue;
        const int lr = l / kBrTW, lc = l % kBrTW;
        if (lc + 1 < kBrTW && lp[l + 1] >= 0) br_lds_union(lp, l, l + 1);
        if (lr + 1 < kBrTH) {
            if (lc > 0 && lp[l + kBrTW - 1] >= 0) br_lds_union(lp, l, l + kBrTW - 1);
            if (lp[l + kBrTW] >= 0) br_lds_union(lp, l, l + kBrTW);
            if (lc + 1 < kBrTW && lp[l + kBrTW + 1] >= 0) br_lds_union(lp, l, l + kBrTW + 1);
        }
    }
    __syncthreads();
    for (int l = threadIdx.x; l < kBrTN; l += 256) {
        const int r = r0 + l / kBrTW, c = c0 + l % kBrTW;
        if (r >= nrow || c >= ncol) continue;
        uint32_t out = kBrNone;
        if (lp[l] >= 0) {
            const int root = br_lds_find(lp, l);
            out = (uint32_t)((size_t)(r0 + root / kBrTW) * ncol + (size_t)(c0 + root % kBrTW));
        }
        par[(size_t)r * ncol + c] = out;
    }
}

// The forward 8-neighbour edges that leave the pixel's tile: E across the
// right border, S across the bottom, SE across either, SW across the left
// border or the bottom.
__global__ void k_br_border(uint32_t *par, int nrow, int ncol) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (size_t)nrow * ncol) return;
    const int r = (int)(i / ncol), c = (int)(i % ncol);
    const bool right = (c % kBrTW) == kBrTW - 1, left = (c % kBrTW) == 0, bottom = (r % kBrTH) == kBrTH - 1;
    if (!right && !left && !bottom) return;
    if (br_load(par + i) == kBrNone) return;
    const uint32_t a = (uint32_t)i;
    if (right && c + 1 < ncol && br_load(par + i + 1) != kBrNone) br_union(par, a, (uint32_t)(i + 1));
    if (r + 1 < nrow) {
        const size_t d = i + (size_t)ncol;
        if (bottom && br_load(par + d) != kBrNone) br_union(par, a, (uint32_t)d);
        if ((right || bottom) && c + 1 < ncol && br_load(par + d + 1) != kBrNone) br_union(par, a, (uint32_t)(d + 1));
        if ((left || bottom) && c > 0 && br_load(par + d - 1) != kBrNone) br_union(par, a, (uint32_t)(d - 1));
    }
}

__global__ void k_br_flatten(uint32_t *par, size_t n, uint32_t *rootflag) {
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    uint32_t f = 0u;
    if (br_load(par + i) != kBrNone) {
        const uint32_t root = br_find(par, (uint32_t)i);
        atomicMin(&par[i], root);   // ancestors only: concurrent finds stay valid
        f = (root == (uint32_t)i) ? 1u : 0u;
    }
    rootflag[i] = f;
}

__global__ void k_br_label(const uint32_t *par, const uint32_t *rank, size_t n, uint32_t *lab, uint32_t *cnt) {
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const uint32_t p = par[i];
    const uint32_t l = (p == kBrNone) ? 0u : rank[p] + 1u;
    lab[i] = l;
    if (l) atomicAdd(&cnt[l], 1u);
}

__global__ void k_br_remap(uint32_t *lab, size_t n, const uint32_t *remap) {
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) lab[i] = remap[lab[i]];
}

// a labelled pixel survives the erosion when every structuring-element
// offset lands on a labelled pixel inside the image
__global__ void k_br_erode(const uint32_t *lab, int nrow, int ncol, const int2 *se, int nse, unsigned char *surv) {
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (size_t)nrow * ncol) return;
    const uint32_t l = lab[i];
    if (!l) return;
    const int r = (int)(i / ncol), c = (int)(i % ncol);
    for (int k = 0; k < nse; ++k) {
        const int rr = r + se[k].x, cc = c + se[k].y;
        if (rr < 0 || rr >= nrow || cc < 0 || cc >= ncol || !lab[(size_t)rr * ncol + cc]) return;
    }
    surv[l] = 1;   // every writer stores the same value
}

__global__ void k_br_boundary(const uint32_t *lab, int nrow, int ncol, uint32_t *flag) {
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (size_t)nrow * ncol) return;
    const int r = (int)(i / ncol), c = (int)(i % ncol);
    const uint32_t ctr = lab[i];
    uint32_t mx = ctr, mn = ctr;
    for (int dr = -1; dr <= 1; ++dr)
        for (int dc = -1; dc <= 1; ++dc) {
            const int rr = r + dr, cc = c + dc;
            const uint32_t v = (rr < 0 || rr >= nrow || cc < 0 || cc >= ncol) ? ctr : lab[(size_t)rr * ncol + cc];
            mx = v > mx ? v : mx;
            mn = v < mn ? v : mn;
        }
    flag[i] = (mx != mn && ctr > 0u) ? 1u : 0u;
}

// ramp samples of the reference region: every step-th row and column at
// offset `half` when step > 1
__global__ void k_br_ramp_flag(const uint32_t *lab, const float *unw, int nrow, int ncol, uint32_t ref, int step,
                               int half, uint32_t *flag) {
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (size_t)nrow * ncol) return;
    const int r = (int)(i / ncol), c = (int)(i % ncol);
    bool take = lab[i] == ref && unw[i] != 0.0f;
    if (step > 1) take = take && (r % step == half) && (c % step == half);
    flag[i] = take ? 1u : 0u;
}

// stream compaction in raster order: idx[rank[i]] = i where flag[i]
__global__ void k_br_gather(const uint32_t *flag, const uint32_t *rank, size_t n, uint32_t *idx) {
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n && flag[i]) idx[rank[i]] = (uint32_t)i;
}

// dst[k] = src[idx[k]]
template <class T>
__global__ void k_br_take(const uint32_t *idx, uint32_t m, const T *src, T *dst) {
    size_t k = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (k < m) dst[k] = src[idx[k]];
}

__global__ void k_br_nonzero(const float *unw, size_t n, unsigned char *nz) {
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) nz[i] = unw[i] != 0.0f;
}

struct RampDev {
    double coef[6];
    int type, nterm;
};
__device__ __forceinline__ int br_basis(int rt, double row, double col, double g[6]) {
    switch (rt) {   // cuphu_solver.cpp:1049-1067
    case CUPHU_RAMP_LINEAR: g[0] = row; g[1] = col; g[2] = 1; return 3;
    case CUPHU_RAMP_QUADRATIC: g[0] = row * row; g[1] = col * col; g[2] = row * col; g[3] = row; g[4] = col; g[5] = 1; return 6;
    case CUPHU_RAMP_LINEAR_RANGE: g[0] = col; g[1] = 1; return 2;
    case CUPHU_RAMP_LINEAR_AZIMUTH: g[0] = row; g[1] = 1; return 2;
    case CUPHU_RAMP_QUADRATIC_RANGE: g[0] = col * col; g[1] = col; g[2] = 1; return 3;
    case CUPHU_RAMP_QUADRATIC_AZIMUTH: g[0] = row * row; g[1] = row; g[2] = 1; return 3;
    default: return 0;
    }
}
// unw += (float)(sign * ramp) where the original value was nonzero
__global__ void k_br_ramp(float *unw, const unsigned char *nz, int nrow, int ncol, RampDev rp, double sign) {
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (size_t)nrow * ncol || !nz[i]) return;
    double g[6];
    br_basis(rp.type, (double)(int)(i / ncol), (double)(int)(i % ncol), g);
    double ramp = 0.0;
    for (int a = 0; a < rp.nterm; ++a) ramp += rp.coef[a] * g[a];
    unw[i] += (float)(sign * ramp);
}

__device__ __forceinline__ uint64_t br_key(float d, uint32_t lp, uint32_t lq) {
    const uint64_t qd = (uint64_t)(d * 64.0f);   // 1/64 px, truncated
    return (qd << 24) | ((uint64_t)lp << 12) | (uint64_t)lq;
}
__device__ __forceinline__ uint64_t min64(uint64_t a, uint64_t b) { return a < b ? a : b; }

// Workgroup per region pair li < lj (grid N x N, the rest exit): minimum
// key over all point pairs, in both orders (the tie-break indices differ).
__global__ __launch_bounds__(256) void k_br_pairs(const int *px, const int *py, const int *start, int N,
                                                  unsigned long long *table) {
    const int li = (int)blockIdx.y + 1, lj = (int)blockIdx.x + 1;
    if (li >= lj || lj > N) return;   // uniform per block
    const int si = start[li], ni = start[li + 1] - si, sj = start[lj], nj = start[lj + 1] - sj;
    uint64_t kij = ~0ull, kji = ~0ull;
    for (int a = threadIdx.x; a < ni; a += 256) {
        const int xa = px[si + a], ya = py[si + a];
        for (int b = 0; b < nj; ++b) {
            const float dx = (float)(px[sj + b] - xa), dy = (float)(py[sj + b] - ya);
            const float dxx = dx * dx, dyy = dy * dy;
            const float d = sqrtf(dxx + dyy);
            kij = min64(kij, br_key(d, (uint32_t)a, (uint32_t)b));
            kji = min64(kji, br_key(d, (uint32_t)b, (uint32_t)a));
        }
    }
    __shared__ uint64_t s1[256], s2[256];
    s1[threadIdx.x] = kij;
    s2[threadIdx.x] = kji;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            s1[threadIdx.x] = min64(s1[threadIdx.x], s1[threadIdx.x + o]);
            s2[threadIdx.x] = min64(s2[threadIdx.x], s2[threadIdx.x + o]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        table[(size_t)li * (N + 1) + lj] = s1[0];
        table[(size_t)lj * (N + 1) + li] = s2[0];
    }
}

struct BrWin {
    int r0, r1, c0, c1;
    uint32_t region;
    int pad_;
};
__device__ __forceinline__ uint32_t br_okey(float v) {
    uint32_t u = __float_as_uint(v + 0.0f);   // -0 -> +0: equal values, one key
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float br_oval(uint32_t k) {
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k);
}
// Workgroup per window: the value of rank count/2 among the window's pixels
// of the given region (ok[w] = 0 when there are none).
__global__ __launch_bounds__(256) void k_br_window_median(const float *unw, const uint32_t *lab, int ncol,
                                                          const BrWin *wins, double *med, unsigned char *ok) {
    __shared__ unsigned cnt[256];
    __shared__ unsigned s_prefix, s_rank, s_stop;
    const BrWin w = wins[blockIdx.x];
    const int wc = w.c1 - w.c0;
    const long long m = (w.r1 > w.r0 && wc > 0) ? (long long)(w.r1 - w.r0) * wc : 0;
    unsigned prefix = 0u, rank = 0u;
    for (int pass = 0; pass < 4; ++pass) {
        const int shift = 24 - 8 * pass;
        cnt[threadIdx.x] = 0u;
        __syncthreads();
        for (long long k = threadIdx.x; k < m; k += 256) {
            const size_t g = (size_t)(w.r0 + (int)(k / wc)) * ncol + (w.c0 + (int)(k % wc));
            if (lab[g] != w.region) continue;
            const uint32_t key = br_okey(unw[g]);
            if (pass > 0 && (key >> (shift + 8)) != prefix) continue;
            atomicAdd(&cnt[(key >> shift) & 255u], 1u);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            unsigned stop = 0u;
            if (pass == 0) {
                unsigned tot = 0u;
                for (int b = 0; b < 256; ++b) tot += cnt[b];
                if (tot == 0u) stop = 1u;
                else rank = tot / 2u;
            }
            if (!stop) {
                unsigned below = 0u;
                int b = 0;
                for (; b < 255; ++b) {
                    if (rank < below + cnt[b]) break;
                    below += cnt[b];
                }
                prefix = (prefix << 8) | (unsigned)b;
                rank -= below;
            }
            s_prefix = prefix;
            s_rank = rank;
            s_stop = stop;
        }
        __syncthreads();
        prefix = s_prefix;
        rank = s_rank;
        if (s_stop) {
            if (threadIdx.x == 0) { med[blockIdx.x] = 0.0; ok[blockIdx.x] = 0; }
            return;
        }
    }
    if (threadIdx.x == 0) {
        med[blockIdx.x] = (double)br_oval(prefix);
        ok[blockIdx.x] = 1;
    }
}

__global__ void k_br_apply(float *unw, uint32_t *cc, const uint32_t *lab, size_t n, const float *corr) {
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const uint32_t l = lab[i];
    cc[i] = l;
    if (l) unw[i] += corr[l];
}

// ── host side ───────────────────────────────────────────────────────────────

unsigned grid1(size_t n) { return (unsigned)ceil_div(std::max<size_t>(n, 1), 256); }

// exclusive scan of flag[n] into rank[n]; returns the total
uint32_t scan_flags(const uint32_t *flag, uint32_t *rank, size_t n, hipStream_t st) {
    size_t tmp_bytes = 0;
    CUPHU_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, flag, rank, (int)n, st));
    DevBuf<unsigned char> tmp(tmp_bytes + 16);
    CUPHU_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(tmp.get(), tmp_bytes, flag, rank, (int)n, st));
    uint32_t last_rank = 0, last_flag = 0;
    CUPHU_HIP_CHECK(hipMemcpyAsync(&last_rank, rank + n - 1, 4, hipMemcpyDeviceToHost, st));
    CUPHU_HIP_CHECK(hipMemcpyAsync(&last_flag, flag + n - 1, 4, hipMemcpyDeviceToHost, st));
    CUPHU_HIP_CHECK(hipStreamSynchronize(st));
    return last_rank + last_flag;
}

// Apply a host remap (regions kept = nonzero) and return the new count.
int remap_regions(uint32_t *d_lab, size_t n, std::vector<uint32_t> &remap, std::vector<long long> &cnt, hipStream_t st) {
    DevBuf<uint32_t> d_rm(remap.size());
    CUPHU_HIP_CHECK(hipMemcpyAsync(d_rm.get(), remap.data(), remap.size() * 4, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_br_remap, dim3(grid1(n)), dim3(256), 0, st, d_lab, n, d_rm.get());
    CUPHU_HIP_CHECK(hipGetLastError());
    int k = 0;
    for (uint32_t v : remap) k = std::max<int>(k, (int)v);
    std::vector<long long> nc((size_t)k + 1, 0);
    for (size_t l = 1; l < remap.size(); ++l)
        if (remap[l]) nc[remap[l]] = cnt[l];
    cnt.swap(nc);
    CUPHU_HIP_CHECK(hipStreamSynchronize(st));   // d_rm is freed on return
    return k;
}

int erode_prune(uint32_t *d_lab, int nrow, int ncol, int N, int e, bool circular, std::vector<long long> &cnt,
                hipStream_t st) {
    if (e <= 0 || N == 0) return N;
    std::vector<int2> se;
    for (int dr = -e; dr <= e; ++dr)
        for (int dc = -e; dc <= e; ++dc)
            if (!circular || dr * dr + dc * dc <= e * e) se.push_back(make_int2(dr, dc));
    const size_t n = (size_t)nrow * ncol;
    DevBuf<int2> d_se(se.size());
    DevBuf<unsigned char> d_surv((size_t)N + 1);
    CUPHU_HIP_CHECK(hipMemcpyAsync(d_se.get(), se.data(), se.size() * sizeof(int2), hipMemcpyHostToDevice, st));
    CUPHU_HIP_CHECK(hipMemsetAsync(d_surv.get(), 0, (size_t)N + 1, st));
    hipLaunchKernelGGL(k_br_erode, dim3(grid1(n)), dim3(256), 0, st, d_lab, nrow, ncol, d_se.get(), (int)se.size(),
                       d_surv.get());
    CUPHU_HIP_CHECK(hipGetLastError());
    std::vector<unsigned char> surv((size_t)N + 1);
    CUPHU_HIP_CHECK(hipMemcpyAsync(surv.data(), d_surv.get(), (size_t)N + 1, hipMemcpyDeviceToHost, st));
    CUPHU_HIP_CHECK(hipStreamSynchronize(st));
    std::vector<uint32_t> remap((size_t)N + 1, 0u);
    uint32_t next = 1;
    for (int l = 1; l <= N; ++l)
        if (surv[(size_t)l]) remap[(size_t)l] = next++;
    return remap_regions(d_lab, n, remap, cnt, st);
}

// Gaussian elimination with partial pivoting (cuphu_solver.cpp:1069-1088)
bool solve_small(double A[6][6], double b[6], int n, double x[6]) {
    for (int p = 0; p < n; ++p) {
        int piv = p;
        for (int r = p + 1; r < n; ++r)
            if (std::fabs(A[r][p]) > std::fabs(A[piv][p])) piv = r;
        if (std::fabs(A[piv][p]) < 1e-12) return false;
        if (piv != p) {
            for (int k = 0; k < 6; ++k) std::swap(A[piv][k], A[p][k]);
            std::swap(b[piv], b[p]);
        }
        for (int r = p + 1; r < n; ++r) {
            const double f = A[r][p] / A[p][p];
            for (int k = p; k < n; ++k) A[r][k] -= f * A[p][k];
            b[r] -= f * b[p];
        }
    }
    for (int r = n - 1; r >= 0; --r) {
        double s = b[r];
        for (int k = r + 1; k < n; ++k) s -= A[r][k] * x[k];
        x[r] = s / A[r][r];
    }
    return true;
}
int host_basis(int rt, double row, double col, double g[6]) {
    switch (rt) {
    case CUPHU_RAMP_LINEAR: g[0] = row; g[1] = col; g[2] = 1; return 3;
    case CUPHU_RAMP_QUADRATIC: g[0] = row * row; g[1] = col * col; g[2] = row * col; g[3] = row; g[4] = col; g[5] = 1; return 6;
    case CUPHU_RAMP_LINEAR_RANGE: g[0] = col; g[1] = 1; return 2;
    case CUPHU_RAMP_LINEAR_AZIMUTH: g[0] = row; g[1] = 1; return 2;
    case CUPHU_RAMP_QUADRATIC_RANGE: g[0] = col * col; g[1] = col; g[2] = 1; return 3;
    case CUPHU_RAMP_QUADRATIC_AZIMUTH: g[0] = row * row; g[1] = row; g[2] = 1; return 3;
    default: return 0;
    }
}

}  // namespace

void bridge_dev(float *d_unw, uint32_t *d_cc, const unsigned char *d_mask, int nrow, int ncol,
                const CuPhuBridgeParams &bp, hipStream_t st) {
    const size_t n = (size_t)nrow * ncol;
    CUPHU_REQUIRE(n < 0x7FFFFFF0ull, "bridging: image too large for 31-bit pixel indices");
    CUPHU_REQUIRE(bp.ramp_type >= CUPHU_RAMP_NONE && bp.ramp_type <= CUPHU_RAMP_QUADRATIC_AZIMUTH,
                  "bridging: invalid ramp type");
    TraceRange tr("cuphu:bridge");
    // ── regions of (unw != 0 and not masked), 8-connected ──
    DevBuf<uint32_t> d_par(n), d_flag(n), d_rank(n), d_lab(n);
    hipLaunchKernelGGL(k_br_local, dim3((unsigned)ceil_div((size_t)ncol, kBrTW), (unsigned)ceil_div((size_t)nrow, kBrTH)),
                       dim3(256), 0, st, d_unw, d_mask, nrow, ncol, d_par.get());
    hipLaunchKernelGGL(k_br_border, dim3(grid1(n)), dim3(256), 0, st, d_par.get(), nrow, ncol);
    hipLaunchKernelGGL(k_br_flatten, dim3(grid1(n)), dim3(256), 0, st, d_par.get(), n, d_flag.get());
    CUPHU_HIP_CHECK(hipGetLastError());
    int N = (int)scan_flags(d_flag.get(), d_rank.get(), n, st);
    if (N <= 1) return;
    std::vector<long long> cnt;
    {
        DevBuf<uint32_t> d_cnt((size_t)N + 1);
        CUPHU_HIP_CHECK(hipMemsetAsync(d_cnt.get(), 0, ((size_t)N + 1) * 4, st));
        hipLaunchKernelGGL(k_br_label, dim3(grid1(n)), dim3(256), 0, st, d_par.get(), d_rank.get(), n, d_lab.get(),
                           d_cnt.get());
        CUPHU_HIP_CHECK(hipGetLastError());
        std::vector<uint32_t> c32((size_t)N + 1);
        CUPHU_HIP_CHECK(hipMemcpyAsync(c32.data(), d_cnt.get(), c32.size() * 4, hipMemcpyDeviceToHost, st));
        CUPHU_HIP_CHECK(hipStreamSynchronize(st));
        cnt.assign(c32.begin(), c32.end());
    }
    d_par.release();
    // ── pruning (:1290-1299) ──
    {
        std::vector<uint32_t> remap((size_t)N + 1, 0u);
        uint32_t next = 1;
        for (int l = 1; l <= N; ++l)
            if (cnt[(size_t)l] >= bp.min_num_pixel) remap[(size_t)l] = next++;
        N = remap_regions(d_lab.get(), n, remap, cnt, st);
    }
    if (N <= 1) return;
    N = erode_prune(d_lab.get(), nrow, ncol, N, bp.erosion_size, false, cnt, st);
    if (N <= 1) return;
    N = erode_prune(d_lab.get(), nrow, ncol, N, bp.erosion_size, true, cnt, st);
    if (N <= 1) return;
    CUPHU_REQUIRE(N <= kBrMaxRegions, "bridging: more than 4096 regions");
    const int n1 = N + 1;
    // reference region: the largest, the lowest label on a tie (:1037-1046)
    int ref = 1;
    for (int l = 2; l <= N; ++l)
        if (cnt[(size_t)l] > cnt[(size_t)ref]) ref = l;
    // ── optional ramp fitted on the reference region (:1094-1141) ──
    RampDev rp{};
    bool ramp = false;
    DevBuf<unsigned char> d_nz;
    if (bp.ramp_type != CUPHU_RAMP_NONE && cnt[(size_t)ref] > 0) {
        double dummy[6];
        rp.type = (int)bp.ramp_type;
        rp.nterm = host_basis(rp.type, 0, 0, dummy);
        const long long count = cnt[(size_t)ref];
        int step = 1;
        if (bp.ramp_max_num_sample > 0 && count > bp.ramp_max_num_sample)
            step = (int)std::ceil(std::sqrt((double)count / bp.ramp_max_num_sample));
        hipLaunchKernelGGL(k_br_ramp_flag, dim3(grid1(n)), dim3(256), 0, st, d_lab.get(), d_unw, nrow, ncol,
                           (uint32_t)ref, step, step / 2, d_flag.get());
        CUPHU_HIP_CHECK(hipGetLastError());
        const uint32_t ns = scan_flags(d_flag.get(), d_rank.get(), n, st);
        std::vector<uint32_t> idx(ns);
        std::vector<float> val(ns);
        if (ns) {
            DevBuf<uint32_t> d_idx(ns);
            DevBuf<float> d_val(ns);
            hipLaunchKernelGGL(k_br_gather, dim3(grid1(n)), dim3(256), 0, st, d_flag.get(), d_rank.get(), n,
                               d_idx.get());
            hipLaunchKernelGGL(k_br_take<float>, dim3(grid1(ns)), dim3(256), 0, st, d_idx.get(), ns,
                               (const float *)d_unw, d_val.get());
            CUPHU_HIP_CHECK(hipGetLastError());
            CUPHU_HIP_CHECK(hipMemcpyAsync(idx.data(), d_idx.get(), (size_t)ns * 4, hipMemcpyDeviceToHost, st));
            CUPHU_HIP_CHECK(hipMemcpyAsync(val.data(), d_val.get(), (size_t)ns * 4, hipMemcpyDeviceToHost, st));
            CUPHU_HIP_CHECK(hipStreamSynchronize(st));
        }
        double A[6][6] = {{0}}, b[6] = {0};
        for (uint32_t k = 0; k < ns; ++k) {   // raster order, as the reference sums
            double g[6];
            host_basis(rp.type, (double)(int)(idx[k] / (uint32_t)ncol), (double)(int)(idx[k] % (uint32_t)ncol), g);
            for (int a = 0; a < rp.nterm; ++a) {
                b[a] += g[a] * (double)val[k];
                for (int c = 0; c < rp.nterm; ++c) A[a][c] += g[a] * g[c];
            }
        }
        ramp = ns > 0 && solve_small(A, b, rp.nterm, rp.coef);
        if (ramp) {
            d_nz.alloc(n);
            hipLaunchKernelGGL(k_br_nonzero, dim3(grid1(n)), dim3(256), 0, st, d_unw, n, d_nz.get());
            hipLaunchKernelGGL(k_br_ramp, dim3(grid1(n)), dim3(256), 0, st, d_unw, d_nz.get(), nrow, ncol, rp, -1.0);
            CUPHU_HIP_CHECK(hipGetLastError());
        }
    }
    // ── boundary points, raster order, subsampled per region (:363-420) ──
    hipLaunchKernelGGL(k_br_boundary, dim3(grid1(n)), dim3(256), 0, st, d_lab.get(), nrow, ncol, d_flag.get());
    CUPHU_HIP_CHECK(hipGetLastError());
    const uint32_t nb = scan_flags(d_flag.get(), d_rank.get(), n, st);
    std::vector<uint32_t> bidx(nb), blab(nb);
    if (nb) {
        DevBuf<uint32_t> d_idx(nb), d_bl(nb);
        hipLaunchKernelGGL(k_br_gather, dim3(grid1(n)), dim3(256), 0, st, d_flag.get(), d_rank.get(), n, d_idx.get());
        hipLaunchKernelGGL(k_br_take<uint32_t>, dim3(grid1(nb)), dim3(256), 0, st, d_idx.get(), nb,
                           (const uint32_t *)d_lab.get(), d_bl.get());
        CUPHU_HIP_CHECK(hipGetLastError());
        CUPHU_HIP_CHECK(hipMemcpyAsync(bidx.data(), d_idx.get(), (size_t)nb * 4, hipMemcpyDeviceToHost, st));
        CUPHU_HIP_CHECK(hipMemcpyAsync(blab.data(), d_bl.get(), (size_t)nb * 4, hipMemcpyDeviceToHost, st));
        CUPHU_HIP_CHECK(hipStreamSynchronize(st));
    }
    const int ms = std::min(bp.max_boundary_samples > 0 ? bp.max_boundary_samples : 1, kBrMaxLocal);
    std::vector<std::vector<uint32_t>> reg((size_t)n1);
    for (uint32_t k = 0; k < nb; ++k)
        if (blab[k] && (int)blab[k] <= N) reg[blab[k]].push_back(bidx[k]);
    std::vector<int> px, py, start((size_t)n1 + 1, 0);
    for (int l = 1; l <= N; ++l) {
        const std::vector<uint32_t> &p = reg[(size_t)l];
        const size_t stride = std::max<size_t>(1, p.size() / (size_t)ms);
        int local = 0;
        for (size_t k = 0; k < p.size() && local < ms; k += stride, ++local) {
            px.push_back((int)(p[k] % (uint32_t)ncol));
            py.push_back((int)(p[k] / (uint32_t)ncol));
        }
        start[(size_t)l + 1] = (int)px.size();
    }
    // ── nearest boundary points per region pair ──
    std::vector<unsigned long long> table((size_t)n1 * n1, ~0ull);
    if (!px.empty()) {
        DevBuf<int> d_px(px.size()), d_py(py.size()), d_start(start.size());
        DevBuf<unsigned long long> d_table(table.size());
        CUPHU_HIP_CHECK(hipMemcpyAsync(d_px.get(), px.data(), px.size() * 4, hipMemcpyHostToDevice, st));
        CUPHU_HIP_CHECK(hipMemcpyAsync(d_py.get(), py.data(), py.size() * 4, hipMemcpyHostToDevice, st));
        CUPHU_HIP_CHECK(hipMemcpyAsync(d_start.get(), start.data(), start.size() * 4, hipMemcpyHostToDevice, st));
        CUPHU_HIP_CHECK(hipMemcpyAsync(d_table.get(), table.data(), table.size() * 8, hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_br_pairs, dim3((unsigned)N, (unsigned)N), dim3(256), 0, st, d_px.get(), d_py.get(),
                           d_start.get(), N, d_table.get());
        CUPHU_HIP_CHECK(hipGetLastError());
        CUPHU_HIP_CHECK(hipMemcpyAsync(table.data(), d_table.get(), table.size() * 8, hipMemcpyDeviceToHost, st));
        CUPHU_HIP_CHECK(hipStreamSynchronize(st));
    }
    // distances (-1 = none) and endpoints (:603-625)
    auto dist = [&](int i, int j) -> float {
        const unsigned long long k = table[(size_t)i * n1 + j];
        if (k == ~0ull) return std::numeric_limits<float>::infinity();
        return (float)(k >> 24) / 64.0f;
    };
    // ── Prim spanning tree from ref, breadth-first bridge order (:1169-1229) ──
    std::vector<char> in((size_t)n1, 0);
    std::vector<float> bd((size_t)n1, std::numeric_limits<float>::infinity());
    std::vector<int> bpar((size_t)n1, -1);
    in[(size_t)ref] = 1;
    for (int j = 1; j <= N; ++j)
        if (j != ref) { bd[(size_t)j] = dist(ref, j); bpar[(size_t)j] = ref; }
    std::vector<std::pair<int, int>> edges;
    for (int it = 1; it < N; ++it) {
        int u = -1;
        float ud = std::numeric_limits<float>::infinity();
        for (int j = 1; j <= N; ++j)
            if (!in[(size_t)j] && bd[(size_t)j] < ud) { ud = bd[(size_t)j]; u = j; }
        if (u < 0) break;
        in[(size_t)u] = 1;
        edges.push_back({bpar[(size_t)u], u});
        for (int j = 1; j <= N; ++j) {
            if (in[(size_t)j]) continue;
            const float d = dist(u, j);
            if (d < bd[(size_t)j]) { bd[(size_t)j] = d; bpar[(size_t)j] = u; }
        }
    }
    std::vector<std::vector<int>> adj((size_t)n1);
    for (auto &e : edges) { adj[(size_t)e.first].push_back(e.second); adj[(size_t)e.second].push_back(e.first); }
    struct Bridge { int u, v; };
    std::vector<Bridge> bridges;
    {
        std::vector<char> vis((size_t)n1, 0);
        std::queue<int> q;
        q.push(ref);
        vis[(size_t)ref] = 1;
        while (!q.empty()) {
            const int u = q.front();
            q.pop();
            for (int v : adj[(size_t)u]) {
                if (vis[(size_t)v]) continue;
                vis[(size_t)v] = 1;
                bridges.push_back({u, v});
                q.push(v);
            }
        }
    }
    // ── window medians at both ends of every bridge (:1232-1253) ──
    const int rad = std::min(bp.radius, std::min(nrow, ncol) / 2);
    std::vector<BrWin> wins;
    std::vector<int> win_of(bridges.size() * 2, -1);
    for (size_t k = 0; k < bridges.size(); ++k) {
        const unsigned long long key = table[(size_t)bridges[k].u * n1 + bridges[k].v];
        if (key == ~0ull) continue;
        const int a = (int)((key >> 12) & 4095u), b = (int)(key & 4095u);
        const int su = start[(size_t)bridges[k].u], sv = start[(size_t)bridges[k].v];
        const int ys[2] = {py[(size_t)(su + a)], py[(size_t)(sv + b)]};
        const int xs[2] = {px[(size_t)(su + a)], px[(size_t)(sv + b)]};
        const int regs[2] = {bridges[k].u, bridges[k].v};
        for (int s = 0; s < 2; ++s) {
            BrWin w{};
            w.r0 = std::max(0, ys[s] - rad); w.r1 = std::min(nrow, ys[s] + rad + 1);
            w.c0 = std::max(0, xs[s] - rad); w.c1 = std::min(ncol, xs[s] + rad + 1);
            w.region = (uint32_t)regs[s];
            if (w.r1 <= w.r0 || w.c1 <= w.c0) continue;   // empty window (radius < 0): no median, as in :1239-1248
            win_of[2 * k + (size_t)s] = (int)wins.size();
            wins.push_back(w);
        }
    }
    std::vector<double> med(wins.size());
    std::vector<unsigned char> mok(wins.size());
    if (!wins.empty()) {
        DevBuf<BrWin> d_w(wins.size());
        DevBuf<double> d_med(wins.size());
        DevBuf<unsigned char> d_ok(wins.size());
        CUPHU_HIP_CHECK(hipMemcpyAsync(d_w.get(), wins.data(), wins.size() * sizeof(BrWin), hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_br_window_median, dim3((unsigned)wins.size()), dim3(256), 0, st, d_unw, d_lab.get(), ncol,
                           d_w.get(), d_med.get(), d_ok.get());
        CUPHU_HIP_CHECK(hipGetLastError());
        CUPHU_HIP_CHECK(hipMemcpyAsync(med.data(), d_med.get(), med.size() * 8, hipMemcpyDeviceToHost, st));
        CUPHU_HIP_CHECK(hipMemcpyAsync(mok.data(), d_ok.get(), mok.size(), hipMemcpyDeviceToHost, st));
        CUPHU_HIP_CHECK(hipStreamSynchronize(st));
    }
    // cumulative whole-cycle jump per region along the bridges (:1326-1339)
    std::vector<double> corr((size_t)n1, 0.0);
    for (size_t k = 0; k < bridges.size(); ++k) {
        const int w0 = win_of[2 * k], w1 = win_of[2 * k + 1];
        if (w0 < 0 || w1 < 0 || !mok[(size_t)w0] || !mok[(size_t)w1]) continue;
        const double diff = med[(size_t)w1] - (med[(size_t)w0] + corr[(size_t)bridges[k].u]);
        double nj = std::floor((std::fabs(diff) + M_PI) / (2.0 * M_PI));
        if (diff > 0) nj *= -1.0;
        corr[(size_t)bridges[k].v] = nj * (2.0 * M_PI);
    }
    // ── apply; conncomp takes the bridging regions (:1345-1352) ──
    std::vector<float> corrf((size_t)n1);
    for (int l = 0; l <= N; ++l) corrf[(size_t)l] = (float)corr[(size_t)l];
    DevBuf<float> d_corr((size_t)n1);
    CUPHU_HIP_CHECK(hipMemcpyAsync(d_corr.get(), corrf.data(), (size_t)n1 * 4, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_br_apply, dim3(grid1(n)), dim3(256), 0, st, d_unw, d_cc, d_lab.get(), n, d_corr.get());
    CUPHU_HIP_CHECK(hipGetLastError());
    if (ramp) {
        hipLaunchKernelGGL(k_br_ramp, dim3(grid1(n)), dim3(256), 0, st, d_unw, d_nz.get(), nrow, ncol, rp, 1.0);
        CUPHU_HIP_CHECK(hipGetLastError());
    }
    CUPHU_HIP_CHECK(hipStreamSynchronize(st));   // scratch buffers are freed on return
}

}  // namespace cuphu
