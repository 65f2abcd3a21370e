// postpass.hip -- optional passes over the stitched scene
// (reference: src/cxx/cuphu/cuphu_solver.cpp:1385-1657).
//
// fix_cycle_spikes (:1385-1441): per-line medians on the GPU (one workgroup
// per row or column, exact rank selection on the order-preserving key, 8
// bits per pass, LDS counts), the reference's sequential spike decision over
// the line medians on the host (a few thousand numbers; each decision sees
// the previous line's corrected median), and the shift of the flagged lines
// on the GPU.  Rows first, then columns on the row-corrected phase.
//
// Neighbour feedback (:1469-1657): for each internal tile seam in the
// reference's order (column seams, then row seams, each on the result of
// the previous), the two pixel lines either side of the seam are read
// back, the correction profile (difference, moving median, bounded gap
// fill) is formed on the host in FP64 as the reference does, and the
// feathered correction over the seam zone is applied on the GPU.
#include <algorithm>
#include <cmath>
#include <vector>

#include "common.hpp"
#include "util.hpp"

namespace cuphu {

namespace {

__device__ __forceinline__ uint32_t pp_key(float v) {
    uint32_t u = __float_as_uint(v + 0.0f);   // -0 -> +0: equal values, one key
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float pp_value(uint32_t k) {
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k);
}

// Line l's valid elements are base(l) + j*estride, j < len, base = l*lstride.
// med[l] = the valid element of rank count/2 (the reference's nth_element
// at size/2), have[l] = count > 20.  Used for rows (contiguous elements;
// k_col_median below takes the columns).  Workgroups are spread so that each
// XCD takes a contiguous run of lines.
__global__ __launch_bounds__(256) void k_line_median(const float *unw, const unsigned char *mask, int nlines, int len,
                                                     long long lstride, long long estride, double *med,
                                                     unsigned char *have) {
    __shared__ unsigned cnt[256];
    __shared__ unsigned s_prefix, s_rank, s_stop;
    const int X = 8, nb = (int)gridDim.x, b = (int)blockIdx.x;
    const int q = nb / X, rr = nb % X, x = b % X, i = b / X;
    const int line = x * q + (x < rr ? x : rr) + i;
    if (line >= nlines) return;   // uniform per block
    const long long base = (long long)line * lstride;
    unsigned prefix = 0u, rank = 0u;
    for (int pass = 0; pass < 4; ++pass) {
        const int shift = 24 - 8 * pass;
        cnt[threadIdx.x] = 0u;
        __syncthreads();
        for (int j = threadIdx.x; j < len; j += 256) {
            const long long g = base + (long long)j * estride;
            if (mask && !mask[g]) continue;
            const uint32_t key = pp_key(unw[g]);
            if (pass > 0 && (key >> (shift + 8)) != prefix) continue;
            atomicAdd(&cnt[(key >> shift) & 255u], 1u);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            unsigned stop = 0u;
            if (pass == 0) {
                unsigned tot = 0u;
                for (int k = 0; k < 256; ++k) tot += cnt[k];
                if (tot <= 20u) stop = 1u;
                else rank = tot / 2u;
            }
            if (!stop) {
                unsigned below = 0u;
                int k = 0;
                for (; k < 255; ++k) {
                    if (rank < below + cnt[k]) break;
                    below += cnt[k];
                }
                prefix = (prefix << 8) | (unsigned)k;
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
            if (threadIdx.x == 0) { med[line] = 0.0; have[line] = 0; }
            return;
        }
    }
    if (threadIdx.x == 0) {
        med[line] = (double)pp_value(prefix);
        have[line] = 1;
    }
}

// The same selection for the columns of an image (elements ncol apart), 32
// adjacent columns per workgroup.  Each wave reads two rows of 32 consecutive
// floats (one 128-byte line each) per step, so every line fetched serves 32
// columns; with one workgroup per column each line fetched served one float,
// and the column pass moved 17x its bytes.  Per-column counts in LDS (rows
// of 257 words: the same bucket of neighbouring columns falls in different
// banks); thread j < 32 selects column c0 + j's next 8 bits.
constexpr int kColBlk = 32;
__global__ __launch_bounds__(256) void k_col_median(const float *unw, const unsigned char *mask, int nrow, int ncol,
                                                    double *med, unsigned char *have) {
    __shared__ unsigned cnt[kColBlk][257];
    __shared__ unsigned s_prefix[kColBlk], s_rank[kColBlk], s_stop[kColBlk];
    const int c0 = (int)blockIdx.x * kColBlk;
    const int j = (int)threadIdx.x % kColBlk, g = (int)threadIdx.x / kColBlk;   // column in the block, row group
    const int c = c0 + j;
    if (threadIdx.x < kColBlk) { s_prefix[threadIdx.x] = 0u; s_rank[threadIdx.x] = 0u; s_stop[threadIdx.x] = 0u; }
    for (int pass = 0; pass < 4; ++pass) {
        const int shift = 24 - 8 * pass;
        for (int k = (int)threadIdx.x; k < kColBlk * 257; k += 256) (&cnt[0][0])[k] = 0u;
        __syncthreads();
        const unsigned prefix = s_prefix[j];
        if (c < ncol && !s_stop[j]) {
            // rows g, g + RG, g + 2 RG, ...; eight of them loaded before any is
            // counted, so eight loads are in flight per thread
            constexpr int RG = 256 / kColBlk, U = 8;
            for (int r0 = g; r0 < nrow; r0 += U * RG) {
                float v[U];
                bool ok[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int r = r0 + u * RG;
                    ok[u] = r < nrow;
                    v[u] = 0.0f;
                    if (ok[u]) {
                        const long long i = (long long)r * ncol + c;
                        v[u] = unw[i];
                        if (mask) ok[u] = mask[i] != 0;
                    }
                }
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    if (!ok[u]) continue;
                    const uint32_t key = pp_key(v[u]);
                    if (pass > 0 && (key >> (shift + 8)) != prefix) continue;
                    atomicAdd(&cnt[j][(key >> shift) & 255u], 1u);
                }
            }
        }
        __syncthreads();
        if ((int)threadIdx.x < kColBlk && c0 + (int)threadIdx.x < ncol && !s_stop[threadIdx.x]) {
            const int jj = (int)threadIdx.x;
            unsigned rank = s_rank[jj];
            bool stop = false;
            if (pass == 0) {
                unsigned tot = 0u;
                for (int k = 0; k < 256; ++k) tot += cnt[jj][k];
                if (tot <= 20u) stop = true;
                else rank = tot / 2u;
            }
            if (stop) {
                s_stop[jj] = 1u;
            } else {
                unsigned below = 0u;
                int k = 0;
                for (; k < 255; ++k) {
                    if (rank < below + cnt[jj][k]) break;
                    below += cnt[jj][k];
                }
                s_prefix[jj] = (s_prefix[jj] << 8) | (unsigned)k;
                s_rank[jj] = rank - below;
            }
        }
        __syncthreads();
    }
    if ((int)threadIdx.x < kColBlk && c0 + (int)threadIdx.x < ncol) {
        const int jj = (int)threadIdx.x;
        med[c0 + jj] = s_stop[jj] ? 0.0 : (double)pp_value(s_prefix[jj]);
        have[c0 + jj] = s_stop[jj] ? 0 : 1;
    }
}

// unw -= shift[line] on the valid pixels of every line with a nonzero shift.
__global__ void k_line_shift(float *unw, const unsigned char *mask, int nrow, int ncol, int by_row, const float *shift) {
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= (long long)nrow * ncol) return;
    const int r = (int)(i / ncol), c = (int)(i % ncol);
    const float s = shift[by_row ? r : c];
    if (s != 0.0f && (!mask || mask[i])) unw[i] -= s;
}

// Seam zone update: by_col, rows [0, nrow) x cols [seam, seam + extent),
// profile indexed by row; else rows [seam, seam + extent) x cols [0, ncol),
// profile indexed by column.  unw += (float)(diff * (1 - d/feather)), d the
// distance past the seam, on valid pixels whose profile entry exists.
__global__ void k_feather(float *unw, const unsigned char *mask, int nrow, int ncol, int seam, int by_col, int feather,
                          int extent, const double *diff, const unsigned char *have) {
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    const long long n = by_col ? (long long)nrow * extent : (long long)extent * ncol;
    if (i >= n) return;
    int r, c, d, k;
    if (by_col) { r = (int)(i / extent); d = (int)(i % extent); c = seam + d; k = r; }
    else { d = (int)(i / ncol); r = seam + d; c = (int)(i % ncol); k = c; }
    if (!have[k]) return;
    const long long g = (long long)r * ncol + c;
    if (mask && !mask[g]) return;
    const float w = 1.0f - (float)d / (float)feather;
    unw[g] += (float)(diff[k] * (double)w);
}

// The two pixel lines either side of a seam, contiguous: by_col, columns
// seam - 1 and seam (n = nrow entries); else rows seam - 1 and seam (n =
// ncol).  lines = [a (n floats), b (n floats)], valid = [ma (n), mb (n)]
// (1 where the pixel is valid; all 1 without a mask).  A strided column
// copied to the host row by row costs a transfer per row.
__global__ void k_seam_lines(const float *unw, const unsigned char *mask, int nrow, int ncol, int seam, int by_col,
                             float *lines, unsigned char *valid) {
    const int n = by_col ? nrow : ncol;
    const int i = (int)(blockIdx.x * 256 + threadIdx.x);
    if (i >= n) return;
    const long long ga = by_col ? (long long)i * ncol + (seam - 1) : (long long)(seam - 1) * ncol + i;
    const long long gb = by_col ? (long long)i * ncol + seam : (long long)seam * ncol + i;
    lines[i] = unw[ga];
    lines[n + i] = unw[gb];
    valid[i] = mask ? (unsigned char)(mask[ga] != 0) : (unsigned char)1;
    valid[n + i] = mask ? (unsigned char)(mask[gb] != 0) : (unsigned char)1;
}

// ── host side of the passes ─────────────────────────────────────────────────

void spikes_1d(float *d_unw, const unsigned char *d_mask, int nrow, int ncol, bool by_row, hipStream_t st) {
    const int nl = by_row ? nrow : ncol, len = by_row ? ncol : nrow;
    if (nl < 3) return;
    DevBuf<double> d_med((size_t)nl);
    DevBuf<unsigned char> d_have((size_t)nl);
    DevBuf<float> d_shift((size_t)nl);
    if (by_row)
        hipLaunchKernelGGL(k_line_median, dim3((unsigned)nl), dim3(256), 0, st, d_unw, d_mask, nl, len, (long long)ncol,
                           1LL, d_med.get(), d_have.get());
    else
        hipLaunchKernelGGL(k_col_median, dim3((unsigned)ceil_div((size_t)ncol, (size_t)kColBlk)), dim3(256), 0, st,
                           d_unw, d_mask, nrow, ncol, d_med.get(), d_have.get());
    CUPHU_HIP_CHECK(hipGetLastError());
    std::vector<double> med((size_t)nl);
    std::vector<unsigned char> have((size_t)nl);
    CUPHU_HIP_CHECK(hipMemcpyAsync(med.data(), d_med.get(), (size_t)nl * 8, hipMemcpyDeviceToHost, st));
    CUPHU_HIP_CHECK(hipMemcpyAsync(have.data(), d_have.get(), (size_t)nl, hipMemcpyDeviceToHost, st));
    CUPHU_HIP_CHECK(hipStreamSynchronize(st));
    // the reference's decision, line by line (:1411-1433)
    const double two_pi = 2.0 * M_PI, tol = 0.5;
    std::vector<float> shift((size_t)nl, 0.0f);
    bool any = false;
    for (int i = 1; i < nl - 1; ++i) {
        if (!have[(size_t)i] || !have[(size_t)i - 1] || !have[(size_t)i + 1]) continue;
        const double up = med[(size_t)i] - med[(size_t)i - 1], down = med[(size_t)i] - med[(size_t)i + 1];
        const double nu = std::round(up / two_pi), nd = std::round(down / two_pi);
        if (nu == 0.0 || nu != nd) continue;
        if (std::fabs(up - nu * two_pi) > tol || std::fabs(down - nd * two_pi) > tol) continue;
        const float s = (float)(nu * two_pi);
        shift[(size_t)i] = s;
        med[(size_t)i] -= s;
        any = true;
    }
    if (!any) return;
    CUPHU_HIP_CHECK(hipMemcpyAsync(d_shift.get(), shift.data(), (size_t)nl * 4, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_line_shift, dim3((unsigned)ceil_div((size_t)nrow * ncol, 256)), dim3(256), 0, st, d_unw, d_mask,
                       nrow, ncol, by_row ? 1 : 0, d_shift.get());
    CUPHU_HIP_CHECK(hipGetLastError());
    CUPHU_HIP_CHECK(hipStreamSynchronize(st));   // the host shift table is freed on return
}

// The reference's profile smoothing: moving median over +-w/2 of the
// measured entries (rank m/2), then linear gap fill between measurements at
// most 2*max_fill apart and a linear fade to 0 over max_fill entries
// elsewhere (:1492-1583).
// The median of each window is the entry of rank m/2 among its m measured
// entries, the value nth_element selects.  The windows only slide forward,
// so a sorted copy of the current window is kept and updated by one removal
// and one insertion per position; a profile with a NaN entry (no order)
// takes the per-window selection instead.
void profile_smooth(std::vector<double> &x, std::vector<unsigned char> &have, int w, int max_fill) {
    const int n = (int)x.size(), half = w / 2;
    std::vector<double> out((size_t)n, 0.0);
    bool has_nan = false;
    for (int i = 0; i < n; ++i)
        if (have[(size_t)i] && std::isnan(x[(size_t)i])) has_nan = true;
    if (has_nan) {
        std::vector<double> seg;
        for (int i = 0; i < n; ++i) {
            if (!have[(size_t)i]) continue;
            seg.clear();
            for (int j = std::max(0, i - half); j < std::min(n, i + half + 1); ++j)
                if (have[(size_t)j]) seg.push_back(x[(size_t)j]);
            std::nth_element(seg.begin(), seg.begin() + (long)(seg.size() / 2), seg.end());
            out[(size_t)i] = seg[seg.size() / 2];
        }
    } else {
        std::vector<double> win;
        win.reserve((size_t)w + 2);
        int lo = 0, hi = -1;   // win holds the measured entries of [lo, hi], sorted
        for (int i = 0; i < n; ++i) {
            const int a = std::max(0, i - half), b = std::min(n - 1, i + half);
            while (hi < b) {
                ++hi;
                if (have[(size_t)hi])
                    win.insert(std::upper_bound(win.begin(), win.end(), x[(size_t)hi]), x[(size_t)hi]);
            }
            while (lo < a) {
                if (have[(size_t)lo]) win.erase(std::lower_bound(win.begin(), win.end(), x[(size_t)lo]));
                ++lo;
            }
            if (have[(size_t)i]) out[(size_t)i] = win[win.size() / 2];
        }
    }
    x.swap(out);
    std::vector<int> idx;
    for (int i = 0; i < n; ++i)
        if (have[(size_t)i]) idx.push_back(i);
    if (idx.empty()) return;
    std::vector<double> f((size_t)n, 0.0);
    std::vector<unsigned char> fh((size_t)n, 0);
    auto fade = [&](int src, int i) {
        const double wt = 1.0 - (double)std::abs(i - src) / (double)max_fill;
        f[(size_t)i] = x[(size_t)src] * wt;
        fh[(size_t)i] = 1;
    };
    for (int i = std::max(0, idx.front() - max_fill); i < idx.front(); ++i) fade(idx.front(), i);
    for (int i = idx.back() + 1; i < std::min(n, idx.back() + 1 + max_fill); ++i) fade(idx.back(), i);
    for (size_t k = 0; k + 1 < idx.size(); ++k) {
        const int a = idx[k], b = idx[k + 1], gap = b - a;
        const double va = x[(size_t)a], vb = x[(size_t)b];
        if (gap <= 2 * max_fill) {
            for (int i = a + 1; i < b; ++i) {
                const double t = (double)(i - a) / (double)gap;
                f[(size_t)i] = va + t * (vb - va);
                fh[(size_t)i] = 1;
            }
        } else {
            for (int i = a + 1; i <= a + max_fill; ++i) fade(a, i);
            for (int i = b - max_fill; i < b; ++i) fade(b, i);
        }
    }
    for (int i : idx) { f[(size_t)i] = x[(size_t)i]; fh[(size_t)i] = 1; }
    x.swap(f);
    have.swap(fh);
}

// Device buffers of one feedback call, sized for the longer seam line and
// shared by its seams (each seam's work ends with a stream sync).
struct SeamBufs {
    DevBuf<float> lines;
    DevBuf<unsigned char> valid, have;
    DevBuf<double> diff;
    explicit SeamBufs(size_t n) : lines(2 * n), valid(2 * n), have(n), diff(n) {}
};

void feedback_seam(float *d_unw, const unsigned char *d_mask, int nrow, int ncol, int seam, bool by_col, int feather,
                   SeamBufs &bufs, hipStream_t st) {
    const int n = by_col ? nrow : ncol, lim = by_col ? ncol : nrow;
    if (seam - 1 < 0 || seam + 1 >= lim) return;   // :1587, :1624
    // the pixel lines just before and at the seam, gathered on the device
    float *d_lines = bufs.lines.get();
    unsigned char *d_valid = bufs.valid.get();
    hipLaunchKernelGGL(k_seam_lines, dim3((unsigned)ceil_div((size_t)n, 256)), dim3(256), 0, st, d_unw, d_mask, nrow,
                       ncol, seam, by_col ? 1 : 0, d_lines, d_valid);
    CUPHU_HIP_CHECK(hipGetLastError());
    std::vector<float> lines(2 * (size_t)n);
    std::vector<unsigned char> valid(2 * (size_t)n);
    CUPHU_HIP_CHECK(hipMemcpyAsync(lines.data(), d_lines, lines.size() * 4, hipMemcpyDeviceToHost, st));
    CUPHU_HIP_CHECK(hipMemcpyAsync(valid.data(), d_valid, valid.size(), hipMemcpyDeviceToHost, st));
    CUPHU_HIP_CHECK(hipStreamSynchronize(st));
    const float *a = lines.data(), *b = lines.data() + n;
    const unsigned char *ma = valid.data(), *mb = valid.data() + n;
    std::vector<double> diff((size_t)n, 0.0);
    std::vector<unsigned char> have((size_t)n, 0);
    for (int i = 0; i < n; ++i) {
        if (!ma[(size_t)i] || !mb[(size_t)i]) continue;
        diff[(size_t)i] = (double)a[(size_t)i] - (double)b[(size_t)i];
        have[(size_t)i] = 1;
    }
    const int smooth = 21;   // :1490-1491, :1539
    profile_smooth(diff, have, smooth, 2 * smooth);
    const int extent = std::min(lim, seam + feather) - seam;
    if (extent <= 0) return;
    double *d_diff = bufs.diff.get();
    unsigned char *d_have = bufs.have.get();
    CUPHU_HIP_CHECK(hipMemcpyAsync(d_diff, diff.data(), (size_t)n * 8, hipMemcpyHostToDevice, st));
    CUPHU_HIP_CHECK(hipMemcpyAsync(d_have, have.data(), (size_t)n, hipMemcpyHostToDevice, st));
    const size_t total = by_col ? (size_t)nrow * extent : (size_t)extent * ncol;
    hipLaunchKernelGGL(k_feather, dim3((unsigned)ceil_div(total, 256)), dim3(256), 0, st, d_unw, d_mask, nrow, ncol, seam,
                       by_col ? 1 : 0, feather, extent, d_diff, d_have);
    CUPHU_HIP_CHECK(hipGetLastError());
    CUPHU_HIP_CHECK(hipStreamSynchronize(st));   // host profile buffers are freed on return
}

}  // namespace

void fix_cycle_spikes_dev(float *d_unw, const unsigned char *d_mask, int nrow, int ncol, hipStream_t st) {
    TraceRange tr("cuphu:fix_cycle_spikes");
    spikes_1d(d_unw, d_mask, nrow, ncol, true, st);
    spikes_1d(d_unw, d_mask, nrow, ncol, false, st);
}

void neighbor_feedback_dev(float *d_unw, const unsigned char *d_mask, int nrow, int ncol, const std::vector<int> &col_seams,
                           const std::vector<int> &row_seams, int feather, hipStream_t st) {
    if (feather <= 0) return;
    TraceRange tr("cuphu:neighbor_feedback");
    if (col_seams.empty() && row_seams.empty()) return;
    SeamBufs bufs((size_t)std::max(nrow, ncol));
    for (int s : col_seams) feedback_seam(d_unw, d_mask, nrow, ncol, s, true, feather, bufs, st);
    for (int s : row_seams) feedback_seam(d_unw, d_mask, nrow, ncol, s, false, feather, bufs, st);
}

}  // namespace cuphu
