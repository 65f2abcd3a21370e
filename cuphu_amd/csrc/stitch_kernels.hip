// stitch_kernels.hip -- tile-pair registration on the GPU
// (reference: src/cxx/cuphu/cuphu_solver.cpp:1867-1959, the coherence-
// weighted median of the overlap differences of two adjacent tiles).
//
// Samples of a pair: every overlap pixel where both tiles have corr >= 0.05
// (and, with a confidence mask, both are confident); value v = unw_A - unw_B
// (float), weight w = min(corr_A, corr_B).  The registration is the smallest
// v whose cumulative weight W(<= v) reaches half of the total weight
// (:1867-1881's sort-and-scan), 0 when the pair has no samples.
//
// GPU scheme: a weighted radix select on the order-preserving 32-bit key of
// v, 8 bits per pass.  Each pass histograms the weights of the samples that
// still match the selected key prefix (FP64, per-block LDS histogram, then
// global atomics) and one thread per pair scans its 256 buckets to pick the
// next 8 bits.  The weights are floats in [0.05, 1] summed in FP64, so every
// partial sum is exact: the histograms, and therefore the result, do not
// depend on the order of the atomic additions, and equal the host stitch's.
#include "common.hpp"
#include "util.hpp"

namespace cuphu {

__device__ __forceinline__ uint32_t order_key(float v) {
    uint32_t u = __float_as_uint(v + 0.0f);   // -0 -> +0: equal values, one key
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float key_value(uint32_t k) {
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k);
}

constexpr int kPairSpt = 16;   // samples per thread per histogram block

// One pass: histogram of the weights of pair p's samples whose key matches
// sel[p].prefix in the bits above `shift + 8`, by the 8 bits at `shift`.
__global__ __launch_bounds__(256) void k_pair_hist(const PairDesc *pairs, const PairSel *sel, const float *unw,
                                                   const float *corr, const unsigned char *conf, int shift,
                                                   double *hist) {
    __shared__ double h[256];
    const int p = blockIdx.y;
    const PairDesc pd = pairs[p];
    const PairSel s = sel[p];
    const long long n = (long long)pd.nr * pd.nc;
    const long long s0 = (long long)blockIdx.x * 256 * kPairSpt;
    if (s0 >= n || s.done) return;   // uniform per block
    h[threadIdx.x] = 0.0;
    __syncthreads();
    for (int k = 0; k < kPairSpt; ++k) {
        const long long si = s0 + (long long)k * 256 + threadIdx.x;
        if (si >= n) break;
        const int r = (int)(si / pd.nc), c = (int)(si % pd.nc);
        const long long ia = pd.a_off + (long long)r * pd.a_pitch + c;
        const long long ib = pd.b_off + (long long)r * pd.b_pitch + c;
        const float ca = corr[ia], cb = corr[ib];
        if (ca < 0.05f || cb < 0.05f) continue;
        if (conf && (!conf[pd.ca_off + (long long)r * pd.ca_stride + c] ||
                     !conf[pd.cb_off + (long long)r * pd.cb_stride + c]))
            continue;
        const uint32_t key = order_key(unw[ia] - unw[ib]);
        if (shift < 24 && (key >> (shift + 8)) != s.prefix) continue;
        const double w = (double)((cb < ca) ? cb : ca);
        atomicAdd(&h[(key >> shift) & 255u], w);
    }
    __syncthreads();
    const double v = h[threadIdx.x];
    if (v != 0.0) atomicAdd(&hist[(size_t)p * 256 + threadIdx.x], v);
}

// Pick the next 8 bits: the first bucket where the cumulative weight
// (weights below the prefix + buckets so far) reaches half of the total.
// The first pass also sets the total.  One thread per pair.
__global__ void k_pair_select(int npairs, const double *hist, int shift, PairSel *sel) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npairs) return;
    PairSel s = sel[p];
    if (s.done) return;
    const double *h = hist + (size_t)p * 256;
    if (shift == 24) {
        double tot = 0.0;
        for (int b = 0; b < 256; ++b) tot += h[b];
        s.half = tot / 2.0;
        s.below = 0.0;
        s.prefix = 0u;
        if (!(tot > 0.0)) {   // no samples: registration 0
            s.done = 1;
            s.value = 0.0f;
            sel[p] = s;
            return;
        }
    }
    double cum = s.below;
    int pick = 255;
    for (int b = 0; b < 256; ++b) {
        if (cum + h[b] >= s.half) { pick = b; break; }
        cum += h[b];
    }
    s.below = cum;
    s.prefix = (s.prefix << 8) | (uint32_t)pick;
    if (shift == 0) {
        s.done = 1;
        s.value = key_value(s.prefix);
    }
    sel[p] = s;
}

void pair_medians(const PairDesc *d_pairs, int npairs, int max_chunks, const float *unw, const float *corr,
                  const unsigned char *conf, PairSel *d_sel, double *d_hist, hipStream_t st) {
    if (npairs <= 0) return;
    CUPHU_HIP_CHECK(hipMemsetAsync(d_sel, 0, (size_t)npairs * sizeof(PairSel), st));
    for (int shift = 24; shift >= 0; shift -= 8) {
        CUPHU_HIP_CHECK(hipMemsetAsync(d_hist, 0, (size_t)npairs * 256 * sizeof(double), st));
        hipLaunchKernelGGL(k_pair_hist, dim3((unsigned)std::max(1, max_chunks), (unsigned)npairs), dim3(256), 0, st,
                           d_pairs, d_sel, unw, corr, conf, shift, d_hist);
        CUPHU_HIP_CHECK(hipGetLastError());
        hipLaunchKernelGGL(k_pair_select, dim3((unsigned)ceil_div((size_t)npairs, 64)), dim3(64), 0, st, npairs,
                           d_hist, shift, d_sel);
        CUPHU_HIP_CHECK(hipGetLastError());
    }
}

int pair_chunks(long long nsamples) {
    return (int)ceil_div((size_t)nsamples, (size_t)256 * kPairSpt);
}

}  // namespace cuphu
