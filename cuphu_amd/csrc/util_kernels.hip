// util_kernels.hip -- pre-processing, statistical costs, residues and flow
// integration (reference: src/cxx/cuphu/cuphu_util.cu, cuphu_cost.cu,
// cuphu_phase.cu).  Reference-layout (unpitched, flat arc) arrays.
#include "common.hpp"
#include "util.hpp"

namespace cuphu {

CostConsts make_cost_consts(const CuPhuParams &p, double ncorrlooks) {
    // upload_smooth_constants cuphu_cost.cu:41-59 (kept per instance).
    CostConsts k{};
    k.rho0 = p.rhosconst1 / ncorrlooks + p.rhosconst2;
    k.thresh = p.defothreshfactor * k.rho0;
    k.rhopow = 2.0 * p.cstd1 + p.cstd2 * std::log(ncorrlooks) + p.cstd3 * ncorrlooks;
    k.sigsqrhoconst = 2.0 / 12.0;
    k.sigsqcorr = p.sigsqcorr;
    k.nsc = p.nshortcycle;
    k.nsc2 = k.nsc * k.nsc;
    k.costscale = p.costscale;
    k.defomax = p.defomax;
    k.sqmin = (int)p.sigsqshortmin;
    return k;
}

// ── WrapPhaseKernel (cuphu_util.cu:43-46) ─────────────────────────────────
__global__ void k_wrap_phase(float *ph, size_t n) {
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) ph[i] = wrap_to_2pi(ph[i]);
}

// ── Range/AzimuthDiffsKernel (cuphu_util.cu:54-92), one launch over both ──
__global__ void k_phase_diffs(const float *ph, float *rng, float *az, int nrow, int ncol) {
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    size_t nr = (size_t)nrow * (ncol - 1), na = (size_t)(nrow - 1) * ncol;
    if (i < nr) {
        size_t r = i / (ncol - 1), c = i % (ncol - 1);
        float d = (ph[r * ncol + c + 1] - ph[r * ncol + c]) / kTwoPiF;
        rng[i] = wrap_cycles(d);
    } else if (i < nr + na) {
        size_t j = i - nr, r = j / ncol, c = j % ncol;
        float d = (ph[r * ncol + c] - ph[(r + 1) * ncol + c]) / kTwoPiF;
        az[j] = wrap_cycles(d);
    }
}

// ── boxcar (cuphu_util.cu:104-181, 237-287) without a padded copy: the
// mirror index is evaluated on the fly; taps are summed in the reference's
// order (k = 0..2k) and divided by the window length. ─────────────────────
__device__ __forceinline__ int mirror_idx(int i, int n) {
    if (n == 1) return 0;
    for (;;) {
        if (i < 0) i = -i;
        else if (i > n - 1) i = 2 * (n - 1) - i;
        else return i;
    }
}
__global__ void k_box_rows(const float *src, float *tmp, int nrow, int ncol, int kr, int kc) {
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    int prow = nrow + 2 * kr;
    if (i >= (size_t)prow * ncol) return;
    int pr = (int)(i / ncol), c = (int)(i % ncol);
    int sr = mirror_idx(pr - kr, nrow), wc = 2 * kc + 1;
    float sum = 0.0f;
    for (int k = 0; k < wc; ++k) sum += src[(size_t)sr * ncol + mirror_idx(c + k - kc, ncol)];
    tmp[i] = sum / (float)wc;
}
__global__ void k_box_cols(const float *tmp, float *dst, int nrow, int ncol, int kr) {
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (size_t)nrow * ncol) return;
    int r = (int)(i / ncol), c = (int)(i % ncol), wr = 2 * kr + 1;
    float sum = 0.0f;
    for (int k = 0; k < wr; ++k) sum += tmp[(size_t)(r + k) * ncol + c];
    dst[i] = sum / (float)wr;
}

// ── SNAPHU smooth/defo costs (cuphu_cost.cu:77-323) ─────────────────────────
// One thread per arc over the flat row-then-column arc layout.  FP64,
// uncontracted.  Smooth offsets truncate (:126), defo offsets round to
// nearest even (:257).  Masked arcs get the MaskCost sentinels (:98-103,
// :225-232).
__global__ void k_costs(const float *rng, const float *az, const float *arng, const float *aaz, const float *corr,
                        const short *weights, int nrow, int ncol, CostConsts k, int defo, short *out) {
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    size_t nra = (size_t)(nrow - 1) * ncol, nca = (size_t)nrow * (ncol - 1);
    if (i >= nra + nca) return;
    float ca, cb, dpsi, avg;
    if (i < nra) {
        size_t r = i / ncol, c = i % ncol;
        ca = corr[r * ncol + c]; cb = corr[(r + 1) * ncol + c];
        dpsi = az[i]; avg = aaz[i];
    } else {
        size_t j = i - nra, r = j / (ncol - 1), c = j % (ncol - 1);
        ca = corr[r * ncol + c]; cb = corr[r * ncol + c + 1];
        dpsi = rng[j]; avg = arng[j];
    }
    short w = weights ? weights[i] : (short)1;
    int ne = defo ? 4 : 2;
    short *o = out + (size_t)ne * i;
    if (w == 0) {
        o[0] = (short)(kLargeShort / 2); o[1] = (short)kLargeShort;
        if (defo) { o[2] = (short)kLargeShort; o[3] = 0; }
        return;
    }
    double rho = 0.5 * ((double)ca + (double)cb);
    if (!(rho >= k.thresh)) rho = 0.0;
    double t = k.sigsqrhoconst * pow(1.0 - rho, k.rhopow);
    double sigsqrho = (t + k.sigsqcorr) * k.nsc2;
    double offset = k.nsc * ((double)dpsi - (rho > 0 ? (double)avg : 0.5 * (double)avg));
    double sf = sigsqrho / (k.costscale * (double)w);
    if (sf > 32767.0) sf = 32767.0;
    int si = (int)sf;
    if (si < k.sqmin) si = k.sqmin;
    if (offset > 32767.0) offset = 32767.0;
    if (offset < -32768.0) offset = -32768.0;
    int oi = defo ? (int)rint(offset) : (int)offset;
    if (si > 32767) si = 32767;
    o[0] = (short)oi; o[1] = (short)si;
    if (defo) {
        double dz = rint(k.nsc * k.defomax);
        int dzmax = dz > 32767.0 ? 32767 : (dz < 1.0 ? 1 : (int)dz);
        int lay = (int)sqrt((double)(dzmax * si));
        if (lay > 32767) lay = 32767;
        o[2] = (short)dzmax; o[3] = (short)lay;
    }
}

// ── residues: plaquette charge of the wrapped-gradient loop ────────────────
__global__ void k_residues(const float *ph, int nrow, int ncol, signed char *res) {
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    size_t n = (size_t)(nrow - 1) * (ncol - 1);
    if (i >= n) return;
    size_t r = i / (ncol - 1), c = i % (ncol - 1);
    const float *p0 = ph + r * ncol + c, *p1 = p0 + ncol;
    float rng0 = wrap_cycles((p0[1] - p0[0]) / kTwoPiF);
    float rng1 = wrap_cycles((p1[1] - p1[0]) / kTwoPiF);
    float az0 = wrap_cycles((p0[0] - p1[0]) / kTwoPiF);
    float az1 = wrap_cycles((p0[1] - p1[1]) / kTwoPiF);
    float s = rng0 - az1;
    s = s - rng1;
    s = s + az0;
    res[i] = (signed char)__builtin_rintf(s);
}

// ── flow integration (cuphu_phase.cu:27-106), parallel ─────────────────────
// Terms: column 0 downward fma(-2pi, hflow, wrap(dphi)); row r rightward
// fma(2pi, vflow, wrap(dphi)); wrap_angle = x - 2pi*rint(x/2pi) as one fma.
// Prefix sums are FP64 workgroup scans; unw = (float)(phi00 + rowoff + coloff).
__device__ __forceinline__ float wrap_angle(float x) { return __builtin_fmaf(-kTwoPiF, __builtin_rintf(x / kTwoPiF), x); }

// Inclusive FP64 scan of one workgroup (256 threads), fixed order.
__device__ double block_scan_incl(double v, double *sh) {
    int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int o = 1; o < 256; o <<= 1) {
        double a = (t >= o) ? sh[t - o] : 0.0;
        __syncthreads();
        sh[t] += a;
        __syncthreads();
    }
    double r = sh[t];
    __syncthreads();
    return r;
}

__global__ __launch_bounds__(256) void k_int_rowoff(const float *ph, const short *hflows, int nrow, int ncol,
                                                    double *rowoff) {
    __shared__ double sh[256];
    double carry = 0.0;
    for (int base = 0; base < nrow; base += 256) {
        int r = base + threadIdx.x;
        double term = 0.0;
        if (r >= 1 && r < nrow) {
            float t = __builtin_fmaf(-kTwoPiF, (float)hflows[(size_t)(r - 1) * ncol],
                                     wrap_angle(ph[(size_t)r * ncol] - ph[(size_t)(r - 1) * ncol]));
            term = (double)t;
        }
        double inc = block_scan_incl(term, sh);
        if (r < nrow) rowoff[r] = carry + inc;
        __shared__ double last;
        if (threadIdx.x == 255) last = inc;
        __syncthreads();
        carry += last;
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void k_int_rows(const float *ph, const short *vflows, const double *rowoff,
                                                  int nrow, int ncol, float *unw) {
    __shared__ double sh[256];
    __shared__ double last;
    int r = blockIdx.x;
    const float *pr = ph + (size_t)r * ncol;
    double ref = (double)ph[0] + rowoff[r];
    double carry = 0.0;
    for (int base = 0; base < ncol; base += 256) {
        int c = base + threadIdx.x;
        double term = 0.0;
        if (c >= 1 && c < ncol) {
            float t = __builtin_fmaf(kTwoPiF, (float)vflows[(size_t)r * (ncol - 1) + c - 1], wrap_angle(pr[c] - pr[c - 1]));
            term = (double)t;
        }
        double inc = block_scan_incl(term, sh);
        if (c < ncol) unw[(size_t)r * ncol + c] = (float)(ref + (carry + inc));
        if (threadIdx.x == 255) last = inc;
        __syncthreads();
        carry += last;
        __syncthreads();
    }
}

__global__ void k_phase_from_igram(const float2 *ig, float *ph, size_t n) {
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) { float2 z = ig[i]; ph[i] = phase_of(z.x, z.y); }
}

// ── launchers ───────────────────────────────────────────────────────────────
static unsigned g1(size_t n) { return (unsigned)ceil_div(n, 256); }

void util_wrap_phase(float *d, size_t n, hipStream_t st) {
    if (!n) return;
    hipLaunchKernelGGL(k_wrap_phase, dim3(g1(n)), dim3(256), 0, st, d, n);
    CUPHU_HIP_CHECK(hipGetLastError());
}
void phase_from_igram(const float2 *ig, float *ph, size_t n, hipStream_t st) {
    if (!n) return;
    hipLaunchKernelGGL(k_phase_from_igram, dim3(g1(n)), dim3(256), 0, st, ig, ph, n);
    CUPHU_HIP_CHECK(hipGetLastError());
}
void util_phase_diffs(const float *ph, float *rng, float *az, int nrow, int ncol, hipStream_t st) {
    size_t n = (size_t)nrow * (ncol - 1) + (size_t)(nrow - 1) * ncol;
    if (!n) return;
    hipLaunchKernelGGL(k_phase_diffs, dim3(g1(n)), dim3(256), 0, st, ph, rng, az, nrow, ncol);
    CUPHU_HIP_CHECK(hipGetLastError());
}
void util_boxcar(const float *src, float *dst, float *tmp, int nrow, int ncol, int kr, int kc, hipStream_t st) {
    if (nrow <= 0 || ncol <= 0) return;
    size_t nt = (size_t)(nrow + 2 * kr) * ncol;
    hipLaunchKernelGGL(k_box_rows, dim3(g1(nt)), dim3(256), 0, st, src, tmp, nrow, ncol, kr, kc);
    CUPHU_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_box_cols, dim3(g1((size_t)nrow * ncol)), dim3(256), 0, st, tmp, dst, nrow, ncol, kr);
    CUPHU_HIP_CHECK(hipGetLastError());
}
void util_costs(const float *ph, const float *corr, const short *weights, int nrow, int ncol, const CostConsts &k,
                int kperp, int kpar, int defo, short *out, hipStream_t st) {
    size_t nra = (size_t)(nrow - 1) * ncol, nca = (size_t)nrow * (ncol - 1);
    DevBuf<float> rng(nca + 1), az(nra + 1), arng(nca + 1), aaz(nra + 1);
    DevBuf<float> tmp((size_t)(std::max(nrow, 1) + 2 * std::max(kperp / 2, kpar / 2)) * std::max(ncol, 1) + 1);
    util_phase_diffs(ph, rng.get(), az.get(), nrow, ncol, st);
    if (ncol > 1) util_boxcar(rng.get(), arng.get(), tmp.get(), nrow, ncol - 1, kperp / 2, kpar / 2, st);
    if (nrow > 1) util_boxcar(az.get(), aaz.get(), tmp.get(), nrow - 1, ncol, kpar / 2, kperp / 2, st);
    if (nra + nca) {
        hipLaunchKernelGGL(k_costs, dim3(g1(nra + nca)), dim3(256), 0, st, rng.get(), az.get(), arng.get(), aaz.get(),
                           corr, weights, nrow, ncol, k, defo, out);
        CUPHU_HIP_CHECK(hipGetLastError());
    }
    CUPHU_HIP_CHECK(hipStreamSynchronize(st));
}
void util_residues(const float *ph, int nrow, int ncol, signed char *res, hipStream_t st) {
    size_t n = (size_t)(nrow - 1) * (ncol - 1);
    if (nrow < 2 || ncol < 2) return;
    hipLaunchKernelGGL(k_residues, dim3(g1(n)), dim3(256), 0, st, ph, nrow, ncol, res);
    CUPHU_HIP_CHECK(hipGetLastError());
}
void util_integrate(const float *ph, const short *hflows, const short *vflows, int nrow, int ncol, float *unw,
                    hipStream_t st) {
    DevBuf<double> rowoff((size_t)nrow);
    hipLaunchKernelGGL(k_int_rowoff, dim3(1), dim3(256), 0, st, ph, hflows, nrow, ncol, rowoff.get());
    CUPHU_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_int_rows, dim3(nrow), dim3(256), 0, st, ph, vflows, rowoff.get(), nrow, ncol, unw);
    CUPHU_HIP_CHECK(hipGetLastError());
    CUPHU_HIP_CHECK(hipStreamSynchronize(st));
}

// ── incremental costs (cuphu_incrcost.cu:55-191, SNAPHU CalcCostSmooth) ────
__device__ __forceinline__ long long ic_ceil_div(long long a, long long b) { return (a > 0) ? (a + b - 1) / b : a / b; }
__device__ __forceinline__ long long ic_floor_div(long long a, long long b) { return (a >= 0) ? a / b : (a - b + 1) / b; }

__global__ void k_fill_i32(int *p, int n, int v) {
    int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n) p[i] = v;
}

// One arc (SNAPHU CalcCostSmooth as cuphu_incrcost.cu:55-191 applies it):
// the squared distances d^2 = (flow * nshortcycle + offset)^2 at the flow
// and at flow +- nflow, c1 = d0^2 / sigsq, pr / nr = d+-^2 / sigsq - c1,
// each rounded away from zero over nflow^2, clamped to +-32000; a masked
// arc (sigsq 32000) costs 0.
//
// Exact FP64 form, taken while every |d| < 2^25 and sigsq >= 1: the
// squares are integers below 2^50, exact in FP64, and so are the
// quotients' remainders.  trunc(a * (1/b)) carries a relative error of at
// most 2^-52 (two roundings) on a quotient below 2^50, so it is within one
// of floor(a / b), and the remainder a - q b, an integer of magnitude below
// 2b formed exactly by one fma, fixes it.  Otherwise (|d| >= 2^25, or
// sigsq <= 0) the 64-bit integer divisions.  Both give the same integers.
__device__ __forceinline__ double ic_floor_div_f64(double a, double b, double rb) {   // a >= 0, b >= 1
    double q = trunc(a * rb);
    const double r = fma(-q, b, a);
    if (r < 0.0) q -= 1.0;
    else if (r >= b) q += 1.0;
    return q;
}
__device__ __forceinline__ double ic_away_div_f64(double a, double b, double rb) {   // a integer-valued
    const double m = fabs(a);
    double q = ic_floor_div_f64(m, b, rb);
    if (fma(-q, b, m) > 0.0) q += 1.0;
    return a < 0.0 ? -q : q;
}

struct IncrArgs {
    long long nsc, nfl;
    double dnsc, dnfl;   // the same as doubles (exact)
    double nsq, rnsq;    // nflow^2 and its reciprocal
    int unit;            // nflow == 1: the division by nflow^2 is the identity
};

// The 64-bit integer form, for the arcs the FP64 form does not cover.
__device__ __noinline__ void incr_arc_int(long long off, long long sig, long long f, const IncrArgs &a,
                                          long long &pc, long long &nc) {
    long long idz1 = f * a.nsc + off, idp = (f + a.nfl) * a.nsc + off, idn = (f - a.nfl) * a.nsc + off;
    idz1 = idz1 < 0 ? -idz1 : idz1;
    idp = idp < 0 ? -idp : idp;
    idn = idn < 0 ? -idn : idn;
    const long long c1 = (idz1 * idz1) / sig;
    const long long pr = (idp * idp) / sig - c1, nr = (idn * idn) / sig - c1;
    const long long nsq = a.nfl * a.nfl;
    pc = pr > 0 ? ic_ceil_div(pr, nsq) : ic_floor_div(pr, nsq);
    nc = nr > 0 ? ic_ceil_div(nr, nsq) : ic_floor_div(nr, nsq);
    pc = pc > 32000 ? 32000 : (pc < -32000 ? -32000 : pc);
    nc = nc > 32000 ? 32000 : (nc < -32000 ? -32000 : nc);
}

// The distances |(flow + k nflow) nshortcycle + offset| are formed as
// fabs(fma(.)) in FP64: exact while the true value is below 2^53, and a
// larger one fails the 2^25 test either way.
__device__ __forceinline__ void incr_arc(int off, int sig, int f, const IncrArgs &a, short &pos, short &neg, int &mp,
                                         int &mn) {
    int pc = 0, nc = 0;
    if (sig != 32000) {
        const double d0 = (double)f, doff = (double)off;
        const double z1 = fabs(fma(d0, a.dnsc, doff));
        const double zp = fabs(fma(d0 + a.dnfl, a.dnsc, doff));
        const double zn = fabs(fma(d0 - a.dnfl, a.dnsc, doff));
        const double lim = 33554432.0;   // 2^25
        if (sig >= 1 && z1 < lim && zp < lim && zn < lim) {
            const double b = (double)sig, rb = 1.0 / b;
            const double c1 = ic_floor_div_f64(z1 * z1, b, rb);
            double pr = ic_floor_div_f64(zp * zp, b, rb) - c1;
            double nr = ic_floor_div_f64(zn * zn, b, rb) - c1;
            if (!a.unit) {
                pr = ic_away_div_f64(pr, a.nsq, a.rnsq);
                nr = ic_away_div_f64(nr, a.nsq, a.rnsq);
            }
            pc = (int)fmin(fmax(pr, -32000.0), 32000.0);
            nc = (int)fmin(fmax(nr, -32000.0), 32000.0);
        } else {
            long long p64, n64;
            incr_arc_int(off, sig, f, a, p64, n64);
            pc = (int)p64;
            nc = (int)n64;
        }
    }
    pos = (short)pc;
    neg = (short)nc;
    mp = min(mp, pc);
    mn = min(mn, nc);
}

// Four consecutive arcs per thread, grid-stride over groups of four; VEC:
// 16-byte cost and 8-byte flow / output accesses (the host checks the
// alignment).  The minima are reduced per workgroup, one atomic each.
template <bool VEC>
__global__ __launch_bounds__(256) void k_incrcost(const short *costs, const short *flows, size_t narcs, IncrArgs a,
                                                  short *pos, short *neg, int *mins) {
    __shared__ int sm[2][4];
    int mp = 32000, mn = 32000;
    const size_t nq = (narcs + 3) / 4;
    for (size_t q = (size_t)blockIdx.x * 256 + threadIdx.x; q < nq; q += (size_t)gridDim.x * 256) {
        const size_t i0 = 4 * q;
        if (VEC && i0 + 4 <= narcs) {
            const short4 c01 = reinterpret_cast<const short4 *>(costs)[2 * q];
            const short4 c23 = reinterpret_cast<const short4 *>(costs)[2 * q + 1];
            const short4 fl = reinterpret_cast<const short4 *>(flows)[q];
            short4 po, ne;
            incr_arc(c01.x, c01.y, fl.x, a, po.x, ne.x, mp, mn);
            incr_arc(c01.z, c01.w, fl.y, a, po.y, ne.y, mp, mn);
            incr_arc(c23.x, c23.y, fl.z, a, po.z, ne.z, mp, mn);
            incr_arc(c23.z, c23.w, fl.w, a, po.w, ne.w, mp, mn);
            reinterpret_cast<short4 *>(pos)[q] = po;
            reinterpret_cast<short4 *>(neg)[q] = ne;
        } else {
            for (size_t i = i0; i < i0 + 4 && i < narcs; ++i)
                incr_arc(costs[2 * i], costs[2 * i + 1], flows[i], a, pos[i], neg[i], mp, mn);
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        mp = min(mp, __shfl_xor(mp, o, 64));
        mn = min(mn, __shfl_xor(mn, o, 64));
    }
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { sm[0][wave] = mp; sm[1][wave] = mn; }
    __syncthreads();
    if (threadIdx.x == 0) {
        atomicMin(mins, min(min(sm[0][0], sm[0][1]), min(sm[0][2], sm[0][3])));
        atomicMin(mins + 1, min(min(sm[1][0], sm[1][1]), min(sm[1][2], sm[1][3])));
    }
}

int util_incrcost(const short *costs, const short *flows, size_t narcs, long nshortcycle, int nflow, short *poscost,
                  short *negcost, int *mins, hipStream_t st) {
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, st, mins, 2, 32000);
    CUPHU_HIP_CHECK(hipGetLastError());
    if (narcs) {
        IncrArgs a;
        a.nsc = (long long)nshortcycle;
        a.nfl = nflow;
        a.dnsc = (double)nshortcycle;
        a.dnfl = (double)nflow;
        a.nsq = (double)nflow * (double)nflow;
        a.rnsq = 1.0 / a.nsq;
        a.unit = nflow == 1 ? 1 : 0;
        // 8 workgroups per CU at most: every thread takes several groups of four
        const unsigned g = (unsigned)std::min<size_t>(ceil_div(ceil_div(narcs, 4), 256), 2048);
        auto al = [](const void *p, size_t b) { return ((uintptr_t)p % b) == 0; };
        if (al(costs, 16) && al(flows, 8) && al(poscost, 8) && al(negcost, 8))
            hipLaunchKernelGGL(k_incrcost<true>, dim3(g), dim3(256), 0, st, costs, flows, narcs, a, poscost, negcost,
                               mins);
        else
            hipLaunchKernelGGL(k_incrcost<false>, dim3(g), dim3(256), 0, st, costs, flows, narcs, a, poscost, negcost,
                               mins);
        CUPHU_HIP_CHECK(hipGetLastError());
    }
    int h[2] = {0, 0};
    CUPHU_HIP_CHECK(hipMemcpyAsync(h, mins, sizeof(h), hipMemcpyDeviceToHost, st));
    CUPHU_HIP_CHECK(hipStreamSynchronize(st));
    return (h[0] >= 0 && h[1] >= 0) ? 1 : 0;
}

}  // namespace cuphu
