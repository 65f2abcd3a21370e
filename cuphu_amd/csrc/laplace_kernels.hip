// laplace_kernels.hip -- gfx950 kernels of the weighted-Laplacian PCG/IRLS
// solve (reference: src/cxx/cuphu/cuphu_laplace.cu).
//
// Layout.  Every instance (a tile or an interferogram) lives in pooled
// planes with a row pitch that is a multiple of 64 elements (256 on the
// fused solver, whose iteration kernel covers only the first wcols, the
// 64-rounded width); the pad columns and every missing arc hold zero, so a
// pad pixel is an isolated node with r = u = 0 that drops out of every dot
// product.  Arc weights are
// stored per pixel as an interleaved {down, right} pair (the pixel's row arc
// (r,c)->(r+1,c) and column arc (r,c)->(r,c+1)), so one 32-byte load gives
// the two outgoing arcs of four pixels.
//
// Tiles.  A 256-thread workgroup (4 waves of 64) owns a 256-column x
// 4*RPW-row tile; each lane owns 4 consecutive columns of RPW rows and
// moves them with 16-byte loads.  The fused solver's kernels also run
// 128-column x 8*RPW-row tiles (their default), where each half-wave owns
// RPW rows (Shape<RPW, LPR> below).  Stencil kernels stage the stencil operand
// of the tile plus a one-pixel halo in LDS, then apply the 5-point weighted
// Laplacian from LDS.  Workgroups are renumbered so each XCD receives a
// contiguous run of tiles (vertically adjacent tiles share halo rows in the
// same L2).
//
// Arithmetic is FP32 (or FP64 in the double instantiation) with -ffp-contract=off;
// every fused multiply-add is written explicitly.  Inner products are
// accumulated in FP64 per lane, reduced by a fixed xor-butterfly per wave,
// in wave order per workgroup, and in block order per instance: the result
// is deterministic run to run.
#include "laplace.hpp"

namespace cuphu {

// ── small vector helpers ────────────────────────────────────────────────────
__device__ __forceinline__ void ld4(const float *p, float (&o)[4]) {
    float4 t = *reinterpret_cast<const float4 *>(p);
    o[0] = t.x; o[1] = t.y; o[2] = t.z; o[3] = t.w;
}
__device__ __forceinline__ void ld4(const double *p, double (&o)[4]) {
    double2 a = *reinterpret_cast<const double2 *>(p);
    double2 b = *reinterpret_cast<const double2 *>(p + 2);
    o[0] = a.x; o[1] = a.y; o[2] = b.x; o[3] = b.y;
}
__device__ __forceinline__ void st4(float *p, const float (&v)[4]) {
    *reinterpret_cast<float4 *>(p) = make_float4(v[0], v[1], v[2], v[3]);
}
__device__ __forceinline__ void st4(double *p, const double (&v)[4]) {
    *reinterpret_cast<double2 *>(p) = make_double2(v[0], v[1]);
    *reinterpret_cast<double2 *>(p + 2) = make_double2(v[2], v[3]);
}
template <class T> __device__ __forceinline__ void zero4(T (&v)[4]) { v[0] = v[1] = v[2] = v[3] = T(0); }
// Streaming (non-temporal) loads and stores: the fused PCG kernel's outputs
// are read again only by the next launch, and x only by its own pixel, so
// they bypass the caches the halo reads use.  Measured on scene8192:
// 0.663 -> 0.598 ms per launch.
typedef float nt_f4 __attribute__((ext_vector_type(4)));
typedef double nt_d2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void st4_nt(float *p, const float (&v)[4]) {
    nt_f4 t = {v[0], v[1], v[2], v[3]};
    __builtin_nontemporal_store(t, reinterpret_cast<nt_f4 *>(p));
}
__device__ __forceinline__ void ld4_nt(const float *p, float (&o)[4]) {
    nt_f4 t = __builtin_nontemporal_load(reinterpret_cast<const nt_f4 *>(p));
    o[0] = t.x; o[1] = t.y; o[2] = t.z; o[3] = t.w;
}
__device__ __forceinline__ void ld4_nt(const double *p, double (&o)[4]) {
    nt_d2 a = __builtin_nontemporal_load(reinterpret_cast<const nt_d2 *>(p));
    nt_d2 b = __builtin_nontemporal_load(reinterpret_cast<const nt_d2 *>(p + 2));
    o[0] = a.x; o[1] = a.y; o[2] = b.x; o[3] = b.y;
}
__device__ __forceinline__ void st4_nt(double *p, const double (&v)[4]) {
    nt_d2 a = {v[0], v[1]}, b = {v[2], v[3]};
    __builtin_nontemporal_store(a, reinterpret_cast<nt_d2 *>(p));
    __builtin_nontemporal_store(b, reinterpret_cast<nt_d2 *>(p + 2));
}

// Arcs of 4 pixels starting at element e: w2 holds {down, right} pairs.
__device__ __forceinline__ void ld_arcs(const float *w2, long long e, float (&wd)[4], float (&wr)[4]) {
    float4 a = *reinterpret_cast<const float4 *>(w2 + 2 * e);
    float4 b = *reinterpret_cast<const float4 *>(w2 + 2 * e + 4);
    wd[0] = a.x; wr[0] = a.y; wd[1] = a.z; wr[1] = a.w;
    wd[2] = b.x; wr[2] = b.y; wd[3] = b.z; wr[3] = b.w;
}
__device__ __forceinline__ void ld_arcs(const double *w2, long long e, double (&wd)[4], double (&wr)[4]) {
    const double2 *q = reinterpret_cast<const double2 *>(w2 + 2 * e);
    double2 a = q[0], b = q[1], c = q[2], d = q[3];
    wd[0] = a.x; wr[0] = a.y; wd[1] = b.x; wr[1] = b.y;
    wd[2] = c.x; wr[2] = c.y; wd[3] = d.x; wr[3] = d.y;
}
__device__ __forceinline__ void st_arcs(float *w2, long long e, const float (&wd)[4], const float (&wr)[4]) {
    float4 *q = reinterpret_cast<float4 *>(w2 + 2 * e);
    q[0] = make_float4(wd[0], wr[0], wd[1], wr[1]);
    q[1] = make_float4(wd[2], wr[2], wd[3], wr[3]);
}
__device__ __forceinline__ void st_arcs(double *w2, long long e, const double (&wd)[4], const double (&wr)[4]) {
    double2 *q = reinterpret_cast<double2 *>(w2 + 2 * e);
    q[0] = make_double2(wd[0], wr[0]); q[1] = make_double2(wd[1], wr[1]);
    q[2] = make_double2(wd[2], wr[2]); q[3] = make_double2(wd[3], wr[3]);
}

// Hardware dispatches workgroup b to XCD b % 8.  Give each XCD a contiguous
// range of logical tiles instead.
__device__ __forceinline__ int xcd_block(int b, int nb) {
    const int X = 8;
    int q = nb / X, r = nb % X, x = b % X, i = b / X;
    return x * q + (x < r ? x : r) + i;
}

struct TileInfo {
    InstDesc d;
    int inst, blk, tx0, ty0;
};
// Tile of this workgroup.  The launch order (after the XCD remap) walks an
// instance in raster order, or, when the instance is wider than `band` tile
// columns, band by band: `band` columns wide, row by row within a band (the
// last band narrower), so that vertically adjacent tiles stay close in the
// order and meet in the same L2.  t.blk is the tile's raster slot, where its
// partial sums go: the reductions add them in raster order whatever the
// launch order, so the order does not change any result.
__device__ __forceinline__ TileInfo get_tile(const InstDesc *descs, const int *blk_inst, int th, int tw = kTileW) {
    TileInfo t;
    const int b = xcd_block((int)blockIdx.x, (int)gridDim.x);
    t.inst = blk_inst[b];
    t.d = descs[t.inst];
    const int local = b - t.d.block_begin;
    int tx, ty;
    if (t.d.band > 0) {
        const int bsz = t.d.band * t.d.tiles_y;          // tiles of a full band
        const int k = local / bsz;
        const int bw = min(t.d.band, t.d.tiles_x - k * t.d.band);
        const int rem = local - k * bsz;
        ty = rem / bw;
        tx = k * t.d.band + rem % bw;
    } else {
        tx = local % t.d.tiles_x;
        ty = local / t.d.tiles_x;
    }
    t.blk = t.d.block_begin + ty * t.d.tiles_x + tx;
    t.tx0 = tx * tw;
    t.ty0 = ty * th;
    return t;
}

// Tile shapes.  LPR lanes cover one row of a tile, 4 pixels each, so a tile
// is 4*LPR columns wide; a wave holds 64/LPR such lane groups, and lane group
// v = wave * (64/LPR) + lane / LPR of the workgroup owns rows v*RPW ..
// v*RPW + RPW - 1.  LPR 64 is the 256 x 4*RPW tile; LPR 32 the 128 x 8*RPW
// tile (same pixels per workgroup, 5 % fewer halo pixels), which the fused
// solver takes on large launches that both shapes tile without idle lanes.
// LPR 16, the 64 x 16*RPW tile, is the iteration kernel's tall tile for an
// instance's last 64 columns (k_iq_mix).
template <int RPW, int LPR>
struct Shape {
    static_assert(LPR == 64 || LPR == 32 || LPR == 16, "lanes per tile row");
    static constexpr int G = 64 / LPR;      // lane groups per wave
    static constexpr int TW = 4 * LPR;      // tile width
    static constexpr int TH = 4 * G * RPW;  // tile height
    static constexpr int S = TW + 8;        // LDS row stride; tile col j -> LDS col kLdsC0 + j
    static constexpr int L0 = (TH + 4) * S; // the iteration kernel's LDS: p_j with a two-pixel halo
    static constexpr int L1 = (TH + 2) * S; // and p_{j+1} with a one-pixel halo
};

// Halo fill: tile rows -1 and TH (lane group 0 of waves 0 and 3, 4 pixels
// per lane) and tile columns -1 and TW for rows -1..TH (waves 1 and 2, one
// pixel per lane).  put4(lr, lc, e, ok) / put1(lr, lc, e, ok) store the
// operand values of pixel element e at LDS (row lr, col lc); ok = false
// stores 0.
template <int TH, int LPR = 64, class P4, class P1>
__device__ __forceinline__ void fill_halo(const TileInfo &t, int wave, int lane, P4 &&put4, P1 &&put1) {
    static_assert(TH + 2 <= 64, "ring columns: one lane per row");
    constexpr int TW = 4 * LPR;
    const InstDesc &d = t.d;
    if (wave == 0 || wave == 3) {
        if (lane < LPR) {
            int lr = (wave == 0) ? 0 : TH + 1;
            int row = t.ty0 + lr - 1;
            int c = t.tx0 + 4 * lane;
            bool ok = row >= 0 && row < d.nrow && c < d.pitch;
            put4(lr, kLdsC0 + 4 * lane, d.off + (long long)row * d.pitch + c, ok);
        }
    } else if (lane < TH + 2) {
        int lr = lane, row = t.ty0 + lr - 1;
        int c = (wave == 1) ? t.tx0 - 1 : t.tx0 + TW;
        int lc = (wave == 1) ? kLdsC0 - 1 : kLdsC0 + TW;
        bool ok = row >= 0 && row < d.nrow && c >= 0 && c < d.pitch;
        put1(lr, lc, d.off + (long long)row * d.pitch + c, ok);
    }
}

// Weighted 5-point Laplacian of 4 pixels from an LDS tile (laplace_matvec
// cuphu_laplace.cu:156-190): y = eps*a, then fma over down, up, right, left.
// lr: tile row; cl: the lane's column group (lane mod LPR).
template <class T, int LPR = 64>
__device__ __forceinline__ void stencil4(const T *lds, int lr, int cl, const T (&wd)[4], const T (&wu)[4],
                                         const T (&wr)[4], const T (&wl)[4], T (&y)[4]) {
    constexpr int S = 4 * LPR + 8;
    const T *rc = lds + (lr + 1) * S + kLdsC0 + 4 * cl;
    T c4[4], u4[4], d4[4];
    ld4(rc, c4);
    ld4(rc - S, u4);
    ld4(rc + S, d4);
    T left = rc[-1], right = rc[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        T a = c4[j];
        T xl = (j == 0) ? left : c4[j - 1];
        T xr = (j == 3) ? right : c4[j + 1];
        T yv = Prec<T>::eps * a;
        yv = Prec<T>::fmaT(wd[j], a - d4[j], yv);
        yv = Prec<T>::fmaT(wu[j], a - u4[j], yv);
        yv = Prec<T>::fmaT(wr[j], a - xr, yv);
        yv = Prec<T>::fmaT(wl[j], a - xl, yv);
        y[j] = yv;
    }
}

// The same stencil accumulated in FP64 (the true residual b - L x of a
// residual replacement: b and L x nearly cancel, so the difference is formed
// from exact products and rounded once).
template <class T, int LPR = 64>
__device__ __forceinline__ void stencil4_wide(const T *lds, int lr, int cl, const T (&wd)[4], const T (&wu)[4],
                                              const T (&wr)[4], const T (&wl)[4], double (&y)[4]) {
    constexpr int S = 4 * LPR + 8;
    const T *rc = lds + (lr + 1) * S + kLdsC0 + 4 * cl;
    T c4[4], u4[4], d4[4];
    ld4(rc, c4);
    ld4(rc - S, u4);
    ld4(rc + S, d4);
    T left = rc[-1], right = rc[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        double a = (double)c4[j];
        double xl = (double)((j == 0) ? left : c4[j - 1]);
        double xr = (double)((j == 3) ? right : c4[j + 1]);
        double yv = (double)Prec<T>::eps * a;
        yv = fma((double)wd[j], a - (double)d4[j], yv);
        yv = fma((double)wu[j], a - (double)u4[j], yv);
        yv = fma((double)wr[j], a - xr, yv);
        yv = fma((double)wl[j], a - xl, yv);
        y[j] = yv;
    }
}

// Arcs of a lane's 4 pixels in one row: own down/right arcs and the left
// arc (the right arc of the previous column: the neighbouring lane's, or a
// load for the first lane of a row).  Executed by every lane (contains a
// shuffle); cl = lane mod LPR.
template <class T, int LPR = 64>
__device__ __forceinline__ void row_arcs(const T *w2, const TileInfo &t, int row, bool cok, int cl,
                                         T (&wd)[4], T (&wr)[4], T (&wl)[4]) {
    long long e = t.d.off + (long long)row * t.d.pitch + t.tx0 + 4 * cl;
    if (cok) ld_arcs(w2, e, wd, wr);
    else { zero4(wd); zero4(wr); }
    T from_left = __shfl_up(wr[3], 1, LPR);
    if (cl == 0) from_left = (t.tx0 > 0) ? w2[2 * (e - 1) + 1] : T(0);
    wl[0] = from_left; wl[1] = wr[0]; wl[2] = wr[1]; wl[3] = wr[2];
}

// Down arcs of the row above (the up arcs of this row); zero on row 0.
template <class T>
__device__ __forceinline__ void up_arcs(const T *w2, const TileInfo &t, int row, bool cok, int cl, T (&wu)[4]) {
    if (row > 0 && cok) {
        T dummy[4];
        ld_arcs(w2, t.d.off + (long long)(row - 1) * t.d.pitch + t.tx0 + 4 * cl, wu, dummy);
    } else {
        zero4(wu);
    }
}

// Deterministic workgroup sum of two doubles; thread 0 writes *out.
__device__ __forceinline__ void block_sum2(double a, double b, double2 *out) {
    __shared__ double sa[4], sb[4];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        a += __shfl_xor(a, o, 64);
        b += __shfl_xor(b, o, 64);
    }
    int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) { sa[wave] = a; sb[wave] = b; }
    __syncthreads();
    if (threadIdx.x == 0) *out = make_double2(((sa[0] + sa[1]) + sa[2]) + sa[3], ((sb[0] + sb[1]) + sb[2]) + sb[3]);
}

// ═══════════════════════════════════════════════════════════════════════════
// Ingest: complex64 (or wrapped phase) + coherence + mask -> pitched phi,
// corr, mask, {down, right} sigsq, initial x.  One block row per image row.
// phi: atan2 evaluated in FP64 and rounded once, +2pi if negative, then
// wrapped to [0, 2pi) (cuphu_solver.cpp:1707-1712, cuphu_util.cu:33-46).
// sigsq: the SNAPHU sigsq of cuphu_cost.cu:105-118 (the only cost field the
// Laplace weights use), mask -> arc weight as cuphu_solver.cpp:410-426.
// ═══════════════════════════════════════════════════════════════════════════
__device__ __forceinline__ short sigsq_arc(const CostConsts &k, float ca, float cb, int w) {
    if (w == 0) return (short)kLargeShort;
    double rho = 0.5 * ((double)ca + (double)cb);
    if (!(rho >= k.thresh)) rho = 0.0;
    double t = k.sigsqrhoconst * pow(1.0 - rho, k.rhopow);
    double sigsqrho = (t + k.sigsqcorr) * k.nsc2;
    double f = sigsqrho / (k.costscale * (double)w);
    if (f > 32767.0) f = 32767.0;
    int si = (int)f;
    if (si < k.sqmin) si = k.sqmin;
    if (si > 32767) si = 32767;
    return (short)si;
}


template <class T>
__global__ __launch_bounds__(256) void k_ingest(const IngestSrc *srcs, const InstDesc *descs, const int *rowinst,
                                                Planes<T> pl, int has_u0, int nxb) {
    // 1-D grid: block b covers row b / nxb, column block b % nxb
    int grow = (int)(blockIdx.x / (unsigned)nxb);
    int inst = rowinst[grow];
    const InstDesc d = descs[inst];
    const IngestSrc &s = srcs[inst];
    int row = grow - d.row_begin;
    int c0 = ((int)(blockIdx.x % (unsigned)nxb) * 256 + threadIdx.x) * 4;
    if (c0 >= d.pitch) return;
    long long e = d.off + (long long)row * d.pitch + c0;
    float ph[4], co[4];
    unsigned char mk[4];
    short sd[4], sr[4];
    T x[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        int c = c0 + j;
        ph[j] = 0.0f; co[j] = 0.0f; mk[j] = 0; sd[j] = 0; sr[j] = 0; x[j] = T(0);
        if (c < d.ncol) {
            long long se = (long long)row * s.src_stride + c;
            if (s.igram) { float2 z = s.igram[se]; ph[j] = phase_of(z.x, z.y); }
            else ph[j] = s.phase[se];
            co[j] = s.corr[se];
            int m = s.mask ? (s.mask[se] != 0) : 1;
            mk[j] = (unsigned char)m;
            if (row < d.nrow - 1) {
                int w = s.mask ? (m && s.mask[se + s.src_stride] != 0) : 1;
                sd[j] = sigsq_arc(s.cc, co[j], s.corr[se + s.src_stride], w);
            }
            if (c < d.ncol - 1) {
                int w = s.mask ? (m && s.mask[se + 1] != 0) : 1;
                sr[j] = sigsq_arc(s.cc, co[j], s.corr[se + 1], w);
            }
            if (has_u0 && s.u0) x[j] = (T)s.u0[se];
        }
    }
    *reinterpret_cast<float4 *>(pl.phi + e) = make_float4(ph[0], ph[1], ph[2], ph[3]);
    *reinterpret_cast<float4 *>(pl.corr + e) = make_float4(co[0], co[1], co[2], co[3]);
    *reinterpret_cast<uchar4 *>(pl.mask + e) = make_uchar4(mk[0], mk[1], mk[2], mk[3]);
    short2 *sq = pl.sigsq + e;
    *reinterpret_cast<int4 *>(sq) = make_int4((int)((unsigned short)sd[0] | ((unsigned)(unsigned short)sr[0] << 16)),
                                              (int)((unsigned short)sd[1] | ((unsigned)(unsigned short)sr[1] << 16)),
                                              (int)((unsigned short)sd[2] | ((unsigned)(unsigned short)sr[2] << 16)),
                                              (int)((unsigned short)sd[3] | ((unsigned)(unsigned short)sr[3] << 16)));
    st4(pl.x + e, x);
}

// ═══════════════════════════════════════════════════════════════════════════
// Arc weights: initial w0 = 1/sigsq (laplace_init_weights :198-205) or the
// IRLS reweight w0 / sqrt(e^2 + delta^2) (laplace_reweight :214-247) for the
// two outgoing arcs of each pixel.
// ═══════════════════════════════════════════════════════════════════════════
template <class T>
__device__ __forceinline__ T reweight1(short s, float ph, float pt, T uh, T ut, T d2) {
    T w0 = Prec<T>::weight(s);
    if (w0 <= T(0)) return T(0);
    T g = Prec<T>::wrap((T)ph - (T)pt);
    T e = (uh - ut) - g;
    return w0 / Prec<T>::sqrtT(Prec<T>::fmaT(e, e, d2));
}

__device__ __forceinline__ void ld_sigsq4(const short2 *p, short (&sd)[4], short (&sr)[4]) {
    int4 v = *reinterpret_cast<const int4 *>(p);
    int a[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int j = 0; j < 4; ++j) { sd[j] = (short)(a[j] & 0xFFFF); sr[j] = (short)((unsigned)a[j] >> 16); }
}
__device__ __forceinline__ void ld4f(const float *p, float (&o)[4]) { ld4(p, o); }

template <class T, int RPW, int LPR>
__global__ __launch_bounds__(256) void k_weights(const InstDesc *descs, const int *blk_inst, Planes<T> pl,
                                                 int reweight, T d2) {
    using Sh = Shape<RPW, LPR>;
    TileInfo t = get_tile(descs, blk_inst, Sh::TH, Sh::TW);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int cl = lane % LPR, v = wave * Sh::G + lane / LPR;
    const InstDesc &d = t.d;
    int c = t.tx0 + 4 * cl;
    bool cok = c < d.pitch;
    for (int k = 0; k < RPW; ++k) {
        int row = t.ty0 + v * RPW + k;
        if (row >= d.nrow) break;
        long long e = d.off + (long long)row * d.pitch + c;
        short sd[4] = {0, 0, 0, 0}, sr[4] = {0, 0, 0, 0};
        T wd[4], wr[4];
        if (!reweight) {
            if (cok) ld_sigsq4(pl.sigsq + e, sd, sr);
#pragma unroll
            for (int j = 0; j < 4; ++j) { wd[j] = Prec<T>::weight(sd[j]); wr[j] = Prec<T>::weight(sr[j]); }
        } else {
            float pa[4] = {0, 0, 0, 0}, pd[4] = {0, 0, 0, 0};
            T xa[4], xd[4];
            zero4(xa); zero4(xd);
            if (cok) {
                ld_sigsq4(pl.sigsq + e, sd, sr);
                ld4f(pl.phi + e, pa);
                ld4(pl.x + e, xa);
                if (row + 1 < d.nrow) { ld4f(pl.phi + e + d.pitch, pd); ld4(pl.x + e + d.pitch, xd); }
            }
            float pr = __shfl_down(pa[0], 1, LPR);
            T xr = __shfl_down(xa[0], 1, LPR);
            if (cl == LPR - 1) {
                bool ok = cok && (c + 4 < d.pitch);
                pr = ok ? pl.phi[e + 4] : 0.0f;
                xr = ok ? pl.x[e + 4] : T(0);
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                wd[j] = reweight1<T>(sd[j], pd[j], pa[j], xd[j], xa[j], d2);
                float phr = (j < 3) ? pa[j + 1] : pr;
                T xrr = (j < 3) ? xa[j + 1] : xr;
                wr[j] = reweight1<T>(sr[j], phr, pa[j], xrr, xa[j], d2);
            }
        }
        if (cok) st_arcs(pl.w2, e, wd, wr);
    }
}

// ═══════════════════════════════════════════════════════════════════════════
// Assemble (laplace_assemble :88-148) fused with the initial residual
// r = b - L x (:431-434) and the Jacobi inverse 1/diag; zeroes p and s.
// ═══════════════════════════════════════════════════════════════════════════
template <class T, int RPW, int LPR>
__global__ __launch_bounds__(256) void k_assemble(const InstDesc *descs, const int *blk_inst, Planes<T> pl) {
    using Sh = Shape<RPW, LPR>;
    constexpr int TH = Sh::TH, S = Sh::S;
    __shared__ float lph[(TH + 2) * S];
    __shared__ T lx[(TH + 2) * S];
    TileInfo t = get_tile(descs, blk_inst, TH, Sh::TW);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int cl = lane % LPR, v = wave * Sh::G + lane / LPR;
    const InstDesc &d = t.d;
    int c = t.tx0 + 4 * cl;
    bool cok = c < d.pitch;
    for (int k = 0; k < RPW; ++k) {
        int lr = v * RPW + k, row = t.ty0 + lr;
        float ph[4] = {0, 0, 0, 0};
        T xv[4];
        zero4(xv);
        if (row < d.nrow && cok) {
            long long e = d.off + (long long)row * d.pitch + c;
            ld4f(pl.phi + e, ph);
            ld4(pl.x + e, xv);
        }
        float *lp = lph + (lr + 1) * S + kLdsC0 + 4 * cl;
        T *lq = lx + (lr + 1) * S + kLdsC0 + 4 * cl;
#pragma unroll
        for (int j = 0; j < 4; ++j) { lp[j] = ph[j]; lq[j] = xv[j]; }
    }
    fill_halo<TH, LPR>(
        t, wave, lane,
        [&](int lr, int lc, long long e, bool ok) {
            float ph[4] = {0, 0, 0, 0};
            T xv[4];
            zero4(xv);
            if (ok) { ld4f(pl.phi + e, ph); ld4(pl.x + e, xv); }
#pragma unroll
            for (int j = 0; j < 4; ++j) { lph[lr * S + lc + j] = ph[j]; lx[lr * S + lc + j] = xv[j]; }
        },
        [&](int lr, int lc, long long e, bool ok) {
            lph[lr * S + lc] = ok ? pl.phi[e] : 0.0f;
            lx[lr * S + lc] = ok ? pl.x[e] : T(0);
        });
    __syncthreads();
    T wu[4];
    for (int k = 0; k < RPW; ++k) {
        int lr = v * RPW + k, row = t.ty0 + lr;
        if (row >= d.nrow) break;
        if (k == 0) up_arcs(pl.w2, t, row, cok, cl, wu);
        T wd[4], wr[4], wl[4];
        row_arcs<T, LPR>(pl.w2, t, row, cok, cl, wd, wr, wl);
        const float *pc = lph + (lr + 1) * S + kLdsC0 + 4 * cl;
        T b[4], di[4], y[4], rr[4], z4[4];
        stencil4<T, LPR>(lx, lr, cl, wd, wu, wr, wl, y);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            T pa = (T)pc[j];
            T bv = T(0), ws = T(0);
            if (wd[j] > T(0)) { bv = Prec<T>::fmaT(-wd[j], Prec<T>::wrap((T)pc[j + S] - pa), bv); ws += wd[j]; }
            if (wu[j] > T(0)) { bv = Prec<T>::fmaT(wu[j], Prec<T>::wrap(pa - (T)pc[j - S]), bv); ws += wu[j]; }
            if (wr[j] > T(0)) { bv = Prec<T>::fmaT(-wr[j], Prec<T>::wrap((T)pc[j + 1] - pa), bv); ws += wr[j]; }
            if (wl[j] > T(0)) { bv = Prec<T>::fmaT(wl[j], Prec<T>::wrap(pa - (T)pc[j - 1]), bv); ws += wl[j]; }
            b[j] = bv;
            di[j] = T(1) / (ws + Prec<T>::eps);
            rr[j] = bv - y[j];
            z4[j] = T(0);
        }
        if (cok) {
            long long e = d.off + (long long)row * d.pitch + c;
            st4(pl.b + e, b);
            st4(pl.dinv + e, di);
            st4(pl.r + e, rr);
            st4(pl.p + e, z4);
            st4(pl.s + e, z4);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) wu[j] = wd[j];
    }
}

// ═══════════════════════════════════════════════════════════════════════════
// Start of a solve: partial sums of r.z with z = r * dinv, the Jacobi-
// preconditioned residual (cuphu_laplace.cu:437-446).  Elementwise over the
// stencil tiles, so its partial sums line up with the other kernels'.
// ═══════════════════════════════════════════════════════════════════════════
template <class T, int RPW>
__global__ __launch_bounds__(256) void k_rz(const InstDesc *descs, const int *blk_inst, Planes<T> pl,
                                            const SolveScal *scal, int force, double2 *partials) {
    constexpr int TH = 4 * RPW;
    TileInfo t = get_tile(descs, blk_inst, TH);
    if (!force && !scal[t.inst].active) return;
    int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const InstDesc &d = t.d;
    int c = t.tx0 + 4 * lane;
    double g = 0.0;
    if (c < d.pitch) {
        for (int k = 0; k < RPW; ++k) {
            int row = t.ty0 + wave * RPW + k;
            if (row >= d.nrow) break;
            long long e = d.off + (long long)row * d.pitch + c;
            T r4[4], di[4];
            ld4(pl.r + e, r4);
            ld4(pl.dinv + e, di);
#pragma unroll
            for (int j = 0; j < 4; ++j) g += (double)r4[j] * (double)(r4[j] * di[j]);
        }
    }
    block_sum2(g, 0.0, partials + t.blk);
}

// ═══════════════════════════════════════════════════════════════════════════
// PCG direction and operator, fused (cuphu_laplace.cu:493-498 then
// :452-458):
//   p = z + beta p   (z = r * dinv; p = z on a solve's first step, beta = 0)
//   q = L p,  partial sums of p.q.
// The halo values of p are recomputed from r, dinv and the old direction of
// the neighbouring pixels with the same operations, so they equal the values
// the neighbouring workgroup writes bit for bit.  The old direction is read
// from one plane and the new one written to the other (parity 0: p -> p2,
// parity 1: p2 -> p), so no workgroup overwrites a value that a neighbour
// still reads.  q is stored in the s plane.
// ═══════════════════════════════════════════════════════════════════════════
template <class T, int RPW>
__global__ __launch_bounds__(256) void k_pq(const InstDesc *descs, const int *blk_inst, Planes<T> pl,
                                            const SolveScal *scal, double2 *partials, int parity) {
    constexpr int TH = 4 * RPW;
    __shared__ T lp[(TH + 2) * kLdsStride];
    TileInfo t = get_tile(descs, blk_inst, TH);
    const SolveScal *sc = scal + t.inst;
    if (!sc->active) return;
    const T be = (T)sc->beta;
    const bool keep = sc->beta != 0.0;   // first step of a solve: p = z
    const T *__restrict__ p_in = parity ? pl.p2 : pl.p;
    T *__restrict__ p_out = parity ? pl.p : pl.p2;
    const T *__restrict__ rg = pl.r;
    const T *__restrict__ dg = pl.dinv;
    const T *__restrict__ w2 = pl.w2;
    int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const InstDesc &d = t.d;
    const int c = t.tx0 + 4 * lane;
    const bool cok = c < d.pitch;
    const int row0 = t.ty0 + wave * RPW;
    // Every load of the own rows is issued before any use: r, 1/diag, the old
    // direction, the {down, right} arcs, the up arcs of the first row and,
    // for lane 0, the left arc from the column before the tile.
    T r4[RPW][4], di[RPW][4], po[RPW][4], wd[RPW][4], wr[RPW][4], wu[4], wlo[RPW];
#pragma unroll
    for (int k = 0; k < RPW; ++k) {
        zero4(r4[k]); zero4(di[k]); zero4(po[k]); zero4(wd[k]); zero4(wr[k]);
        wlo[k] = T(0);
        const int row = row0 + k;
        if (row < d.nrow) {
            long long e = d.off + (long long)row * d.pitch + c;
            if (cok) {
                ld4(rg + e, r4[k]);
                ld4(dg + e, di[k]);
                if (keep) ld4(p_in + e, po[k]);
                ld_arcs(w2, e, wd[k], wr[k]);
            }
            if (lane == 0 && t.tx0 > 0) wlo[k] = w2[2 * (e - 1) + 1];
        }
    }
    up_arcs(w2, t, row0, cok && row0 < d.nrow, lane, wu);
    fill_halo<TH>(
        t, wave, lane,
        [&](int lr, int lc, long long e, bool ok) {
            T pn[4];
            zero4(pn);
            if (ok) {
                T a[4], b[4], o[4];
                ld4(rg + e, a);
                ld4(dg + e, b);
                if (keep) ld4(p_in + e, o);
                else zero4(o);
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    T z = a[j] * b[j];
                    pn[j] = keep ? Prec<T>::fmaT(be, o[j], z) : z;
                }
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) lp[lr * kLdsStride + lc + j] = pn[j];
        },
        [&](int lr, int lc, long long e, bool ok) {
            T v = T(0);
            if (ok) {
                T z = rg[e] * dg[e];
                v = keep ? Prec<T>::fmaT(be, p_in[e], z) : z;
            }
            lp[lr * kLdsStride + lc] = v;
        });
#pragma unroll
    for (int k = 0; k < RPW; ++k) {
        const int lr = wave * RPW + k, row = row0 + k;
        T pn[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            T z = r4[k][j] * di[k][j];
            pn[j] = keep ? Prec<T>::fmaT(be, po[k][j], z) : z;
        }
        if (row < d.nrow && cok) st4(p_out + d.off + (long long)row * d.pitch + c, pn);
        T *lq = lp + (lr + 1) * kLdsStride + kLdsC0 + 4 * lane;
#pragma unroll
        for (int j = 0; j < 4; ++j) lq[j] = pn[j];
    }
    __syncthreads();
    double pq = 0.0;
#pragma unroll
    for (int k = 0; k < RPW; ++k) {
        const int lr = wave * RPW + k, row = row0 + k;
        if (row >= d.nrow) break;
        T wl[4], y[4];
        T from_left = __shfl_up(wr[k][3], 1, 64);
        wl[0] = (lane == 0) ? wlo[k] : from_left;
        wl[1] = wr[k][0]; wl[2] = wr[k][1]; wl[3] = wr[k][2];
        stencil4(lp, lr, lane, wd[k], (k == 0) ? wu : wd[k > 0 ? k - 1 : 0], wr[k], wl, y);
        if (cok) {
            st4(pl.s + d.off + (long long)row * d.pitch + c, y);
            const T *pc = lp + (lr + 1) * kLdsStride + kLdsC0 + 4 * lane;
#pragma unroll
            for (int j = 0; j < 4; ++j) pq += (double)pc[j] * (double)y[j];
        }
    }
    block_sum2(pq, 0.0, partials + t.blk);
}

// ═══════════════════════════════════════════════════════════════════════════
// PCG update (cuphu_laplace.cu:466-476): x += alpha p, r -= alpha q,
// z = r * dinv, partial sums of r.z.  p is the direction k_pq of the same
// parity wrote.  All loads of a lane's rows are issued before the first use.
// ═══════════════════════════════════════════════════════════════════════════
template <class T, int RPW>
__global__ __launch_bounds__(256) void k_xr(const InstDesc *descs, const int *blk_inst, Planes<T> pl,
                                            const SolveScal *scal, double2 *partials, int parity) {
    constexpr int TH = 4 * RPW;
    TileInfo t = get_tile(descs, blk_inst, TH);
    const SolveScal *sc = scal + t.inst;
    if (!sc->active) return;
    const T al = (T)sc->alpha;
    const T *__restrict__ pd = parity ? pl.p : pl.p2;
    T *__restrict__ xg = pl.x;
    T *__restrict__ rg = pl.r;
    const T *__restrict__ qg = pl.s;
    const T *__restrict__ dg = pl.dinv;
    int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const InstDesc &d = t.d;
    const int c = t.tx0 + 4 * lane;
    const int row0 = t.ty0 + wave * RPW;
    double g = 0.0;
    if (c < d.pitch) {
        T x4[RPW][4], p4[RPW][4], r4[RPW][4], q4[RPW][4], di[RPW][4];
#pragma unroll
        for (int k = 0; k < RPW; ++k) {
            if (row0 + k < d.nrow) {
                long long e = d.off + (long long)(row0 + k) * d.pitch + c;
                ld4(xg + e, x4[k]);
                ld4(pd + e, p4[k]);
                ld4(rg + e, r4[k]);
                ld4(qg + e, q4[k]);
                ld4(dg + e, di[k]);
            }
        }
#pragma unroll
        for (int k = 0; k < RPW; ++k) {
            if (row0 + k < d.nrow) {
                long long e = d.off + (long long)(row0 + k) * d.pitch + c;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    x4[k][j] = Prec<T>::fmaT(al, p4[k][j], x4[k][j]);
                    r4[k][j] = Prec<T>::fmaT(-al, q4[k][j], r4[k][j]);
                    g += (double)r4[k][j] * (double)(r4[k][j] * di[k][j]);
                }
                st4(xg + e, x4[k]);
                st4(rg + e, r4[k]);
            }
        }
    }
    block_sum2(g, 0.0, partials + t.blk);
}

// ═══════════════════════════════════════════════════════════════════════════
// True residual r_t = b - L x into the p2 plane (free at the end of an even
// chunk: the current direction is in p), with partial sums of r_t.z_t.
// ═══════════════════════════════════════════════════════════════════════════
template <class T, int RPW>
__global__ __launch_bounds__(256) void k_rr(const InstDesc *descs, const int *blk_inst, Planes<T> pl,
                                            const SolveScal *scal, double2 *partials) {
    constexpr int TH = 4 * RPW;
    __shared__ T lx[(TH + 2) * kLdsStride];
    TileInfo t = get_tile(descs, blk_inst, TH);
    if (!scal[t.inst].active) return;
    int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const InstDesc &d = t.d;
    int c = t.tx0 + 4 * lane;
    bool cok = c < d.pitch;
    for (int k = 0; k < RPW; ++k) {
        int lr = wave * RPW + k, row = t.ty0 + lr;
        T xv[4];
        zero4(xv);
        if (row < d.nrow && cok) ld4(pl.x + d.off + (long long)row * d.pitch + c, xv);
        T *lq = lx + (lr + 1) * kLdsStride + kLdsC0 + 4 * lane;
#pragma unroll
        for (int j = 0; j < 4; ++j) lq[j] = xv[j];
    }
    fill_halo<TH>(
        t, wave, lane,
        [&](int lr, int lc, long long e, bool ok) {
            T xv[4];
            zero4(xv);
            if (ok) ld4(pl.x + e, xv);
#pragma unroll
            for (int j = 0; j < 4; ++j) lx[lr * kLdsStride + lc + j] = xv[j];
        },
        [&](int lr, int lc, long long e, bool ok) { lx[lr * kLdsStride + lc] = ok ? pl.x[e] : T(0); });
    __syncthreads();
    double g = 0.0;
    T wu[4];
    for (int k = 0; k < RPW; ++k) {
        int lr = wave * RPW + k, row = t.ty0 + lr;
        if (row >= d.nrow) break;
        if (k == 0) up_arcs(pl.w2, t, row, cok, lane, wu);
        T wd[4], wr[4], wl[4];
        double y[4];
        row_arcs(pl.w2, t, row, cok, lane, wd, wr, wl);
        stencil4_wide(lx, lr, lane, wd, wu, wr, wl, y);
        if (cok) {
            long long e = d.off + (long long)row * d.pitch + c;
            T b4[4], di[4], r4[4];
            ld4(pl.b + e, b4);
            ld4(pl.dinv + e, di);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                r4[j] = (T)((double)b4[j] - y[j]);
                g += (double)r4[j] * (double)(r4[j] * di[j]);
            }
            st4(pl.p2 + e, r4);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) wu[j] = wd[j];
    }
    block_sum2(g, 0.0, partials + t.blk);
}

// Copy the true residual over r where the reduction chose to replace it.
template <class T, int RPW>
__global__ __launch_bounds__(256) void k_rr_commit(const InstDesc *descs, const int *blk_inst, Planes<T> pl,
                                                   const SolveScal *scal) {
    constexpr int TH = 4 * RPW;
    TileInfo t = get_tile(descs, blk_inst, TH);
    if (!scal[t.inst].active || !scal[t.inst].replace) return;
    int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const InstDesc &d = t.d;
    int c = t.tx0 + 4 * lane;
    if (c >= d.pitch) return;
    for (int k = 0; k < RPW; ++k) {
        int row = t.ty0 + wave * RPW + k;
        if (row >= d.nrow) break;
        long long e = d.off + (long long)row * d.pitch + c;
        T v[4];
        ld4(pl.p2 + e, v);
        st4(pl.r + e, v);
    }
}

// ═══════════════════════════════════════════════════════════════════════════
// Fused PCG iteration: one stencil kernel and one reduction per iteration.
//
// The reference iteration j (cuphu_laplace.cu:452-498) is
//   q_j = A p_j; alpha_j = rz_j / p_j.q_j; x += alpha_j p_j;
//   r_{j+1} = r_j - alpha_j q_j; z_{j+1} = r_{j+1} * dinv;
//   rz_{j+1} = r_{j+1}.z_{j+1}; beta_j = rz_{j+1} / rz_j;
//   p_{j+1} = z_{j+1} + beta_j p_j.
// k_it fuses the vector updates of iteration j with the operator of
// iteration j+1: from (x, r_j, p_j, q_j) and (alpha_j, beta_j) it forms
// x_{j+1}, r_{j+1}, z_{j+1}, p_{j+1}, q_{j+1} = A p_{j+1} and the partial
// sums of p.q, r.z, q.(dinv r) and q.(dinv q) of step j+1.  beta_j is needed
// before r_{j+1}.z_{j+1} is known, so the reduction after step j takes it
// from the expansion
//   rz_{j+1} = rz_j - 2 alpha_j q_j.(dinv r_j) + alpha_j^2 q_j.(dinv q_j),
// which equals the reference's sum up to the rounding of r_{j+1}; alpha, the
// stopping test (rz_{j+1} < tol^2 rz_0) and the breakdown test (p.q <= 0)
// use the sums of the vectors actually formed, as in the reference.
//
// Vector operations are the reference's, in the oracle's FP32 order: x =
// fma(a, p, x); r = fma(-a, q, r); z = r * dinv; p = fma(b, p, z).  dinv of
// the own pixels is recomputed from their arcs (bit-identical to the stored
// plane); halo pixels read the plane.
// r, p and q are read at halo pixels, so they alternate between two plane
// sets (parity 0: (r, p, s) -> (r2, p2, q2); parity 1: the reverse); x is
// only touched at its own pixel.  start = 1 forms p_0 = z_0 from r in set 0
// (no update) and writes set 1.  A residual replacement (replace flag set by
// the reduction) takes r_{j+1} from the wv plane instead of r_j - a q_j.
// ═══════════════════════════════════════════════════════════════════════════
__device__ __forceinline__ void block_sum4(double a, double b, double c, double e, double4 *out) {
    __shared__ double s4[4][4];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        a += __shfl_xor(a, o, 64);
        b += __shfl_xor(b, o, 64);
        c += __shfl_xor(c, o, 64);
        e += __shfl_xor(e, o, 64);
    }
    int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) { s4[0][wave] = a; s4[1][wave] = b; s4[2][wave] = c; s4[3][wave] = e; }
    __syncthreads();
    if (threadIdx.x == 0)
        *out = make_double4(((s4[0][0] + s4[0][1]) + s4[0][2]) + s4[0][3], ((s4[1][0] + s4[1][1]) + s4[1][2]) + s4[1][3],
                            ((s4[2][0] + s4[2][1]) + s4[2][2]) + s4[2][3], ((s4[3][0] + s4[3][1]) + s4[3][2]) + s4[3][3]);
}

// Arguments of the folded iteration kernel: the previous launch's partial
// sums (reduced here), where the new scalar state goes, and the stopping
// rule of the reduction.
struct FoldArgs {
    const double4 *part_prev;
    SolveScal *scal_next;
    int max_iter;
    double tol2;
};

// Sum of nb workgroup partials in the reductions' fixed order: thread t adds
// partials t, t + 256, t + 512, ... (four loads in flight before their
// sums), then the block butterfly; *res holds the sums for every thread
// after the closing barrier.
__device__ __forceinline__ void sum_partials4(const double4 *pp, int nb, double4 *res) {
    double a = 0.0, b = 0.0, c = 0.0, e = 0.0;
    int i = threadIdx.x;
    for (; i + 768 < nb; i += 1024) {
        const double4 v0 = pp[i], v1 = pp[i + 256], v2 = pp[i + 512], v3 = pp[i + 768];
        a += v0.x; b += v0.y; c += v0.z; e += v0.w;
        a += v1.x; b += v1.y; c += v1.z; e += v1.w;
        a += v2.x; b += v2.y; c += v2.z; e += v2.w;
        a += v3.x; b += v3.y; c += v3.z; e += v3.w;
    }
    for (; i < nb; i += 256) {
        const double4 v = pp[i];
        a += v.x; b += v.y; c += v.z; e += v.w;
    }
    block_sum4(a, b, c, e, res);
    __syncthreads();
}

// alpha_j = rz / p.q and beta from the rz expansion (see k_it); stops on
// p.q <= 0 (cuphu_laplace.cu:460).  v = (p.q, r.z, q.(D r), q.(D q)).
__device__ __forceinline__ void fused_alpha_beta(SolveScal &s, const double4 v) {
    const double pq = v.x, rz = v.y, qr = v.z, qq = v.w;
    if (!(pq > 0.0)) {
        s.status = isfinite(pq) ? CUPHU_SOLVE_BREAKDOWN : CUPHU_SOLVE_NONFINITE;
        s.active = 0;
        return;
    }
    const double al = rz / pq;
    double rzn = rz - 2.0 * al * qr + al * al * qq;
    if (!(rzn > 0.0)) rzn = 0.0;
    s.alpha = al;
    s.beta = rzn / rz;
    s.status = CUPHU_SOLVE_MAXITER;
    s.active = 1;
}

// The scalar recurrence after one fused iteration (the reduction's RED_IT
// mode; the folded iteration kernel runs the same code): one iteration
// counted, converged if rz < tol^2 rz_0 (:478), the cap, then alpha / beta.
__device__ __forceinline__ void fused_it_scalars(SolveScal &s, const double4 v, int max_iter, double tol2) {
    const double pq = v.x, rz = v.y;
    s.replace = 0;
    s.iters += 1;
    // the launch just done applied xmode 0 (deferred alpha_j p_j) or 1
    // (applied both); alpha_prev = alpha_j for the finalize / next launch
    s.xmode = (s.xmode == 0) ? 1 : (s.xmode == 1 ? 0 : 2);
    s.alpha_prev = s.alpha;
    s.relres = (s.gamma0 > 0.0) ? sqrt(fabs(rz) / s.gamma0) : 0.0;
    s.gamma_prev = s.gamma;
    s.gamma = rz;
    s.delta = pq;
    if (rz < tol2 * s.gamma0) { s.status = CUPHU_SOLVE_CONVERGED; s.active = 0; return; }
    if (s.iters >= max_iter) { s.status = CUPHU_SOLVE_MAXITER; s.active = 0; return; }
    fused_alpha_beta(s, v);
}


template <class T>
struct ItSets {
    const T *r_in, *p_in, *q_in;
    T *r_out, *p_out, *q_out;
};
template <class T>
__device__ __forceinline__ ItSets<T> it_sets(const Planes<T> &pl, int parity) {
    ItSets<T> s;
    s.r_in = parity ? pl.r2 : pl.r;
    s.p_in = parity ? pl.p2 : pl.p;
    s.q_in = parity ? pl.q2 : pl.s;
    s.r_out = parity ? pl.r : pl.r2;
    s.p_out = parity ? pl.p : pl.p2;
    s.q_out = parity ? pl.s : pl.q2;
    return s;
}

template <class T, int RPW>
__global__ __launch_bounds__(256) void k_it(const InstDesc *descs, const int *blk_inst, Planes<T> pl,
                                            const SolveScal *scal, double4 *partials, int parity, int start) {
    constexpr int TH = 4 * RPW;
    __shared__ T lp[(TH + 2) * kLdsStride];
    TileInfo t = get_tile(descs, blk_inst, TH);
    const SolveScal *sc = scal + t.inst;
    if (!start && !sc->active) return;
    const bool upd = !start;                        // x += a p, r -= a q
    const bool keep = !start && sc->beta != 0.0;    // p = z + b p (else p = z)
    const bool rep = !start && sc->replace != 0;    // r from the true-residual plane
    const T al = upd ? (T)sc->alpha : T(0);
    const T be = (T)sc->beta;
    const ItSets<T> S = it_sets(pl, parity);
    const T *__restrict__ r_in = rep ? pl.wv : S.r_in;
    const T *__restrict__ p_in = S.p_in;
    const T *__restrict__ q_in = S.q_in;
    const T *__restrict__ dg = pl.dinv;
    const T *__restrict__ w2 = pl.w2;
    T *__restrict__ xg = pl.x;
    const bool need_p = upd || keep, need_q = upd && !rep;
    int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const InstDesc &d = t.d;
    const int c = t.tx0 + 4 * lane;
    const bool cok = c < d.pitch;
    const int row0 = t.ty0 + wave * RPW;
    // the pixel update shared by own and halo pixels
    auto step1 = [&](T r, T q, T p, T di, T &rn, T &pn) {
        rn = (upd && !rep) ? Prec<T>::fmaT(-al, q, r) : r;
        T z = rn * di;
        pn = keep ? Prec<T>::fmaT(be, p, z) : z;
    };
    T x4[RPW][4], r4[RPW][4], p4[RPW][4], q4[RPW][4], di[RPW][4], wd[RPW][4], wr[RPW][4], wl[RPW][4], wu[4],
        wlo[RPW];
#pragma unroll
    for (int k = 0; k < RPW; ++k) {
        zero4(x4[k]); zero4(r4[k]); zero4(p4[k]); zero4(q4[k]); zero4(wd[k]); zero4(wr[k]);
        wlo[k] = T(0);
        const int row = row0 + k;
        if (row < d.nrow) {
            long long e = d.off + (long long)row * d.pitch + c;
            if (cok) {
                // x is touched by this pixel only: stream it past the caches;
                // r, p and q stay cached for the neighbours' halo reads
                if (upd) ld4_nt(xg + e, x4[k]);
                ld4(r_in + e, r4[k]);
                if (need_p) ld4(p_in + e, p4[k]);
                if (need_q) ld4(q_in + e, q4[k]);
                ld_arcs(w2, e, wd[k], wr[k]);
            }
            if (lane == 0 && t.tx0 > 0) wlo[k] = w2[2 * (e - 1) + 1];
        }
    }
    up_arcs(w2, t, row0, cok && row0 < d.nrow, lane, wu);
    fill_halo<TH>(
        t, wave, lane,
        [&](int lr, int lc, long long e, bool ok) {
            T pn[4];
            zero4(pn);
            if (ok) {
                T a[4], b[4], q[4], p[4];
                ld4(r_in + e, a);
                ld4(dg + e, b);
                if (need_q) ld4(q_in + e, q); else zero4(q);
                if (keep) ld4(p_in + e, p); else zero4(p);
#pragma unroll
                for (int j = 0; j < 4; ++j) { T rn; step1(a[j], q[j], p[j], b[j], rn, pn[j]); }
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) lp[lr * kLdsStride + lc + j] = pn[j];
        },
        [&](int lr, int lc, long long e, bool ok) {
            T v = T(0);
            if (ok) {
                T rn;
                step1(r_in[e], need_q ? q_in[e] : T(0), keep ? p_in[e] : T(0), dg[e], rn, v);
            }
            lp[lr * kLdsStride + lc] = v;
        });
    // Own pixels' Jacobi inverse from their four arcs, summed in the
    // assemble kernel's order (down, up, right, left; a zero arc adds
    // nothing), so it equals the stored 1/diag bit for bit without reading
    // it.  Left arcs: the right arc of the previous column (neighbouring
    // lane, or the load for lane 0).
#pragma unroll
    for (int k = 0; k < RPW; ++k) {
        T from_left = __shfl_up(wr[k][3], 1, 64);
        wl[k][0] = (lane == 0) ? wlo[k] : from_left;
        wl[k][1] = wr[k][0]; wl[k][2] = wr[k][1]; wl[k][3] = wr[k][2];
        const T *wuk = (k == 0) ? wu : wd[k > 0 ? k - 1 : 0];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            T ws = ((wd[k][j] + wuk[j]) + wr[k][j]) + wl[k][j];
            di[k][j] = T(1) / (ws + Prec<T>::eps);
        }
    }
    // own pixels: updates, stores, LDS
#pragma unroll
    for (int k = 0; k < RPW; ++k) {
        const int lr = wave * RPW + k, row = row0 + k;
        T pn[4], rn[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (upd) x4[k][j] = Prec<T>::fmaT(al, p4[k][j], x4[k][j]);
            step1(r4[k][j], q4[k][j], p4[k][j], di[k][j], rn[j], pn[j]);
            r4[k][j] = rn[j];
        }
        if (row < d.nrow && cok) {
            long long e = d.off + (long long)row * d.pitch + c;
            if (upd) st4_nt(xg + e, x4[k]);
            st4_nt(S.r_out + e, rn);
            st4_nt(S.p_out + e, pn);
        }
        T *lq = lp + (lr + 1) * kLdsStride + kLdsC0 + 4 * lane;
#pragma unroll
        for (int j = 0; j < 4; ++j) lq[j] = pn[j];
    }
    __syncthreads();
    double s_pq = 0.0, s_rz = 0.0, s_qr = 0.0, s_qq = 0.0;
#pragma unroll
    for (int k = 0; k < RPW; ++k) {
        const int lr = wave * RPW + k, row = row0 + k;
        if (row >= d.nrow) break;
        T y[4];
        stencil4(lp, lr, lane, wd[k], (k == 0) ? wu : wd[k > 0 ? k - 1 : 0], wr[k], wl[k], y);
        if (cok) {
            st4_nt(S.q_out + d.off + (long long)row * d.pitch + c, y);
            const T *pc = lp + (lr + 1) * kLdsStride + kLdsC0 + 4 * lane;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                double rd = (double)r4[k][j], dd = (double)di[k][j], qd = (double)y[j];
                s_pq += (double)pc[j] * qd;
                s_rz += rd * (double)(r4[k][j] * di[k][j]);
                double qdi = qd * dd;
                s_qr += qdi * rd;
                s_qq += qdi * qd;
            }
        }
    }
    block_sum4(s_pq, s_rz, s_qr, s_qq, partials + t.blk);
}

// True residual for the fused solver, at the end of an even chunk (state in
// set 1 with alpha_j pending): x' = x + alpha_j p_j on own and halo pixels,
// r_t = b - A x' into the wv plane, partial sums of r_t.(dinv r_t) and of
// r'.(dinv r') for the recursive r' = r_j - alpha_j q_j.
template <class T, int RPW>
__global__ __launch_bounds__(256) void k_rr_it(const InstDesc *descs, const int *blk_inst, Planes<T> pl,
                                               const SolveScal *scal, double4 *partials) {
    constexpr int TH = 4 * RPW;
    __shared__ T lx[(TH + 2) * kLdsStride];
    TileInfo t = get_tile(descs, blk_inst, TH);
    const SolveScal *sc = scal + t.inst;
    if (!sc->active) return;
    const T al = (T)sc->alpha;
    const T *__restrict__ pg = pl.p2;
    const T *__restrict__ xg = pl.x;
    int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const InstDesc &d = t.d;
    const int c = t.tx0 + 4 * lane;
    const bool cok = c < d.pitch;
    for (int k = 0; k < RPW; ++k) {
        int lr = wave * RPW + k, row = t.ty0 + lr;
        T xv[4], pv[4];
        zero4(xv); zero4(pv);
        if (row < d.nrow && cok) {
            long long e = d.off + (long long)row * d.pitch + c;
            ld4(xg + e, xv);
            ld4(pg + e, pv);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) lx[(lr + 1) * kLdsStride + kLdsC0 + 4 * lane + j] = Prec<T>::fmaT(al, pv[j], xv[j]);
    }
    fill_halo<TH>(
        t, wave, lane,
        [&](int lr, int lc, long long e, bool ok) {
            T xv[4], pv[4];
            zero4(xv); zero4(pv);
            if (ok) { ld4(xg + e, xv); ld4(pg + e, pv); }
#pragma unroll
            for (int j = 0; j < 4; ++j) lx[lr * kLdsStride + lc + j] = Prec<T>::fmaT(al, pv[j], xv[j]);
        },
        [&](int lr, int lc, long long e, bool ok) {
            lx[lr * kLdsStride + lc] = ok ? Prec<T>::fmaT(al, pg[e], xg[e]) : T(0);
        });
    __syncthreads();
    double s_t = 0.0, s_r = 0.0;
    T wu[4];
    for (int k = 0; k < RPW; ++k) {
        int lr = wave * RPW + k, row = t.ty0 + lr;
        if (row >= d.nrow) break;
        if (k == 0) up_arcs(pl.w2, t, row, cok, lane, wu);
        T wd[4], wr[4], wl[4];
        double y[4];
        row_arcs(pl.w2, t, row, cok, lane, wd, wr, wl);
        stencil4_wide(lx, lr, lane, wd, wu, wr, wl, y);
        if (cok) {
            long long e = d.off + (long long)row * d.pitch + c;
            T b4[4], di[4], rt[4], r4[4], q4[4];
            ld4(pl.b + e, b4);
            ld4(pl.dinv + e, di);
            ld4(pl.r2 + e, r4);
            ld4(pl.q2 + e, q4);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                rt[j] = (T)((double)b4[j] - y[j]);
                T rr = Prec<T>::fmaT(-al, q4[j], r4[j]);
                s_t += (double)rt[j] * (double)(rt[j] * di[j]);
                s_r += (double)rr * (double)(rr * di[j]);
            }
            st4(pl.wv + e, rt);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) wu[j] = wd[j];
    }
    block_sum4(s_t, s_r, 0.0, 0.0, partials + t.blk);
}

// ═══════════════════════════════════════════════════════════════════════════
// Fused PCG iteration without a stored q (the default fused kernel).
//
// Same iteration as k_it, but q_j = A p_j is recomputed from p_j instead of
// being written by the previous launch and read back: the stencil of the
// previous launch and this recomputation see the same p_j and arcs and use
// the same operation order, so q_j is bit-identical.  Per pixel this moves
// 24 B (reads r, p and the arc pair; writes r, p), plus 12 B (x and the
// previous p in, x out) on every other launch, which applies the deferred x
// updates: 30 B on average instead of k_it's 40.
//
// LDS: lp0 holds p_j with a two-pixel halo (rows -2..TH+1, columns
// -2..TW+1), lp1 holds p_{j+1} with a one-pixel halo.  Phase 1 loads the own
// rows and stages p_j; the first lane group loads the ring row above the
// tile (and p_j two rows up), the last lane group the ring row below, waves
// 1 and 2 the ring columns left and right (one pixel per lane).  Phase 2 forms q_j, r_{j+1}, z, p_{j+1} at the own pixels and on
// the ring (whose arcs come from the loads of phase 1, the ring's 1/diag
// recomputed from its arcs like the own pixels') and stores x, r, p.
// Phase 3 applies the stencil to p_{j+1} for the sums.  Plane sets: (r, p)
// and (r2, p2); start and replacement modes as in k_it.
// ═══════════════════════════════════════════════════════════════════════════
template <class T>
__device__ __forceinline__ T stencil1(T a, T dn, T up, T rt, T lf, T wd, T wu, T wr, T wl) {
    T y = Prec<T>::eps * a;
    y = Prec<T>::fmaT(wd, a - dn, y);
    y = Prec<T>::fmaT(wu, a - up, y);
    y = Prec<T>::fmaT(wr, a - rt, y);
    y = Prec<T>::fmaT(wl, a - lf, y);
    return y;
}
template <class T>
__device__ __forceinline__ T diag_inv(T wd, T wu, T wr, T wl) {
    return T(1) / ((((wd + wu) + wr) + wl) + Prec<T>::eps);
}

// Workgroups that start together on one CU run identical instruction streams
// and, left alone, stay in step for the whole launch: all halo loads in flight
// at once, then all ALU work, then all stores.  Holding back two thirds of the
// first resident round (3 workgroups x 256 CUs) by 0.85 / 1.7 us puts them
// out of phase, and the workgroups that replace them inherit the offsets:
// 0.446 -> 0.433 ms per k_iq launch over 8192^2 pixels.  Any offset from 0.4
// to 3.4 us gives the same; s_sleep(32) is 2048 clocks.  A grid of a few
// rounds (a single 2048^2 or 1024^2 image) lasts too short a time to repay
// the delay: the host launches the unstaggered instantiation for it.
constexpr unsigned kStaggerBlocks = 768;
constexpr int kStaggerMinGrid = 4 * (int)kStaggerBlocks;

__device__ __forceinline__ void stagger_first_round() {
    if (blockIdx.x < kStaggerBlocks) {
        const unsigned ph = (blockIdx.x / 8u) % 3u;   // consecutive ids alternate over the 8 XCDs
        for (unsigned i = 0; i < ph; ++i) __builtin_amdgcn_s_sleep(32);
    }
}

// The iteration body; FOLD: the previous launch's reduction first (see
// launch_iq_fold).  Three kernels wrap it (k_iq, k_iq_fold, k_iq_mix) so that
// each keeps its own register budget.  lp0 (Sh::L0 elements) holds p_j, LDS
// row = row + 2, col = col + kLdsC0; lp1 (Sh::L1) p_{j+1}, LDS row = row + 1.
// EXT (k_iq_mix, two tile shapes in one kernel): the tile is *ext_t and the
// LDS the caller's lx0 / lx1, one allocation for both shapes; otherwise the
// body declares its LDS and looks its tile up itself (a tile passed in from
// the wrapper lost the compiler some index range facts: 0.5 % longer k_iq
// launches).
template <class T, int RPW, int LPR, bool STAGGER, bool FOLD, bool EXT = false>
__device__ __forceinline__ void iq_body(const InstDesc *descs, const int *blk_inst, const TileInfo *ext_t,
                                        T *__restrict__ lx0, T *__restrict__ lx1, const Planes<T> &pl,
                                        const SolveScal *scal, double4 *partials, int parity, int start,
                                        const FoldArgs &fa) {
    using Sh = Shape<RPW, LPR>;
    constexpr int TH = Sh::TH, S = Sh::S, TW = Sh::TW;
    __shared__ __align__(16) T own0[EXT ? 4 : Sh::L0];
    __shared__ __align__(16) T own1[EXT ? 4 : Sh::L1];
    T *__restrict__ lp0 = EXT ? lx0 : own0;
    T *__restrict__ lp1 = EXT ? lx1 : own1;
    TileInfo t;
    if constexpr (EXT) t = *ext_t;
    else t = get_tile(descs, blk_inst, TH, TW);
    const SolveScal *sc;
    if constexpr (FOLD) {
        // the previous launch's reduction, done here (every workgroup of the
        // instance forms the same state; its first tile stores it)
        __shared__ SolveScal ls;
        __shared__ double4 sums;
        const SolveScal prev = scal[t.inst];
        const bool lead = t.blk == t.d.block_begin;
        if (!prev.active) {
            if (lead && threadIdx.x == 0) fa.scal_next[t.inst] = prev;
            return;
        }
        sum_partials4(fa.part_prev + t.d.block_begin, t.d.tiles_x * t.d.tiles_y, &sums);
        if (threadIdx.x == 0) {
            SolveScal sn = prev;
            fused_it_scalars(sn, sums, fa.max_iter, fa.tol2);
            ls = sn;
            if (lead) fa.scal_next[t.inst] = sn;
        }
        __syncthreads();
        if (!ls.active) return;
        sc = &ls;
    } else {
        sc = scal + t.inst;
        if (!start && !sc->active) return;
    }
    if (STAGGER) stagger_first_round();
    const bool upd = !start;                          // x += a p
    const bool rep = upd && sc->replace != 0;         // r from the true-residual plane
    const bool useq = upd && !rep;                    // r -= a q_j, q_j = A p_j
    const bool keep = upd && sc->beta != 0.0;         // p = z + b p (else p = z)
    const T al = upd ? (T)sc->alpha : T(0);
    const T be = (T)sc->beta;
    // x update of this launch (see SolveScal::xmode): ax = apply alpha_j p_j,
    // ax2 = first apply the deferred alpha_{j-1} p_{j-1} (own pixel of the
    // output set, read before this launch overwrites it)
    const int xm = upd ? sc->xmode : 0;
    const bool ax = upd && xm != 0, ax2 = upd && xm == 1;
    const T alp = ax2 ? (T)sc->alpha_prev : T(0);
    const T *__restrict__ r_src = rep ? pl.wv : (parity ? pl.r2 : pl.r);
    const T *__restrict__ p_in = parity ? pl.p2 : pl.p;
    T *__restrict__ r_out = parity ? pl.r : pl.r2;
    T *__restrict__ p_out = parity ? pl.p : pl.p2;
    const T *__restrict__ w2 = pl.w2;
    T *__restrict__ xg = pl.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int cl = lane % LPR, v = wave * Sh::G + lane / LPR;   // column group, lane group
    const InstDesc &d = t.d;
    const int c = t.tx0 + 4 * cl;
    const bool cok = c < d.wcols;
    const int lrow0 = v * RPW;               // tile-local first own row of this lane group
    const bool top = v == 0, bot = v == 4 * Sh::G - 1;   // lane groups next to ring rows -1 and TH
    auto grow = [&](int lr) { return t.ty0 + lr; };
    auto elem = [&](int lr, int col) { return d.off + (long long)grow(lr) * d.pitch + col; };
    auto row_ok = [&](int lr) { return grow(lr) >= 0 && grow(lr) < d.nrow; };
    auto upd_px = [&](T r, T q, T p, T di, T &rn, T &pn) {
        rn = useq ? Prec<T>::fmaT(-al, q, r) : r;
        T z = rn * di;
        pn = keep ? Prec<T>::fmaT(be, p, z) : z;
    };

    // ── phase 1: loads; stage p_j ──
    T x4[RPW][4], r4[RPW][4], p4[RPW][4], wd[RPW][4], wr[RPW][4], wu[4], wlo[RPW], pp4[RPW][4];
    // p_j and the x-update operands first: the x update (own pixels only)
    // then waits for these loads alone and retires x and the previous p
    // before the stencil phases
#pragma unroll
    for (int k = 0; k < RPW; ++k) {
        zero4(x4[k]); zero4(p4[k]); zero4(pp4[k]);
        const int lr = lrow0 + k;
        if (row_ok(lr) && cok) {
            const long long e = elem(lr, c);
            if (upd) ld4(p_in + e, p4[k]);
            if (ax) ld4_nt(xg + e, x4[k]);
            if (ax2) ld4_nt(p_out + e, pp4[k]);
        }
    }
#pragma unroll
    for (int k = 0; k < RPW; ++k) {
        zero4(r4[k]); zero4(wd[k]); zero4(wr[k]);
        wlo[k] = T(0);
        const int lr = lrow0 + k;
        if (row_ok(lr)) {
            const long long e = elem(lr, c);
            if (cok) {
                ld4(r_src + e, r4[k]);
                ld_arcs(w2, e, wd[k], wr[k]);
            }
            if (cl == 0 && t.tx0 > 0) wlo[k] = w2[2 * (e - 1) + 1];
        }
    }
    up_arcs(w2, t, grow(lrow0), cok && row_ok(lrow0), cl, wu);
    // ring rows (top lane group: row -1 and p_j of row -2; bottom lane group:
    // row TH and p_j of row TH+1): r, p, arcs; ring row arcs {down, right} +
    // up + left
    T rr4[4], rp4[4], rwd[4], rwr[4], rwu[4], rwlo = T(0);
    zero4(rr4); zero4(rp4); zero4(rwd); zero4(rwr); zero4(rwu);
    if (top || bot) {
        const int lr = top ? -1 : TH;
        const int lr2 = top ? -2 : TH + 1;
        T p2v[4];
        zero4(p2v);
        if (row_ok(lr)) {
            const long long e = elem(lr, c);
            if (cok) {
                ld4(r_src + e, rr4);
                if (upd) ld4(p_in + e, rp4);
                ld_arcs(w2, e, rwd, rwr);
            }
            if (cl == 0 && t.tx0 > 0) rwlo = w2[2 * (e - 1) + 1];
        }
        if (top) {
            if (row_ok(lr2) && cok) {
                T dummy[4];
                ld_arcs(w2, elem(lr2, c), rwu, dummy);   // down arcs of row -2 = up arcs of row -1
                if (useq) ld4(p_in + elem(lr2, c), p2v);
            }
        } else {
#pragma unroll
            for (int j = 0; j < 4; ++j) rwu[j] = wd[RPW - 1][j];   // up arcs of row TH
            if (useq && row_ok(lr2) && cok) ld4(p_in + elem(lr2, c), p2v);
        }
        if (upd) {
            T *a = lp0 + (lr + 2) * S + kLdsC0 + 4 * cl;
            T *b = lp0 + (lr2 + 2) * S + kLdsC0 + 4 * cl;
#pragma unroll
            for (int j = 0; j < 4; ++j) { a[j] = rp4[j]; b[j] = p2v[j]; }
        }
    }
    // ring columns (wave 1: column -1, wave 2: column TW): lane l < TH owns
    // pixel (l, col); lanes l < TH + 2 stage p_j at (l - 1, col) (the 64-row
    // tall tile stages its last two rows in a second pass) and lanes l < TH
    // p_j at (l, col -/+ 1) beyond it
    static_assert(TH <= 64, "ring columns: one lane per own row");
    T cr = T(0), cp = T(0), cwd = T(0), cwr = T(0), cwu = T(0), cwl = T(0);
    if (wave == 1 || wave == 2) {
        const int col = (wave == 1) ? t.tx0 - 1 : t.tx0 + TW;
        const int col2 = (wave == 1) ? col - 1 : col + 1;
        const bool colok = col >= 0 && col < d.wcols, col2ok = col2 >= 0 && col2 < d.wcols;
        const int lcol = (wave == 1) ? -1 : TW, lcol2 = (wave == 1) ? -2 : TW + 1;
        if constexpr (TH + 2 <= 64) {
            if (lane < TH + 2) {
                const int lr = lane - 1;
                T v = T(0);
                if (upd && colok && row_ok(lr)) v = p_in[elem(lr, col)];
                if (upd) lp0[(lr + 2) * S + kLdsC0 + lcol] = v;
            }
        } else {
            for (int lr = lane - 1; lr <= TH; lr += 64) {
                T v = T(0);
                if (upd && colok && row_ok(lr)) v = p_in[elem(lr, col)];
                if (upd) lp0[(lr + 2) * S + kLdsC0 + lcol] = v;
            }
        }
        if (lane < TH) {
            const int lr = lane;
            T v2 = T(0);
            if (colok && row_ok(lr)) {
                const long long e = elem(lr, col);
                cr = r_src[e];
                if (upd) cp = p_in[e];
                cwd = w2[2 * e];
                cwr = w2[2 * e + 1];
                if (grow(lr) > 0) cwu = w2[2 * (e - d.pitch)];
                // left arc: the right arc of the pixel before (column -2
                // for the left ring, column 255 for the right ring)
                if (wave == 2 || col2ok) cwl = w2[2 * (e - 1) + 1];
                if (useq && col2ok) v2 = p_in[elem(lr, col2)];
            }
            if (useq) lp0[(lr + 2) * S + kLdsC0 + lcol2] = v2;
        }
    }
    // x update of the own pixels
    if (ax) {
#pragma unroll
        for (int k = 0; k < RPW; ++k) {
            const int lr = lrow0 + k;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                if (ax2) x4[k][j] = Prec<T>::fmaT(alp, pp4[k][j], x4[k][j]);
                x4[k][j] = Prec<T>::fmaT(al, p4[k][j], x4[k][j]);
            }
            if (row_ok(lr) && cok) st4_nt(xg + elem(lr, c), x4[k]);
        }
    }
    // own p_j
    if (upd) {
#pragma unroll
        for (int k = 0; k < RPW; ++k) {
            T *a = lp0 + (lrow0 + k + 2) * S + kLdsC0 + 4 * cl;
#pragma unroll
            for (int j = 0; j < 4; ++j) a[j] = p4[k][j];
        }
    }
    __syncthreads();

    // ── phase 2: q_j, r_{j+1}, p_{j+1} on own pixels and the ring ──
    T wl[RPW][4], di[RPW][4];
#pragma unroll
    for (int k = 0; k < RPW; ++k) {
        T from_left = __shfl_up(wr[k][3], 1, LPR);
        wl[k][0] = (cl == 0) ? wlo[k] : from_left;
        wl[k][1] = wr[k][0]; wl[k][2] = wr[k][1]; wl[k][3] = wr[k][2];
        const T *wuk = (k == 0) ? wu : wd[k > 0 ? k - 1 : 0];
#pragma unroll
        for (int j = 0; j < 4; ++j) di[k][j] = diag_inv(wd[k][j], wuk[j], wr[k][j], wl[k][j]);
    }
#pragma unroll
    for (int k = 0; k < RPW; ++k) {
        const int lr = lrow0 + k;
        T q[4], rn[4], pn[4];
        zero4(q);
        if (useq) stencil4<T, LPR>(lp0 + S, lr, cl, wd[k], (k == 0) ? wu : wd[k > 0 ? k - 1 : 0], wr[k], wl[k], q);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            upd_px(r4[k][j], q[j], p4[k][j], di[k][j], rn[j], pn[j]);
            r4[k][j] = rn[j];
        }
        if (row_ok(lr) && cok) {
            const long long e = elem(lr, c);
            st4_nt(r_out + e, rn);
            st4_nt(p_out + e, pn);
        }
        T *a = lp1 + (lr + 1) * S + kLdsC0 + 4 * cl;
#pragma unroll
        for (int j = 0; j < 4; ++j) a[j] = pn[j];
    }
    if (top || bot) {
        const int lr = top ? -1 : TH;
        T rwl[4], q[4], pn[4];
        T from_left = __shfl_up(rwr[3], 1, LPR);
        rwl[0] = (cl == 0) ? rwlo : from_left;
        rwl[1] = rwr[0]; rwl[2] = rwr[1]; rwl[3] = rwr[2];
        zero4(q);
        if (useq) stencil4<T, LPR>(lp0 + S, lr, cl, rwd, rwu, rwr, rwl, q);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            T rn;
            upd_px(rr4[j], q[j], rp4[j], diag_inv(rwd[j], rwu[j], rwr[j], rwl[j]), rn, pn[j]);
        }
        T *a = lp1 + (lr + 1) * S + kLdsC0 + 4 * cl;
#pragma unroll
        for (int j = 0; j < 4; ++j) a[j] = pn[j];
    }
    if ((wave == 1 || wave == 2) && lane < TH) {
        const int lr = lane;
        const int lcol = (wave == 1) ? -1 : TW;
        const T *m = lp0 + (lr + 2) * S + kLdsC0 + lcol;
        T q = T(0);
        if (useq) q = stencil1(m[0], m[S], m[-S], m[1], m[-1], cwd, cwu, cwr, cwl);
        T rn, pn;
        upd_px(cr, q, cp, diag_inv(cwd, cwu, cwr, cwl), rn, pn);
        lp1[(lr + 1) * S + kLdsC0 + lcol] = pn;
    }
    __syncthreads();

    // ── phase 3: q_{j+1} = A p_{j+1} on own pixels, sums ──
    double s_pq = 0.0, s_rz = 0.0, s_qr = 0.0, s_qq = 0.0;
#pragma unroll
    for (int k = 0; k < RPW; ++k) {
        const int lr = lrow0 + k;
        if (!row_ok(lr)) break;
        T y[4];
        stencil4<T, LPR>(lp1, lr, cl, wd[k], (k == 0) ? wu : wd[k > 0 ? k - 1 : 0], wr[k], wl[k], y);
        if (cok) {
            const T *pc = lp1 + (lr + 1) * S + kLdsC0 + 4 * cl;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                double rd = (double)r4[k][j], dd = (double)di[k][j], qd = (double)y[j];
                s_pq += (double)pc[j] * qd;
                s_rz += rd * (double)(r4[k][j] * di[k][j]);
                double qdi = qd * dd;
                s_qr += qdi * rd;
                s_qq += qdi * qd;
            }
        }
    }
    block_sum4(s_pq, s_rz, s_qr, s_qq, partials + t.blk);
}

template <class T, int RPW, int LPR, bool STAGGER>
__global__ __launch_bounds__(256) void k_iq(const InstDesc *descs, const int *blk_inst, Planes<T> pl,
                                            const SolveScal *scal, double4 *partials, int parity, int start) {
    const FoldArgs none{nullptr, nullptr, 0, 0.0};
    iq_body<T, RPW, LPR, STAGGER, false>(descs, blk_inst, nullptr, nullptr, nullptr, pl, scal, partials, parity,
                                         start, none);
}
// The folded FP32 kernel with 4 rows per lane group is held to 3 resident
// workgroups per CU (168 registers), like the unfolded one.
template <class T, int RPW, int LPR>
__global__ __launch_bounds__(256, (RPW == 4 && sizeof(T) == 4) ? 3 : 1) void k_iq_fold(
    const InstDesc *descs, const int *blk_inst, Planes<T> pl, const SolveScal *scal_prev, double4 *partials,
    int parity, FoldArgs fa) {
    iq_body<T, RPW, LPR, false, true>(descs, blk_inst, nullptr, nullptr, nullptr, pl, scal_prev, partials, parity,
                                      0, fa);
}

// The iteration on 256 x TH tiles (TH = 16 RPW) where an instance's last tile
// column holds 64 or 128 work columns (a width of 256 k + 64 columns, as the
// reference tile formula's 1843-column tiles have, or 256 k + 128): that
// column runs in 64 x 4TH (or 128 x 2TH) tall tiles, one workgroup for four
// (two) of the 256-column tiles, which would each keep three quarters (half)
// of their lanes idle.  Workgroups carry a fixed cost (tile lookup, halo
// ring, LDS phases, block sums) that the other workgroups' memory traffic
// does not hide, so four quarter-filled ones cost about 2.2 full ones and the
// tall one about 1.2.  tiles[b] = {instance, tile column, tile row, kind} in
// the launch order (b after the XCD remap), tile coordinates in 256 x TH
// units, kind 0 = 256 x TH, 1 = 64 x 4TH, 2 = 128 x 2TH.  A tall tile's sums
// go to the raster slot of the first 256 x TH tile it covers and the slots of
// the others hold zeros, so the reductions read the same slots as after k_iq.
template <class T, int RPW, bool STAGGER, bool FOLD>
__device__ __forceinline__ void iq_mix_body(const InstDesc *descs, const int4 *tiles, const Planes<T> &pl,
                                            const SolveScal *scal, double4 *partials, int parity, int start,
                                            const FoldArgs &fa) {
    using Sw = Shape<RPW, 64>;
    using Sn = Shape<RPW, 16>;
    using Sh = Shape<RPW, 32>;
    static_assert(Sn::TH == 4 * Sw::TH && Sn::TW == Sw::TW / 4, "a tall tile covers four wide tiles' rows");
    static_assert(Sh::TH == 2 * Sw::TH && Sh::TW == Sw::TW / 2, "a 128-column tile covers two wide tiles' rows");
    constexpr int kWide = Sw::L0 + Sw::L1, kTall = Sn::L0 + Sn::L1, kHalf = Sh::L0 + Sh::L1;
    static_assert(kTall <= kWide && kHalf <= kWide, "the tall shapes fit the wide tile's LDS");
    __shared__ __align__(16) T lds[kWide];
    const int b = xcd_block((int)blockIdx.x, (int)gridDim.x);
    const int4 q = tiles[b];
    TileInfo t;
    t.inst = q.x;
    t.d = descs[q.x];
    t.blk = t.d.block_begin + q.z * t.d.tiles_x + q.y;
    t.tx0 = q.y * Sw::TW;
    t.ty0 = q.z * Sw::TH;
    if (q.w) {
        if (q.w == 1)
            iq_body<T, RPW, 16, STAGGER, FOLD, true>(descs, nullptr, &t, lds, lds + Sn::L0, pl, scal, partials,
                                                     parity, start, fa);
        else
            iq_body<T, RPW, 32, STAGGER, FOLD, true>(descs, nullptr, &t, lds, lds + Sh::L0, pl, scal, partials,
                                                     parity, start, fa);
        const int k = (int)threadIdx.x, cover = q.w == 1 ? 4 : 2;
        if (k >= 1 && k < cover && q.z + k < t.d.tiles_y)
            partials[t.blk + k * t.d.tiles_x] = make_double4(0.0, 0.0, 0.0, 0.0);
    } else {
        iq_body<T, RPW, 64, STAGGER, FOLD, true>(descs, nullptr, &t, lds, lds + Sw::L0, pl, scal, partials, parity,
                                                 start, fa);
    }
}
template <class T, int RPW, bool STAGGER>
__global__ __launch_bounds__(256) void k_iq_mix(const InstDesc *descs, const int4 *tiles, Planes<T> pl,
                                                const SolveScal *scal, double4 *partials, int parity, int start) {
    const FoldArgs none{nullptr, nullptr, 0, 0.0};
    iq_mix_body<T, RPW, STAGGER, false>(descs, tiles, pl, scal, partials, parity, start, none);
}
// The merged tiling with the previous launch's reduction folded in (grids of
// at most kFoldMaxBlocks workgroups of the plain tiling).  Every instance's
// first workgroup is its tile (0, 0), in either tiling, and stores the state.
template <class T, int RPW>
__global__ __launch_bounds__(256, (RPW == 4 && sizeof(T) == 4) ? 3 : 1) void k_iq_fold_mix(
    const InstDesc *descs, const int4 *tiles, Planes<T> pl, const SolveScal *scal_prev, double4 *partials,
    int parity, FoldArgs fa) {
    iq_mix_body<T, RPW, false, true>(descs, tiles, pl, scal_prev, partials, parity, 0, fa);
}

// True residual for the q-free fused solver at the end of an even chunk
// (state in set 1 = (r2, p2) with alpha_j pending): x' = x + alpha_j p_j and
// q_j = A p_j from LDS tiles with a one-pixel halo; r_t = b - A x' into wv;
// partial sums of r_t.(dinv r_t) and of r'.(dinv r') for r' = r_j - alpha_j q_j.
template <class T, int RPW, int LPR>
__global__ __launch_bounds__(256) void k_rr_iq(const InstDesc *descs, const int *blk_inst, Planes<T> pl,
                                               const SolveScal *scal, double4 *partials) {
    using Sh = Shape<RPW, LPR>;
    constexpr int TH = Sh::TH, S = Sh::S;
    __shared__ T lx[(TH + 2) * S];
    __shared__ T lq[(TH + 2) * S];
    TileInfo t = get_tile(descs, blk_inst, TH, Sh::TW);
    const SolveScal *sc = scal + t.inst;
    if (!sc->active) return;
    const T al = (T)sc->alpha;
    const T *__restrict__ pg = pl.p2;
    const T *__restrict__ xg = pl.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int cl = lane % LPR, v = wave * Sh::G + lane / LPR;
    const InstDesc &d = t.d;
    const int c = t.tx0 + 4 * cl;
    const bool cok = c < d.pitch;
    for (int k = 0; k < RPW; ++k) {
        int lr = v * RPW + k, row = t.ty0 + lr;
        T xv[4], pv[4];
        zero4(xv); zero4(pv);
        if (row < d.nrow && cok) {
            long long e = d.off + (long long)row * d.pitch + c;
            ld4(xg + e, xv);
            ld4(pg + e, pv);
        }
        const int o = (lr + 1) * S + kLdsC0 + 4 * cl;
#pragma unroll
        for (int j = 0; j < 4; ++j) { lx[o + j] = Prec<T>::fmaT(al, pv[j], xv[j]); lq[o + j] = pv[j]; }
    }
    fill_halo<TH, LPR>(
        t, wave, lane,
        [&](int lr, int lc, long long e, bool ok) {
            T xv[4], pv[4];
            zero4(xv); zero4(pv);
            if (ok) { ld4(xg + e, xv); ld4(pg + e, pv); }
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                lx[lr * S + lc + j] = Prec<T>::fmaT(al, pv[j], xv[j]);
                lq[lr * S + lc + j] = pv[j];
            }
        },
        [&](int lr, int lc, long long e, bool ok) {
            lx[lr * S + lc] = ok ? Prec<T>::fmaT(al, pg[e], xg[e]) : T(0);
            lq[lr * S + lc] = ok ? pg[e] : T(0);
        });
    __syncthreads();
    double s_t = 0.0, s_r = 0.0;
    T wu[4];
    for (int k = 0; k < RPW; ++k) {
        int lr = v * RPW + k, row = t.ty0 + lr;
        if (row >= d.nrow) break;
        if (k == 0) up_arcs(pl.w2, t, row, cok, cl, wu);
        T wd[4], wr[4], wl[4], q[4];
        double y[4];
        row_arcs<T, LPR>(pl.w2, t, row, cok, cl, wd, wr, wl);
        stencil4_wide<T, LPR>(lx, lr, cl, wd, wu, wr, wl, y);
        stencil4<T, LPR>(lq, lr, cl, wd, wu, wr, wl, q);
        if (cok) {
            long long e = d.off + (long long)row * d.pitch + c;
            T b4[4], rt[4], r4[4];
            ld4(pl.b + e, b4);
            ld4(pl.r2 + e, r4);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                T di = diag_inv(wd[j], wu[j], wr[j], wl[j]);
                rt[j] = (T)((double)b4[j] - y[j]);
                T rr = Prec<T>::fmaT(-al, q[j], r4[j]);
                s_t += (double)rt[j] * (double)(rt[j] * di);
                s_r += (double)rr * (double)(rr * di);
            }
            st4(pl.wv + e, rt);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) wu[j] = wd[j];
    }
    block_sum4(s_t, s_r, 0.0, 0.0, partials + t.blk);
}

// Deferred x update still pending when a q-free solve stopped (xmode 1 after
// an odd number of launches): x += alpha_j p_j, p_j in plane set 1 (the
// input set of the odd-numbered launches).
template <class T, int RPW, int LPR>
__global__ __launch_bounds__(256) void k_x_finalize(const InstDesc *descs, const int *blk_inst, Planes<T> pl,
                                                    const SolveScal *scal) {
    using Sh = Shape<RPW, LPR>;
    TileInfo t = get_tile(descs, blk_inst, Sh::TH, Sh::TW);
    const SolveScal *sc = scal + t.inst;
    if (sc->xmode != 1) return;
    const T al = (T)sc->alpha_prev;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int cl = lane % LPR, v = wave * Sh::G + lane / LPR;
    const InstDesc &d = t.d;
    const int c = t.tx0 + 4 * cl;
    if (c >= d.pitch) return;
    for (int k = 0; k < RPW; ++k) {
        const int row = t.ty0 + v * RPW + k;
        if (row >= d.nrow) break;
        const long long e = d.off + (long long)row * d.pitch + c;
        T x4[4], p4[4];
        ld4(pl.x + e, x4);
        ld4(pl.p2 + e, p4);
#pragma unroll
        for (int j = 0; j < 4; ++j) x4[j] = Prec<T>::fmaT(al, p4[j], x4[j]);
        st4(pl.x + e, x4);
    }
}

// Reduction and scalar recurrences of the fused solver (one workgroup per
// instance).
//   start: rz_0; stop if not positive (:446-450) or if max_iter <= 0; else
//          as `it` without counting an iteration.
//   it:    one iteration done (iters++); converged if rz < tol^2 rz_0
//          (:478); max_iter; p.q <= 0 breakdown (:460); alpha = rz / p.q
//          (:462); beta from the rz expansion (see k_it).
//   rr:    a monitor: the true residual (FP64-evaluated) replaces r only
//          when its r.z already meets the tolerance, so the next iteration
//          stops on an honest residual; otherwise the recursion continues
//          with beta = (recursive r.z) / rz_j computed from the vectors
//          themselves.  Replacing an unconverged FP32 residual perturbs the
//          short recurrence: on nlooks-5 interferograms stopped by the
//          1000-iteration cap it left up to 5 % of pixels a cycle away from
//          the FP64 solve, whose iterates the unreplaced FP32 recursion
//          tracks to 99.98 %.
__global__ __launch_bounds__(256) void k_reduce_it(const InstDesc *descs, const double4 *partials, const SolveScal *scal_in,
                                                   SolveScal *scal, int mode, int max_iter, double tol2, int defer_x) {
    int inst = blockIdx.x;
    const InstDesc d = descs[inst];
    const SolveScal *sp_in = scal_in + inst;
    SolveScal *sp = scal + inst;
    if (mode != RED_START && !sp_in->active) {
        if (threadIdx.x == 0 && sp_in != sp) *sp = *sp_in;   // a frozen state moves to the output buffer
        return;
    }
    __shared__ double4 res;
    sum_partials4(partials + d.block_begin, d.tiles_x * d.tiles_y, &res);
    if (threadIdx.x != 0) return;
    SolveScal s = *sp_in;
    const double4 v = res;
    if (mode == RED_RR) {
        s.replace = 0;
        const double rt = v.x, rr = v.y;
        s.delta = rt / rr;   // true over recursive r.z, kept until the next reduction
        if (isfinite(rt) && rt < tol2 * s.gamma0) {
            s.replace = 1;
            s.beta = rt / s.gamma;
        } else if (isfinite(rr)) {
            s.beta = rr / s.gamma;
        }
        *sp = s;
        return;
    }
    if (mode == RED_START) {
        const double pq = v.x, rz = v.y;
        s.replace = 0;
        s.iters = 0; s.gamma0 = rz; s.gamma_prev = rz; s.relres = 1.0; s.active = 0;
        s.alpha = 0.0; s.alpha_prev = 0.0; s.beta = 0.0; s.delta = pq; s.gamma = rz;
        s.xmode = defer_x ? 0 : 2;
        if (!(rz > 0.0)) { s.status = (rz == 0.0) ? CUPHU_SOLVE_EMPTY : CUPHU_SOLVE_NONFINITE; s.relres = 0.0; *sp = s; return; }
        if (max_iter <= 0) { s.status = CUPHU_SOLVE_MAXITER; *sp = s; return; }
        fused_alpha_beta(s, v);
    } else {
        fused_it_scalars(s, v, max_iter, tol2);
    }
    *sp = s;
}

// ═══════════════════════════════════════════════════════════════════════════
// Per-instance reduction of the workgroup partial sums (fixed order) and
// the scalar recurrences of the reference loop (one workgroup per
// instance, cuphu_laplace.cu:446-499):
//   start: rz0 = r.z; stop if rz0 is not positive (:446-450)
//   pq:    stop on p.q <= 0 (:460); alpha = rz / p.q (:462)
//   rz:    iteration done; converged if rz' < tol^2 rz0 (:478);
//          beta = rz' / rz (:493)
//   rr:    true residual r_t.z_t (FP64-evaluated): converged if below
//          tol^2 rz0; otherwise the recursion continues unreplaced (a
//          monitor, as in the fused solver's reduction).
// ═══════════════════════════════════════════════════════════════════════════
__global__ __launch_bounds__(256) void k_reduce(const InstDesc *descs, const double2 *partials, SolveScal *scal,
                                                int mode, int max_iter, double tol2) {
    int inst = blockIdx.x;
    const InstDesc d = descs[inst];
    SolveScal *sp = scal + inst;
    if (mode != RED_START && !sp->active) return;
    int nb = d.tiles_x * d.tiles_y;
    double a = 0.0;
    const double2 *pp = partials + d.block_begin;
    int i = threadIdx.x;
    for (; i + 768 < nb; i += 1024) {                 // four loads before their sums (same order)
        const double v0 = pp[i].x, v1 = pp[i + 256].x, v2 = pp[i + 512].x, v3 = pp[i + 768].x;
        a += v0; a += v1; a += v2; a += v3;
    }
    for (; i < nb; i += 256) a += pp[i].x;
    __shared__ double2 res;
    block_sum2(a, 0.0, &res);
    __syncthreads();
    if (threadIdx.x != 0) return;
    SolveScal s = *sp;
    const double v = res.x;
    s.replace = 0;
    if (mode == RED_START) {
        s.iters = 0;
        s.gamma0 = v; s.gamma = v; s.gamma_prev = v; s.delta = 0.0; s.alpha = 0.0; s.beta = 0.0;
        s.relres = 1.0;
        s.active = 0;
        if (!(v > 0.0)) { s.status = (v == 0.0) ? CUPHU_SOLVE_EMPTY : CUPHU_SOLVE_NONFINITE; s.relres = 0.0; }
        else if (max_iter <= 0) s.status = CUPHU_SOLVE_MAXITER;
        else { s.active = 1; s.status = CUPHU_SOLVE_MAXITER; }
    } else if (mode == RED_PQ) {
        s.delta = v;
        if (!(v > 0.0)) { s.status = isfinite(v) ? CUPHU_SOLVE_BREAKDOWN : CUPHU_SOLVE_NONFINITE; s.active = 0; }
        else s.alpha = s.gamma / v;
    } else if (mode == RED_RZ) {
        s.iters += 1;
        s.relres = (s.gamma0 > 0.0) ? sqrt(fabs(v) / s.gamma0) : 0.0;
        if (v < tol2 * s.gamma0) { s.status = CUPHU_SOLVE_CONVERGED; s.active = 0; s.gamma = v; }
        else if (s.iters >= max_iter) { s.status = CUPHU_SOLVE_MAXITER; s.active = 0; s.gamma = v; }
        else { s.beta = v / s.gamma; s.gamma_prev = s.gamma; s.gamma = v; }
    } else {   // RED_RR
        if (!isfinite(v)) { s.status = CUPHU_SOLVE_NONFINITE; s.active = 0; }
        else if (v < tol2 * s.gamma0) {
            s.status = CUPHU_SOLVE_CONVERGED; s.active = 0; s.gamma = v;
            s.relres = (s.gamma0 > 0.0) ? sqrt(v / s.gamma0) : 0.0;
        }
    }
    *sp = s;
}

// ═══════════════════════════════════════════════════════════════════════════
// Anchor to phi[0,0] (:534-539) and congruence (laplace_congruence :256-266)
// into a pitched float output plane.
// ═══════════════════════════════════════════════════════════════════════════
template <class T>
__global__ __launch_bounds__(256) void k_final(const InstDesc *descs, const int *rowinst, Planes<T> pl, float *out,
                                               int nxb) {
    int grow = (int)(blockIdx.x / (unsigned)nxb);
    const InstDesc d = descs[rowinst[grow]];
    int row = grow - d.row_begin;
    int c0 = ((int)(blockIdx.x % (unsigned)nxb) * 256 + threadIdx.x) * 4;
    if (c0 >= d.pitch) return;
    long long e = d.off + (long long)row * d.pitch + c0;
    T off = (T)pl.phi[d.off] - pl.x[d.off];
    T xv[4];
    float ph[4], o[4];
    ld4(pl.x + e, xv);
    ld4f(pl.phi + e, ph);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        T v = (off != T(0)) ? xv[j] + off : xv[j];
        if constexpr (sizeof(T) == 4) {
            float dd = v - ph[j];
            o[j] = __builtin_fmaf(kTwoPiF, __builtin_rintf(dd * kInvTwoPiF), ph[j]);
        } else {
            double dd = v - (double)ph[j];
            o[j] = (float)__builtin_fma(kTwoPiD, __builtin_rint(dd / kTwoPiD), (double)ph[j]);
        }
    }
    *reinterpret_cast<float4 *>(out + e) = make_float4(o[0], o[1], o[2], o[3]);
}

// ═══════════════════════════════════════════════════════════════════════════
// Plain path: the reference's per-iteration launch sequence
// (cuphu_laplace.cu:455-484) on one instance, with FP64 partial sums.
// ═══════════════════════════════════════════════════════════════════════════
template <class T>
__global__ __launch_bounds__(256) void k_plain_matvec(InstDesc d, const T *w2, const T *x, T *y, int nxb) {
    int row = (int)(blockIdx.x / (unsigned)nxb);
    int lane = threadIdx.x & 63;
    int c = ((int)(blockIdx.x % (unsigned)nxb) * 256 + threadIdx.x) * 4;
    bool cok = c < d.pitch;
    long long e = d.off + (long long)row * d.pitch + c;
    T xa[4], xu[4], xd[4], wd[4], wr[4], wu[4];
    zero4(xa); zero4(xu); zero4(xd); zero4(wd); zero4(wr); zero4(wu);
    if (cok) {
        ld4(x + e, xa);
        ld_arcs(w2, e, wd, wr);
        if (row > 0) { T dummy[4]; ld4(x + e - d.pitch, xu); ld_arcs(w2, e - d.pitch, wu, dummy); }
        if (row + 1 < d.nrow) ld4(x + e + d.pitch, xd);
    }
    T xl = __shfl_up(xa[3], 1, 64), wl0 = __shfl_up(wr[3], 1, 64), xr = __shfl_down(xa[0], 1, 64);
    if (lane == 0) {
        bool ok = cok && c > 0;
        xl = ok ? x[e - 1] : T(0);
        wl0 = ok ? w2[2 * (e - 1) + 1] : T(0);
    }
    if (lane == 63) xr = (cok && c + 4 < d.pitch) ? x[e + 4] : T(0);
    T wl[4] = {wl0, wr[0], wr[1], wr[2]};
    T yv[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        T a = xa[j];
        T left = (j == 0) ? xl : xa[j - 1];
        T right = (j == 3) ? xr : xa[j + 1];
        T v = Prec<T>::eps * a;
        v = Prec<T>::fmaT(wd[j], a - xd[j], v);
        v = Prec<T>::fmaT(wu[j], a - xu[j], v);
        v = Prec<T>::fmaT(wr[j], a - right, v);
        v = Prec<T>::fmaT(wl[j], a - left, v);
        yv[j] = v;
    }
    if (cok) st4(y + e, yv);
}

template <class T> __global__ void k_plain_axpy(T *y, const T *x, T alpha, size_t n) {
    size_t i = ((size_t)blockIdx.x * 256 + threadIdx.x) * 4;
    if (i >= n) return;
    T a[4], b[4];
    ld4(x + i, a);
    ld4(y + i, b);
#pragma unroll
    for (int j = 0; j < 4; ++j) b[j] = Prec<T>::fmaT(alpha, a[j], b[j]);
    st4(y + i, b);
}
template <class T> __global__ void k_plain_vmul(T *z, const T *r, const T *dinv, size_t n) {
    size_t i = ((size_t)blockIdx.x * 256 + threadIdx.x) * 4;
    if (i >= n) return;
    T a[4], b[4];
    ld4(r + i, a);
    ld4(dinv + i, b);
#pragma unroll
    for (int j = 0; j < 4; ++j) a[j] = a[j] * b[j];
    st4(z + i, a);
}
template <class T> __global__ void k_plain_pz(T *p, T *z, const T *r, const T *dinv, size_t n) {
    size_t i = ((size_t)blockIdx.x * 256 + threadIdx.x) * 4;
    if (i >= n) return;
    T a[4], b[4];
    ld4(r + i, a);
    ld4(dinv + i, b);
#pragma unroll
    for (int j = 0; j < 4; ++j) a[j] = a[j] * b[j];
    st4(z + i, a);
    st4(p + i, a);
}
template <class T> __global__ void k_plain_axpby(T *p, const T *z, T beta, size_t n) {
    size_t i = ((size_t)blockIdx.x * 256 + threadIdx.x) * 4;
    if (i >= n) return;
    T a[4], b[4];
    ld4(z + i, a);
    ld4(p + i, b);
#pragma unroll
    for (int j = 0; j < 4; ++j) b[j] = Prec<T>::fmaT(beta, b[j], a[j]);
    st4(p + i, b);
}
template <class T> __global__ __launch_bounds__(256) void k_plain_dot_part(const T *a, const T *b, size_t n, double *parts) {
    double s = 0.0;
    for (size_t i = ((size_t)blockIdx.x * 256 + threadIdx.x) * 4; i < n; i += (size_t)gridDim.x * 1024) {
        T x[4], y[4];
        ld4(a + i, x);
        ld4(b + i, y);
#pragma unroll
        for (int j = 0; j < 4; ++j) s += (double)x[j] * (double)y[j];
    }
    __shared__ double2 res;
    block_sum2(s, 0.0, &res);
    if (threadIdx.x == 0) parts[blockIdx.x] = res.x;
}
__global__ __launch_bounds__(256) void k_plain_dot_final(const double *parts, int nparts, double *out) {
    double s = 0.0;
    for (int i = threadIdx.x; i < nparts; i += 256) s += parts[i];
    __shared__ double2 res;
    block_sum2(s, 0.0, &res);
    if (threadIdx.x == 0) *out = res.x;
}

// ── launchers ───────────────────────────────────────────────────────────────
#define CUPHU_LAUNCH_CHECK() CUPHU_HIP_CHECK(hipGetLastError())

template <class T>
void launch_ingest(const IngestSrc *d_src, const InstDesc *d_desc, const int *d_rowinst, int nrows_total,
                   const Planes<T> &pl, int has_u0, int max_pitch, hipStream_t st) {
    int nxb = (int)ceil_div((size_t)max_pitch, 1024);
    dim3 grid((unsigned)((size_t)nxb * (size_t)nrows_total));
    hipLaunchKernelGGL(k_ingest<T>, grid, dim3(256), 0, st, d_src, d_desc, d_rowinst, pl, has_u0, nxb);
    CUPHU_LAUNCH_CHECK();
}

#define CUPHU_RPW_DISPATCH(rpw, KERNEL, ...)                                      \
    switch (rpw) {                                                                \
    case 2: hipLaunchKernelGGL((KERNEL<T, 2>), __VA_ARGS__); break;               \
    case 8: hipLaunchKernelGGL((KERNEL<T, 8>), __VA_ARGS__); break;               \
    default: hipLaunchKernelGGL((KERNEL<T, 4>), __VA_ARGS__); break;              \
    }
// Tile shapes of the fused solver's kernels: (rpw, lpr) in {2, 4, 8} x {64}
// and {2, 4} x {32}; KERNEL<T, RPW, LPR, extra...>.
static void check_shape(TileShape sh) {
    const bool ok = (sh.lpr == 64 && (sh.rpw == 2 || sh.rpw == 4 || sh.rpw == 8)) ||
                    (sh.lpr == 32 && (sh.rpw == 2 || sh.rpw == 4));
    CUPHU_REQUIRE(ok, "unsupported tile shape");
}
#define CUPHU_SHAPE_DISPATCH(sh, KERNEL, ...)                                      \
    check_shape(sh);                                                               \
    if (sh.lpr == 32) {                                                            \
        if (sh.rpw == 2) hipLaunchKernelGGL((KERNEL<T, 2, 32>), __VA_ARGS__);      \
        else hipLaunchKernelGGL((KERNEL<T, 4, 32>), __VA_ARGS__);                  \
    } else {                                                                       \
        switch (sh.rpw) {                                                          \
        case 2: hipLaunchKernelGGL((KERNEL<T, 2, 64>), __VA_ARGS__); break;        \
        case 8: hipLaunchKernelGGL((KERNEL<T, 8, 64>), __VA_ARGS__); break;        \
        default: hipLaunchKernelGGL((KERNEL<T, 4, 64>), __VA_ARGS__); break;       \
        }                                                                          \
    }
#define CUPHU_SHAPE_DISPATCH_X(sh, KERNEL, X, ...)                                 \
    check_shape(sh);                                                               \
    if (sh.lpr == 32) {                                                            \
        if (sh.rpw == 2) hipLaunchKernelGGL((KERNEL<T, 2, 32, X>), __VA_ARGS__);   \
        else hipLaunchKernelGGL((KERNEL<T, 4, 32, X>), __VA_ARGS__);               \
    } else {                                                                       \
        switch (sh.rpw) {                                                          \
        case 2: hipLaunchKernelGGL((KERNEL<T, 2, 64, X>), __VA_ARGS__); break;     \
        case 8: hipLaunchKernelGGL((KERNEL<T, 8, 64, X>), __VA_ARGS__); break;     \
        default: hipLaunchKernelGGL((KERNEL<T, 4, 64, X>), __VA_ARGS__); break;    \
        }                                                                          \
    }

template <class T>
void launch_weights(const InstDesc *d_desc, const int *d_blk, int nblk, const Planes<T> &pl, int reweight,
                    T delta2, TileShape sh, hipStream_t st) {
    CUPHU_SHAPE_DISPATCH(sh, k_weights, dim3(nblk), dim3(256), 0, st, d_desc, d_blk, pl, reweight, delta2);
    CUPHU_LAUNCH_CHECK();
}
template <class T>
void launch_assemble(const InstDesc *d_desc, const int *d_blk, int nblk, const Planes<T> &pl, TileShape sh,
                     hipStream_t st) {
    CUPHU_SHAPE_DISPATCH(sh, k_assemble, dim3(nblk), dim3(256), 0, st, d_desc, d_blk, pl);
    CUPHU_LAUNCH_CHECK();
}
template <class T>
void launch_rz(const InstDesc *d_desc, const int *d_blk, int nblk, const Planes<T> &pl, const SolveScal *scal,
               int force, double2 *partials, int rpw, hipStream_t st) {
    CUPHU_RPW_DISPATCH(rpw, k_rz, dim3(nblk), dim3(256), 0, st, d_desc, d_blk, pl, scal, force, partials);
    CUPHU_LAUNCH_CHECK();
}
template <class T>
void launch_pq(const InstDesc *d_desc, const int *d_blk, int nblk, const Planes<T> &pl, const SolveScal *scal,
               double2 *partials, int rpw, int parity, hipStream_t st) {
    CUPHU_RPW_DISPATCH(rpw, k_pq, dim3(nblk), dim3(256), 0, st, d_desc, d_blk, pl, scal, partials, parity);
    CUPHU_LAUNCH_CHECK();
}
template <class T>
void launch_xr(const InstDesc *d_desc, const int *d_blk, int nblk, const Planes<T> &pl, const SolveScal *scal,
               double2 *partials, int rpw, int parity, hipStream_t st) {
    CUPHU_RPW_DISPATCH(rpw, k_xr, dim3(nblk), dim3(256), 0, st, d_desc, d_blk, pl, scal, partials, parity);
    CUPHU_LAUNCH_CHECK();
}
template <class T>
void launch_rr(const InstDesc *d_desc, const int *d_blk, int nblk, const Planes<T> &pl, const SolveScal *scal,
               double2 *partials, int rpw, hipStream_t st) {
    CUPHU_RPW_DISPATCH(rpw, k_rr, dim3(nblk), dim3(256), 0, st, d_desc, d_blk, pl, scal, partials);
    CUPHU_LAUNCH_CHECK();
}
template <class T>
void launch_rr_commit(const InstDesc *d_desc, const int *d_blk, int nblk, const Planes<T> &pl, const SolveScal *scal,
                      int rpw, hipStream_t st) {
    CUPHU_RPW_DISPATCH(rpw, k_rr_commit, dim3(nblk), dim3(256), 0, st, d_desc, d_blk, pl, scal);
    CUPHU_LAUNCH_CHECK();
}
template <class T>
void launch_it(const InstDesc *d_desc, const int *d_blk, int nblk, const Planes<T> &pl, const SolveScal *scal,
               double4 *partials, int rpw, int parity, int start, hipStream_t st) {
    CUPHU_RPW_DISPATCH(rpw, k_it, dim3(nblk), dim3(256), 0, st, d_desc, d_blk, pl, scal, partials, parity, start);
    CUPHU_LAUNCH_CHECK();
}
template <class T>
void launch_rr_it(const InstDesc *d_desc, const int *d_blk, int nblk, const Planes<T> &pl, const SolveScal *scal,
                  double4 *partials, int rpw, hipStream_t st) {
    CUPHU_RPW_DISPATCH(rpw, k_rr_it, dim3(nblk), dim3(256), 0, st, d_desc, d_blk, pl, scal, partials);
    CUPHU_LAUNCH_CHECK();
}
template <class T>
void launch_iq(const InstDesc *d_desc, const int *d_blk, int nblk, const Planes<T> &pl, const SolveScal *scal,
               double4 *partials, TileShape sh, int parity, int start, hipStream_t st) {
    if (nblk >= kStaggerMinGrid) {
        CUPHU_SHAPE_DISPATCH_X(sh, k_iq, true, dim3(nblk), dim3(256), 0, st, d_desc, d_blk, pl, scal, partials,
                               parity, start);
    } else {
        CUPHU_SHAPE_DISPATCH_X(sh, k_iq, false, dim3(nblk), dim3(256), 0, st, d_desc, d_blk, pl, scal, partials,
                               parity, start);
    }
    CUPHU_LAUNCH_CHECK();
}
template <class T>
void launch_iq_mix(const InstDesc *d_desc, const int4 *d_tiles, int ntiles, const Planes<T> &pl,
                   const SolveScal *scal, double4 *partials, int rpw, int parity, int start, hipStream_t st) {
    CUPHU_REQUIRE(rpw == 2 || rpw == 4, "the merged iteration launch runs 2 or 4 rows per lane group");
    if (ntiles >= kStaggerMinGrid) {
        if (rpw == 2) hipLaunchKernelGGL((k_iq_mix<T, 2, true>), dim3(ntiles), dim3(256), 0, st, d_desc, d_tiles, pl,
                                         scal, partials, parity, start);
        else hipLaunchKernelGGL((k_iq_mix<T, 4, true>), dim3(ntiles), dim3(256), 0, st, d_desc, d_tiles, pl, scal,
                                partials, parity, start);
    } else {
        if (rpw == 2) hipLaunchKernelGGL((k_iq_mix<T, 2, false>), dim3(ntiles), dim3(256), 0, st, d_desc, d_tiles,
                                         pl, scal, partials, parity, start);
        else hipLaunchKernelGGL((k_iq_mix<T, 4, false>), dim3(ntiles), dim3(256), 0, st, d_desc, d_tiles, pl, scal,
                                partials, parity, start);
    }
    CUPHU_LAUNCH_CHECK();
}
template <class T>
void launch_iq_fold_mix(const InstDesc *d_desc, const int4 *d_tiles, int ntiles, const Planes<T> &pl,
                        const SolveScal *scal_prev, SolveScal *scal_next, const double4 *part_prev,
                        double4 *partials, int rpw, int parity, int max_iter, double tol2, hipStream_t st) {
    CUPHU_REQUIRE(rpw == 2 || rpw == 4, "the merged iteration launch runs 2 or 4 rows per lane group");
    const FoldArgs fa{part_prev, scal_next, max_iter, tol2};
    if (rpw == 2) hipLaunchKernelGGL((k_iq_fold_mix<T, 2>), dim3(ntiles), dim3(256), 0, st, d_desc, d_tiles, pl,
                                     scal_prev, partials, parity, fa);
    else hipLaunchKernelGGL((k_iq_fold_mix<T, 4>), dim3(ntiles), dim3(256), 0, st, d_desc, d_tiles, pl, scal_prev,
                            partials, parity, fa);
    CUPHU_LAUNCH_CHECK();
}
template <class T>
void launch_iq_fold(const InstDesc *d_desc, const int *d_blk, int nblk, const Planes<T> &pl,
                    const SolveScal *scal_prev, SolveScal *scal_next, const double4 *part_prev, double4 *partials,
                    TileShape sh, int parity, int max_iter, double tol2, hipStream_t st) {
    const FoldArgs fa{part_prev, scal_next, max_iter, tol2};
    CUPHU_SHAPE_DISPATCH(sh, k_iq_fold, dim3(nblk), dim3(256), 0, st, d_desc, d_blk, pl, scal_prev, partials, parity,
                         fa);
    CUPHU_LAUNCH_CHECK();
}
template <class T>
void launch_rr_iq(const InstDesc *d_desc, const int *d_blk, int nblk, const Planes<T> &pl, const SolveScal *scal,
                  double4 *partials, TileShape sh, hipStream_t st) {
    CUPHU_SHAPE_DISPATCH(sh, k_rr_iq, dim3(nblk), dim3(256), 0, st, d_desc, d_blk, pl, scal, partials);
    CUPHU_LAUNCH_CHECK();
}
template <class T>
void launch_x_finalize(const InstDesc *d_desc, const int *d_blk, int nblk, const Planes<T> &pl, const SolveScal *scal,
                       TileShape sh, hipStream_t st) {
    CUPHU_SHAPE_DISPATCH(sh, k_x_finalize, dim3(nblk), dim3(256), 0, st, d_desc, d_blk, pl, scal);
    CUPHU_LAUNCH_CHECK();
}
void launch_reduce_it(const InstDesc *d_desc, int ninst, const double4 *partials, SolveScal *scal, int mode,
                      int max_iter, double tol2, hipStream_t st, int defer_x, const SolveScal *scal_in) {
    hipLaunchKernelGGL(k_reduce_it, dim3(ninst), dim3(256), 0, st, d_desc, partials, scal_in ? scal_in : scal, scal,
                       mode, max_iter, tol2, defer_x);
    CUPHU_LAUNCH_CHECK();
}
void launch_reduce(const InstDesc *d_desc, int ninst, const double2 *partials, SolveScal *scal, int mode,
                   int max_iter, double tol2, hipStream_t st) {
    hipLaunchKernelGGL(k_reduce, dim3(ninst), dim3(256), 0, st, d_desc, partials, scal, mode, max_iter, tol2);
    CUPHU_LAUNCH_CHECK();
}
template <class T>
void launch_final(const InstDesc *d_desc, const int *d_rowinst, int nrows_total, const Planes<T> &pl, float *out,
                  int max_pitch, hipStream_t st) {
    int nxb = (int)ceil_div((size_t)max_pitch, 1024);
    dim3 grid((unsigned)((size_t)nxb * (size_t)nrows_total));
    hipLaunchKernelGGL(k_final<T>, grid, dim3(256), 0, st, d_desc, d_rowinst, pl, out, nxb);
    CUPHU_LAUNCH_CHECK();
}

template <class T> void launch_plain_matvec(const InstDesc &d, const T *w2, const T *x, T *y, hipStream_t st) {
    int nxb = (int)ceil_div((size_t)d.pitch, 1024);
    dim3 grid((unsigned)((size_t)nxb * (size_t)d.nrow));
    hipLaunchKernelGGL(k_plain_matvec<T>, grid, dim3(256), 0, st, d, w2, x, y, nxb);
    CUPHU_LAUNCH_CHECK();
}
static inline unsigned grid1(size_t n) { return (unsigned)ceil_div(n, 1024); }
template <class T> void launch_plain_axpy(T *y, const T *x, T alpha, size_t n, hipStream_t st) {
    hipLaunchKernelGGL(k_plain_axpy<T>, dim3(grid1(n)), dim3(256), 0, st, y, x, alpha, n);
    CUPHU_LAUNCH_CHECK();
}
template <class T> void launch_plain_vmul(T *z, const T *r, const T *dinv, size_t n, hipStream_t st) {
    hipLaunchKernelGGL(k_plain_vmul<T>, dim3(grid1(n)), dim3(256), 0, st, z, r, dinv, n);
    CUPHU_LAUNCH_CHECK();
}
template <class T> void launch_plain_pz(T *p, T *z, const T *r, const T *dinv, size_t n, hipStream_t st) {
    hipLaunchKernelGGL(k_plain_pz<T>, dim3(grid1(n)), dim3(256), 0, st, p, z, r, dinv, n);
    CUPHU_LAUNCH_CHECK();
}
template <class T> void launch_plain_axpby(T *p, const T *z, T beta, size_t n, hipStream_t st) {
    hipLaunchKernelGGL(k_plain_axpby<T>, dim3(grid1(n)), dim3(256), 0, st, p, z, beta, n);
    CUPHU_LAUNCH_CHECK();
}
template <class T>
void launch_plain_dot(const T *a, const T *b, size_t n, double *partials, double *out, hipStream_t st) {
    hipLaunchKernelGGL(k_plain_dot_part<T>, dim3(kDotParts), dim3(256), 0, st, a, b, n, partials);
    CUPHU_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_plain_dot_final, dim3(1), dim3(256), 0, st, partials, kDotParts, out);
    CUPHU_LAUNCH_CHECK();
}

#define CUPHU_INSTANTIATE(T)                                                                                    \
    template void launch_ingest<T>(const IngestSrc *, const InstDesc *, const int *, int, const Planes<T> &, int, \
                                   int, hipStream_t);                                                            \
    template void launch_weights<T>(const InstDesc *, const int *, int, const Planes<T> &, int, T, TileShape,    \
                                    hipStream_t);                                                                \
    template void launch_assemble<T>(const InstDesc *, const int *, int, const Planes<T> &, TileShape,           \
                                     hipStream_t);                                                               \
    template void launch_rz<T>(const InstDesc *, const int *, int, const Planes<T> &, const SolveScal *, int,    \
                               double2 *, int, hipStream_t);                                                     \
    template void launch_pq<T>(const InstDesc *, const int *, int, const Planes<T> &, const SolveScal *,         \
                               double2 *, int, int, hipStream_t);                                                \
    template void launch_xr<T>(const InstDesc *, const int *, int, const Planes<T> &, const SolveScal *,         \
                               double2 *, int, int, hipStream_t);                                                \
    template void launch_rr<T>(const InstDesc *, const int *, int, const Planes<T> &, const SolveScal *,         \
                               double2 *, int, hipStream_t);                                                     \
    template void launch_rr_commit<T>(const InstDesc *, const int *, int, const Planes<T> &, const SolveScal *,  \
                                      int, hipStream_t);                                                         \
    template void launch_it<T>(const InstDesc *, const int *, int, const Planes<T> &, const SolveScal *,         \
                               double4 *, int, int, int, hipStream_t);                                           \
    template void launch_rr_it<T>(const InstDesc *, const int *, int, const Planes<T> &, const SolveScal *,      \
                                  double4 *, int, hipStream_t);                                                  \
    template void launch_iq<T>(const InstDesc *, const int *, int, const Planes<T> &, const SolveScal *,         \
                               double4 *, TileShape, int, int, hipStream_t);                                     \
    template void launch_iq_mix<T>(const InstDesc *, const int4 *, int, const Planes<T> &, const SolveScal *,    \
                                   double4 *, int, int, int, hipStream_t);                                       \
    template void launch_iq_fold_mix<T>(const InstDesc *, const int4 *, int, const Planes<T> &,                \
                                        const SolveScal *, SolveScal *, const double4 *, double4 *, int, int,   \
                                        int, double, hipStream_t);                                               \
    template void launch_iq_fold<T>(const InstDesc *, const int *, int, const Planes<T> &, const SolveScal *,    \
                                    SolveScal *, const double4 *, double4 *, TileShape, int, int, double,        \
                                    hipStream_t);                                                                \
    template void launch_rr_iq<T>(const InstDesc *, const int *, int, const Planes<T> &, const SolveScal *,      \
                                  double4 *, TileShape, hipStream_t);                                            \
    template void launch_x_finalize<T>(const InstDesc *, const int *, int, const Planes<T> &, const SolveScal *,  \
                                       TileShape, hipStream_t);                                                  \
    template void launch_final<T>(const InstDesc *, const int *, int, const Planes<T> &, float *, int,           \
                                  hipStream_t);                                                                  \
    template void launch_plain_matvec<T>(const InstDesc &, const T *, const T *, T *, hipStream_t);              \
    template void launch_plain_axpy<T>(T *, const T *, T, size_t, hipStream_t);                                  \
    template void launch_plain_vmul<T>(T *, const T *, const T *, size_t, hipStream_t);                          \
    template void launch_plain_pz<T>(T *, T *, const T *, const T *, size_t, hipStream_t);                       \
    template void launch_plain_axpby<T>(T *, const T *, T, size_t, hipStream_t);                                \
    template void launch_plain_dot<T>(const T *, const T *, size_t, double *, double *, hipStream_t);

CUPHU_INSTANTIATE(float)
CUPHU_INSTANTIATE(double)

}  // namespace cuphu
