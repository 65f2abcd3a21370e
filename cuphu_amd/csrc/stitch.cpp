// stitch.cpp -- tile geometry and tile stitching on the host
// (reference: src/cxx/cuphu/cuphu_solver.cpp:1755-1791 and :1846-2031).
//
// The weighted median is computed by 3-way selection instead of a full
// sort.  It returns the same value as the reference's sort-and-scan
// (:1867-1881): the smallest overlap difference v whose cumulative weight
// W(<= v) reaches half of the total.  The weights are floats >= 0.05 summed
// in double, so every partial sum is exact and the result does not depend
// on summation order.  Pair registrations are computed concurrently.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <future>
#include <queue>
#include <thread>
#include <vector>

#include "stitch.hpp"

namespace cuphu {

int tile_axis(int n, int tile_size, int overlap, int ntiles, int *first, int *size, int max_tiles) {
    int k = 0;
    if (n <= 0 || overlap < 0) return -1;
    if (tile_size <= 0) {
        if (ntiles < 1) return -1;
        for (int t = 0; t < ntiles; ++t) {
            int f = (int)((long)t * (n - overlap) / ntiles);
            int l = (t == ntiles - 1) ? n : (int)((long)(t + 1) * (n - overlap) / ntiles) + overlap;
            if (k < max_tiles) { first[k] = f; size[k] = l - f; }
            k++;
        }
        return k;
    }
    int step = tile_size - overlap;
    if (step <= 0) return -1;
    for (int s = 0;; s += step) {
        int e = std::min(s + tile_size, n);
        if (k < max_tiles) { first[k] = s; size[k] = e - s; }
        k++;
        if (e >= n) break;
    }
    return k;
}

struct Sample {
    float v, w;
};

double weighted_median(std::vector<Sample> &d) {
    if (d.empty()) return 0.0;
    double total = 0.0;
    for (const Sample &s : d) total += s.w;
    const double half = total / 2.0;
    size_t lo = 0, hi = d.size();
    double before = 0.0;
    while (lo < hi) {
        float a = d[lo].v, b = d[(lo + hi) / 2].v, c = d[hi - 1].v;
        float piv = std::max(std::min(a, b), std::min(std::max(a, b), c));
        // 3-way partition [lo,hi) into < piv | == piv | > piv
        size_t lt = lo, i = lo, gt = hi;
        while (i < gt) {
            if (d[i].v < piv) std::swap(d[lt++], d[i++]);
            else if (d[i].v > piv) std::swap(d[i], d[--gt]);
            else ++i;
        }
        double wl = 0.0, we = 0.0;
        for (size_t j = lo; j < lt; ++j) wl += d[j].w;
        for (size_t j = lt; j < gt; ++j) we += d[j].w;
        if (before + wl >= half) { hi = lt; continue; }
        if (before + wl + we >= half) return (double)piv;
        before += wl + we;
        lo = gt;
    }
    // Only reachable if rounding kept the running sum below half.
    float mx = d[0].v;
    for (const Sample &s : d) mx = std::max(mx, s.v);
    return (double)mx;
}

// Registration of one tile pair: weighted median of A - B over the overlap
// (horizontal: A's rightmost col_ovrlp columns vs B's leftmost; vertical:
// A's bottom row_ovrlp rows vs B's top), in cycles.
static double pair_cycles(const TileView &A, const TileView &B, bool horizontal, int rovl, int covl, bool laplace) {
    std::vector<Sample> d;
    auto take = [&](int ar, int ac, int br, int bc) {
        size_t ia = (size_t)ar * A.stride + ac, ib = (size_t)br * B.stride + bc;
        if (A.corr[ia] < 0.05f || B.corr[ib] < 0.05f) return;
        if (A.conf && !A.conf[(size_t)ar * A.conf_stride + ac]) return;
        if (B.conf && !B.conf[(size_t)br * B.conf_stride + bc]) return;
        d.push_back({A.unw[ia] - B.unw[ib], std::min(A.corr[ia], B.corr[ib])});
    };
    if (horizontal) {
        if (covl <= 0 || covl > A.ncol || covl > B.ncol) return 0.0;
        int nr = std::min(A.nrow, B.nrow);
        d.reserve((size_t)nr * covl);
        for (int r = 0; r < nr; ++r)
            for (int c = 0; c < covl; ++c) take(r, A.ncol - covl + c, r, c);
    } else {
        if (rovl <= 0 || rovl > A.nrow || rovl > B.nrow) return 0.0;
        int nc = std::min(A.ncol, B.ncol);
        d.reserve((size_t)rovl * nc);
        for (int r = 0; r < rovl; ++r)
            for (int c = 0; c < nc; ++c) take(A.nrow - rovl + r, c, r, c);
    }
    double cyc = weighted_median(d) / (2.0 * M_PI);
    return laplace ? cyc : std::round(cyc);
}

std::vector<double> stitch_offsets(int ntr, int ntc, const std::vector<TileView> &tiles, int rovl, int covl,
                                   bool laplace) {
    int nt = ntr * ntc;
    std::vector<double> tk((size_t)nt, 0.0);
    if (nt <= 1 || (rovl <= 0 && covl <= 0)) return tk;
    // All candidate pair registrations, concurrently.
    std::vector<double> hk((size_t)nt, 0.0), vk((size_t)nt, 0.0);
    std::vector<std::pair<int, bool>> jobs;
    for (int t = 0; t < nt; ++t) {
        int br = t / ntc, bc = t % ntc;
        if (bc + 1 < ntc) jobs.push_back({t, true});
        if (br + 1 < ntr) jobs.push_back({t, false});
    }
    unsigned nth = std::max(1u, std::min<unsigned>(std::thread::hardware_concurrency(), 16u));
    std::vector<std::thread> pool;
    std::atomic<size_t> next(0);
    for (unsigned w = 0; w < nth; ++w)
        pool.emplace_back([&]() {
            for (size_t j = next.fetch_add(1); j < jobs.size(); j = next.fetch_add(1)) {
                int t = jobs[j].first;
                bool h = jobs[j].second;
                int nb = h ? t + 1 : t + ntc;
                double k = pair_cycles(tiles[(size_t)t], tiles[(size_t)nb], h, rovl, covl, laplace);
                (h ? hk : vk)[(size_t)t] = k;
            }
        });
    for (auto &th : pool) th.join();
    return bfs_offsets(ntr, ntc, hk, vk);
}

std::vector<double> bfs_offsets(int ntr, int ntc, const std::vector<double> &hk, const std::vector<double> &vk) {
    int nt = ntr * ntc;
    std::vector<double> tk((size_t)nt, 0.0);
    // BFS from tile 0: right neighbour first, then bottom (:1969-1999).
    std::vector<char> vis((size_t)nt, 0);
    std::queue<int> q;
    q.push(0);
    vis[0] = 1;
    while (!q.empty()) {
        int idx = q.front();
        q.pop();
        int br = idx / ntc, bc = idx % ntc;
        if (bc + 1 < ntc) {
            int nb = idx + 1;
            if (!vis[(size_t)nb]) { tk[(size_t)nb] = tk[(size_t)idx] + hk[(size_t)idx]; vis[(size_t)nb] = 1; q.push(nb); }
        }
        if (br + 1 < ntr) {
            int nb = idx + ntc;
            if (!vis[(size_t)nb]) { tk[(size_t)nb] = tk[(size_t)idx] + vk[(size_t)idx]; vis[(size_t)nb] = 1; q.push(nb); }
        }
    }
    return tk;
}

void trim_range(int t, int nt, int size, int ovl, int *a, int *b) {
    *a = (t == 0) ? 0 : ovl / 2;
    *b = (t == nt - 1) ? size : size - ovl / 2;
}

}  // namespace cuphu

// ── C ABI: host-only stitch (merge :2003-2031) ───────────────────────────────
extern "C" int cuphu_tile_axis(int n, int tile_size, int overlap, int ntiles, int *first, int *size, int max_tiles) {
    if (!first || !size) return CUPHU_ERR_INVALID;
    int k = cuphu::tile_axis(n, tile_size, overlap, ntiles, first, size, max_tiles);
    return k < 0 ? CUPHU_ERR_INVALID : k;
}

extern "C" int cuphu_stitch_tiles_host(int nrow, int ncol, int ntr, int ntc, const int *row_first, const int *row_size,
                                       const int *col_first, const int *col_size, int row_ovrlp, int col_ovrlp,
                                       int laplace, const float *const *tile_unw, const uint32_t *const *tile_cc,
                                       const float *const *tile_corr, const unsigned char *const *tile_conf,
                                       float *unw_out, uint32_t *cc_out, double *tile_k_out) {
    if (nrow <= 0 || ncol <= 0 || ntr <= 0 || ntc <= 0 || !row_first || !row_size || !col_first || !col_size ||
        !tile_unw || !tile_cc || !tile_corr || !unw_out || !cc_out)
        return CUPHU_ERR_INVALID;
    int nt = ntr * ntc;
    std::vector<cuphu::TileView> tv((size_t)nt);
    for (int t = 0; t < nt; ++t) {
        cuphu::TileView &v = tv[(size_t)t];
        v.nrow = row_size[t / ntc];
        v.ncol = col_size[t % ntc];
        v.stride = v.ncol;
        v.conf_stride = v.ncol;
        v.unw = tile_unw[t];
        v.corr = tile_corr[t];
        v.conf = tile_conf ? tile_conf[t] : nullptr;
        if (!v.unw || !v.corr || !tile_cc[t]) return CUPHU_ERR_INVALID;
    }
    std::vector<double> tk = cuphu::stitch_offsets(ntr, ntc, tv, row_ovrlp, col_ovrlp, laplace != 0);
    std::memset(unw_out, 0, (size_t)nrow * ncol * sizeof(float));
    std::memset(cc_out, 0, (size_t)nrow * ncol * sizeof(uint32_t));
    uint32_t off = 0;
    for (int tr = 0; tr < ntr; ++tr)
        for (int tc = 0; tc < ntc; ++tc) {
            int t = tr * ntc + tc;
            const cuphu::TileView &v = tv[(size_t)t];
            uint32_t lmax = 0;
            for (size_t i = 0; i < (size_t)v.nrow * v.ncol; ++i) lmax = std::max(lmax, tile_cc[t][i]);
            float add = (float)(tk[(size_t)t] * (2.0 * M_PI));
            int r0, r1, c0, c1;
            cuphu::trim_range(tr, ntr, v.nrow, row_ovrlp, &r0, &r1);
            cuphu::trim_range(tc, ntc, v.ncol, col_ovrlp, &c0, &c1);
            for (int r = r0; r < r1; ++r)
                for (int c = c0; c < c1; ++c) {
                    size_t ti = (size_t)r * v.ncol + c, gi = (size_t)(row_first[tr] + r) * ncol + (col_first[tc] + c);
                    unw_out[gi] = v.unw[ti] + add;
                    cc_out[gi] = (tile_cc[t][ti] == 0) ? 0u : tile_cc[t][ti] + off;
                }
            off += lmax;
        }
    if (tile_k_out) std::memcpy(tile_k_out, tk.data(), (size_t)nt * sizeof(double));
    return 0;
}
