// stitch.hpp -- host tile geometry and stitching.
#pragma once
#include <atomic>
#include <cstdint>
#include <vector>

#include "../../include/cuphu.h"

namespace cuphu {

// A solved tile as seen by the stitcher (host memory, row stride in elements).
struct TileView {
    int nrow = 0, ncol = 0, stride = 0, conf_stride = 0;
    const float *unw = nullptr;
    const float *corr = nullptr;
    const unsigned char *conf = nullptr;   // nullable
};

int tile_axis(int n, int tile_size, int overlap, int ntiles, int *first, int *size, int max_tiles);
// Per-tile cycle offsets (fractional when laplace) from the overlap
// registrations and the BFS over the tile grid.
std::vector<double> stitch_offsets(int ntr, int ntc, const std::vector<TileView> &tiles, int rovl, int covl,
                                   bool laplace);
// Tile offsets from the pair registrations (hk[t]: tile t -> t+1,
// vk[t]: t -> t+ntc) by the reference's BFS from tile 0, right neighbour
// first, then bottom (cuphu_solver.cpp:1969-1999).
std::vector<double> bfs_offsets(int ntr, int ntc, const std::vector<double> &hk, const std::vector<double> &vk);
// Rows/cols [a, b) of a tile kept by the merge (interior sides trimmed by ovl/2).
void trim_range(int t, int nt, int size, int ovl, int *a, int *b);

}  // namespace cuphu
