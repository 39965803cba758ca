// tile_plan.cpp -- see tile_plan.hpp.
#include "tile_plan.hpp"

#include <algorithm>

namespace cuopt_amd {

namespace {

// Cost model of a row-run tile processed by `block` lanes in groups of G:
// passes over the rows times the per-row work (strided LDS reads plus the
// log2(G) shuffle stages of the group reduction).
int pick_lanes_log2(int rows, int maxlen, int block)
{
  long best_cost = -1;
  int best       = 0;
  for (int lg = 0; lg <= 6; ++lg) {
    const int G       = 1 << lg;
    const long passes = ((long)rows * G + block - 1) / block;
    const long cost   = passes * ((maxlen + G - 1) / G + lg);
    if (best_cost < 0 || cost < best_cost) {
      best_cost = cost;
      best      = lg;
    }
  }
  return best;
}

}  // namespace

TilePlan plan_tiles(const int* rowptr, int nrows, int tile_nnz, int max_rows, int block)
{
  TilePlan p;
  int r = 0;
  while (r < nrows) {
    const int len = rowptr[r + 1] - rowptr[r];
    if (len > tile_nnz) {
      const int nch = (len + tile_nnz - 1) / tile_nnz;
      const int lid = (int)p.longrows.size();
      p.longrows.push_back(LongRow{r, (int)p.tiles.size(), nch, 0});
      for (int c = 0; c < nch; ++c) {
        const int z0 = rowptr[r] + c * tile_nnz;
        const int z1 = std::min(z0 + tile_nnz, rowptr[r + 1]);
        p.tiles.push_back(Tile{r, r + 1, z0, z1});
        p.meta.push_back(-(lid + 1));
      }
      ++r;
      continue;
    }
    const int r0 = r, z0 = rowptr[r];
    int maxlen   = 0;
    while (r < nrows && r - r0 < max_rows) {
      const int l = rowptr[r + 1] - rowptr[r];
      if (l > tile_nnz || rowptr[r + 1] - z0 > tile_nnz) break;
      maxlen = std::max(maxlen, l);
      ++r;
    }
    p.tiles.push_back(Tile{r0, r, z0, rowptr[r]});
    p.meta.push_back(pick_lanes_log2(r - r0, maxlen, block));
  }
  return p;
}

}  // namespace cuopt_amd
