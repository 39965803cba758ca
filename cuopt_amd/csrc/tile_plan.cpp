// tile_plan.cpp -- see tile_plan.hpp.
#include "tile_plan.hpp"

#include <algorithm>
#include <stdexcept>

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

// The planner; col_at(k) is the column of nonzero k (read only at chunk
// midpoints, and only when aware).
template <class ColAt>
static TilePlan plan_impl(const int* rowptr, int nrows, int tile_nnz, int max_rows, int block, bool aware, ColAt col_at,
                   int ncols, int xcds)
{
  TilePlan p;
  std::vector<int> chunk_xcd;  // per chunk tile: XCD owning its columns (-1 for row-run tiles)
  int r     = 0;
  int slots = 0;
  while (r < nrows) {
    const int len = rowptr[r + 1] - rowptr[r];
    if (len > tile_nnz) {
      const int nch = (len + tile_nnz - 1) / tile_nnz;
      const int lid = (int)p.longrows.size();
      p.longrows.push_back(LongRow{r, slots, nch, 0});
      for (int c = 0; c < nch; ++c) {
        const int z0 = rowptr[r] + c * tile_nnz;
        const int z1 = std::min(z0 + tile_nnz, rowptr[r + 1]);
        p.tiles.push_back(Tile{r, slots + c, z0, z1});
        p.meta.push_back(-(lid + 1));
        int x = -1;
        if (aware) {
          const long long mid = col_at((z0 + z1 - 1) / 2);
          x                   = (int)std::min<long long>(xcds - 1, mid * xcds / ncols);
        }
        chunk_xcd.push_back(x);
      }
      slots += nch;
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
    chunk_xcd.push_back(-1);
  }
  if (!aware || p.longrows.empty()) return p;
  // Interleave: position t prefers a chunk owned by XCD t % xcds.
  std::vector<std::vector<int>> queue(xcds);
  std::vector<int> free_tiles;
  for (int t = 0; t < (int)p.tiles.size(); ++t) {
    if (chunk_xcd[t] >= 0) queue[chunk_xcd[t]].push_back(t);
    else free_tiles.push_back(t);
  }
  std::vector<size_t> head(xcds, 0);
  size_t fh = 0;
  std::vector<int> order;
  order.reserve(p.tiles.size());
  for (size_t t = 0; order.size() < p.tiles.size(); ++t) {
    const int x = (int)(t % xcds);
    if (head[x] < queue[x].size()) {
      order.push_back(queue[x][head[x]++]);
    } else if (fh < free_tiles.size()) {
      order.push_back(free_tiles[fh++]);
    } else {  // only other XCDs' chunks are left
      for (int y = 1; y < xcds; ++y) {
        const int z = (x + y) % xcds;
        if (head[z] < queue[z].size()) {
          order.push_back(queue[z][head[z]++]);
          break;
        }
      }
    }
  }
  TilePlan q;
  q.longrows = std::move(p.longrows);
  q.tiles.reserve(order.size());
  q.meta.reserve(order.size());
  for (int t : order) {
    q.tiles.push_back(p.tiles[t]);
    q.meta.push_back(p.meta[t]);
  }
  return q;
}

TilePlan plan_tiles(const int* rowptr, int nrows, int tile_nnz, int max_rows, int block, const int* colidx, int ncols,
                    int xcds)
{
  return plan_impl(rowptr, nrows, tile_nnz, max_rows, block, colidx && ncols > 0 && xcds > 1,
                   [colidx](int k) { return colidx[k]; }, ncols, xcds);
}

void chunk_midpoints(const int* rowptr, int nrows, int tile_nnz, std::vector<int>& out)
{
  for (int r = 0; r < nrows; ++r) {
    const int b = rowptr[r], e = rowptr[r + 1];
    if (e - b <= tile_nnz) continue;
    for (int z0 = b; z0 < e; z0 += tile_nnz) out.push_back((z0 + std::min(z0 + tile_nnz, e) - 1) / 2);
  }
}

TilePlan plan_tiles(const int* rowptr, int nrows, int tile_nnz, int max_rows, int block, const ColSample& cols,
                    int ncols, int xcds)
{
  auto col_at = [&cols](int k) {
    const auto it = std::lower_bound(cols.pos.begin(), cols.pos.end(), k);
    if (it == cols.pos.end() || *it != k) throw std::logic_error("tile plan: column sample lacks a chunk midpoint");
    return cols.col[it - cols.pos.begin()];
  };
  return plan_impl(rowptr, nrows, tile_nnz, max_rows, block, ncols > 0 && xcds > 1 && !cols.pos.empty(), col_at,
                   ncols, xcds);
}

}  // namespace cuopt_amd
