// Host check of the SpMV tile planner (cuopt_amd/csrc/tile_plan.cpp), built by
// tests/test_tile_plan.py: the plan made from a sample of column indices at the
// chunk midpoints equals the plan made from the whole index array, the tiles
// cover every nonzero once, and a sample lacking a midpoint is refused.
#include <algorithm>
#include <cstdio>
#include <random>
#include <stdexcept>
#include <vector>

#include "tile_plan.hpp"

using namespace cuopt_amd;

static int fails = 0;
#define CHECK(c)                                                  \
  do {                                                            \
    if (!(c)) {                                                   \
      std::printf("FAIL line %d: %s\n", __LINE__, #c);            \
      ++fails;                                                    \
    }                                                             \
  } while (0)

static bool same(const TilePlan& a, const TilePlan& b)
{
  if (a.tiles.size() != b.tiles.size() || a.meta != b.meta || a.longrows.size() != b.longrows.size()) return false;
  for (size_t t = 0; t < a.tiles.size(); ++t)
    if (a.tiles[t].r0 != b.tiles[t].r0 || a.tiles[t].r1 != b.tiles[t].r1 || a.tiles[t].z0 != b.tiles[t].z0 ||
        a.tiles[t].z1 != b.tiles[t].z1)
      return false;
  for (size_t l = 0; l < a.longrows.size(); ++l)
    if (a.longrows[l].row != b.longrows[l].row || a.longrows[l].first != b.longrows[l].first ||
        a.longrows[l].nchunks != b.longrows[l].nchunks)
      return false;
  return true;
}

int main()
{
  const int ncols = 100000;
  std::mt19937 gen(7);
  // Short rows of 0..40 nonzeros, and long rows whose lengths sit on and around
  // the chunk sizes (multiples, one over, one under).
  std::vector<int> lens;
  for (int r = 0; r < 3000; ++r) lens.push_back((int)(gen() % 41));
  for (int L : {509, 1021, 2045, 1020 * 3, 1020 * 3 + 1, 2044 * 2 - 1, 50000, 12345})
    lens.insert(lens.begin() + (gen() % lens.size()), L);
  std::vector<int> ptr{0}, idx;
  for (int L : lens) {
    std::vector<int> row;
    std::uniform_int_distribution<int> col(0, ncols - 1);
    while ((int)row.size() < L) row.push_back(col(gen));
    std::sort(row.begin(), row.end());
    idx.insert(idx.end(), row.begin(), row.end());
    ptr.push_back((int)idx.size());
  }
  const int nrows = (int)lens.size();
  const int sizes[3] = {508, 1020, 2044};
  std::vector<int> pos;
  for (int t : sizes) chunk_midpoints(ptr.data(), nrows, t, pos);
  std::sort(pos.begin(), pos.end());
  pos.erase(std::unique(pos.begin(), pos.end()), pos.end());
  ColSample smp;
  smp.pos = pos;
  for (int p : pos) smp.col.push_back(idx[p]);
  CHECK(!pos.empty());
  for (int t : sizes) {
    const TilePlan full = plan_tiles(ptr.data(), nrows, t, 1024, 256, idx.data(), ncols);
    const TilePlan sampled = plan_tiles(ptr.data(), nrows, t, 1024, 256, smp, ncols);
    const TilePlan plain = plan_tiles(ptr.data(), nrows, t, 1024, 256);
    CHECK(same(full, sampled));
    CHECK(!same(full, plain) || full.longrows.empty());  // the XCD-aware order moved chunks
    // every nonzero in exactly one tile
    std::vector<int> cover(idx.size(), 0);
    for (const Tile& tl : full.tiles)
      for (int k = tl.z0; k < tl.z1; ++k) ++cover[k];
    CHECK(std::all_of(cover.begin(), cover.end(), [](int c) { return c == 1; }));
    int chunks = 0;
    for (int m : full.meta) chunks += m < 0;
    int want = 0;
    for (int L : lens)
      if (L > t) want += (L + t - 1) / t;
    CHECK(chunks == want);
  }
  // drop one position that the 1 020-entry plan reads
  std::vector<int> need;
  chunk_midpoints(ptr.data(), nrows, 1020, need);
  ColSample missing = smp;
  const size_t at = std::lower_bound(missing.pos.begin(), missing.pos.end(), need[need.size() / 2]) - missing.pos.begin();
  CHECK(at < missing.pos.size() && missing.pos[at] == need[need.size() / 2]);
  missing.pos.erase(missing.pos.begin() + at);
  missing.col.erase(missing.col.begin() + at);
  bool threw = false;
  try {
    (void)plan_tiles(ptr.data(), nrows, 1020, 1024, 256, missing, ncols);
  } catch (const std::logic_error&) {
    threw = true;
  }
  CHECK(threw);
  if (fails) return 1;
  std::printf("OK\n");
  return 0;
}
