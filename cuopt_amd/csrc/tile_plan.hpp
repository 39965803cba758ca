// tile_plan.hpp -- host-side partition of a CSR matrix into SpMV work tiles.
//
// A tile is one workgroup's share of a sparse matrix-vector product: a run of
// consecutive rows whose nonzeros fit in the workgroup's LDS product buffer
// (kTileNnz entries), or one fixed-size chunk of a row too long for that.
// Tiles are balanced by nonzeros, so the extreme row skew of Zipf-distributed
// constraint matrices (a single row holding ~10 % of all nonzeros next to
// hundreds of thousands of empty rows) becomes an even grid of workgroups.
//
// Per tile the planner also picks how many lanes cooperate on one row
// (1..64, a power of two) from the tile's row count and longest row, so that
// short rows use one lane each and long rows use whole wavefronts.
#pragma once

#include <cstdint>
#include <vector>

namespace cuopt_amd {

struct alignas(16) Tile {
  int r0, r1;  // rows [r0, r1); a chunk of a long row has r0 = the row, r1 = its partial-sum slot
  int z0, z1;  // nonzeros [z0, z1)
};

struct alignas(16) LongRow {
  int row;      // the long row
  int first;    // slot of its first chunk's partial sum (chunks occupy consecutive slots)
  int nchunks;  // number of chunk tiles
  int pad;
};

struct TilePlan {
  std::vector<Tile> tiles;
  // meta[t] in [0, 6]: log2(lanes per row) for a row-run tile;
  // meta[t] <  0: chunk of long row number -(meta[t] + 1).
  std::vector<int> meta;
  std::vector<LongRow> longrows;
};

// Partition rows [0, nrows) of a CSR matrix with offsets rowptr.
//
// With colidx (column indices, sorted within each row) and ncols given, the
// tile order is made XCD-aware: workgroups are dispatched to the XCDs round
// robin (tile t runs on XCD t % xcds), and a chunk of a long row covers a
// contiguous column range, so each chunk is placed on the XCD that owns the
// 1/xcds slice of columns it gathers from; that XCD's L2 then holds its slice
// of the dense operand.  Row-run tiles, whose columns are scattered, fill the
// remaining positions.  The order of tiles does not change any sum.
TilePlan plan_tiles(const int* rowptr, int nrows, int tile_nnz, int max_rows, int block, const int* colidx = nullptr,
                    int ncols = 0, int xcds = 8);

// The column indices the XCD-aware order reads -- the middle nonzero of every
// chunk of a long row -- at chosen positions, for a matrix whose index array is
// not on the host.  pos is ascending; col[i] is the column of nonzero pos[i].
struct ColSample {
  std::vector<int> pos;
  std::vector<int> col;
};

// Appends the positions plan_tiles reads for tile_nnz (one per chunk of every row
// longer than tile_nnz); sort and deduplicate before building a ColSample.
void chunk_midpoints(const int* rowptr, int nrows, int tile_nnz, std::vector<int>& out);

// plan_tiles with the columns taken from a sample that holds every position
// chunk_midpoints(rowptr, nrows, tile_nnz) lists.
TilePlan plan_tiles(const int* rowptr, int nrows, int tile_nnz, int max_rows, int block, const ColSample& cols,
                    int ncols, int xcds = 8);

}  // namespace cuopt_amd
