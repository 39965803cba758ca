// spmv_tiles.hpp -- tiled row-block CSR SpMV with fused per-row epilogues (gfx950).
//
// One workgroup (kBlock = 256 lanes, four 64-wide wavefronts) owns one tile
// (tile_plan.hpp).  It streams the tile's value/column-index arrays with
// coalesced non-temporal loads, gathers the dense operand, and parks the
// products in LDS; then groups of G lanes (G chosen per tile) reduce each row
// from LDS and hand the row sum to an epilogue functor, which does the
// projection / Halpern / residual arithmetic of that row in the same kernel.
// Rows longer than a tile are split into chunks; every chunk writes a partial
// sum and a second small launch (long_rows_kernel, one workgroup per long row)
// merges them in a fixed tree and runs the epilogue.  No workgroup waits on or
// signals another inside a launch: on a multi-XCD part every device-scope
// fence would write back and invalidate the XCD's L2.  All summation orders
// are fixed by the tile plan, so results are bitwise reproducible run to run.
//
// Epilogues may also accumulate up to kMaxRed FP64 sums (norms, dots); each
// workgroup writes its sums to a partials array, reduced afterwards in a fixed
// order by reduce_partials_kernel.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "tile_plan.hpp"

namespace cuopt_amd {

constexpr int kBlock   = 256;
constexpr int kWave    = 64;
constexpr int kMaxRowsPerTile = 1024;
constexpr int kMaxRed  = 8;
// Rows per lane group whose operands are loaded before the matrix stream, unless
// TilePipelining<Epi, TILE>::pre_rows says otherwise.
constexpr int kPreRows = 2;
// LDS product capacity of a tile (template parameter TILE of the kernel).  The
// planner caps a tile at TILE - kVecPad nonzeros so that the aligned vector
// loads covering [z0, z1) never need more than TILE slots.
constexpr int kVecPad = 4;
// Value / index arrays read by the tiled kernel carry this many trailing
// elements so the last aligned vector load stays inside the allocation.
constexpr int kArrayPad = 16;

// 16-byte vector types for the value and index streams.
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));
typedef int i32x4 __attribute__((ext_vector_type(4)));
typedef int i32x2 __attribute__((ext_vector_type(2)));

template <typename Mat>
struct VecOf;
template <>
struct VecOf<float> {
  static constexpr int N = 4;
  using V = f32x4;
  using I = i32x4;
};
template <>
struct VecOf<double> {
  static constexpr int N = 2;
  using V = f64x2;
  using I = i32x2;
};

struct PlanDev {
  const Tile* tiles;
  const int* meta;
  const LongRow* longrows;
  int ntiles;
};

// ---- wavefront helpers ------------------------------------------------------

__device__ inline double shfl_xor_d(double v, int m) { return __shfl_xor(v, m, kWave); }
__device__ inline float shfl_xor_f(float v, int m) { return __shfl_xor(v, m, kWave); }

// Knuth's TwoSum: s + e == a + b exactly.
__device__ inline void two_sum(float a, float b, float& s, float& e)
{
  s        = a + b;
  float bb = s - a;
  e        = (a - (s - bb)) + (b - bb);
}

// ---- accumulators -----------------------------------------------------------

// FP64 sum.
struct AccSum64 {
  using Elem = double;
  using Val  = double;
  double s;
  __device__ static AccSum64 zero() { return AccSum64{0.0}; }
  __device__ static Elem prod(double a, double b) { return a * b; }
  __device__ void add(Elem e) { s += e; }
  __device__ void merge(const AccSum64& o) { s += o.s; }
  __device__ AccSum64 shfl(int m) const { return AccSum64{shfl_xor_d(s, m)}; }
  __device__ Val value() const { return s; }
  __device__ Val rsub(Val x) const { return x - s; }  // x - sum
  __device__ uint64_t pack() const { return __double_as_longlong(s); }
  __device__ static AccSum64 unpack(uint64_t u) { return AccSum64{__longlong_as_double(u)}; }
};

// FP64 max of non-negative terms (row / column infinity norms).
struct AccMax64 {
  using Elem = double;
  using Val  = double;
  double s;
  __device__ static AccMax64 zero() { return AccMax64{0.0}; }
  __device__ void add(Elem e) { s = (s < e) ? e : s; }
  __device__ void merge(const AccMax64& o) { s = (s < o.s) ? o.s : s; }
  __device__ AccMax64 shfl(int m) const { return AccMax64{shfl_xor_d(s, m)}; }
  __device__ Val value() const { return s; }
  __device__ Val rsub(Val x) const { return x - s; }
  __device__ uint64_t pack() const { return __double_as_longlong(s); }
  __device__ static AccMax64 unpack(uint64_t u) { return AccMax64{__longlong_as_double(u)}; }
};

// Compensated FP32 sum: each product carries its exact rounding error
// (fma(a, b, -p)), and the running sum carries a Neumaier compensation term.
struct AccComp32 {
  using Elem = float2;  // (product, product rounding error)
  using Val  = float;
  float s, c;
  __device__ static AccComp32 zero() { return AccComp32{0.f, 0.f}; }
  __device__ static Elem prod(float a, float b)
  {
    const float p = a * b;
    return make_float2(p, __builtin_fmaf(a, b, -p));
  }
  __device__ void add(Elem e)
  {
    float t, err;
    two_sum(s, e.x, t, err);
    s = t;
    c += err + e.y;
  }
  __device__ void merge(const AccComp32& o)
  {
    float t, err;
    two_sum(s, o.s, t, err);
    s = t;
    c += o.c + err;
  }
  __device__ AccComp32 shfl(int m) const { return AccComp32{shfl_xor_f(s, m), shfl_xor_f(c, m)}; }
  __device__ Val value() const { return s + c; }
  // x - (s + c) without rounding s + c first: near a fixed point x and the sum
  // cancel, and the rounded sum would leave an error of eps |x| in place of
  // eps |x - sum|.
  __device__ Val rsub(Val x) const { return (x - s) - c; }
  __device__ uint64_t pack() const
  {
    return (uint64_t)__float_as_uint(s) | ((uint64_t)__float_as_uint(c) << 32);
  }
  __device__ static AccComp32 unpack(uint64_t u)
  {
    return AccComp32{__uint_as_float((uint32_t)u), __uint_as_float((uint32_t)(u >> 32))};
  }
};

// ---- block reductions -------------------------------------------------------

// Sum NR per-lane doubles over the workgroup; lane 0 writes them to out[0..NR).
template <int NR>
__device__ inline void block_sum_store(double (&v)[kMaxRed], double* out, double (*lds)[kMaxRed])
{
  const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
#pragma unroll
  for (int k = 0; k < NR; ++k)
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v[k] += shfl_xor_d(v[k], off);
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < NR; ++k) lds[wid][k] = v[k];
  __syncthreads();
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < NR; ++k) {
      double s = lds[0][k];
      for (int w = 1; w < kBlock / kWave; ++w) s += lds[w][k];
      out[k] = s;
    }
}

// Workgroup merge in a fixed order (wave butterflies, then waves 0..3).
// Callers separate two uses of the same lds with a __syncthreads().
template <class Acc>
__device__ inline Acc block_merge(Acc a, uint64_t* lds)
{
  const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) a.merge(a.shfl(off));
  if (lane == 0) lds[wid] = a.pack();
  __syncthreads();
  Acc r = Acc::unpack(lds[0]);
  for (int w = 1; w < kBlock / kWave; ++w) r.merge(Acc::unpack(lds[w]));
  return r;  // valid in every lane
}

// ---- the kernel -------------------------------------------------------------

// Epi must provide:
//   using Mat, Acc;  static constexpr int NRED (0..kMaxRed);
//   struct Pre;                                         per-row operands of the epilogue
//   __device__ bool skip() const;                       uniform early exit
//   __device__ typename Acc::Elem product(Mat v, int col, int k) const;
//   __device__ Pre load(int r) const;                   issued before the products
//   __device__ void row(int r, const Acc& sum, const Pre& p, double (&red)[kMaxRed]) const;
template <class Epi, bool ATOMIC>
__device__ inline void merge_long_row(const PlanDev& plan, int lid, uint64_t* long_part, double* red_part,
                                      const Epi& epi);

// Ints between the chunk counters of consecutive long rows: each counter has a
// 256-byte line of its own, so the atomics of different rows do not queue on one line.
constexpr int kCntStride = 64;

// How a workgroup of tiled_spmv_kernel<Epi, TILE> overlaps its tiles: with
// prefetch it loads the next tile's stream into registers while the current tile
// reduces its rows and runs their epilogues (holding that stream costs 10-20
// VGPRs); with lds_stream it loads it into LDS instead (FP32 only: no registers,
// but the tile's LDS grows by 8 bytes per slot); min_waves is the waves per SIMD
// the kernel is built for; pre_rows the rows per lane group whose row pointers
// and epilogue operands are loaded before the stream.  FP64 kernels prefetch into
// registers at their four waves; FP32 kernels load at the top of each tile,
// except where an epilogue specialises this (the FP32 step kernels, in
// pdlp_kernels.hpp).
template <class Epi, int TILE>
struct TilePipelining {
  static constexpr bool prefetch   = std::is_same<typename Epi::Mat, double>::value;
  static constexpr bool lds_stream = false;
  static constexpr int min_waves   = 1;
  static constexpr int pre_rows    = kPreRows;
};

// A tile's matrix stream in registers: each lane's 16-byte vectors of values and
// indices from the aligned base below z0 (elements outside [z0, z1) are loaded
// but dropped), with the tile's header.
template <class Epi, int TILE>
struct TileStream {
  using VT             = VecOf<typename Epi::Mat>;
  static constexpr int VN  = VT::N;
  static constexpr int kIt = (TILE + kBlock * VN - 1) / (kBlock * VN);
  Tile tl;
  int mt;
  typename VT::V v[kIt];
  typename VT::I c[kIt];

  __device__ void load(int t, const PlanDev& plan, const int* __restrict__ colidx,
                       const typename Epi::Mat* __restrict__ vals)
  {
    tl = plan.tiles[t];
    mt = plan.meta[t];
    load_data(colidx, vals);
  }
  __device__ void load_data(const int* __restrict__ colidx, const typename Epi::Mat* __restrict__ vals)
  {
    const int base = tl.z0 & ~(VN - 1);
#pragma unroll
    for (int it = 0; it < kIt; ++it) {
      const int e0 = base + (it * kBlock + (int)threadIdx.x) * VN;
      if (e0 < tl.z1) {
        v[it] = __builtin_nontemporal_load(reinterpret_cast<const typename VT::V*>(vals + e0));
        c[it] = __builtin_nontemporal_load(reinterpret_cast<const typename VT::I*>(colidx + e0));
      }
    }
  }
};

// FP32 stream through LDS: every lane's 16-byte vectors of values and indices
// are loaded straight into LDS (global_load_lds_dwordx4: lane l of a wavefront
// writes its 16 bytes at the wavefront's base + 16 l), in the layout of the
// register stream, so that a workgroup can have the next tile's stream in
// flight without holding it in registers.  The buffer is free once the tile's
// products are staged.
template <int TILE>
struct LdsStream {
  static constexpr int kIt    = (TILE + kBlock * 4 - 1) / (kBlock * 4);
  static constexpr int kElems = kIt * kBlock * 4;  // floats (and ints) per buffer

  __device__ static void issue(const Tile& tl, const int* __restrict__ colidx, const float* __restrict__ vals,
                               float* vbuf, int* ibuf)
  {
    const int base = tl.z0 & ~3;
    const int w    = (int)threadIdx.x / kWave;
#pragma unroll
    for (int it = 0; it < kIt; ++it) {
      const int e0  = base + (it * kBlock + (int)threadIdx.x) * 4;
      const int off = (it * kBlock + w * kWave) * 4;  // this wavefront's slot
      if (e0 < tl.z1) {
        // Default cache policy.  With the register stream's non-temporal hint (aux
        // 2) the 500 k x 1 M dual step fetched 85 MB instead of 89 MB per launch,
        // but took 35.5 us instead of 34.1 us (primal 25.8 against 24.4 us).
        __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)(vals + e0),
                                         (__attribute__((address_space(3))) void*)(vbuf + off), 16, 0, 0);
        __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)(colidx + e0),
                                         (__attribute__((address_space(3))) void*)(ibuf + off), 16, 0, 0);
      }
    }
  }
};

// One tile.  MODE says where its stream comes from: kStreamOwn loads it into
// registers here, after the row pointers and epilogue operands below; kStreamRegs
// takes it from S (loaded by the caller, maybe still in flight); kStreamLds reads
// it from the LDS buffers sv / si, filled by LdsStream::issue.  S is read only in
// kStreamRegs, sv / si only in kStreamLds.  prefetch() is called once by every
// lane right after the tile's products are staged in LDS: the kernel issues the
// next tile's stream there, so that it is in flight while this tile reduces its
// rows and runs their epilogues.
constexpr int kStreamOwn = 0, kStreamRegs = 1, kStreamLds = 2;

template <class Epi, int TILE, int MODE, class Prefetch>
__device__ inline void tile_body(int t, const TileStream<Epi, TILE>& S, const float* sv, const int* si,
                                 const int* __restrict__ rowptr, const int* __restrict__ colidx,
                                 const typename Epi::Mat* __restrict__ vals, const PlanDev& plan,
                                 uint64_t* __restrict__ long_part, double* __restrict__ red_part, const Epi& epi,
                                 int* __restrict__ long_cnt, typename Epi::Acc::Elem* prod,
                                 double (*red_lds)[kMaxRed], uint64_t* acc_lds, Prefetch&& prefetch)
{
  using Acc  = typename Epi::Acc;
  const Tile tl  = MODE == kStreamRegs ? S.tl : plan.tiles[t];
  const int mt   = MODE == kStreamRegs ? S.mt : plan.meta[t];
  const int tid  = threadIdx.x;
  const int nnz  = tl.z1 - tl.z0;

  // Row-run tiles: load each lane's first kPre row pointers and epilogue
  // operands now, so their latency overlaps the matrix stream below instead of
  // being paid once per row after the barrier.
  const int G    = mt >= 0 ? 1 << mt : 1;
  const int sub  = tid & (G - 1);
  const int grp  = mt >= 0 ? tid >> mt : 0;
  const int ngrp = mt >= 0 ? kBlock >> mt : 0;
  constexpr int kPre = TilePipelining<Epi, TILE>::pre_rows;
  int rb[kPre], re[kPre];
  typename Epi::Pre pre[kPre];
  if (mt >= 0) {
#pragma unroll
    for (int i = 0; i < kPre; ++i) {
      const int r = tl.r0 + grp + i * ngrp;
      if (r < tl.r1) {
        rb[i]  = rowptr[r] - tl.z0;
        re[i]  = rowptr[r + 1] - tl.z0;
        pre[i] = epi.load(r);
      }
    }
  }

  // Stage the tile's products in LDS from the stream (all of a lane's stream
  // loads are issued before its gathers).
  {
    constexpr int VN  = TileStream<Epi, TILE>::VN;
    constexpr int kIt = TileStream<Epi, TILE>::kIt;
    const int base    = tl.z0 & ~(VN - 1);
    auto stage = [&](const TileStream<Epi, TILE>& X) {
#pragma unroll
      for (int it = 0; it < kIt; ++it) {
        const int e0 = base + (it * kBlock + tid) * VN;
        if (e0 < tl.z1) {
#pragma unroll
          for (int i = 0; i < VN; ++i) {
            const int e = e0 + i;
            if (e >= tl.z0 && e < tl.z1) prod[e - tl.z0] = epi.product(X.v[it][i], X.c[it][i], e);
          }
        }
      }
    };
    if constexpr (MODE == kStreamRegs) {
      stage(S);
    } else if constexpr (MODE == kStreamLds) {
      // This lane's vectors, where its own global_load_lds put them (the wait
      // covers the wavefront's loads into LDS).
      __builtin_amdgcn_s_waitcnt(0);
      TileStream<Epi, TILE> X;
#pragma unroll
      for (int it = 0; it < kIt; ++it) {
        const int o = (it * kBlock + tid) * 4;
        X.v[it]     = *reinterpret_cast<const typename TileStream<Epi, TILE>::VT::V*>(sv + o);
        X.c[it]     = *reinterpret_cast<const typename TileStream<Epi, TILE>::VT::I*>(si + o);
      }
      stage(X);
    } else {
      TileStream<Epi, TILE> X;
      X.tl = tl;
      X.load_data(colidx, vals);
      stage(X);
    }
  }
  __syncthreads();
  prefetch();

  double red[kMaxRed];
#pragma unroll
  for (int k = 0; k < kMaxRed; ++k) red[k] = 0.0;

  if (mt >= 0) {
#pragma unroll
    for (int i = 0; i < kPre; ++i) {
      const int r = tl.r0 + grp + i * ngrp;
      if (r < tl.r1) {
        Acc a = Acc::zero();
        for (int k = rb[i] + sub; k < re[i]; k += G) a.add(prod[k]);
        for (int off = G >> 1; off > 0; off >>= 1) a.merge(a.shfl(off));
        if (sub == 0) epi.row(r, a, pre[i], red);
      }
    }
    // Tiles with more rows than kPre per lane group: the rest load as they go.
    for (int r = tl.r0 + grp + kPre * ngrp; r < tl.r1; r += ngrp) {
      const int b = rowptr[r] - tl.z0, e = rowptr[r + 1] - tl.z0;
      Acc a = Acc::zero();
      for (int k = b + sub; k < e; k += G) a.add(prod[k]);
      for (int off = G >> 1; off > 0; off >>= 1) a.merge(a.shfl(off));
      if (sub == 0) epi.row(r, a, epi.load(r), red);
    }
  } else {
    // Chunk of a long row: the workgroup partial goes to the chunk's slot, and
    // wave 0 alone finishes the chunk (the other waves retire).  Without long_cnt
    // the row's epilogue runs in long_rows_kernel after this launch.  With it,
    // the chunk that finishes last runs it here: every chunk publishes its
    // partial and then counts itself with device-scope atomics (read-modify-
    // writes, performed where all XCDs see them, the second issued after the
    // first completed); the last one reads the partials back the same way, so
    // no stale L2 copy of another XCD's partial can be read.
    Acc a = Acc::zero();
    for (int k = tid; k < nnz; k += kBlock) a.add(prod[k]);
    a = block_merge(a, acc_lds);
    if (tid >= kWave) return;
    const int lid = -(mt + 1);
    if constexpr (Epi::NRED > 0)  // a chunk runs no row epilogue: its reduction slot is 0
      if (tid == 0)
#pragma unroll
        for (int k = 0; k < Epi::NRED; ++k) red_part[(size_t)t * Epi::NRED + k] = 0.0;
    if (!long_cnt) {
      if (tid == 0) long_part[tl.r1] = a.pack();  // the chunk's slot
      return;
    }
    int before = 0;
    if (tid == 0) {
      (void)__hip_atomic_exchange(&long_part[tl.r1], a.pack(), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __builtin_amdgcn_s_waitcnt(0);
      before = __hip_atomic_fetch_add(&long_cnt[lid * kCntStride], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    before = __shfl(before, 0, kWave);
    if (before != plan.longrows[lid].nchunks - 1) return;
    merge_long_row<Epi, true>(plan, lid, long_part, red_part, epi);
    if (tid == 0) (void)__hip_atomic_exchange(&long_cnt[lid * kCntStride], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return;
  }
  if constexpr (Epi::NRED > 0) block_sum_store<Epi::NRED>(red, red_part + (size_t)t * Epi::NRED, red_lds);
}

// Workgroup b processes tiles b, b + gridDim.x, ... (the launch sizes the grid
// to the resident workgroups, a multiple of 8, so every tile of workgroup b runs
// on XCD b mod 8 as the plan placed it); in FP64 the next tile's stream is
// loaded while the current one reduces its rows.  A barrier before every later tile
// orders the reuse of the LDS buffers: after a row-run tile without reductions,
// or a long-row chunk that wave 0 finishes alone, other waves may reach the next
// tile first.
template <class Epi, int TILE>
__global__ __launch_bounds__(kBlock, (TilePipelining<Epi, TILE>::min_waves)) void tiled_spmv_kernel(const int* __restrict__ rowptr,
                                                            const int* __restrict__ colidx,
                                                            const typename Epi::Mat* __restrict__ vals,
                                                            PlanDev plan,
                                                            uint64_t* __restrict__ long_part,
                                                            double* __restrict__ red_part,
                                                            Epi epi,
                                                            int* __restrict__ long_cnt)
{
  using Elem = typename Epi::Acc::Elem;
  constexpr bool kLds  = TilePipelining<Epi, TILE>::lds_stream;
  constexpr int kLdsN  = kLds ? LdsStream<TILE>::kElems : 1;
  __shared__ Elem prod[TILE];
  __shared__ double red_lds[kBlock / kWave][kMaxRed];
  __shared__ uint64_t acc_lds[kBlock / kWave];
  __shared__ __attribute__((aligned(16))) float sv[kLdsN];
  __shared__ __attribute__((aligned(16))) int si[kLdsN];
  if (epi.skip()) return;
  if constexpr (kLds) {
    // The next tile's stream goes into the LDS buffers as soon as this tile's
    // products are staged, and is in flight while this tile reduces its rows.
    static_assert(std::is_same<typename Epi::Mat, float>::value, "the LDS stream carries FP32 values");
    const TileStream<Epi, TILE> none{};  // not read
    if ((int)blockIdx.x < plan.ntiles) LdsStream<TILE>::issue(plan.tiles[blockIdx.x], colidx, vals, sv, si);
    for (int t = blockIdx.x; t < plan.ntiles; t += gridDim.x) {
      if (t != (int)blockIdx.x) __syncthreads();
      const int tn = t + gridDim.x;
      tile_body<Epi, TILE, kStreamLds>(t, none, sv, si, rowptr, colidx, vals, plan, long_part, red_part, epi,
                                       long_cnt, prod, red_lds, acc_lds, [&]() {
                                         if (tn < plan.ntiles) LdsStream<TILE>::issue(plan.tiles[tn], colidx, vals, sv, si);
                                       });
    }
  } else if constexpr (TilePipelining<Epi, TILE>::prefetch) {
    // The next tile's stream is loaded while this one reduces.  Measured at
    // 500 k x 1 M in FP64: dual step 43.5 -> 41.3 us (tuned plan 42.2 -> 39.1 us),
    // primal 33.9 -> 32.7 us, at the same occupancy.
    TileStream<Epi, TILE> cur, nxt;
    if ((int)blockIdx.x < plan.ntiles) cur.load(blockIdx.x, plan, colidx, vals);
    for (int t = blockIdx.x; t < plan.ntiles; t += gridDim.x) {
      if (t != (int)blockIdx.x) __syncthreads();
      const int tn = t + gridDim.x;
      tile_body<Epi, TILE, kStreamRegs>(t, cur, nullptr, nullptr, rowptr, colidx, vals, plan, long_part, red_part,
                                        epi, long_cnt, prod, red_lds, acc_lds, [&]() {
                                   if (tn < plan.ntiles) nxt.load(tn, plan, colidx, vals);
                                 });
      cur = nxt;
    }
  } else {
    // The stream is loaded at the top of each tile.  In FP32, holding the next
    // tile's stream raised the step kernels from 78 to 89-99 VGPRs (six waves per
    // SIMD to five or four), and the dual step slowed: banded 4 M x 8 M 167 ->
    // 199 us, the tuned 500 k x 1 M plan 32.4 -> 34 us (built for six waves, with
    // a few spilled registers, 36.7 us).
    const TileStream<Epi, TILE> none{};  // not read: each tile loads its own stream
    for (int t = blockIdx.x; t < plan.ntiles; t += gridDim.x) {
      if (t != (int)blockIdx.x) __syncthreads();
      tile_body<Epi, TILE, kStreamOwn>(t, none, nullptr, nullptr, rowptr, colidx, vals, plan, long_part, red_part,
                                       epi, long_cnt, prod, red_lds, acc_lds, []() {});
    }
  }
}

// Epilogue of long row lid, run by one whole wave: lane l merges chunk partials
// l, l + 64, l + 128, ... in that order (eight loads in flight), a butterfly
// merges the lanes, and lane 0 runs the row epilogue and writes the row's
// reduction slot red_part[(ntiles + lid) * NRED ...].  ATOMIC reads the partials
// with device-scope atomics (the in-launch merge); without it they are plain
// loads, made visible by the kernel boundary before long_rows_kernel.
template <class Epi, bool ATOMIC>
__device__ inline void merge_long_row(const PlanDev& plan, int lid, uint64_t* long_part, double* red_part,
                                      const Epi& epi)
{
  using Acc        = typename Epi::Acc;
  constexpr int kU = 8;
  const int lane   = threadIdx.x & (kWave - 1);
  const LongRow lr = plan.longrows[lid];
  typename Epi::Pre pre{};
  if (lane == 0) pre = epi.load(lr.row);  // in flight during the merge
  Acc tot = Acc::zero();
  for (int base = lane; base < lr.nchunks; base += kU * kWave) {
    uint64_t v[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const int ch = base + u * kWave;
      v[u]         = 0;
      if (ch < lr.nchunks) {
        uint64_t* slot = long_part + lr.first + ch;
        v[u] = ATOMIC ? __hip_atomic_fetch_add(slot, (uint64_t)0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : *slot;
      }
    }
#pragma unroll
    for (int u = 0; u < kU; ++u)
      if (base + u * kWave < lr.nchunks) tot.merge(Acc::unpack(v[u]));
  }
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) tot.merge(tot.shfl(off));
  if (lane == 0) {
    double red[kMaxRed];
#pragma unroll
    for (int k = 0; k < kMaxRed; ++k) red[k] = 0.0;
    epi.row(lr.row, tot, pre, red);
    if constexpr (Epi::NRED > 0)
#pragma unroll
      for (int k = 0; k < Epi::NRED; ++k) red_part[(size_t)(plan.ntiles + lid) * Epi::NRED + k] = red[k];
  }
}

// One wave per long row, after tiled_spmv_kernel on the same stream.
template <class Epi>
__global__ __launch_bounds__(kWave) void long_rows_kernel(PlanDev plan,
                                                          uint64_t* __restrict__ long_part,
                                                          double* __restrict__ red_part,
                                                          Epi epi)
{
  if (epi.skip()) return;
  merge_long_row<Epi, false>(plan, blockIdx.x, long_part, red_part, epi);
}

// out[k] = sum over blocks b (in order of a fixed tree) of part[b * nred + k].
// One workgroup; `skip` (optional device flag) makes it a no-op when nonzero.
__global__ __launch_bounds__(1024) void reduce_partials_kernel(const double* __restrict__ part,
                                                               int nblocks,
                                                               int nred,
                                                               double* __restrict__ out,
                                                               const double* skip);

}  // namespace cuopt_amd
