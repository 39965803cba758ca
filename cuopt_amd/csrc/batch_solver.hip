// batch_solver.hip -- many independent LPs solved in one persistent launch (gfx950).
//
// The heterogeneous batch of the reference (call_batch_solve, cython_solve.cu:286-352,
// which runs independent solve_lp calls on a few streams) becomes a single kernel
// launch: workgroup g owns LP order[g] for its whole life -- CSR transpose,
// Ruiz / Pock-Chambolle scaling, bound / objective rescaling, the step size
// (the reference's power iteration in FP64; in FP32-mixed, Lanczos as the
// single-LP solver's), and the Stable3 loop with FP64 termination
// checks, cuPDLPx restarts and the PID primal weight -- so the batch needs no
// host synchronisation and no per-iteration launch at all.  LPs are handed out
// largest first, so the hardware dispatcher's in-order assignment of
// workgroups to free CUs balances the uneven sizes.
//
// Per LP the arithmetic is the single-LP solver's (pdlp_solver.hip): the same
// fused epilogues (pdlp_kernels.hpp), the same scaling order of operations and
// restart logic; only the order of the row reductions differs (workgroup
// lane groups instead of tiles), so FP64 batch results agree with single-LP
// results to rounding.
//
// SpMV inside a workgroup: groups of G lanes (G = 2^lg chosen per LP and matrix
// by the host from the mean row length) own a row, read its nonzeros with a
// stride of G (consecutive groups own consecutive rows, so a wavefront reads a
// contiguous run of the value / index arrays), merge by shuffles, and run the
// row epilogue.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cmath>
#include <cstring>
#include <memory>
#include <numeric>
#include <random>
#include <exception>
#include <mutex>
#include <thread>
#include <vector>

#include "batch_solver.hpp"
#include "common.hpp"
#include "pdlp_kernels.hpp"

namespace cuopt_amd {

namespace {

constexpr int kBT = 256;  // lanes per LP workgroup
// Workgroups per CU the kernel is built for (its register budget, __launch_bounds__)
// and the LDS share of each.  Three (at most 168 VGPRs) spilled ~100 registers
// to scratch and ran 10-40 % slower than two.
constexpr int kWgPerCu = 2;

struct BatchDesc {
  int m, n, nnz;
  int row0, col0;     // offsets in m- and n-sized arrays
  int aptr0, tptr0;   // offsets of the row-pointer arrays of A (m+1) and A^T (n+1)
  int lga, lgt;       // log2 lanes per row for A and A^T
  int ablk0, nablk;   // row blocks of A: start rows at blk[ablk0 .. ablk0 + nablk]
  int tblk0, ntblk;   // row blocks of A^T
  long long nz0;      // offset in nonzero arrays
  double obj_sf, obj_off;
};

struct BatchOut {
  double l2p, l2d, pobj, dobj, gap, absobj, b_norm, c_norm;
  double eta, omega, sigma_max_sq, beta, gamma;
  int status, iters, restarts, checks, power_iters, recentrings, pad[2];
  unsigned long long t0, t1;  // wall clock of the workgroup's start and end (device ticks)
};

template <typename V>
struct BatchArgs {
  using V2 = typename Vec2<V>::type;
  const BatchDesc* desc;
  const int* order;
  // structure: A from the host; A^T built in the kernel
  const int* a_ptr;
  const int* a_idx;
  const double* a_orig;
  int* t_ptr;
  int* t_idx;
  int* t_pos;
  double* t_orig;
  int* t_fill;
  V* a_w;
  V* t_w;
  // unscaled problem (minimisation form)
  const double* c;
  const double* lb;
  const double* ub;
  const double* lc;
  const double* uc;
  // scaling and working vectors
  double *Dr, *Dc, *rowv, *colv;
  V* c_w;
  V2* bnd_w;
  V2* cb_w;
  double2* cb_orig;
  V *x0, *x1, *y0v, *y1, *xa, *ya, *xbar, *ybar, *xhat, *yhat, *slack, *aty;
  double *xu, *yu, *su;
  double *pz, *pz2, *patq;
  const double* z0;
  BatchOut* out;
  // recentring of FP32-mixed solves (FP64 refinement), as in the single-LP solver
  int refine;
  double *xstar, *ystar, *ax64, *aty64;
  V* kink;
  double tol_ap, tol_rp, tol_ad, tol_rd, tol_ag, tol_rg;
  long long iter_limit;
  // Per-LP time limit in wall-clock ticks (0 = none), counted from the moment the
  // LP's workgroup starts: each LP gets the whole limit from its own dispatch, as
  // each of the reference's per-LP solve_lp calls starts its own timer.  With more
  // LPs than resident workgroups, later LPs start later, so the launch can take
  // longer than the limit (at most about the number of dispatch waves times it).
  unsigned long long tick_limit;
  double* lz;                     // FP32-mixed: Lanczos coefficients, kLzStride per LP (by LP index)
  const int* blk;                 // row-block starts of every LP's A and A^T (plan_row_blocks)
  int lds_bytes;                  // dynamic LDS of the launch: the stage, then the gathered vectors
};

// Lanczos steps at most (and their coefficient slots per LP), as in the single-LP solver.
constexpr int kLzMax    = 240;
constexpr int kLzStride = 2 * kLzMax + 2;

constexpr int kMajorB = 200;
constexpr double kArtificialB = 0.36, kSufficientB = 0.2, kNecessaryB = 0.8;
constexpr double kKpB = 0.99, kKiB = 0.01, kKdB = 0.0, kISmoothB = 0.3;

__device__ inline long long cond_major(long long k)
{
  long long step = 10, threshold = 1000;
  while (k >= threshold) {
    step *= 10;
    threshold *= 10;
  }
  return step;
}
__device__ inline bool check_at(long long k)
{
  const bool major = ((k % kMajorB) == 0 && k > 0) || k <= 0;
  return major || (k % cond_major(k)) == 0;
}

// ---- workgroup reductions (fixed order) -------------------------------------

// x holds the same value in every lane: say so to the compiler, which then keeps
// it (and what is computed from it: the restart and termination scalars) in
// scalar registers instead of one vector register pair per value.
__device__ inline double uniform_d(double x)
{
  const unsigned long long b = (unsigned long long)__double_as_longlong(x);
  const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(b & 0xffffffffu));
  const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(b >> 32));
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

template <int NR>
__device__ inline void wg_sum(double (&v)[kMaxRed], double (*lds)[kMaxRed])
{
  const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
#pragma unroll
  for (int k = 0; k < NR; ++k)
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v[k] += __shfl_xor(v[k], off, kWave);
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < NR; ++k) lds[wid][k] = v[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NR; ++k) {
    double s = lds[0][k];
    for (int w = 1; w < kBT / kWave; ++w) s += lds[w][k];
    v[k] = uniform_d(s);
  }
  __syncthreads();
}

__device__ inline void zero_red(double (&r)[kMaxRed])
{
#pragma unroll
  for (int k = 0; k < kMaxRed; ++k) r[k] = 0.0;
}

// ---- row passes ---------------------------------------------------------------

// Row passes walk an LP's CSR in row blocks planned on the host (plan_row_blocks):
// a block holds at most kBlockNnz nonzeros and at most one row per lane group,
// or is a single row longer than that.  The stage has two buffers; block b uses
// buffer b & 1.  While the workgroup reduces block b from LDS it
//   - copies block b + 1's indices / values into the other buffer with
//     global_load_lds (memory to LDS without passing through registers),
//   - loads the row bounds of its lane group's row in block b + 1,
// so the copy's round trip overlaps the reduction (one barrier per block).
// Each group of G lanes reduces one row and runs the epilogue.
// Each lane adds its row's products in increasing k and groups merge by a
// fixed shuffle tree, so the sums do not depend on the blocking.
constexpr int kStage      = 2560;                // entries per stage buffer
constexpr int kBlockNnz   = kStage - 8;          // nonzeros per staged block (room for the copy's slack)
constexpr int kStageBytes = 2 * kStage * 12;     // two buffers of 8-byte values and 4-byte indices
constexpr int kNnzPad     = 4;                   // extra elements after each nonzero array (16 bytes or more)

// The stage occupies the start of the workgroup's dynamic LDS.  A pass over
// values of VB bytes (8 for the FP64 matrices, 4 for FP32 working ones) lays
// out its buffers by VB: values of buffer b at base + b * kStage * VB, then the
// indices; it double-buffers when dbl8 / dbl4 says the two buffers fit.
struct Stage {
  unsigned char* base;
  uint64_t* merge;  // kBT / kWave accumulator slots
  bool dbl8, dbl4;  // two buffers for passes over 8- / 4-byte values
};

struct RowBlocks {
  const int* start;  // first row of each block; start[nblk] = number of rows
  int nblk;
};

// Lanes per row and row blocks of one matrix (rows given by their lengths).
// G keeps a row's share of a lane to about 16 nonzeros read from LDS.
int plan_lanes(int nrows, long long nnz)
{
  const double avg = nrows > 0 ? (double)nnz / nrows : 0.0;
  int lg           = 0;
  while (lg < 6 && (double)(16 << lg) < avg) ++lg;
  return lg;
}

template <class Len>
void plan_row_blocks(int nrows, int lg, Len len, std::vector<int>& out)
{
  const int max_rows = kBT >> lg;
  int r              = 0;
  out.push_back(0);
  while (r < nrows) {
    long long z = len(r);
    int e       = r + 1;
    if (z <= kBlockNnz)
      while (e < nrows && e - r < max_rows && z + len(e) <= kBlockNnz) z += len(e++);
    out.push_back(e);
    r = e;
  }
}

// Dword offset of p within its 16-byte line.
__device__ inline int dword_mis(const void* p)
{
  return (int)((reinterpret_cast<uintptr_t>(p) >> 2) & 3);
}

// Copies the dwords [d0, d1) of src into LDS in 16-byte pieces: the copy starts
// at the 16-byte boundary at or below src + d0, so dword d0 lands at
// dst[dword_mis(src + d0)], and up to 3 dwords past d1 come along (the nonzero
// arrays carry kNnzPad elements of padding).  Wave w moves pieces
// w*64 + i*256 + lane; global_load_lds with 16 bytes writes lane l's piece to
// the wave-uniform LDS address given plus 16 l.  One 16-byte request moves four
// times the data of a dword request, and the copy is bound by requests.  The
// caller's barrier waits for it.
__device__ inline void copy_to_lds(const int* __restrict__ src, long long d0, long long d1, int* dst)
{
  if (d1 <= d0) return;
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  const int mis  = dword_mis(src + d0);
  const int* s0  = src + d0 - mis;
  const int npc  = (int)((d1 - d0 + mis + 3) >> 2);
  for (int base = w * kWave; base < npc; base += kBT)
    if (base + lane < npc)
      __builtin_amdgcn_global_load_lds(s0 + 4 * (base + lane), dst + 4 * base, 16, 0, 0);
}

// Fixed-order workgroup merge of one accumulator (wave butterflies, then waves 0..3).
template <class Acc>
__device__ inline Acc wg_merge(Acc a, uint64_t* slots)
{
  const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) a.merge(a.shfl(off));
  if (lane == 0) slots[wid] = a.pack();
  __syncthreads();
  Acc r = Acc::unpack(slots[0]);
  for (int w = 1; w < kBT / kWave; ++w) r.merge(Acc::unpack(slots[w]));
  __syncthreads();
  return r;
}

// Rows [0, nrows) of one LP's CSR; Epi as in spmv_tiles.hpp, with the row index
// passed as the third argument of product().
template <class Epi>
__device__ inline void rows_pass(const int* __restrict__ ptr,
                                 const int* __restrict__ idx,
                                 const typename Epi::Mat* __restrict__ vals,
                                 int lg,
                                 RowBlocks blocks,
                                 const Stage& st,
                                 const Epi& epi,
                                 double (&red)[kMaxRed])
{
  using Acc        = typename Epi::Acc;
  using Mat        = typename Epi::Mat;
  using Pre        = typename Epi::Pre;
  constexpr int W  = (int)(sizeof(Mat) / 4);  // dwords per value
  constexpr int VB = 4 * W;                   // bytes per staged value
  const int G = 1 << lg, sub = threadIdx.x & (G - 1), grp = threadIdx.x >> lg;
  const int nb = blocks.nblk;
  if (nb <= 0) return;
  const bool dbl       = W == 1 ? st.dbl4 : st.dbl8;
  int* const sidx_base = reinterpret_cast<int*>(st.base + (dbl ? 2 : 1) * kStage * VB);
  // Buffer addresses are computed, not indexed from an array: a runtime-indexed
  // pointer array would live in scratch, and its load would wait for the copy
  // in flight.
  auto stage = [&](int za, int zb, int buf) {
    copy_to_lds(reinterpret_cast<const int*>(vals), (long long)W * za, (long long)W * zb,
                reinterpret_cast<int*>(st.base + buf * (kStage * VB)));
    copy_to_lds(idx, za, zb, sidx_base + buf * kStage);
  };
  // (r0, r1) are block bi's rows, (r1, r2) block bi + 1's, z the matching
  // nonzero offsets; r3 / z3 (block bi + 2's end) are loaded during block bi.
  int r0 = blocks.start[0], r1 = blocks.start[1];
  int z0 = ptr[r0], z1 = ptr[r1];
  int r2 = r1, z2 = z1;
  if (nb >= 2) {
    r2 = blocks.start[2];
    z2 = ptr[r2];
  }
  if (dbl && z1 - z0 <= kBlockNnz) stage(z0, z1, 0);
  // This lane group's row of the current block: its ptr entries and its
  // epilogue operands, both loaded one block ahead so that the barrier that
  // waits for a block's copy also covers them.  The loads are clamped into range
  // instead of guarded, so that no branch waits for them (a wait for a load also
  // waits for every copy issued before it); a clamped row's operands go unused.
  auto load_pre = [&](int r, int rend) { return epi.load(min(r, max(rend - 1, 0))); };
  int pb = ptr[min(r0 + grp, r1)], pe = ptr[min(r0 + grp + 1, r1)];
  Pre pre = load_pre(r0 + grp, r1);
  for (int bi = 0; bi < nb; ++bi) {
    const int cur = dbl ? bi & 1 : 0;
    // Every wave's copy of block bi has landed (the barrier waits for each
    // wave's memory operations), and block bi - 1's reads of the other buffer
    // are done, so it may be refilled.
    __syncthreads();
    const int b = pb - z0, e = pe - z0;
    const int r3          = bi + 3 <= nb ? blocks.start[bi + 3] : r2;
    const bool stage_next = bi + 1 < nb && z2 - z1 <= kBlockNnz;
    if (dbl && stage_next) stage(z1, z2, cur ^ 1);
    if (!dbl && z1 - z0 <= kBlockNnz) {  // one buffer: copy this block now and wait for it
      stage(z0, z1, 0);
      __syncthreads();
    }
    const int npb = ptr[min(r1 + grp, r2)], npe = ptr[min(r1 + grp + 1, r2)];
    const Pre npre = load_pre(r1 + grp, r2);
    if (z1 - z0 > kBlockNnz) {  // one row longer than a block: read it in place (lane 0's pre is row r0's)
      Acc a = Acc::zero();
      for (int k = z0 + (int)threadIdx.x; k < z1; k += kBT) a.add(epi.product(vals[k], idx[k], r0));
      a = wg_merge(a, st.merge);
      if (threadIdx.x == 0) epi.row(r0, a, pre, red);
    } else if (r0 + grp < r1) {
      const int r       = r0 + grp;
      // where the copy put nonzero z0 (values are 8-byte aligned when W = 2,
      // so their dword offset is even)
      const int* sidx  = sidx_base + cur * kStage + dword_mis(idx + z0);
      const Mat* svals = reinterpret_cast<const Mat*>(st.base + cur * (kStage * VB)) +
                         dword_mis(reinterpret_cast<const int*>(vals) + (long long)W * z0) / W;
      Acc a = Acc::zero();
      int k = b + sub;
      for (; k + 3 * G < e; k += 4 * G) {  // four LDS reads in flight, added in order
        const int c0 = sidx[k], c1 = sidx[k + G], c2 = sidx[k + 2 * G], c3 = sidx[k + 3 * G];
        const Mat v0 = svals[k], v1 = svals[k + G], v2 = svals[k + 2 * G], v3 = svals[k + 3 * G];
        const auto p0 = epi.product(v0, c0, r), p1 = epi.product(v1, c1, r);
        const auto p2 = epi.product(v2, c2, r), p3 = epi.product(v3, c3, r);
        a.add(p0);
        a.add(p1);
        a.add(p2);
        a.add(p3);
      }
      for (; k < e; k += G) a.add(epi.product(svals[k], sidx[k], r));
      for (int off = G >> 1; off > 0; off >>= 1) a.merge(a.shfl(off));
      if (sub == 0) epi.row(r, a, pre, red);
    }
    const int z3 = bi + 3 <= nb ? ptr[r3] : z2;
    r0 = r1;
    z0 = z1;
    r1 = r2;
    z1 = z2;
    r2 = r3;
    z2 = z3;
    pb  = npb;
    pe  = npe;
    pre = npre;
  }
  __syncthreads();  // the pass's stores are visible to the whole workgroup
}

// A PDHG step epilogue whose gathered vector lives in LDS, read through an
// LDS-typed pointer.  Through the generic pointer each gather is a flat load,
// which the compiler waits for with both counters: one gather in flight at a
// time, and the first wait of a block also waits for the next block's copy.
// Typed LDS reads wait on the LDS counter only, so the products keep several
// reads in flight and overlap the copy.
template <class Base>
struct LdsGather : Base {
  using Acc = typename Base::Acc;
  using Mat = typename Base::Mat;
  const __attribute__((address_space(3))) Mat* g;
  __device__ typename Acc::Elem product(Mat v, int col, int) const { return Acc::prod(v, g[col]); }
};

template <typename T>
__device__ inline const __attribute__((address_space(3))) T* as_lds(const T* p)
{
  return (const __attribute__((address_space(3))) T*)p;
}

// The same with the gathered vector in global memory (an LP whose vectors do not
// fit in LDS): typed global loads are waited for one counter at a time, in
// order, so a lane keeps its four gathers in flight instead of one.
template <class Base>
struct GlobalGather : Base {
  using Acc = typename Base::Acc;
  using Mat = typename Base::Mat;
  const __attribute__((address_space(1))) Mat* g;
  __device__ typename Acc::Elem product(Mat v, int col, int) const { return Acc::prod(v, g[col]); }
};

template <typename T>
__device__ inline const __attribute__((address_space(1))) T* as_global(const T* p)
{
  return (const __attribute__((address_space(1))) T*)p;
}

// Ruiz / Pock-Chambolle statistics of |(a * D_row-side) * D_col-side| in the
// reference operand order; TRANSPOSED passes A^T (its rows are columns of A).
template <bool TRANSPOSED, bool MAXNORM>
struct BScaleEpi {
  using Mat = double;
  using Acc = typename std::conditional<MAXNORM, AccMax64, AccSum64>::type;
  struct Pre {};
  const double* Dr;
  const double* Dc;
  double* out;
  __device__ double product(double v, int col, int r) const
  {
    const double dr = TRANSPOSED ? Dr[col] : Dr[r];
    const double dc = TRANSPOSED ? Dc[r] : Dc[col];
    return fabs((v * dr) * dc);
  }
  __device__ Pre load(int) const { return Pre{}; }
  __device__ void row(int r, const Acc& s, const Pre&, double (&)[kMaxRed]) const { out[r] = s.value(); }
};

// Power iteration on the FP64 scaled matrix, values formed on the fly as in
// the scaled copy of the single-LP solver: A~^T q, then A~ (A~^T q).
struct BPowerAtEpi {
  using Mat = double;
  using Acc = AccSum64;
  struct Pre {};
  const double* Dr;
  const double* Dc;
  const double* z;
  double nrm;
  double* atq;
  __device__ double product(double v, int col, int r) const { return ((v * Dc[r]) * Dr[col]) * (z[col] / nrm); }
  __device__ Pre load(int) const { return Pre{}; }
  __device__ void row(int r, const Acc& s, const Pre&, double (&)[kMaxRed]) const { atq[r] = s.value(); }
};
struct BPowerAEpi {
  using Mat = double;
  using Acc = AccSum64;
  struct Pre {
    double zold;
  };
  const double* Dr;
  const double* Dc;
  const double* atq;
  const double* zold;
  double nrm;
  double* znew;
  __device__ double product(double v, int col, int r) const { return ((v * Dr[r]) * Dc[col]) * atq[col]; }
  __device__ Pre load(int r) const { return Pre{zold[r]}; }
  __device__ void row(int r, const Acc& sum, const Pre& p, double (&red)[kMaxRed]) const
  {
    const double s = sum.value();
    znew[r]        = s;
    red[0] += (p.zold / nrm) * s;
  }
};

// Lanczos on A~ A~^T (FP32-mixed): A~^T q, then A~ (A~^T q), FP64 values formed
// on the fly in the operand order of the scaled copies.
struct BLzAtEpi {
  using Mat = double;
  using Acc = AccSum64;
  struct Pre {};
  const double* Dr;
  const double* Dc;
  const double* q;
  double* atq;
  __device__ double product(double v, int col, int r) const { return ((v * Dc[r]) * Dr[col]) * q[col]; }
  __device__ Pre load(int) const { return Pre{}; }
  __device__ void row(int r, const Acc& s, const Pre&, double (&)[kMaxRed]) const { atq[r] = s.value(); }
};
struct BLzAEpi {
  using Mat = double;
  using Acc = AccSum64;
  struct Pre {};
  const double* Dr;
  const double* Dc;
  const double* atq;
  double* z;
  __device__ double product(double v, int col, int r) const { return ((v * Dr[r]) * Dc[col]) * atq[col]; }
  __device__ Pre load(int) const { return Pre{}; }
  __device__ void row(int r, const Acc& s, const Pre&, double (&)[kMaxRed]) const { z[r] = s.value(); }
};

// Largest eigenvalue of the symmetric tridiagonal matrix of k Lanczos steps,
// ab[2i] = alpha_i (diagonal) and ab[2i+1] = beta_i^2 (squared off-diagonal),
// by Sturm counts: each round every lane counts the eigenvalues below one of kBT
// points spread evenly over the bracket (Gershgorin to start), and the bracket
// becomes the gap between the last point with fewer than k below and the first
// with all k below, kBT + 1 times narrower.  Stops at the single-LP solver's
// bisection tolerance (relative width 1e-15).  Whole workgroup; slot is one LDS int.
__device__ inline double wg_tridiag_max_eig(const double* ab, int k, int* slot)
{
  double lo = 0.0, hi = 0.0;
  for (int i = 0; i < k; ++i) {
    const double r = (i > 0 ? sqrt(ab[2 * i - 1]) : 0.0) + (i + 1 < k ? sqrt(ab[2 * i + 1]) : 0.0);
    lo = fmin(lo, ab[2 * i] - r);
    hi = fmax(hi, ab[2 * i] + r);
  }
  for (int round = 0; round < 64 && hi - lo > 1e-15 * fmax(fabs(hi), 1e-300); ++round) {
    const double step = (hi - lo) / (double)(kBT + 1);
    const double x    = lo + step * (double)(threadIdx.x + 1);
    int c             = 0;
    double d          = 1.0;
    for (int i = 0; i < k; ++i) {
      d = ab[2 * i] - x - (i > 0 ? ab[2 * i - 1] / d : 0.0);
      if (d == 0.0) d = -1e-300;
      if (d < 0.0) ++c;
    }
    if (threadIdx.x == 0) *slot = kBT;
    __syncthreads();
    if (c >= k) atomicMin(slot, (int)threadIdx.x);
    __syncthreads();
    const int f = *slot;
    __syncthreads();
    const double nhi = f < kBT ? lo + step * (double)(f + 1) : hi;
    const double nlo = f > 0 ? lo + step * (double)f : lo;
    lo               = uniform_d(nlo);
    hi               = uniform_d(nhi);
  }
  return hi;
}

// y = A x (FP64), row index passed to product() as in the other batch passes.
struct BPlainEpi {
  using Mat = double;
  using Acc = AccSum64;
  struct Pre {};
  const double* x;
  double* y;
  __device__ double product(double v, int col, int) const { return v * x[col]; }
  __device__ Pre load(int) const { return Pre{}; }
  __device__ void row(int r, const Acc& s, const Pre&, double (&)[kMaxRed]) const { y[r] = s.value(); }
};

__device__ inline double bound_sq(double l, double u)
{
  double t = 0.0;
  if (isfinite(l) && l != u) t += l * l;
  if (isfinite(u)) t += u * u;
  return t;
}

// ---- the kernel ----------------------------------------------------------------

template <typename V>
__global__ __launch_bounds__(kBT, kWgPerCu) void batch_pdlp_kernel(BatchArgs<V> A)
{
  using V2 = typename Vec2<V>::type;
  __shared__ double lds[kBT / kWave][kMaxRed];
  __shared__ StepScalars sc;
  __shared__ int scan_lds[kBT];
  __shared__ uint64_t merge_lds[kBT / kWave];
  extern __shared__ __align__(16) unsigned char dyn_lds[];
  const unsigned long long t_start = wall_clock64();
  const int tid = threadIdx.x;
  const BatchDesc d = A.desc[A.order[blockIdx.x]];
  const int m = d.m, n = d.n, nnz = d.nnz;
  // this LP's slices
  const int* aptr     = A.a_ptr + d.aptr0;
  const int* aidx     = A.a_idx + d.nz0;
  const double* aorig = A.a_orig + d.nz0;
  int* tptr           = A.t_ptr + d.tptr0;
  int* tidx           = A.t_idx + d.nz0;
  int* tpos           = A.t_pos + d.nz0;
  double* torig       = A.t_orig + d.nz0;
  int* tfill          = A.t_fill + d.col0;
  V* aw               = A.a_w + d.nz0;
  V* tw               = A.t_w + d.nz0;
  const double *c = A.c + d.col0, *lb = A.lb + d.col0, *ub = A.ub + d.col0;
  const double *lc = A.lc + d.row0, *uc = A.uc + d.row0;
  double *Dr = A.Dr + d.row0, *Dc = A.Dc + d.col0, *rowv = A.rowv + d.row0, *colv = A.colv + d.col0;
  V* c_w = A.c_w + d.col0;
  V2* bnd_w = A.bnd_w + d.col0;
  V2* cb_w = A.cb_w + d.row0;
  double2* cb_orig = A.cb_orig + d.row0;
  V* xs[2] = {A.x0 + d.col0, A.x1 + d.col0};
  V* ys[2] = {A.y0v + d.row0, A.y1 + d.row0};
  V *xanc = A.xa + d.col0, *yanc = A.ya + d.row0, *xbar = A.xbar + d.col0, *ybar = A.ybar + d.row0;
  // The two vectors the PDHG step gathers -- y (primal step, both buffers) and
  // x-bar (dual step) -- live in LDS when the LP fits: each gather is then an
  // LDS read instead of a round trip to L2 / MALL.  Nothing reads them after
  // the kernel (the solution is the unscaled xu / yu / su).
  // They sit at the end of the dynamic LDS when a single-buffered 8-byte stage
  // still fits before them (gathers from LDS outweigh the copy overlap); each
  // pass of the PDHG loop then double-buffers when its two buffers fit in what
  // is left.  Until the loop starts the vectors are unused, so the setup passes
  // may double-buffer over the whole of it.
  const long long lds_bytes = A.lds_bytes;
  const long long vec_bytes = (long long)(2 * m + n) * (long long)sizeof(V);
  const long long vec_off   = (lds_bytes - vec_bytes) & ~15LL;
  const bool vec_in_lds     = vec_off >= (long long)kStage * 12;
  const long long region    = vec_in_lds ? vec_off : lds_bytes;  // the stage's LDS in the loop
  Stage st{dyn_lds, merge_lds, 2LL * kStage * 12 <= lds_bytes, 2LL * kStage * 8 <= lds_bytes};
  const RowBlocks ablk{A.blk + d.ablk0, d.nablk}, tblk{A.blk + d.tblk0, d.ntblk};
  if (vec_in_lds) {
    V* base = reinterpret_cast<V*>(dyn_lds + vec_off);
    ys[0]   = base;
    ys[1]   = base + m;
    xbar    = base + 2 * m;
  }
  V *xhat = A.xhat + d.col0, *yhat = A.yhat + d.row0, *slack = A.slack + d.col0, *aty = A.aty + d.col0;
  double *xu = A.xu + d.col0, *yu = A.yu + d.row0, *su = A.su + d.col0;
  double *pz = A.pz + d.row0, *pz2 = A.pz2 + d.row0, *patq = A.patq + d.col0;
  double *xstar = A.xstar + d.col0, *ystar = A.ystar + d.row0, *ax64 = A.ax64 + d.row0, *aty64 = A.aty64 + d.col0;
  V* kink = A.kink + d.row0;
  const double* lbv = lb;
  const double* ubv = ub;
  bool refine_active = false;
  int recentrings    = 0;
  double rdp = 0.0, rdd = 0.0;  // scales of the current correction problem (0: none yet)
  double red[kMaxRed];

  // ---- A^T: counting sort by column, then each column segment ordered by source position
  for (int j = tid; j < n; j += kBT) tfill[j] = 0;
  __syncthreads();
  for (int k = tid; k < nnz; k += kBT) atomicAdd(&tfill[aidx[k]], 1);
  __syncthreads();
  {
    const int per = (n + kBT - 1) / kBT, j0 = min(n, tid * per), j1 = min(n, j0 + per);
    int s = 0;
    for (int j = j0; j < j1; ++j) s += tfill[j];
    scan_lds[tid] = s;
    __syncthreads();
    if (tid == 0) {
      int acc = 0;
      for (int t = 0; t < kBT; ++t) {
        const int v = scan_lds[t];
        scan_lds[t] = acc;
        acc += v;
      }
    }
    __syncthreads();
    int acc = scan_lds[tid];
    for (int j = j0; j < j1; ++j) {
      const int v = tfill[j];
      tptr[j]     = acc;
      tfill[j]    = acc;
      acc += v;
    }
    if (tid == kBT - 1) tptr[n] = acc;
  }
  __syncthreads();
  for (int i = tid; i < m; i += kBT)
    for (int k = aptr[i]; k < aptr[i + 1]; ++k) {
      const int dst = atomicAdd(&tfill[aidx[k]], 1);
      tpos[dst]     = k;
      tidx[dst]     = i;
    }
  __syncthreads();
  for (int j = tid; j < n; j += kBT) {  // insertion sort of the (short) column by source position
    const int b = tptr[j], e = tptr[j + 1];
    for (int p = b + 1; p < e; ++p) {
      const int kp = tpos[p], ip = tidx[p];
      int q = p - 1;
      while (q >= b && tpos[q] > kp) {
        tpos[q + 1] = tpos[q];
        tidx[q + 1] = tidx[q];
        --q;
      }
      tpos[q + 1] = kp;
      tidx[q + 1] = ip;
    }
    for (int p = b; p < e; ++p) torig[p] = aorig[tpos[p]];
  }
  __syncthreads();

  // ---- Ruiz (x10) and Pock-Chambolle (alpha = 1) equilibration
  for (int i = tid; i < m; i += kBT) Dr[i] = 1.0;
  for (int j = tid; j < n; j += kBT) Dc[j] = 1.0;
  __syncthreads();
  for (int it = 0; it <= 10; ++it) {
    const bool pc = it == 10;
    zero_red(red);
    if (pc) {
      rows_pass(aptr, aidx, aorig, d.lga, ablk, st, BScaleEpi<false, false>{Dr, Dc, rowv}, red);
      rows_pass(tptr, tidx, torig, d.lgt, tblk, st, BScaleEpi<true, false>{Dr, Dc, colv}, red);
    } else {
      rows_pass(aptr, aidx, aorig, d.lga, ablk, st, BScaleEpi<false, true>{Dr, Dc, rowv}, red);
      rows_pass(tptr, tidx, torig, d.lgt, tblk, st, BScaleEpi<true, true>{Dr, Dc, colv}, red);
    }
    __syncthreads();
    for (int i = tid; i < m; i += kBT) Dr[i] = rowv[i] > 0.0 ? Dr[i] / sqrt(rowv[i]) : Dr[i];
    for (int j = tid; j < n; j += kBT) Dc[j] = colv[j] > 0.0 ? Dc[j] / sqrt(colv[j]) : Dc[j];
    __syncthreads();
  }
  // scaled working matrices, reference operand order (vals * first[row]) * second[col]
  for (int i = tid; i < m; i += kBT)
    for (int k = aptr[i]; k < aptr[i + 1]; ++k) aw[k] = (V)((aorig[k] * Dr[i]) * Dc[aidx[k]]);
  for (int j = tid; j < n; j += kBT)
    for (int k = tptr[j]; k < tptr[j + 1]; ++k) tw[k] = (V)((torig[k] * Dc[j]) * Dr[tidx[k]]);
  // vectors and norms (bound_objective_rescaling)
  zero_red(red);
  for (int j = tid; j < n; j += kBT) {
    const double cs = c[j] * Dc[j];
    red[0] += cs * cs * 1.0;
    red[1] += c[j] * c[j];
  }
  for (int i = tid; i < m; i += kBT) {
    red[2] += bound_sq(lc[i] * Dr[i], uc[i] * Dr[i]);
    red[3] += bound_sq(lc[i], uc[i]);
  }
  wg_sum<4>(red, lds);
  double beta = 1.0 / (sqrt(red[2]) + 1.0), gamma = 1.0 / (sqrt(red[0]) + 1.0);
  const double b_norm = sqrt(red[3]), c_norm = sqrt(red[1]);
  for (int j = tid; j < n; j += kBT) {
    const double dc = Dc[j];
    double l = dc == 0.0 ? 0.0 : lb[j] / dc, u = dc == 0.0 ? 0.0 : ub[j] / dc;
    l = l * beta;
    u = u * beta;
    c_w[j] = (V)((c[j] * dc) * gamma);
    V2 bb;
    bb.x     = (V)l;
    bb.y     = (V)u;
    bnd_w[j] = bb;
  }
  for (int i = tid; i < m; i += kBT) {
    V2 bb;
    bb.x       = (V)((lc[i] * Dr[i]) * beta);
    bb.y       = (V)((uc[i] * Dr[i]) * beta);
    cb_w[i]    = bb;
    cb_orig[i] = make_double2(lc[i], uc[i]);
  }
  double s2  = 0.0;
  int piters = 0;
  if constexpr (std::is_same<V, float>::value) {
    // ---- FP32-mixed: Lanczos on A~ A~^T, as the single-LP solver's
    // lanczos_sigma_max_sq (hashed start, at most 240 steps, stop when the Ritz
    // value moves by at most 1e-6 relative or beta^2 vanishes).  Scratch vectors
    // not yet in use: q, q_prev in pz / pz2, w in rowv (scaling is done), A~ A~^T q
    // in yu (first written by the first termination check).
    double* q  = pz;
    double* qp = pz2;
    double* w  = rowv;
    double* zb = yu;
    double* ab = A.lz + (size_t)A.order[blockIdx.x] * kLzStride;
    zero_red(red);
    for (int i = tid; i < m; i += kBT) {
      const double v = hashed_start(i);
      q[i]           = v;
      qp[i]          = 0.0;
      red[0] += v * v;
    }
    wg_sum<1>(red, lds);
    const double inv = 1.0 / sqrt(red[0]);
    for (int i = tid; i < m; i += kBT) q[i] = q[i] * inv;
    __syncthreads();
    const int kmax = min(m, kLzMax);
    double theta = 0.0, prev = -1.0;
    while (piters < kmax) {
      const int j = piters;
      zero_red(red);
      rows_pass(tptr, tidx, torig, d.lgt, tblk, st, BLzAtEpi{Dr, Dc, q, patq}, red);
      __syncthreads();
      rows_pass(aptr, aidx, aorig, d.lga, ablk, st, BLzAEpi{Dr, Dc, patq, zb}, red);
      const double bp = j > 0 ? sqrt(ab[2 * j - 1]) : 0.0;
      zero_red(red);
      for (int i = tid; i < m; i += kBT) {
        const double v = zb[i] - bp * qp[i];
        w[i]           = v;
        red[0] += q[i] * v;
      }
      wg_sum<1>(red, lds);
      const double alpha = red[0];
      zero_red(red);
      for (int i = tid; i < m; i += kBT) {
        const double v = w[i] - alpha * q[i];
        w[i]           = v;
        red[0] += v * v;
      }
      wg_sum<1>(red, lds);
      const double b2 = red[0];
      if (tid == 0) {
        ab[2 * j]     = alpha;
        ab[2 * j + 1] = b2;
      }
      __syncthreads();
      piters = j + 1;
      theta  = wg_tridiag_max_eig(ab, j + 1, scan_lds);
      const bool stop = (j > 2 && fabs(theta - prev) <= 1e-6 * theta) || !(b2 > 1e-26 * theta * theta);
      prev            = theta;
      if (stop) break;
      const double sb = sqrt(b2);
      for (int i = tid; i < m; i += kBT) {
        qp[i] = q[i];
        q[i]  = w[i] / sb;
      }
      __syncthreads();
    }
    s2 = theta;
  } else {
    // ---- power iteration (compute_initial_step_size, pdlp.cu:2622-2738)
    for (int i = tid; i < m; i += kBT) pz[i] = A.z0[i];
    zero_red(red);
    for (int i = tid; i < m; i += kBT) red[0] += A.z0[i] * A.z0[i];
    wg_sum<1>(red, lds);
    double nrm = sqrt(red[0]);
    double* zo = pz;
    double* zn = pz2;
    while (piters < 5000) {
      zero_red(red);
      rows_pass(tptr, tidx, torig, d.lgt, tblk, st, BPowerAtEpi{Dr, Dc, zo, nrm, patq}, red);
      __syncthreads();
      rows_pass(aptr, aidx, aorig, d.lga, ablk, st, BPowerAEpi{Dr, Dc, patq, zo, nrm, zn}, red);
      wg_sum<1>(red, lds);
      s2 = red[0];
      zero_red(red);
      for (int i = tid; i < m; i += kBT) {
        const double q = zo[i] / nrm;
        const double r = q * -s2 + zn[i];
        red[0] += r * r;
        red[1] += zn[i] * zn[i];
      }
      wg_sum<2>(red, lds);
      ++piters;
      if (sqrt(red[0]) < 1e-4) break;
      nrm = sqrt(red[1]);
      double* t = zo;
      zo        = zn;
      zn        = t;
    }
  }
  const double eta = 0.998 / sqrt(s2);

  // ---- the loop's stage: beside the vectors
  st.dbl8 = 2LL * kStage * 12 <= region;
  st.dbl4 = 2LL * kStage * 8 <= region;

  // ---- initial iterates (zeros, projected)
  for (int j = tid; j < n; j += kBT) {
    const V2 bb = bnd_w[j];
    xhat[j]     = V(0);
    xanc[j]     = V(0);
    xs[0][j]    = mn(mx(V(0), bb.x), bb.y);
    slack[j]    = V(0);
  }
  for (int i = tid; i < m; i += kBT) {
    yhat[i]  = V(0);
    yanc[i]  = V(0);
    ys[0][i] = V(0);
  }
  __syncthreads();

  // ---- Stable3 loop (pdlp_solver_t::run_solver, pdlp.cu:2013-2458)
  const double nan = __longlong_as_double(0x7ff8000000000000LL), inf = __longlong_as_double(0x7ff0000000000000LL);
  double omega = 1.0, best_omega = 1.0, tau = eta, sigma = eta;
  double fpe = nan, fpe0 = nan, last_trial = inf, err_sum = 0.0, last_err = 0.0, best_gap = inf;
  double fpe_t[3] = {0, 0, 0}, fpe0_t[3] = {0, 0, 0};  // (||dy||^2, interaction, ||dx||^2)
  double l2p = 0, l2d = 0, pobj = 0, dobj = 0, gap = 0, absobj = 0;
  long long k = 0, kr = 0;
  int cur = 0, restarts = 0, checks = 0, status = 0;
  bool restarted_prev = false, periodic_prev = false;
  while (true) {
    // termination check on the unscaled problem (FP64)
    const bool is_major = ((k % kMajorB) == 0 && k > 0) || k <= 0;
    {
      zero_red(red);
      const double idp = refine_active ? 1.0 / rdp : 1.0, idd = refine_active ? 1.0 / rdd : 1.0;
      // xu carries the gathered primal point until the column pass writes x there.
      const KktPointF<V> point{xhat, Dc, 1.0 / beta, refine_active ? xstar : nullptr, idp, xu};
      double unused[kMaxRed];
      for (int j = tid; j < n; j += kBT) point(j, unused);
      __syncthreads();
      rows_pass(aptr, aidx, aorig, d.lga, ablk, st,
                KktRowsEpi<V>{xu, yhat, Dr, 1.0 / gamma, gamma, cb_orig, yanc, yu, refine_active ? xstar : nullptr,
                              refine_active ? ystar : nullptr, idp, idd},
                red);
      wg_sum<3>(red, lds);
      const double rp2 = red[0], psum = red[1], dd = red[2];  // dd: ||y0 - yhat||^2
      zero_red(red);
      rows_pass(tptr, tidx, torig, d.lgt, tblk, st,
                KktColsEpi<V>{yu, xhat, slack, Dc, 1.0 / beta, beta, 1.0 / gamma, gamma, c, xanc, xu, su,
                              refine_active ? xstar : nullptr, idp, idd, bnd_w, lbv, ubv},
                red);
      wg_sum<4>(red, lds);
      const double dp = red[3];  // ||x0 - xhat||^2
      l2p        = sqrt(rp2);
      l2d        = sqrt(red[0]);
      double po  = red[1], du = red[2] + psum;
      if (d.obj_sf != 1.0 || d.obj_off != 0.0) {
        po = d.obj_sf * (po + d.obj_off);
        du = d.obj_sf * (du + d.obj_off);
      }
      pobj   = po;
      dobj   = du;
      gap    = fabs(po - du);
      absobj = fabs(po) + fabs(du);
      ++checks;
      if (restarted_prev) {
        fpe  = sqrt((fpe0_t[2] * omega + fpe0_t[0] / omega) + 2.0 * fpe0_t[1] * eta);
        fpe0 = fpe;
      }
      if (periodic_prev) fpe = sqrt((fpe_t[2] * omega + fpe_t[0] / omega) + 2.0 * fpe_t[1] * eta);
      const bool optimal = l2d <= A.tol_ad + A.tol_rd * c_norm && l2p <= A.tol_ap + A.tol_rp * b_norm &&
                           gap <= A.tol_ag + A.tol_rg * absobj;
      if (k > 1 && optimal) {
        status = CUOPT_TERIMINATION_STATUS_OPTIMAL;
        break;
      }
      if (isnan(l2p) || isnan(l2d) || isnan(gap)) {
        status = CUOPT_TERIMINATION_STATUS_NUMERICAL_ERROR;
        break;
      }
      if (A.tick_limit && wall_clock64() - t_start >= A.tick_limit) {
        status = CUOPT_TERIMINATION_STATUS_TIME_LIMIT;
        break;
      }
      if (k >= A.iter_limit) {
        status = CUOPT_TERIMINATION_STATUS_ITERATION_LIMIT;
        break;
      }
      // cuPDLPx restart and PID primal weight (pdlp_restart_strategy.cu:688-985)
      bool restarted = false;
      if (is_major) {
        bool do_restart = false;
        if (k == kMajorB) {
          do_restart = true;
        } else if (k > kMajorB) {
          if (fpe <= kSufficientB * fpe0)
            do_restart = true;
          else if (fpe <= kNecessaryB * fpe0 && fpe > last_trial)
            do_restart = true;
          else if ((double)kr >= kArtificialB * (double)k)
            do_restart = true;
        }
        last_trial = fpe;
        if (do_restart) {
          const double l2pd = sqrt(dp), l2dd = sqrt(dd);
          const double rel_d = l2d / (1.0 + c_norm), rel_p = l2p / (1.0 + b_norm);
          const double ratio = (rel_p == 0.0) ? inf : rel_d / rel_p;
          if (l2pd > 1e-16 && l2dd > 1e-16 && l2pd < 1e12 && l2dd < 1e12 && ratio > 1e-8 && ratio < 1e8) {
            const double cw  = omega;
            const double err = log(l2dd) - log(l2pd) - log(cw);
            const double s   = err_sum * kISmoothB + err;
            err_sum          = s;
            const double de  = err - last_err;
            const double w   = bound_primal_weight<V>(exp(kKpB * err + kKiB * s + kKdB * de) * cw);
            omega            = w;
            last_err         = err;
            tau              = eta / w;
            sigma            = eta * w;
          } else {
            omega    = best_omega;
            err_sum  = 0.0;
            last_err = 0.0;
            tau      = eta / best_omega;
            sigma    = eta * best_omega;
          }
          const double g = fabs(log10(ratio));
          if (g < best_gap) {
            best_gap   = g;
            best_omega = omega;
          }
          const double rel_err = fmax(fmax(l2p / (1.0 + b_norm), l2d / (1.0 + c_norm)), gap / (1.0 + absobj));
          if (A.refine && (refine_active || rel_err <= 1e-2)) {
            // Recentre at the restart point (see PdlpSolver::recentre in pdlp_solver.hip).
            double sdp = 1.0 / fmax(l2p, 1e-14 * (1.0 + b_norm));
            double sdd = 1.0 / fmax(l2d, 1e-14 * (1.0 + c_norm));
            sdp        = fmax(sdp, 1.0);
            sdd        = fmax(sdd, 1.0);
            if (sdp * sdd * gap > 1.0) {
              const double f = sqrt(1.0 / (sdp * sdd * gap));
              sdp            = fmax(sdp * f, 1.0);
              sdd            = fmax(sdd * f, 1.0);
            }
            const double p_prev = rdp > 0.0 ? rdp : beta, d_prev = rdd > 0.0 ? rdd : gamma;
            const double wr     = (sdd / d_prev) / (sdp / p_prev);
            omega               = bound_primal_weight<V>(omega * wr);
            best_omega          = bound_primal_weight<V>(best_omega * wr);
            tau                 = eta / omega;
            sigma               = eta * omega;
            rdp = sdp, rdd = sdd;
            for (int j = tid; j < n; j += kBT) xstar[j] = fmin(fmax(xu[j], lbv[j]), ubv[j]);
            for (int i = tid; i < m; i += kBT) ystar[i] = yu[i];
            __syncthreads();
            double rr[kMaxRed];
            rows_pass(aptr, aidx, aorig, d.lga, ablk, st, BPlainEpi{xstar, ax64}, rr);
            rows_pass(tptr, tidx, torig, d.lgt, tblk, st, BPlainEpi{ystar, aty64}, rr);
            __syncthreads();
            for (int j = tid; j < n; j += kBT) {
              const double dc = Dc[j], x = xstar[j];
              c_w[j]          = (V)((sdd * (c[j] - aty64[j])) * dc);
              V2 bb;
              bb.x      = dc != 0.0 ? (V)((sdp * (lbv[j] - x)) / dc) : V(0);
              bb.y      = dc != 0.0 ? (V)((sdp * (ubv[j] - x)) / dc) : V(0);
              bnd_w[j]  = bb;
              xs[cur][j] = mn(mx(V(0), bb.x), bb.y);
              xhat[j]    = V(0);
              xanc[j]    = V(0);
              slack[j]   = V(0);
            }
            for (int i = tid; i < m; i += kBT) {
              const double r = Dr[i];
              V2 bb;
              bb.x       = (V)((sdp * (lc[i] - ax64[i])) * r);
              bb.y       = (V)((sdp * (uc[i] - ax64[i])) * r);
              cb_w[i]    = bb;
              const double kk = -sdd * ystar[i];
              kink[i]    = (V)(r != 0.0 ? kk / r : kk);
              ys[cur][i] = V(0);
              yhat[i]    = V(0);
              yanc[i]    = V(0);
            }
            beta          = 1.0;
            gamma         = 1.0;
            refine_active = true;
            ++recentrings;
          } else {
            for (int j = tid; j < n; j += kBT) {
              const V v  = xhat[j];
              xanc[j]    = v;
              xs[cur][j] = v;
            }
            for (int i = tid; i < m; i += kBT) {
              const V v  = yhat[i];
              yanc[i]    = v;
              ys[cur][i] = v;
            }
          }
          kr         = 0;
          last_trial = inf;
          restarted  = true;
          ++restarts;
        }
      }
      if (tid == 0) sc = StepScalars{tau, sigma, 0, 0};
      __syncthreads();

      // PDHG iterations up to the next check
      long long next = k + 1;
      while (!check_at(next)) ++next;
      periodic_prev = false;
      for (long long s = k; s < next; ++s) {
        const long long s2i     = s + 2;
        const bool should_major = ((s + 1) % kMajorB == 0) || (s2i % cond_major(s2i)) == 0;
        const bool periodic     = ((s + 1) % kMajorB) == 0;
        const bool post_restart = restarted && s == k;
        const bool fpe_step     = periodic || post_restart;
        const int ci = cur, nx = cur ^ 1;
        const int jj = (int)kr;
        zero_red(red);
#define CUOPT_BPRIMAL(MAJ, FP)                                                                              \
  do {                                                                                                        \
    const PrimalStepEpi<V, MAJ, FP> e{ys[ci], xs[ci], c_w, bnd_w, xanc, xbar, xs[nx], xhat, slack, aty, &sc, jj,   \
                                      refine_active};                                                         \
    if (vec_in_lds) rows_pass(tptr, tidx, tw, d.lgt, tblk, st, LdsGather<decltype(e)>{e, as_lds(ys[ci])}, red);  \
    else rows_pass(tptr, tidx, tw, d.lgt, tblk, st, GlobalGather<decltype(e)>{e, as_global(ys[ci])}, red);        \
  } while (0)
#define CUOPT_BDUAL(MAJ, FP)                                                                                 \
  do {                                                                                                        \
    const DualStepEpi<V, MAJ, FP> e{xbar, ys[ci], cb_w, yanc, ys[nx], yhat, ybar, &sc, jj,                         \
                                    refine_active ? (const V*)kink : nullptr};                                  \
    if (vec_in_lds) rows_pass(aptr, aidx, aw, d.lga, ablk, st, LdsGather<decltype(e)>{e, as_lds((const V*)xbar)}, red); \
    else rows_pass(aptr, aidx, aw, d.lga, ablk, st, GlobalGather<decltype(e)>{e, as_global((const V*)xbar)}, red); \
  } while (0)
        if (should_major) {
          if (fpe_step) CUOPT_BPRIMAL(true, true);
          else CUOPT_BPRIMAL(true, false);
        } else {
          if (fpe_step) CUOPT_BPRIMAL(false, true);
          else CUOPT_BPRIMAL(false, false);
        }
        __syncthreads();
        if (should_major) {
          if (fpe_step) CUOPT_BDUAL(true, true);
          else CUOPT_BDUAL(true, false);
        } else {
          if (fpe_step) CUOPT_BDUAL(false, true);
          else CUOPT_BDUAL(false, false);
        }
#undef CUOPT_BPRIMAL
#undef CUOPT_BDUAL
        if (fpe_step) {
          wg_sum<1>(red, lds);
          const double ny = red[0];
          zero_red(red);
          rows_pass(tptr, tidx, tw, d.lgt, tblk, st, FpeEpi<V>{ybar, aty, xbar, xs[ci]}, red);
          wg_sum<2>(red, lds);
          if (post_restart) {
            fpe0_t[0] = ny;
            fpe0_t[1] = red[0];
            fpe0_t[2] = red[1];
          } else {
            fpe_t[0] = ny;
            fpe_t[1] = red[0];
            fpe_t[2] = red[1];
          }
        } else {
          __syncthreads();
        }
        if (periodic) periodic_prev = true;
        cur = nx;
        ++kr;
      }
      k              = next;
      restarted_prev = restarted;
    }
  }
  if (tid == 0) {
    BatchOut o;
    o.l2p = l2p;
    o.l2d = l2d;
    o.pobj = pobj;
    o.dobj = dobj;
    o.gap = gap;
    o.absobj = absobj;
    o.b_norm = b_norm;
    o.c_norm = c_norm;
    o.eta = eta;
    o.omega = omega;
    o.sigma_max_sq = s2;
    o.beta = beta;
    o.gamma = gamma;
    o.status = status;
    o.iters = (int)k;
    o.restarts = restarts;
    o.checks = checks;
    o.power_iters = piters;
    o.recentrings = recentrings;
    o.t0 = t_start;
    o.t1 = wall_clock64();
    A.out[A.order[blockIdx.x]] = o;
  }
}

// f(q) for q in [0, count) on up to 16 host threads (q strided over the threads);
// the first exception thrown by any call is rethrown after all have joined.
template <class F>
void host_parallel_for(int count, F f)
{
  const int nt = std::max(1, std::min({16, count, (int)std::thread::hardware_concurrency()}));
  std::exception_ptr err;
  std::mutex mu;
  auto work = [&](int t) {
    try {
      for (int q = t; q < count; q += nt) f(q);
    } catch (...) {
      std::lock_guard<std::mutex> lock(mu);
      if (!err) err = std::current_exception();
    }
  };
  std::vector<std::thread> th;
  for (int t = 1; t < nt; ++t) th.emplace_back(work, t);
  work(0);
  for (auto& x : th) x.join();
  if (err) std::rethrow_exception(err);
}

template <typename V>
std::vector<SolveResult> run_batch(const std::vector<const LpHost*>& lps,
                                   const std::vector<int>& active,
                                   const std::vector<std::vector<double>>& lcs,
                                   const std::vector<std::vector<double>>& ucs,
                                   const PdlpSettings& st,
                                   std::vector<SolveResult> R)
{
  using V2 = typename Vec2<V>::type;
  const auto t0 = std::chrono::steady_clock::now();
  const int nb  = (int)active.size();
  std::vector<BatchDesc> desc(lps.size());
  long long tot_nnz = 0;
  int tot_m = 0, tot_n = 0, max_m = 0, tot_ap = 0, tot_tp = 0;
  for (int a : active) {
    const LpHost& lp = *lps[a];
    BatchDesc& d     = desc[a];
    d.m = lp.m;
    d.n = lp.n;
    d.nnz = lp.rowptr[lp.m];
    d.row0 = tot_m;
    d.col0 = tot_n;
    d.aptr0 = tot_ap;
    d.tptr0 = tot_tp;
    d.nz0 = tot_nnz;
    d.lga = plan_lanes(lp.m, d.nnz);
    d.lgt = plan_lanes(lp.n, d.nnz);
    d.obj_sf  = lp.maximize ? -1.0 : 1.0;
    d.obj_off = lp.maximize ? -lp.offset : lp.offset;
    tot_m += lp.m;
    tot_n += lp.n;
    tot_ap += lp.m + 1;
    tot_tp += lp.n + 1;
    tot_nnz += d.nnz;
    max_m = std::max(max_m, lp.m);
  }
  // host arrays (rows sorted by column, as problem_t keeps them)
  // Left uninitialised (every element is written below): zero-filling half a
  // gigabyte on one thread took longer than the threaded copy itself.
  std::unique_ptr<int[]> h_ptr(new int[tot_ap]), h_idx(new int[tot_nnz]);
  std::unique_ptr<double[]> h_val(new double[tot_nnz]), h_c(new double[tot_n]), h_lb(new double[tot_n]),
    h_ub(new double[tot_n]), h_lc(new double[tot_m]), h_uc(new double[tot_m]);
  auto fill_one = [&](int q) {
    {
      const int a      = active[q];
      const LpHost& lp = *lps[a];
      const BatchDesc& d = desc[a];
      std::copy(lp.rowptr.begin(), lp.rowptr.begin() + lp.m + 1, h_ptr.get() + d.aptr0);
      std::copy(lp.colidx.begin(), lp.colidx.begin() + d.nnz, h_idx.get() + d.nz0);
      std::copy(lp.vals.begin(), lp.vals.begin() + d.nnz, h_val.get() + d.nz0);
      for (int i = 0; i < lp.m; ++i) {
        const int b = lp.rowptr[i], e = lp.rowptr[i + 1];
        bool sorted = true;
        for (int k = b + 1; k < e && sorted; ++k) sorted = lp.colidx[k - 1] <= lp.colidx[k];
        if (sorted) continue;
        std::vector<std::pair<int, double>> row;
        for (int k = b; k < e; ++k) row.emplace_back(lp.colidx[k], lp.vals[k]);
        std::stable_sort(row.begin(), row.end(), [](const auto& p, const auto& q) { return p.first < q.first; });
        for (int k = b; k < e; ++k) {
          h_idx[d.nz0 + k] = row[k - b].first;
          h_val[d.nz0 + k] = row[k - b].second;
        }
      }
      for (int j = 0; j < lp.n; ++j) {
        h_c[d.col0 + j]  = lp.maximize ? -lp.c[j] : lp.c[j];
        h_lb[d.col0 + j] = lp.lb[j];
        h_ub[d.col0 + j] = lp.ub[j];
      }
      std::copy(lcs[a].begin(), lcs[a].end(), h_lc.get() + d.row0);
      std::copy(ucs[a].begin(), ucs[a].end(), h_uc.get() + d.row0);
    }
  };
  // row blocks of A and A^T for each LP
  std::vector<std::vector<int>> ablocks(lps.size()), tblocks(lps.size());
  {
    auto plan = [&](int t, int nt) {
      std::vector<int> cnt;
      for (int q = t; q < nb; q += nt) {
        const int a      = active[q];
        const LpHost& lp = *lps[a];
        plan_row_blocks(lp.m, desc[a].lga, [&](int r) { return lp.rowptr[r + 1] - lp.rowptr[r]; }, ablocks[a]);
        cnt.assign(lp.n, 0);
        for (int k = 0; k < desc[a].nnz; ++k) ++cnt[lp.colidx[k]];
        plan_row_blocks(lp.n, desc[a].lgt, [&](int j) { return cnt[j]; }, tblocks[a]);
      }
    };
    const int nt = std::max(1, std::min<int>(16, (int)std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (int t = 0; t < nt; ++t) th.emplace_back(plan, t, nt);
    for (auto& t : th) t.join();
  }
  std::vector<int> h_blk;
  for (int a : active) {
    BatchDesc& d = desc[a];
    d.ablk0      = (int)h_blk.size();
    d.nablk      = (int)ablocks[a].size() - 1;
    h_blk.insert(h_blk.end(), ablocks[a].begin(), ablocks[a].end());
    d.tblk0 = (int)h_blk.size();
    d.ntblk = (int)tblocks[a].size() - 1;
    h_blk.insert(h_blk.end(), tblocks[a].begin(), tblocks[a].end());
  }

  // processing order: largest first
  std::vector<int> order(active);
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return desc[x].nnz > desc[y].nnz; });
  std::vector<double> z0(std::max(max_m, 1));
  {
    std::mt19937 gen(1);
    std::normal_distribution<double> dist(0.0, 1.0);
    for (double& v : z0) v = dist(gen);
  }

  hipStream_t s;
  CUOPT_HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  struct StreamGuard {
    hipStream_t s;
    ~StreamGuard() { (void)hipStreamDestroy(s); }
  } guard{s};
  const size_t M = std::max(tot_m, 1), N = std::max(tot_n, 1), Z = std::max<long long>(tot_nnz, 1);
  DevArena arena;  // every working array below: one allocation
  auto d_desc = arena.take<BatchDesc>(std::max<size_t>(desc.size(), 1));
  auto d_blk = arena.take<int>(std::max<size_t>(h_blk.size(), 1));
  // The nonzero arrays read by row passes carry kNnzPad elements of padding:
  // a block's 16-byte copy may read up to 3 dwords past its last nonzero.
  const size_t Zp = Z + kNnzPad;
  auto d_order = arena.take<int>(std::max(nb, 1));
  auto a_ptr = arena.take<int>(std::max(tot_ap, 1));
  auto a_idx = arena.take<int>(Zp);
  auto t_ptr = arena.take<int>(std::max(tot_tp, 1));
  auto t_idx = arena.take<int>(Zp);
  auto t_pos = arena.take<int>(Z);
  auto t_fill = arena.take<int>(N);
  auto a_orig = arena.take<double>(Zp);
  auto t_orig = arena.take<double>(Zp);
  auto c = arena.take<double>(N);
  auto lb = arena.take<double>(N);
  auto ub = arena.take<double>(N);
  auto lc = arena.take<double>(M);
  auto uc = arena.take<double>(M);
  auto Dr = arena.take<double>(M);
  auto Dc = arena.take<double>(N);
  auto rowv = arena.take<double>(M);
  auto colv = arena.take<double>(N);
  auto a_w = arena.take<V>(Zp);
  auto t_w = arena.take<V>(Zp);
  auto c_w = arena.take<V>(N);
  auto x0 = arena.take<V>(N);
  auto x1 = arena.take<V>(N);
  auto y0 = arena.take<V>(M);
  auto y1 = arena.take<V>(M);
  auto xa = arena.take<V>(N);
  auto ya = arena.take<V>(M);
  auto xbar = arena.take<V>(N);
  auto ybar = arena.take<V>(M);
  auto xhat = arena.take<V>(N);
  auto yhat = arena.take<V>(M);
  auto slack = arena.take<V>(N);
  auto aty = arena.take<V>(N);
  auto bnd_w = arena.take<V2>(N);
  auto cb_w = arena.take<V2>(M);
  auto cb_orig = arena.take<double2>(M);
  auto xu = arena.take<double>(N);
  auto yu = arena.take<double>(M);
  auto su = arena.take<double>(N);
  auto pz = arena.take<double>(M);
  auto pz2 = arena.take<double>(M);
  auto patq = arena.take<double>(N);
  auto dz0 = arena.take<double>(z0.size());
  const bool refine = std::is_same<V, float>::value && st.refinement != 0;
  auto xstar = arena.take<double>(refine ? N : 1);
  auto ystar = arena.take<double>(refine ? M : 1);
  auto ax64 = arena.take<double>(refine ? M : 1);
  auto aty64 = arena.take<double>(refine ? N : 1);
  auto kink = arena.take<V>(refine ? M : 1);
  auto out = arena.take<BatchOut>(std::max<size_t>(lps.size(), 1));
  auto lz = arena.take<double>(std::is_same<V, float>::value ? std::max<size_t>(lps.size(), 1) * kLzStride : 1);
  arena.commit();
  d_desc.upload(desc, s);
  if (!h_blk.empty()) d_blk.upload(h_blk, s);
  d_order.upload(order, s);
  dz0.upload(z0, s);
  // The LPs' arrays are concatenated on host threads in kGroups contiguous groups
  // (by position in `active`, which is also their order in the device layout) by a
  // helper thread, while this thread uploads each group as soon as it is complete:
  // the pageable uploads overlap the copying of the next group.
  {
    constexpr int kGroups = 8;
    const int ng = std::max(1, std::min(kGroups, nb));
    auto gbeg = [&](int g) { return (int)((long long)nb * g / ng); };
    std::mutex mu;
    std::condition_variable cv;
    int done = 0;  // groups complete
    std::exception_ptr err;
    std::thread filler([&]() {
      try {
        for (int g = 0; g < ng; ++g) {
          const int q0 = gbeg(g), q1 = gbeg(g + 1);
          host_parallel_for(q1 - q0, [&](int i) { fill_one(q0 + i); });
          std::lock_guard<std::mutex> lock(mu);
          ++done;
          cv.notify_one();
        }
      } catch (...) {
        std::lock_guard<std::mutex> lock(mu);
        err  = std::current_exception();
        done = ng;
        cv.notify_one();
      }
    });
    try {
      for (int g = 0; g < ng && nb > 0; ++g) {
        {
          std::unique_lock<std::mutex> lock(mu);
          cv.wait(lock, [&]() { return done > g; });
          if (err) break;
        }
        const BatchDesc& f = desc[active[gbeg(g)]];
        const BatchDesc& l = desc[active[gbeg(g + 1) - 1]];
        const size_t ap0 = f.aptr0, ap1 = l.aptr0 + l.m + 1, z0g = f.nz0, z1g = l.nz0 + l.nnz;
        const size_t c0 = f.col0, c1 = l.col0 + l.n, r0 = f.row0, r1 = l.row0 + l.m;
        a_ptr.upload_at(ap0, h_ptr.get() + ap0, ap1 - ap0, s);
        a_idx.upload_at(z0g, h_idx.get() + z0g, z1g - z0g, s);
        a_orig.upload_at(z0g, h_val.get() + z0g, z1g - z0g, s);
        c.upload_at(c0, h_c.get() + c0, c1 - c0, s);
        lb.upload_at(c0, h_lb.get() + c0, c1 - c0, s);
        ub.upload_at(c0, h_ub.get() + c0, c1 - c0, s);
        lc.upload_at(r0, h_lc.get() + r0, r1 - r0, s);
        uc.upload_at(r0, h_uc.get() + r0, r1 - r0, s);
      }
    } catch (...) {
      filler.join();
      throw;
    }
    filler.join();
    if (err) std::rethrow_exception(err);
  }
  // Marks the end of the uploads, so that the host staging can be released while
  // the persistent launch runs.
  hipEvent_t uploaded = nullptr;
  struct UploadGuard {
    hipEvent_t* e;
    ~UploadGuard()
    {
      if (*e) (void)hipEventDestroy(*e);
    }
  } up_guard{&uploaded};
  CUOPT_HIP_CHECK(hipEventCreateWithFlags(&uploaded, hipEventDisableTiming));
  CUOPT_HIP_CHECK(hipEventRecord(uploaded, s));

  BatchArgs<V> args{};
  args.desc = d_desc.get();
  args.blk  = d_blk.get();
  args.order = d_order.get();
  args.a_ptr = a_ptr.get();
  args.a_idx = a_idx.get();
  args.a_orig = a_orig.get();
  args.t_ptr = t_ptr.get();
  args.t_idx = t_idx.get();
  args.t_pos = t_pos.get();
  args.t_orig = t_orig.get();
  args.t_fill = t_fill.get();
  args.a_w = a_w.get();
  args.t_w = t_w.get();
  args.c = c.get();
  args.lb = lb.get();
  args.ub = ub.get();
  args.lc = lc.get();
  args.uc = uc.get();
  args.Dr = Dr.get();
  args.Dc = Dc.get();
  args.rowv = rowv.get();
  args.colv = colv.get();
  args.c_w = c_w.get();
  args.bnd_w = bnd_w.get();
  args.cb_w = cb_w.get();
  args.cb_orig = cb_orig.get();
  args.x0 = x0.get();
  args.x1 = x1.get();
  args.y0v = y0.get();
  args.y1 = y1.get();
  args.xa = xa.get();
  args.ya = ya.get();
  args.xbar = xbar.get();
  args.ybar = ybar.get();
  args.xhat = xhat.get();
  args.yhat = yhat.get();
  args.slack = slack.get();
  args.aty = aty.get();
  args.xu = xu.get();
  args.yu = yu.get();
  args.su = su.get();
  args.pz = pz.get();
  args.pz2 = pz2.get();
  args.patq = patq.get();
  args.z0 = dz0.get();
  args.out = out.get();
  args.lz  = lz.get();
  args.tol_ap = st.abs_primal;
  args.tol_rp = st.rel_primal;
  args.tol_ad = st.abs_dual;
  args.tol_rd = st.rel_dual;
  args.tol_ag = st.abs_gap;
  args.tol_rg = st.rel_gap;
  args.iter_limit = st.iteration_limit;
  args.tick_limit = 0;
  args.refine     = refine ? 1 : 0;
  args.xstar      = xstar.get();
  args.ystar      = ystar.get();
  args.ax64       = ax64.get();
  args.aty64      = aty64.get();
  args.kink       = kink.get();
  if (std::isfinite(st.time_limit)) {
    int dev = 0, khz = 0;
    CUOPT_HIP_CHECK(hipGetDevice(&dev));
    CUOPT_HIP_CHECK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
    const double ticks = std::max(1.0, st.time_limit * 1e3 * (double)khz);
    args.tick_limit    = ticks >= 1.8e19 ? 0ULL : (unsigned long long)ticks;
  }
  {
    // LDS: the row-block stage and, at the end, each LP's gathered vectors when
    // they fit; as much as the per-workgroup limit and a CU's LDS divided by
    // kWgPerCu allow.
    int dev = 0, per_block = 0, per_cu = 0;
    CUOPT_HIP_CHECK(hipGetDevice(&dev));
    CUOPT_HIP_CHECK(hipDeviceGetAttribute(&per_block, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    CUOPT_HIP_CHECK(hipDeviceGetAttribute(&per_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev));
    const long long static_lds = 4096;  // the kernel's own __shared__ arrays, rounded up
    long long cap = std::min<long long>(per_block, per_cu / kWgPerCu) - static_lds;
    if (cap < kStageBytes) throw std::runtime_error("batch solve: the device has too little LDS per workgroup");
    args.lds_bytes = (int)cap;  // two workgroups per CU either way (their register budget)
  }
  // HIP events around the persistent launch: its device time is the batch's loop time.
  hipEvent_t ev[2] = {nullptr, nullptr};
  struct EventGuard {
    hipEvent_t* e;
    ~EventGuard()
    {
      for (int i = 0; i < 2; ++i)
        if (e[i]) (void)hipEventDestroy(e[i]);
    }
  } ev_guard{ev};
  CUOPT_HIP_CHECK(hipEventCreate(&ev[0]));
  CUOPT_HIP_CHECK(hipEventCreate(&ev[1]));
  CUOPT_HIP_CHECK(hipEventRecord(ev[0], s));
  if (nb > 0) {
    hipLaunchKernelGGL(batch_pdlp_kernel<V>, dim3(nb), dim3(kBT), (unsigned)args.lds_bytes, s, args);
    CUOPT_HIP_CHECK(hipGetLastError());
  }
  CUOPT_HIP_CHECK(hipEventRecord(ev[1], s));
  // Host work while the launch runs: release the input staging once its uploads
  // are done, page in the download buffers and size every result's arrays, so
  // that the downloads and the assembly after the launch only copy.
  CUOPT_HIP_CHECK(hipEventSynchronize(uploaded));
  for (auto* p : {&h_ptr, &h_idx}) p->reset();
  for (auto* p : {&h_val, &h_c, &h_lb, &h_ub, &h_lc, &h_uc}) p->reset();
  std::vector<BatchOut> h_out(std::max<size_t>(lps.size(), 1));
  std::unique_ptr<double[]> h_xu(new double[N]), h_yu(new double[M]), h_su(new double[N]);
  host_parallel_for(16, [&](int t) {
    auto touch = [&](double* p, size_t len) {
      const size_t b = len * t / 16, e = len * (t + 1) / 16;
      std::fill(p + b, p + e, 0.0);
    };
    touch(h_xu.get(), N);
    touch(h_yu.get(), M);
    touch(h_su.get(), N);
  });
  host_parallel_for((int)active.size(), [&](int q) {
    const int a      = active[q];
    const LpHost& lp = *lps[a];
    R[a].x.resize(lp.n);
    R[a].y.resize(lp.m);
    R[a].rc.resize(lp.n);
  });
  out.download(h_out.data(), h_out.size(), s);
  xu.download(h_xu.get(), N, s);
  yu.download(h_yu.get(), M, s);
  su.download(h_su.get(), N, s);
  CUOPT_HIP_CHECK(hipStreamSynchronize(s));
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  float kernel_ms   = 0.f;
  CUOPT_HIP_CHECK(hipEventElapsedTime(&kernel_ms, ev[0], ev[1]));
  double tick_hz = 1.0;
  unsigned long long first_tick = ~0ULL;
  {
    int dev = 0, khz = 0;
    CUOPT_HIP_CHECK(hipGetDevice(&dev));
    CUOPT_HIP_CHECK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
    tick_hz = std::max(1, khz) * 1e3;
    for (int a : active) first_tick = std::min(first_tick, h_out[a].t0);
  }
  host_parallel_for((int)active.size(), [&](int q) {
    const int a        = active[q];
    const LpHost& lp   = *lps[a];
    const BatchDesc& d = desc[a];
    const BatchOut& o  = h_out[a];
    SolveResult& r     = R[a];
    r.status = o.status;
    std::copy(h_xu.get() + d.col0, h_xu.get() + d.col0 + lp.n, r.x.begin());
    std::copy(h_yu.get() + d.row0, h_yu.get() + d.row0 + lp.m, r.y.begin());
    std::copy(h_su.get() + d.col0, h_su.get() + d.col0 + lp.n, r.rc.begin());
    if (lp.maximize) {  // solve.cu:583-588
      for (double& v : r.y) v = -v;
      for (double& v : r.rc) v = -v;
    }
    r.primal_objective                  = o.pobj;
    r.dual_objective                    = o.dobj;
    r.solve_time                        = secs;
    r.info[CUOPT_AMD_INFO_STEPS]        = o.iters;
    r.info[CUOPT_AMD_INFO_L2_PRIMAL]    = o.l2p;
    r.info[CUOPT_AMD_INFO_REL_PRIMAL]   = o.l2p / (1.0 + o.b_norm);
    r.info[CUOPT_AMD_INFO_L2_DUAL]      = o.l2d;
    r.info[CUOPT_AMD_INFO_REL_DUAL]     = o.l2d / (1.0 + o.c_norm);
    r.info[CUOPT_AMD_INFO_PRIMAL_OBJ]   = o.pobj;
    r.info[CUOPT_AMD_INFO_DUAL_OBJ]     = o.dobj;
    r.info[CUOPT_AMD_INFO_GAP]          = o.gap;
    r.info[CUOPT_AMD_INFO_REL_GAP]      = o.gap / (1.0 + o.absobj);
    r.info[CUOPT_AMD_INFO_SOLVE_TIME]   = secs;
    r.info[CUOPT_AMD_INFO_LOOP_TIME]    = 1e-3 * kernel_ms;  // the whole persistent launch
    r.info[CUOPT_AMD_INFO_ETA0]         = o.eta;
    r.info[CUOPT_AMD_INFO_OMEGA0]       = 1.0;
    r.info[CUOPT_AMD_INFO_OMEGA]        = o.omega;
    r.info[CUOPT_AMD_INFO_RESTARTS]     = o.restarts;
    r.info[CUOPT_AMD_INFO_POWER_ITERS]  = o.power_iters;
    r.info[CUOPT_AMD_INFO_SIGMA_MAX_SQ] = o.sigma_max_sq;
    r.info[CUOPT_AMD_INFO_BETA]         = o.beta;
    r.info[CUOPT_AMD_INFO_GAMMA]        = o.gamma;
    r.info[CUOPT_AMD_INFO_PRECISION]    = std::is_same<V, double>::value ? 0 : 1;
    r.info[CUOPT_AMD_INFO_CHECKS]       = o.checks;
    r.info[CUOPT_AMD_INFO_REFINE_ROUNDS] = o.recentrings;
    r.info[CUOPT_AMD_INFO_LP_TIME]      = (double)(o.t1 - o.t0) / tick_hz;
    r.info[CUOPT_AMD_INFO_LP_START]     = (double)(o.t0 - first_tick) / tick_hz;
  });
  return R;
}

}  // namespace

std::vector<SolveResult> batch_solve(const std::vector<const LpHost*>& lps, const PdlpSettings& st)
{
  const auto t0 = std::chrono::steady_clock::now();
  std::vector<SolveResult> R(lps.size());
  std::vector<int> active;
  std::vector<std::vector<double>> lcs(lps.size()), ucs(lps.size());
  std::vector<char> on_device(lps.size(), 0);
  host_parallel_for((int)lps.size(), [&](int a) {
    const LpHost& lp = *lps[a];
    constraint_bounds(lp, lcs[a], ucs[a]);
    bool crossing = false;
    for (int j = 0; j < lp.n && !crossing; ++j) crossing = !(lp.ub[j] >= lp.lb[j]);
    for (int i = 0; i < lp.m && !crossing; ++i) crossing = !(ucs[a][i] >= lcs[a][i]);
    if (crossing || lp.m == 0) {  // as solve_lp: infeasible bounds, or PDLP without constraints
      R[a].status = crossing ? CUOPT_TERIMINATION_STATUS_INFEASIBLE : CUOPT_TERIMINATION_STATUS_NUMERICAL_ERROR;
      R[a].x.assign(lp.n, 0.0);
      R[a].y.assign(lp.m, 0.0);
      R[a].rc.assign(lp.n, 0.0);
      R[a].solve_time = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      return;
    }
    on_device[a] = 1;
  });
  for (size_t a = 0; a < lps.size(); ++a)
    if (on_device[a]) active.push_back((int)a);
  // LPs answered on the host above keep their results if the device part fails.
  std::vector<SolveResult> keep = R;
  try {
    if (st.precision == CUOPT_AMD_PRECISION_FP32_MIXED)
      return run_batch<float>(lps, active, lcs, ucs, st, std::move(R));
    return run_batch<double>(lps, active, lcs, ucs, st, std::move(R));
  } catch (const std::exception& e) {
    const std::string msg = e.what();
    int code              = CUOPT_RUNTIME_ERROR;
    if (dynamic_cast<const validation_error*>(&e)) code = CUOPT_VALIDATION_ERROR;
    if (dynamic_cast<const std::bad_alloc*>(&e) || msg.find("out of memory") != std::string::npos)
      code = CUOPT_OUT_OF_MEMORY;
    for (int a : active) {
      keep[a].error         = code;
      keep[a].error_message = msg;
      keep[a].status        = CUOPT_TERIMINATION_STATUS_NO_TERMINATION;
    }
    return keep;
  }
}

}  // namespace cuopt_amd
