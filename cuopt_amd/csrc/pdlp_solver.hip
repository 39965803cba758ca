// pdlp_solver.hip -- MI355X PDLP (Stable3) LP solver: setup, PDHG loop, termination.
//
// Host orchestration follows pdlp_solver_t::run_solver (cpp/src/pdlp/pdlp.cu:2013-2458):
//   problem_t construction   (mip_heuristics/problem/problem.cu:47-100) ....... host CSR/CSC
//   Ruiz + Pock-Chambolle     (initial_scaling.cu:151-364) ................... tiled max / sum passes
//   scale_problem + bound/objective rescaling (initial_scaling.cu:416-568) .. elementwise kernels
//   power iteration step size (pdlp.cu:2587-2739) ............................ device loop, early exit
//   PDHG reflected step + Halpern (pdhg.cu:785-1030, pdlp.cu:2505-2563) ...... 2 fused tiled kernels / iteration
//   fixed-point error         (pdlp.cu:1748-1855) ............................. 1 tiled kernel on FPE steps
//   termination on the unscaled problem (convergence_information.cu:326-810) 2 fused tiled kernels / check
//   cuPDLPx restart + PID primal weight (pdlp_restart_strategy.cu:688-985) .. host scalar logic, copy kernels
//
// Iterations between two termination checks are issued back to back on one
// stream with no host synchronisation; the host reads one small record per check.
#include <algorithm>
#include <climits>
#include <cstdio>
#include <chrono>
#include <thread>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <random>
#include <string>
#include <type_traits>

#include <rocprim/device/device_radix_sort.hpp>

#include "mps_io.hpp"
#include "presolve.hpp"

#include "common.hpp"
#include "lp_types.hpp"
#include "pdlp_kernels.hpp"
#include "tile_plan.hpp"

namespace cuopt_amd {

// ---------------------------------------------------------------------------
// Generic kernels
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(1024) void reduce_partials_kernel(const double* __restrict__ part,
                                                               int nblocks,
                                                               int nred,
                                                               double* __restrict__ out,
                                                               const double* skip)
{
  __shared__ double lds[1024];
  if (skip && *skip != 0.0) return;
  for (int k = 0; k < nred; ++k) {
    double s = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += blockDim.x) s += part[(size_t)b * nred + k];
    lds[threadIdx.x] = s;
    __syncthreads();
    for (int off = blockDim.x / 2; off > 0; off >>= 1) {
      if ((int)threadIdx.x < off) lds[threadIdx.x] += lds[threadIdx.x + off];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[k] = lds[0];
    __syncthreads();
  }
}

// per_constraint_residual (convergence_information.cu:390-416, 453-470, utils.cuh:576-590):
// the largest r_i - rel * w_i of a signed residual vector, with w_i the larger
// finite bound magnitude of row i (cb given, combine_finite_abs_bounds) or the
// signed c_i, and 0 as the neutral element.  absolute != 0 takes |r_i| and |c_i|
// instead (an extension).  One workgroup; the order of the max is fixed.
__global__ __launch_bounds__(1024) void linf_rel_kernel(const double* __restrict__ r,
                                                        const double2* __restrict__ cb,
                                                        const double* __restrict__ c,
                                                        int n,
                                                        double rel,
                                                        int absolute,
                                                        double* __restrict__ out)
{
  __shared__ double lds[1024];
  double v = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    double w = 0.0;
    if (cb) {
      const double2 b = cb[i];
      if (isfinite(b.y)) w = fmax(w, fabs(b.y));
      if (isfinite(b.x)) w = fmax(w, fabs(b.x));
    } else {
      w = absolute ? fabs(c[i]) : c[i];
    }
    const double ri = absolute ? fabs(r[i]) : r[i];
    v = fmax(v, ri - rel * w);
  }
  lds[threadIdx.x] = v;
  __syncthreads();
  for (int off = blockDim.x / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) lds[threadIdx.x] = fmax(lds[threadIdx.x], lds[threadIdx.x + off]);
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = lds[0];
}

template <class F>
__global__ __launch_bounds__(kBlock) void elementwise_kernel(int n, double* __restrict__ red_part, F f)
{
  __shared__ double lds[kBlock / kWave][kMaxRed];
  if (f.skip()) return;
  double red[kMaxRed];
#pragma unroll
  for (int k = 0; k < kMaxRed; ++k) red[k] = 0.0;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) f(i, red);
  if constexpr (F::NRED > 0) block_sum_store<F::NRED>(red, red_part + (size_t)blockIdx.x * F::NRED, lds);
}

// Power iteration finalisers (one workgroup).
__global__ __launch_bounds__(1024) void power_finalize_s2(const double* __restrict__ part, int nb, PowerState* st)
{
  __shared__ double lds[1024];
  if (st->done != 0.0) return;
  double s = 0.0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) s += part[b];
  lds[threadIdx.x] = s;
  __syncthreads();
  for (int off = blockDim.x / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) lds[threadIdx.x] += lds[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) st->s2 = lds[0];
}

__global__ __launch_bounds__(1024) void power_finalize_res(const double* __restrict__ part,
                                                           int nb,
                                                           PowerState* st,
                                                           double tol)
{
  __shared__ double lds[2][1024];
  if (st->done != 0.0) return;
  double s0 = 0.0, s1 = 0.0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) {
    s0 += part[2 * (size_t)b];
    s1 += part[2 * (size_t)b + 1];
  }
  lds[0][threadIdx.x] = s0;
  lds[1][threadIdx.x] = s1;
  __syncthreads();
  for (int off = blockDim.x / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) {
      lds[0][threadIdx.x] += lds[0][threadIdx.x + off];
      lds[1][threadIdx.x] += lds[1][threadIdx.x + off];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double res = sqrt(lds[0][0]);
    st->res          = res;
    st->iters += 1.0;
    if (res < tol)
      st->done = 1.0;
    else
      st->nrm = sqrt(lds[1][0]);
  }
}

// Problems of at most kPowerSmallRows rows end a power step in one workgroup:
// s2 from the A-side partials (in power_finalize_s2's order), then the residual
// ||znew - s2 q|| and ||znew|| over the rows, and the state update of
// power_finalize_res.  One launch instead of three: on small problems the
// launches, not the arithmetic, set the time of a step.
constexpr int kPowerSmallRows = 16384;

__global__ __launch_bounds__(1024) void power_step_small(const double* __restrict__ part, int nb,
                                                         const double* __restrict__ zold,
                                                         const double* __restrict__ znew, int m, PowerState* st,
                                                         double tol)
{
  __shared__ double lds[2][1024];
  if (st->done != 0.0) return;
  const int tid = threadIdx.x;
  double s = 0.0;
  for (int b = tid; b < nb; b += blockDim.x) s += part[b];
  lds[0][tid] = s;
  __syncthreads();
  for (int off = blockDim.x / 2; off > 0; off >>= 1) {
    if (tid < off) lds[0][tid] += lds[0][tid + off];
    __syncthreads();
  }
  const double s2  = lds[0][0];
  const double nrm = st->nrm;
  __syncthreads();
  double r2 = 0.0, z2 = 0.0;
  for (int i = tid; i < m; i += blockDim.x) {
    const double q = zold[i] / nrm;
    const double r = q * -s2 + znew[i];
    r2 += r * r;
    z2 += znew[i] * znew[i];
  }
  lds[0][tid] = r2;
  lds[1][tid] = z2;
  __syncthreads();
  for (int off = blockDim.x / 2; off > 0; off >>= 1) {
    if (tid < off) {
      lds[0][tid] += lds[0][tid + off];
      lds[1][tid] += lds[1][tid + off];
    }
    __syncthreads();
  }
  if (tid == 0) {
    st->s2           = s2;
    const double res = sqrt(lds[0][0]);
    st->res          = res;
    st->iters += 1.0;
    if (res < tol)
      st->done = 1.0;
    else
      st->nrm = sqrt(lds[1][0]);
  }
}

// ---------------------------------------------------------------------------
// Elementwise functors
// ---------------------------------------------------------------------------

// Lanczos steps on A~ A~^T (step size of FP32-mixed solves).  The
// coefficients stay on the device (ab[2j] = alpha_j, ab[2j+1] = beta_j^2), so
// the host reads them once per group of steps.
struct LanczosAlphaF {  // w = z - beta_{j-1} q_{j-1}; red0 = q_j . w
  static constexpr int NRED = 1;
  const double* z;
  const double* q;
  const double* qprev;
  const double* ab;
  int j;
  double* w;
  __device__ bool skip() const { return false; }
  __device__ void operator()(int i, double (&red)[kMaxRed]) const
  {
    const double bp = j > 0 ? sqrt(ab[2 * j - 1]) : 0.0;
    const double v  = z[i] - bp * qprev[i];
    w[i]            = v;
    red[0] += q[i] * v;
  }
};
struct LanczosOrthF {  // w -= alpha_j q_j; red0 = ||w||^2
  static constexpr int NRED = 1;
  const double* q;
  const double* ab;
  int j;
  double* w;
  __device__ bool skip() const { return false; }
  __device__ void operator()(int i, double (&red)[kMaxRed]) const
  {
    const double v = w[i] - ab[2 * j] * q[i];
    w[i]           = v;
    red[0] += v * v;
  }
};
struct LanczosNextF {  // q_{j-1} = q_j; q_{j+1} = w / beta_j (left as is once beta_j is zero)
  static constexpr int NRED = 0;
  double* q;
  double* qprev;
  const double* w;
  const double* ab;
  int j;
  __device__ bool skip() const { return !(ab[2 * j + 1] > 0.0); }
  __device__ void operator()(int i, double (&)[kMaxRed]) const
  {
    qprev[i] = q[i];
    q[i]     = w[i] / sqrt(ab[2 * j + 1]);
  }
};
struct HashStartF {  // deterministic start vector in [-1, 1) (integer hash of the index)
  static constexpr int NRED = 0;
  double* v;
  __device__ bool skip() const { return false; }
  __device__ void operator()(int i, double (&)[kMaxRed]) const { v[i] = hashed_start(i); }
};
struct ScaleByF {  // v *= s
  static constexpr int NRED = 0;
  double* v;
  double s;
  __device__ bool skip() const { return false; }
  __device__ void operator()(int i, double (&)[kMaxRed]) const { v[i] *= s; }
};

struct SumSqF {  // red0 = sum v^2
  static constexpr int NRED = 1;
  const double* v;
  __device__ bool skip() const { return false; }
  __device__ void operator()(int i, double (&red)[kMaxRed]) const { red[0] += v[i] * v[i]; }
};

struct PowerResF {  // residual of the power step and ||znew||^2
  static constexpr int NRED = 2;
  const double* zold;
  const double* znew;
  const PowerState* st;
  __device__ bool skip() const { return st->done != 0.0; }
  __device__ void operator()(int i, double (&red)[kMaxRed]) const
  {
    const double q = zold[i] / st->nrm;
    const double r = q * -st->s2 + znew[i];
    red[0] += r * r;
    red[1] += znew[i] * znew[i];
  }
};

struct FillF {
  static constexpr int NRED = 0;
  double* v;
  double val;
  __device__ bool skip() const { return false; }
  __device__ void operator()(int i, double (&)[kMaxRed]) const { v[i] = val; }
};

struct DivSqrtF {  // a_divides_sqrt_b_bounded (utils.cuh:154-161)
  static constexpr int NRED = 0;
  double* d;
  const double* v;
  __device__ bool skip() const { return false; }
  __device__ void operator()(int i, double (&)[kMaxRed]) const
  {
    d[i] = v[i] > 0.0 ? d[i] / sqrt(v[i]) : d[i];
  }
};

struct ScaleValsF {  // a~ = (a * first) * second, in the reference operand order
  static constexpr int NRED = 0;
  const double* vals;
  const int* rowid;
  const int* colidx;
  const double* first;   // indexed by row of the stored matrix
  const double* second;  // indexed by column of the stored matrix
  double* out;
  __device__ bool skip() const { return false; }
  __device__ void operator()(int k, double (&)[kMaxRed]) const
  {
    out[k] = vals[k] * first[rowid[k]] * second[colidx[k]];
  }
};

// Column vectors: c~ = c * Dc, bounds / Dc (divide_check_zero); red0 = ||c~||^2;
// red1 = ||c||^2 of the unscaled objective.
struct ScaleColsF {
  static constexpr int NRED = 2;
  const double* c;
  const double* lb;
  const double* ub;
  const double* Dc;
  double* cs;
  double* lbs;
  double* ubs;
  __device__ bool skip() const { return false; }
  __device__ void operator()(int j, double (&red)[kMaxRed]) const
  {
    const double d = Dc[j];
    const double v = c[j] * d;
    cs[j]          = v;
    if (d == 0.0) {
      lbs[j] = 0.0;
      ubs[j] = 0.0;
    } else {
      lbs[j] = lb[j] / d;
      ubs[j] = ub[j] / d;
    }
    red[0] += v * v * 1.0;
    red[1] += c[j] * c[j];
  }
};

__device__ inline double bound_sq_term(double l, double u)
{
  double t = 0.0;
  if (isfinite(l) && l != u) t += l * l;
  if (isfinite(u)) t += u * u;
  return t;
}

// Row vectors: scaled constraint bounds; red0 = scaled bound norm^2, red1 = unscaled.
struct ScaleRowsF {
  static constexpr int NRED = 2;
  const double* lc;
  const double* uc;
  const double* Dr;
  double* lcs;
  double* ucs;
  __device__ bool skip() const { return false; }
  __device__ void operator()(int i, double (&red)[kMaxRed]) const
  {
    const double l = lc[i] * Dr[i], u = uc[i] * Dr[i];
    lcs[i]         = l;
    ucs[i]         = u;
    red[0] += bound_sq_term(l, u);
    red[1] += bound_sq_term(lc[i], uc[i]);
  }
};

// Multiply bounds by beta and c by gamma, then emit the working-precision copies.
template <typename V>
struct FinishColsF {
  using V2 = typename Vec2<V>::type;
  static constexpr int NRED = 0;
  double* cs;
  double* lbs;
  double* ubs;
  double beta, gamma;
  bool rescale;
  V* c_w;
  V2* bnd_w;
  __device__ bool skip() const { return false; }
  __device__ void operator()(int j, double (&)[kMaxRed]) const
  {
    double l = lbs[j], u = ubs[j], c = cs[j];
    if (rescale) {
      l = l * beta;
      u = u * beta;
      c = c * gamma;
    }
    lbs[j] = l;
    ubs[j] = u;
    cs[j]  = c;
    c_w[j] = (V)c;
    V2 b;
    b.x      = (V)l;
    b.y      = (V)u;
    bnd_w[j] = b;
  }
};

template <typename V>
struct FinishRowsF {
  using V2 = typename Vec2<V>::type;
  static constexpr int NRED = 0;
  double* lcs;
  double* ucs;
  double beta;
  bool rescale;
  V2* cb_w;
  double2* cb_orig;
  const double* lc;
  const double* uc;
  __device__ bool skip() const { return false; }
  __device__ void operator()(int i, double (&)[kMaxRed]) const
  {
    double l = lcs[i], u = ucs[i];
    if (rescale) {
      l = l * beta;
      u = u * beta;
    }
    lcs[i] = l;
    ucs[i] = u;
    V2 b;
    b.x     = (V)l;
    b.y     = (V)u;
    cb_w[i] = b;
    cb_orig[i] = make_double2(lc[i], uc[i]);
  }
};

template <typename V>
struct ConvertF {
  static constexpr int NRED = 0;
  const double* in;
  V* out;
  __device__ bool skip() const { return false; }
  __device__ void operator()(int i, double (&)[kMaxRed]) const { out[i] = (V)in[i]; }
};

// Initial iterate: scale (scale_solutions, initial_scaling.cu:570-650) then
// seed current, x-hat and anchor (update_primal_dual_solutions, pdlp.cu:1236-1393).
template <typename V, bool PRIMAL>
struct InitIterateF {
  using V2 = typename Vec2<V>::type;
  static constexpr int NRED = 0;
  const double* init;  // may be null (zeros)
  const double* D;
  double factor;
  bool rescale;
  const V2* bnd;  // primal only: project the current iterate
  V* cur;
  V* hat;
  V* anchor;
  __device__ bool skip() const { return false; }
  __device__ void operator()(int i, double (&)[kMaxRed]) const
  {
    double v = 0.0;
    if (init) {
      v = init[i];
      v = D[i] != 0.0 ? v / D[i] : v;
      if (rescale) v = v * factor;
    }
    const V vw = (V)v;
    hat[i]     = vw;
    anchor[i]  = vw;
    if constexpr (PRIMAL) {
      const V2 b = bnd[i];
      cur[i]     = mn(mx(vw, b.x), b.y);  // clamp functor: min(max(v, l), u)
    } else {
      cur[i] = vw;
    }
  }
};

template <typename V>
struct ZeroF {
  static constexpr int NRED = 0;
  V* v;
  __device__ bool skip() const { return false; }
  __device__ void operator()(int i, double (&)[kMaxRed]) const { v[i] = V(0); }
};

template <typename V>
struct RestartCopyF {  // anchor = current = hat (cupdlpx_restart, pdlp_restart_strategy.cu:955-970)
  static constexpr int NRED = 0;
  const V* hat;
  V* cur;
  V* anchor;
  __device__ bool skip() const { return false; }
  __device__ void operator()(int i, double (&)[kMaxRed]) const
  {
    const V v = hat[i];
    anchor[i] = v;
    cur[i]    = v;
  }
};

// ---- recentring (FP32-mixed solves refined in FP64) -----------------------

// x* = the last FP64 primal point, clamped into the variable bounds so that the
// zero correction is feasible for the shifted bounds.
struct CaptureColsF {
  static constexpr int NRED = 0;
  const double* xu;
  const double* lb;
  const double* ub;
  double* xstar;
  __device__ bool skip() const { return false; }
  __device__ void operator()(int j, double (&)[kMaxRed]) const { xstar[j] = fmin(fmax(xu[j], lb[j]), ub[j]); }
};

// Correction problem, columns: objective Dd * (c - A^T y*), bounds Dp * ([l, u] - x*),
// both in the scaled coordinates of the solver (x~ = x / Dc, c~ = c * Dc).
template <typename V>
struct CorrectionColsF {
  using V2 = typename Vec2<V>::type;
  static constexpr int NRED = 0;
  const double* xstar;
  const double* aty;
  const double* c;
  const double* lb;
  const double* ub;
  const double* Dc;
  double dp, dd;
  V* c_w;
  V2* bnd_w;
  __device__ bool skip() const { return false; }
  __device__ void operator()(int j, double (&)[kMaxRed]) const
  {
    const double d = Dc[j], x = xstar[j];
    c_w[j]         = (V)((dd * (c[j] - aty[j])) * d);
    V2 b;
    b.x      = d != 0.0 ? (V)((dp * (lb[j] - x)) / d) : V(0);
    b.y      = d != 0.0 ? (V)((dp * (ub[j] - x)) / d) : V(0);
    bnd_w[j] = b;
  }
};

// Correction problem, rows: constraint bounds Dp * ([l_c, u_c] - A x*) (scaled by
// Dr) and the kink of the dual term at y = 0, i.e. at -Dd * y* / Dr.
template <typename V>
struct CorrectionRowsF {
  using V2 = typename Vec2<V>::type;
  static constexpr int NRED = 0;
  const double* ystar;
  const double* ax;
  const double* lc;
  const double* uc;
  const double* Dr;
  double dp, dd;
  V2* cb_w;
  V* kink;
  __device__ bool skip() const { return false; }
  __device__ void operator()(int i, double (&)[kMaxRed]) const
  {
    const double r = Dr[i];
    V2 b;
    b.x     = (V)((dp * (lc[i] - ax[i])) * r);
    b.y     = (V)((dp * (uc[i] - ax[i])) * r);
    cb_w[i] = b;
    const double k = -dd * ystar[i];
    kink[i] = (V)(r != 0.0 ? k / r : k);
  }
};

// ---------------------------------------------------------------------------
// Host helpers
// ---------------------------------------------------------------------------

namespace {

inline int ew_grid(int n)
{
  const int b = (n + kBlock - 1) / kBlock;
  return std::max(1, std::min(b, 2048));
}

template <class F>
void launch_ew(int n, double* red, F f, hipStream_t s)
{
  if (n <= 0) return;
  hipLaunchKernelGGL(elementwise_kernel<F>, dim3(ew_grid(n)), dim3(kBlock), 0, s, n, red, f);
  CUOPT_HIP_CHECK(hipGetLastError());
}

struct PlanBufs {
  DevBuf<Tile> tiles;
  DevBuf<int> meta;
  DevBuf<LongRow> longrows;
  DevBuf<uint64_t> long_part;
  DevBuf<int> long_cnt;  // chunks of each long row done in the running launch (0 between launches)
  int max_chunks = 0;    // most chunks of one long row
  bool fused     = false;  // long rows merged by their last chunk inside the tiled launch
  int ntiles = 0, nlong = 0;
  // workgroups that write reduction partials: the tiles, then one per long row
  int red_blocks() const { return ntiles + nlong; }
  PlanDev dev() const { return PlanDev{tiles.get(), meta.get(), longrows.get(), ntiles}; }
};

struct MatStruct {  // CSR structure shared by the FP64 and working-precision values
  int nrows = 0, ncols = 0, nnz = 0;
  int tile = 1024;  // LDS capacity of the tiles (512, 1024 or 2048 nonzeros)
  std::vector<int> h_ptr;
  DevBuf<int> ptr, idx, rowid;
  PlanBufs plan;
};

// Tile capacity override for experiments (512, 1024 or 2048; anything else keeps the default).
int tile_from_env(const char* name, int def)
{
  const char* v = std::getenv(name);
  if (!v) return def;
  const int t = std::atoi(v);
  return (t == 512 || t == 1024 || t == 2048) ? t : def;
}

// ---- device CSR transpose -------------------------------------------------

__global__ __launch_bounds__(kBlock) void rowid_kernel(const int* __restrict__ ptr, int m, int nnz,
                                                       int* __restrict__ rowid)
{
  for (int k = blockIdx.x * kBlock + threadIdx.x; k < nnz; k += gridDim.x * kBlock) {
    int lo = 0, hi = m;  // last row i with ptr[i] <= k
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (ptr[mid] <= k) lo = mid;
      else hi = mid;
    }
    rowid[k] = lo;
  }
}

__global__ __launch_bounds__(kBlock) void iota_kernel(int n, int* __restrict__ v)
{
  for (int k = blockIdx.x * kBlock + threadIdx.x; k < n; k += gridDim.x * kBlock) v[k] = k;
}

__global__ __launch_bounds__(kBlock) void fill_int_kernel(int n, int value, int* __restrict__ v)
{
  for (int k = blockIdx.x * kBlock + threadIdx.x; k < n; k += gridDim.x * kBlock) v[k] = value;
}

__global__ __launch_bounds__(kBlock) void transpose_gather_kernel(int nnz, const int* __restrict__ perm,
                                                                  const int* __restrict__ rowid_a,
                                                                  const double* __restrict__ val,
                                                                  int* __restrict__ tidx, double* __restrict__ tval)
{
  for (int d = blockIdx.x * kBlock + threadIdx.x; d < nnz; d += gridDim.x * kBlock) {
    const int k = perm[d];
    tidx[d]     = rowid_a[k];
    tval[d]     = val[k];
  }
}

// key[j] <- the least position k of a nonzero of column j in a row of at most
// short_len nonzeros (one lane per row; key starts at INT_MAX).
__global__ __launch_bounds__(kBlock) void first_touch_kernel(const int* __restrict__ ptr, const int* __restrict__ idx,
                                                             int m, int short_len, int* __restrict__ key)
{
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < m; i += gridDim.x * kBlock) {
    const int b = ptr[i], e = ptr[i + 1];
    if (e - b > short_len) continue;
    for (int k = b; k < e; ++k) atomicMin(&key[idx[k]], k);
  }
}

// label[perm[k]] = k; *moved = 1 when some perm[k] != k.
__global__ __launch_bounds__(kBlock) void label_kernel(int n, const int* __restrict__ perm, int* __restrict__ label,
                                                       int* __restrict__ moved)
{
  for (int k = blockIdx.x * kBlock + threadIdx.x; k < n; k += gridDim.x * kBlock) {
    const int j = perm[k];
    label[j]    = k;
    if (j != k) atomicOr(moved, 1);
  }
}

// dst[k] = src[perm[k]] (to the internal column order) and dst[perm[k]] = src[k] (back).
// out[i] = src[pos[i]]: index entries at chosen positions.
__global__ __launch_bounds__(kBlock) void gather_idx_kernel(int n, const int* __restrict__ pos,
                                                            const int* __restrict__ src, int* __restrict__ out)
{
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) out[i] = src[pos[i]];
}

__global__ __launch_bounds__(kBlock) void gather_kernel(int n, const int* __restrict__ perm,
                                                        const double* __restrict__ src, double* __restrict__ dst)
{
  for (int k = blockIdx.x * kBlock + threadIdx.x; k < n; k += gridDim.x * kBlock) dst[k] = src[perm[k]];
}
__global__ __launch_bounds__(kBlock) void scatter_kernel(int n, const int* __restrict__ perm,
                                                         const double* __restrict__ src, double* __restrict__ dst)
{
  for (int k = blockIdx.x * kBlock + threadIdx.x; k < n; k += gridDim.x * kBlock) dst[perm[k]] = src[k];
}

// idx[k] <- label[idx[k]] (column relabelling).
__global__ __launch_bounds__(kBlock) void relabel_kernel(int nnz, const int* __restrict__ label, int* __restrict__ idx)
{
  for (int k = blockIdx.x * kBlock + threadIdx.x; k < nnz; k += gridDim.x * kBlock) idx[k] = label[idx[k]];
}

// tptr[j] = first position of column j among the sorted column keys (j in [0, n]).
__global__ __launch_bounds__(kBlock) void column_ptr_kernel(int n, int nnz, const int* __restrict__ cols,
                                                            int* __restrict__ tptr)
{
  for (int j = blockIdx.x * kBlock + threadIdx.x; j <= n; j += gridDim.x * kBlock) {
    int lo = 0, hi = nnz;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (cols[mid] < j) lo = mid + 1;
      else hi = mid;
    }
    tptr[j] = lo;
  }
}

inline int grid_for(long long n) { return (int)std::max<long long>(1, std::min<long long>((n + kBlock - 1) / kBlock, 4096)); }

// A (CSR, device) -> A^T (CSR, device) plus the row index of every nonzero of both.
void device_transpose(int m, int n, int nnz, const int* ptr, const int* idx, const double* val, int* tptr, int* tidx,
                      double* tval, int* rowid_a, int* rowid_at, hipStream_t s)
{
  if (nnz == 0) {
    hipLaunchKernelGGL(column_ptr_kernel, dim3(grid_for(n + 1)), dim3(kBlock), 0, s, n, 0, rowid_at, tptr);
    CUOPT_HIP_CHECK(hipGetLastError());
    return;
  }
  DevBuf<int> perm_in(nnz), perm(nnz);
  hipLaunchKernelGGL(rowid_kernel, dim3(grid_for(nnz)), dim3(kBlock), 0, s, ptr, m, nnz, rowid_a);
  hipLaunchKernelGGL(iota_kernel, dim3(grid_for(nnz)), dim3(kBlock), 0, s, nnz, perm_in.get());
  CUOPT_HIP_CHECK(hipGetLastError());
  int bits = 1;
  while (bits < 31 && (1LL << bits) < n) ++bits;
  size_t temp_bytes = 0;
  CUOPT_HIP_CHECK(rocprim::radix_sort_pairs(nullptr, temp_bytes, idx, rowid_at, perm_in.get(), perm.get(),
                                            (size_t)nnz, 0, bits, s));
  DevBuf<unsigned char> temp(std::max<size_t>(temp_bytes, 1));
  CUOPT_HIP_CHECK(rocprim::radix_sort_pairs(temp.get(), temp_bytes, idx, rowid_at, perm_in.get(), perm.get(),
                                            (size_t)nnz, 0, bits, s));
  hipLaunchKernelGGL(transpose_gather_kernel, dim3(grid_for(nnz)), dim3(kBlock), 0, s, nnz, perm.get(), rowid_a,
                     val, tidx, tval);
  hipLaunchKernelGGL(column_ptr_kernel, dim3(grid_for(n + 1)), dim3(kBlock), 0, s, n, nnz, rowid_at, tptr);
  CUOPT_HIP_CHECK(hipGetLastError());
  CUOPT_HIP_CHECK(hipStreamSynchronize(s));  // temporaries are freed on return
}

// Column order for the gathers of A x-bar: columns are numbered in the order in
// which rows of at most `short_len` nonzeros first touch them (rows in index
// order, columns in row order), then the untouched columns in index order.  The
// nonzeros of consecutive short rows, which one workgroup reduces together, then
// gather from nearby addresses of x-bar; rows longer than a tile gather over
// sorted columns either way.  On the device: the first position touching each
// column by an atomic minimum, then a stable radix sort of the columns by it.
// label (n) receives each original column's new number; perm (n) the original
// column of each new number.  Returns false when the order is the identity.
bool first_touch_order(int m, int n, const int* ptr, const int* idx, int short_len, int* label, int* perm,
                       hipStream_t s)
{
  DevBuf<int> key(n), key_sorted(n), cols(n), moved(1);
  hipLaunchKernelGGL(fill_int_kernel, dim3(grid_for(n)), dim3(kBlock), 0, s, n, INT_MAX, key.get());
  hipLaunchKernelGGL(first_touch_kernel, dim3(grid_for(m)), dim3(kBlock), 0, s, ptr, idx, m, short_len, key.get());
  hipLaunchKernelGGL(iota_kernel, dim3(grid_for(n)), dim3(kBlock), 0, s, n, cols.get());
  CUOPT_HIP_CHECK(hipGetLastError());
  size_t temp_bytes = 0;
  CUOPT_HIP_CHECK(rocprim::radix_sort_pairs(nullptr, temp_bytes, key.get(), key_sorted.get(), cols.get(), perm,
                                            (size_t)n, 0, 31, s));
  DevBuf<unsigned char> temp(std::max<size_t>(temp_bytes, 1));
  CUOPT_HIP_CHECK(rocprim::radix_sort_pairs(temp.get(), temp_bytes, key.get(), key_sorted.get(), cols.get(), perm,
                                            (size_t)n, 0, 31, s));
  hipLaunchKernelGGL(fill_int_kernel, dim3(1), dim3(kBlock), 0, s, 1, 0, moved.get());
  hipLaunchKernelGGL(label_kernel, dim3(grid_for(n)), dim3(kBlock), 0, s, n, perm, label, moved.get());
  CUOPT_HIP_CHECK(hipGetLastError());
  int h_moved = 0;
  moved.download(&h_moved, 1, s);
  CUOPT_HIP_CHECK(hipStreamSynchronize(s));  // temporaries are freed on return
  return h_moved != 0;
}

// colidx (host, sorted within rows) enables the XCD-aware tile order
// (tile_plan.hpp); CUOPT_AMD_XCD_ORDER=0 turns it off for comparisons.
// Long rows (more than a tile) are merged either by a launch of their own after
// the tiled one, or inside the tiled launch by their last chunk.  Setting
// pdlp_long_row_merge: 0 / 1 force one form; -1 (default) merges inside the launch
// when no row has more than kFusedMaxChunks chunks, since every chunk of a row
// counts itself on one device-scope counter and a few hundred such atomics on one
// address take longer than the merge launch they replace (measured at 500 k x 1 M,
// 465 chunks: FP32 dual step 34.9 -> 50.4 us; at 20 k x 40 k, 22 chunks, the
// solve loop 0.19 -> 0.16 s).  With -1 the SpMV tuner times both forms.
constexpr int kFusedMaxChunks = 64;

void set_merge(PlanBufs& P, int mode)
{
  P.fused = P.nlong > 0 && (mode == 1 || (mode < 0 && P.max_chunks <= kFusedMaxChunks));
}

void upload_plan(PlanBufs& P, const TilePlan& p, hipStream_t s, int merge_mode = -1)
{
  P.ntiles = (int)p.tiles.size();
  P.tiles.resize(p.tiles.size());
  P.tiles.upload(p.tiles, s);
  P.meta.resize(p.meta.size());
  P.meta.upload(p.meta, s);
  P.longrows.resize(std::max<size_t>(1, p.longrows.size()));
  if (!p.longrows.empty()) P.longrows.upload(p.longrows, s);
  P.long_part.resize(std::max<size_t>(1, p.tiles.size()));
  P.nlong = (int)p.longrows.size();
  P.long_cnt.resize(std::max<size_t>(1, p.longrows.size()) * kCntStride);
  P.long_cnt.zero(s);
  P.max_chunks = 0;
  for (const LongRow& lr : p.longrows) P.max_chunks = std::max(P.max_chunks, lr.nchunks);
  set_merge(P, merge_mode);
}

// Lanes per row of every row-run tile shifted by bias powers of two (clamped to
// 1..64 lanes); chunk tiles are unchanged.
void shift_lanes(TilePlan& p, int bias)
{
  for (int& v : p.meta)
    if (v >= 0) v = std::min(6, std::max(0, v + bias));
}

// Where the planner finds the column indices of its XCD-aware order: the whole
// host array, a sample holding the positions it reads, or neither (no reordering).
struct ColSource {
  const int* host         = nullptr;
  const ColSample* sample = nullptr;
};

TilePlan make_plan(const MatStruct& M, int tile, int bias, ColSource cols)
{
  if (const char* e = std::getenv("CUOPT_AMD_XCD_ORDER"))
    if (std::atoi(e) == 0) cols = ColSource{};
  TilePlan p = cols.sample
                 ? plan_tiles(M.h_ptr.data(), M.nrows, tile - kVecPad, kMaxRowsPerTile, kBlock, *cols.sample, M.ncols)
                 : plan_tiles(M.h_ptr.data(), M.nrows, tile - kVecPad, kMaxRowsPerTile, kBlock, cols.host, M.ncols);
  if (bias) shift_lanes(p, bias);
  return p;
}

void build_plan(MatStruct& M, hipStream_t s, ColSource cols = {}, int merge_mode = -1)
{
  upload_plan(M.plan, make_plan(M, M.tile, 0, cols), s, merge_mode);
}

// The tiled SpMV launch (plus the long-row merge) over raw device pointers:
// ptr / idx the CSR structure, plan the tile plan, long_part one partial per tile.
// long_cnt (one zeroed counter per long row) merges each long row in the launch's
// last chunk to finish; null runs the separate long_rows_kernel.
// Workgroups of a tiled launch: as many as are resident at once (the occupancy
// of this instantiation times the CUs), rounded down to a multiple of 8 so that
// workgroup b -- and every tile b + k * grid it loops over -- runs on XCD b mod 8,
// where the plan placed that tile.  A resident grid keeps each workgroup's waves
// alive across tiles instead of retiring and dispatching one workgroup per tile:
// measured at 500 k x 1 M, FP32 dual step 37.8 -> 35.7 us, FP64 44.6 -> 43.2 us
// (twice the resident count: 36.9 us).  CUOPT_AMD_PERSISTENT=0 (read at each
// launch, so that tests can compare both forms) launches one workgroup per tile.
template <class Epi, int TILE>
int tiled_grid(int ntiles)
{
  if (const char* e = std::getenv("CUOPT_AMD_PERSISTENT"))
    if (std::atoi(e) == 0) return ntiles;
  static const int resident = []() {
    int per_cu = 0, cus = 0, dev = 0;
    CUOPT_HIP_CHECK(hipGetDevice(&dev));
    CUOPT_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, tiled_spmv_kernel<Epi, TILE>, kBlock, 0));
    CUOPT_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    return std::max(8, (per_cu * cus) & ~7);
  }();
  return std::min(ntiles, resident);
}

template <class Epi>
void launch_tiled_raw(int tile, const int* ptr, const int* idx, const typename Epi::Mat* vals, PlanDev plan,
                      int nlong, uint64_t* long_part, double* red, Epi epi, hipStream_t s, int* long_cnt = nullptr)
{
  if (plan.ntiles == 0) return;
  if (nlong == 0) long_cnt = nullptr;
  if (tile == 512)
    hipLaunchKernelGGL((tiled_spmv_kernel<Epi, 512>), dim3(tiled_grid<Epi, 512>(plan.ntiles)), dim3(kBlock), 0, s,
                       ptr, idx, vals, plan, long_part, red, epi, long_cnt);
  else if (tile == 1024)
    hipLaunchKernelGGL((tiled_spmv_kernel<Epi, 1024>), dim3(tiled_grid<Epi, 1024>(plan.ntiles)), dim3(kBlock), 0, s,
                       ptr, idx, vals, plan, long_part, red, epi, long_cnt);
  else
    hipLaunchKernelGGL((tiled_spmv_kernel<Epi, 2048>), dim3(tiled_grid<Epi, 2048>(plan.ntiles)), dim3(kBlock), 0, s,
                       ptr, idx, vals, plan, long_part, red, epi, long_cnt);
  CUOPT_HIP_CHECK(hipGetLastError());
  if (nlong > 0 && !long_cnt) {
    hipLaunchKernelGGL(long_rows_kernel<Epi>, dim3(nlong), dim3(kWave), 0, s, plan, long_part, red, epi);
    CUOPT_HIP_CHECK(hipGetLastError());
  }
}

template <class Epi>
void launch_tiled(const MatStruct& M, const typename Epi::Mat* vals, double* red, Epi epi, hipStream_t s)
{
  launch_tiled_raw(M.tile, M.ptr.get(), M.idx.get(), vals, M.plan.dev(), M.plan.nlong, M.plan.long_part.get(), red,
                   epi, s, M.plan.fused ? M.plan.long_cnt.get() : nullptr);
}

void launch_reduce(const double* part, int nblocks, int nred, double* out, hipStream_t s,
                   const double* skip = nullptr)
{
  hipLaunchKernelGGL(reduce_partials_kernel, dim3(1), dim3(1024), 0, s, part, nblocks, nred, out, skip);
  CUOPT_HIP_CHECK(hipGetLastError());
}

// ---- SpMV autotuner ----------------------------------------------------------
//
// Candidates for one matrix: tile capacities of 512, 1 024 and 2 048 nonzeros,
// each with the planner's lanes per row G and with G / 2 and 2 G, and both
// long-row merges.  Each is timed the way the reference's batch-size handler
// times its candidates (optimal_batch_size_handler.cu:196-435): one untimed SpMV,
// then a run of them between two HIP events; the least time wins, but the
// untuned plan is kept unless the winner beats it by kTuneMargin, so that
// candidates within timing noise do not change the choice from run to run.  The
// timed launch is the step kernel over that matrix itself (the dual step's
// epilogue for A, the primal step's for A^T) on zeroed operands, since the plans
// that are fastest for a plain y = A x are not always fastest for the steps: with
// a plain epilogue the tuner kept 1 024-nonzero tiles for A^T of a banded
// 4 M x 8 M LP, whose FP64 primal step then took 365 us against 300 us with
// 2 048.  Choices are cached per matrix for the process.
constexpr int kTuneReps        = 10;
constexpr float kTuneMargin    = 0.03f;

struct SpmvChoice {
  int tile = 1024, bias = 0;
  int merge = -1;  // long rows: 1 merged in the tiled launch, 0 in their own launch, -1 none
};

std::mutex g_tune_mu;
std::map<std::vector<uint64_t>, SpmvChoice> g_tune_cache;

uint64_t fnv1a(const int* p, size_t n)
{
  uint64_t h = 1469598103934665603ULL;
  for (size_t i = 0; i < n; ++i) {
    h ^= (uint32_t)p[i];
    h *= 1099511628211ULL;
  }
  return h;
}

// Picks (and installs in M) the fastest candidate; returns whether it timed anything.
// merge_mode as pdlp_long_row_merge: -1 times both long-row merges, 0 / 1 only that one.
template <typename V>
bool autotune_spmv(MatStruct& M, const V* vals, ColSource cols, hipStream_t s, int merge_mode, bool primal,
                   SpmvChoice& out)
{
  const std::vector<uint64_t> key{(uint64_t)M.nrows, (uint64_t)M.ncols, (uint64_t)M.nnz, sizeof(V),
                                  fnv1a(M.h_ptr.data(), M.h_ptr.size()), (uint64_t)(merge_mode + 1),
                                  (uint64_t)primal};
  {
    std::lock_guard<std::mutex> lock(g_tune_mu);
    auto it = g_tune_cache.find(key);
    if (it != g_tune_cache.end()) {
      out = it->second;
      if (out.tile != M.tile || out.bias != 0) {  // else the plan in place is the cached one
        M.tile = out.tile;
        upload_plan(M.plan, make_plan(M, out.tile, out.bias, cols), s);
      }
      M.plan.fused = out.merge == 1;
      return false;
    }
  }
  // Operands of the timed launches: the gathered vector and zeroed per-row vectors
  // of the step over this matrix (distinct buffers, so that every read streams).
  using V2          = typename Vec2<V>::type;
  const size_t rows = (size_t)std::max(M.nrows, 1);
  DevBuf<V> x(std::max(M.ncols, 1)), r0(rows), r1(rows), r2(rows), w0(rows), w1(rows);
  DevBuf<V2> bnd(rows);
  DevBuf<StepScalars> sc(1);
  for (DevBuf<V>* b : {&x, &r0, &r1, &r2}) b->zero(s);
  bnd.zero(s);
  const StepScalars h_sc{1.0, 1.0, 0, 0};
  sc.upload(&h_sc, 1, s);
  hipEvent_t ev[2] = {nullptr, nullptr};
  struct EventGuard {
    hipEvent_t* e;
    ~EventGuard()
    {
      for (int i = 0; i < 2; ++i)
        if (e[i]) (void)hipEventDestroy(e[i]);
    }
  } guard{ev};
  CUOPT_HIP_CHECK(hipEventCreate(&ev[0]));
  CUOPT_HIP_CHECK(hipEventCreate(&ev[1]));
  float best_ms = -1.f, def_ms = -1.f;
  TilePlan best_tp, def_tp;
  SpmvChoice def_choice;
  const int def_tile = M.tile;
  for (int tile : {512, 1024, 2048}) {
    const TilePlan base = make_plan(M, tile, 0, cols);
    for (int bias : {0, -1, 1}) {
      TilePlan p = base;
      if (bias) {
        shift_lanes(p, bias);
        if (p.meta == base.meta) continue;  // no row-run tile changes: the same candidate
      }
      PlanBufs P;
      upload_plan(P, p, s);
      // long rows merged by their own launch (0) and / or by their last chunk (1)
      const int f0 = P.nlong > 0 && merge_mode == 1 ? 1 : 0;
      const int f1 = P.nlong > 0 && merge_mode != 0 ? 1 : 0;
      const int def_fused = merge_mode >= 0 ? merge_mode : (P.fused ? 1 : 0);  // the untuned plan's merge
      for (int fused = f0; fused <= f1; ++fused) {
        auto run = [&]() {
          int* cnt = fused ? P.long_cnt.get() : nullptr;
          if (primal)
            launch_tiled_raw(tile, M.ptr.get(), M.idx.get(), vals, P.dev(), P.nlong, P.long_part.get(), nullptr,
                             PrimalStepEpi<V, false, false>{x.get(), r0.get(), r1.get(), bnd.get(), r2.get(),
                                                            w0.get(), w1.get(), nullptr, nullptr, nullptr,
                                                            sc.get(), 0, false},
                             s, cnt);
          else
            launch_tiled_raw(tile, M.ptr.get(), M.idx.get(), vals, P.dev(), P.nlong, P.long_part.get(), nullptr,
                             DualStepEpi<V, false, false>{x.get(), r0.get(), bnd.get(), r1.get(), w0.get(), nullptr,
                                                          nullptr, sc.get(), 0, nullptr},
                             s, cnt);
        };
        run();
        CUOPT_HIP_CHECK(hipEventRecord(ev[0], s));
        for (int r = 0; r < kTuneReps; ++r) run();
        CUOPT_HIP_CHECK(hipEventRecord(ev[1], s));
        CUOPT_HIP_CHECK(hipEventSynchronize(ev[1]));
        float ms = 0.f;
        CUOPT_HIP_CHECK(hipEventElapsedTime(&ms, ev[0], ev[1]));
        const SpmvChoice c{tile, bias, P.nlong > 0 ? fused : -1};
        if (best_ms < 0.f || ms < best_ms) {
          best_ms = ms;
          out     = c;
          best_tp = p;
        }
        if (tile == def_tile && bias == 0 && (P.nlong == 0 || fused == def_fused)) {
          def_ms     = ms;
          def_choice = c;
          def_tp     = p;
        }
      }
    }
  }
  if (def_ms >= 0.f && best_ms > (1.f - kTuneMargin) * def_ms) {
    out     = def_choice;
    best_tp = std::move(def_tp);
  }
  M.tile = out.tile;
  upload_plan(M.plan, best_tp, s);
  M.plan.fused = out.merge == 1;
  std::lock_guard<std::mutex> lock(g_tune_mu);
  g_tune_cache[key] = out;
  return true;
}

// conditional_major (cpp/src/pdlp/utils.cuh:62-81)
long long conditional_major(long long k)
{
  long long step = 10, threshold = 1000;
  while (k >= threshold) {
    step *= 10;
    threshold *= 10;
  }
  return step;
}

constexpr int kMajor              = 200;    // major_iteration (Stable3)
constexpr double kArtificial      = 0.36;   // default_artificial_restart_threshold
constexpr double kSufficient      = 0.2;    // sufficient_reduction_for_restart
constexpr double kNecessary       = 0.8;    // necessary_reduction_for_restart
constexpr double kKp = 0.99, kKi = 0.01, kKd = 0.0, kISmooth = 0.3;  // restart PID (pdlp_hyper_params.cuh:51-54)
constexpr int kRuizIters          = 10;
// FP64 refinement of FP32-mixed solves: restarts recentre the iterates once
// the relative KKT error is below this.
constexpr double kRecentreFrom    = 1e-2;

struct Tolerances {
  double ap, rp, ad, rd, ag, rg;  // absolute / relative primal, dual, gap
};
constexpr double kPcAlpha         = 1.0;

inline bool is_check(long long k)
{
  const bool major = ((k % kMajor) == 0 && k > 0) || k <= 0;
  return major || (k % conditional_major(k)) == 0;
}

// Slots of the device scalar record.
enum Slot {
  S_KKT_RP2 = 0, S_KKT_PSUM, S_KKT_DD,           // from KktRowsEpi
  S_KKT_RD2, S_KKT_XC, S_KKT_SX, S_KKT_DP,       // from KktColsEpi
  S_FPE_NY, S_FPE_I, S_FPE_NX,                   // periodic fixed-point error terms
  S_FPE0_NY, S_FPE0_I, S_FPE0_NX,                // post-restart fixed-point error terms
  S_SETUP0, S_SETUP1,                            // setup reductions
  S_LINF_P, S_LINF_D,                            // per_constraint_residual maxima
  S_COUNT
};

}  // namespace

// ---------------------------------------------------------------------------
// The solver
// ---------------------------------------------------------------------------

template <typename V>
class PdlpSolver {
  using V2 = typename Vec2<V>::type;

 public:
  PdlpSolver(const LpHost& lp,
             const std::vector<double>& lc,
             const std::vector<double>& uc,
             const std::vector<double>& lb,
             const std::vector<double>& ub,
             const PdlpSettings& st)
    : lp_(lp), lc_(lc), uc_(uc), lb_(lb), ub_(ub), st_(st), m_(lp.m), n_(lp.n)
  {
  }

  ~PdlpSolver()
  {
    for (auto& g : graphs_) (void)hipGraphExecDestroy(g.second);
    if (stream_) (void)hipStreamDestroy(stream_);
    if (h_scal_) (void)hipHostFree(h_scal_);
    if (h_step_) (void)hipHostFree(h_step_);
  }

  SolveResult run();

 private:
  using clock = std::chrono::steady_clock;
  double elapsed() const { return std::chrono::duration<double>(clock::now() - t0_).count(); }

  void build_structures();
  void setup_scaling();
  void power_iteration();
  double lanczos_sigma_max_sq(int& iters);
  void init_iterates();
  void read_scalars()
  {
    CUOPT_HIP_CHECK(hipMemcpyAsync(h_scal_, scal_.get(), S_COUNT * sizeof(double), hipMemcpyDeviceToHost, stream_));
    CUOPT_HIP_CHECK(hipStreamSynchronize(stream_));
  }
  void upload_step_scalars()
  {
    // Pinned staging record: it is rewritten only after the next check's
    // synchronising read, by which time this copy has executed.
    *h_step_ = StepScalars{tau_, sigma_, kr_, 0};
    CUOPT_HIP_CHECK(hipMemcpyAsync(step_sc_.get(), h_step_, sizeof(StepScalars), hipMemcpyHostToDevice, stream_));
  }
  void evaluate_kkt();
  void launch_step(int j, bool major, bool fpe, double* ny_slot, double* fpe_slot);
  void finish(SolveResult& R, int status);
  int pdhg_loop(SolveResult& R, const Tolerances& tol, bool refine);
  void recentre();
  // Termination tests of the last evaluation (termination_strategy.cu:311-352); with
  // per_constraint_residual the residual tests are the relative L-inf ones
  // (:327-345): max_i |r_i| - rel w_i <= the absolute tolerance.
  bool primal_feasible(const Tolerances& t) const
  {
    return st_.per_constraint_residual ? linf_p_ <= t.ap : l2p_ <= t.ap + t.rp * b_norm_;
  }
  bool is_optimal(const Tolerances& t) const
  {
    const bool dual_ok = st_.per_constraint_residual ? linf_d_ <= t.ad : l2d_ <= t.ad + t.rd * c_norm_;
    return dual_ok && primal_feasible(t) && gap_ <= t.ag + t.rg * absobj_;
  }
  // save_best_primal_so_far (pdlp.cu:508-625): the best unscaled point met at a
  // check -- primal feasible beats infeasible, then the better objective, then
  // the smaller L2 primal residual -- is returned at a time or iteration limit.
  void record_best(const Tolerances& t);
  void restore_best();
  double rel_kkt_error() const
  {
    return std::max({l2p_ / (1.0 + b_norm_), l2d_ / (1.0 + c_norm_), gap_ / (1.0 + absobj_)});
  }

  const LpHost& lp_;
  const std::vector<double>&lc_, &uc_, &lb_, &ub_;
  const PdlpSettings& st_;
  const int m_, n_;
  clock::time_point t0_;
  hipStream_t stream_ = nullptr;
  double* h_scal_     = nullptr;
  StepScalars* h_step_ = nullptr;

  MatStruct A_, AT_;
  DevBuf<double> a_orig_, at_orig_, a_s_, at_s_;
  DevBuf<V> a_w_, at_w_;
  // Host column indices of A sorted within rows (the tile planner's XCD order).
  std::vector<int> a_sorted_idx_;
  ColSample a_cols_;     // A's columns at the planner's chunk midpoints (relabelled matrices)
  ColSource a_src_;      // where plans of A read columns: the host indices or a_cols_
  // Column relabelling (unallocated: columns as given): col_perm_[k] is the
  // original column of internal column k.  Per-column inputs are gathered through
  // it on the device at upload and the returned x and reduced costs scattered back.
  DevBuf<int> col_perm_;
  // dst (n, internal order) <- host vector v in the caller's column order.
  void upload_columns(const double* v, DevBuf<double>& dst)
  {
    if (!col_perm_.get()) {
      dst.upload(v, n_, stream_);
      return;
    }
    DevBuf<double> tmp(n_);
    tmp.upload(v, n_, stream_);
    hipLaunchKernelGGL(gather_kernel, dim3(grid_for(n_)), dim3(kBlock), 0, stream_, n_, col_perm_.get(),
                       tmp.get(), dst.get());
    CUOPT_HIP_CHECK(hipGetLastError());
    CUOPT_HIP_CHECK(hipStreamSynchronize(stream_));  // tmp is freed on return
  }
  // host v (caller's column order) <- src (n, internal order); synchronous.
  void download_columns(const DevBuf<double>& src, double* v)
  {
    if (!col_perm_.get()) {
      src.download(v, n_, stream_);
      CUOPT_HIP_CHECK(hipStreamSynchronize(stream_));
      return;
    }
    DevBuf<double> tmp(n_);
    hipLaunchKernelGGL(scatter_kernel, dim3(grid_for(n_)), dim3(kBlock), 0, stream_, n_, col_perm_.get(),
                       src.get(), tmp.get());
    CUOPT_HIP_CHECK(hipGetLastError());
    tmp.download(v, n_, stream_);
    CUOPT_HIP_CHECK(hipStreamSynchronize(stream_));
  }

  DevBuf<double> Dr_, Dc_, rowv_, colv_;
  DevBuf<double> c_orig_, lb_orig_, ub_orig_, lc_orig_, uc_orig_;
  DevBuf<double> cs_, lbs_, ubs_, lcs_, ucs_;
  DevBuf<double2> cb_orig_;
  DevBuf<V> c_w_;
  DevBuf<V2> bnd_w_, cb_w_;

  DevBuf<V> x_[2], y_[2], x0_, y0_, xbar_, ybar_, xhat_, yhat_, slack_, aty_;
  DevBuf<double> xu_, yu_, su_;
  DevBuf<double> red_, scal_;
  DevBuf<double> rp_res_, rd_res_;          // per_constraint_residual: per-row / per-column residuals
  double linf_p_ = 0.0, linf_d_ = 0.0;
  DevBuf<double> best_x_, best_y_, best_s_;  // save_best_primal_so_far
  bool have_best_ = false, best_feasible_ = false;
  double best_l2p_ = 0, best_l2d_ = 0, best_pobj_ = 0, best_dobj_ = 0, best_gap_ = 0, best_absobj_ = 0;
  DevBuf<StepScalars> step_sc_;
  int cur_ = 0;
  hipEvent_t* ev_ = nullptr;  // events of the current timed iteration (benchmark window)
  bool use_graphs_ = true;
  std::map<std::string, hipGraphExec_t> graphs_;  // captured iteration blocks by pattern

  // recentring: the FP64 point (x*, y*), A x*, A^T y*, dual kinks
  DevBuf<double> xstar_, ystar_, ax_, aty64_;
  DevBuf<V> kink_;
  struct {
    double dp = 0.0, dd = 0.0;  // primal / dual scale of the current correction
  } refine_;
  bool refine_active_ = false;
  int recentrings_    = 0;
  static double bound_w(double w) { return bound_primal_weight<V>(w); }
  long long k_done_   = 0;  // iterations of all finished loops
  int restarts_ = 0, checks_ = 0;
  double setup_time_ = 0.0;

  double beta_ = 1.0, gamma_ = 1.0, b_norm_ = 0.0, c_norm_ = 0.0;
  double eta_ = 1.0, omega_ = 1.0, best_omega_ = 1.0, tau_ = 1.0, sigma_ = 1.0;
  double sigma_max_sq_ = 0.0;
  int power_iters_     = 0;
  long long kr_        = 0;
  double obj_sf_ = 1.0, obj_off_ = 0.0;

  // last KKT evaluation
  double l2p_ = 0, l2d_ = 0, pobj_ = 0, dobj_ = 0, gap_ = 0, absobj_ = 0, dp_ = 0, dd_ = 0;
};

template <typename V>
void PdlpSolver<V>::build_structures()
{
  const int nnz = lp_.rowptr[m_];
  // Row-sorted CSR (problem_t keeps sorted column indices); rows that arrive
  // unsorted are sorted here on the host, which callers rarely need.
  const int* idx_src    = lp_.colidx.data();
  const double* val_src = lp_.vals.data();
  std::vector<int> sorted_idx;
  std::vector<double> sorted_val;
  // The check runs on host threads over row ranges (it reads every index once);
  // the rare unsorted rows are then sorted in order.
  std::vector<int> unsorted;
  {
    const int nt = nnz >= (1 << 20) ? std::max(1, std::min(16, (int)std::thread::hardware_concurrency())) : 1;
    std::vector<std::vector<int>> found(nt);
    auto check = [&](int t) {
      const int r0 = (int)((long long)m_ * t / nt), r1 = (int)((long long)m_ * (t + 1) / nt);
      for (int i = r0; i < r1; ++i) {
        const int b = lp_.rowptr[i], e = lp_.rowptr[i + 1];
        for (int k = b + 1; k < e; ++k)
          if (idx_src[k - 1] > idx_src[k]) {
            found[t].push_back(i);
            break;
          }
      }
    };
    std::vector<std::thread> th;
    for (int t = 1; t < nt; ++t) th.emplace_back(check, t);
    check(0);
    for (auto& x : th) x.join();
    for (const auto& f : found) unsorted.insert(unsorted.end(), f.begin(), f.end());
  }
  for (int i : unsorted) {
    const int b = lp_.rowptr[i], e = lp_.rowptr[i + 1];
    if (sorted_idx.empty()) {
      sorted_idx.assign(lp_.colidx.begin(), lp_.colidx.begin() + nnz);
      sorted_val.assign(lp_.vals.begin(), lp_.vals.begin() + nnz);
    }
    std::vector<std::pair<int, double>> row;
    for (int k = b; k < e; ++k) row.emplace_back(sorted_idx[k], sorted_val[k]);
    std::stable_sort(row.begin(), row.end(), [](const auto& p, const auto& q) { return p.first < q.first; });
    for (int k = b; k < e; ++k) {
      sorted_idx[k] = row[k - b].first;
      sorted_val[k] = row[k - b].second;
    }
  }
  if (!sorted_idx.empty()) {
    a_sorted_idx_ = std::move(sorted_idx);
    idx_src       = a_sorted_idx_.data();
    val_src       = sorted_val.data();
  }
  A_.nrows = m_;
  A_.ncols = n_;
  A_.nnz   = nnz;
  A_.h_ptr.assign(lp_.rowptr.begin(), lp_.rowptr.begin() + m_ + 1);
  AT_.nrows = n_;
  AT_.ncols = m_;
  AT_.nnz   = nnz;
  A_.tile  = tile_from_env("CUOPT_AMD_SPMV_TILE_A", 1024);
  AT_.tile = tile_from_env("CUOPT_AMD_SPMV_TILE_AT", 1024);
  for (MatStruct* M : {&A_, &AT_}) {
    M->ptr.resize(M->nrows + 1);
    M->idx.resize(nnz + kArrayPad);
    M->rowid.resize(std::max(nnz, 1));
    CUOPT_HIP_CHECK(hipMemsetAsync(M->idx.get() + nnz, 0, kArrayPad * sizeof(int), stream_));
  }
  a_orig_.resize(nnz + kArrayPad);
  at_orig_.resize(nnz + kArrayPad);
  CUOPT_HIP_CHECK(hipMemsetAsync(a_orig_.get() + nnz, 0, kArrayPad * sizeof(double), stream_));
  CUOPT_HIP_CHECK(hipMemsetAsync(at_orig_.get() + nnz, 0, kArrayPad * sizeof(double), stream_));
  A_.ptr.upload(A_.h_ptr, stream_);
  A_.idx.upload(idx_src, nnz, stream_);
  a_orig_.upload(val_src, nnz, stream_);
  // Column order: as given, or first-touch (first_touch_order); automatic for
  // matrices of at least 2^20 nonzeros.
  const int order = st_.column_order;
  col_perm_.release();
  if (n_ > 0 && (order == 1 || (order < 0 && nnz >= (1 << 20)))) {
    DevBuf<int> label(n_), perm(n_);
    if (first_touch_order(m_, n_, A_.ptr.get(), A_.idx.get(), A_.tile - kVecPad, label.get(), perm.get(), stream_)) {
      hipLaunchKernelGGL(relabel_kernel, dim3(grid_for(nnz)), dim3(kBlock), 0, stream_, nnz, label.get(),
                         A_.idx.get());
      CUOPT_HIP_CHECK(hipGetLastError());
      CUOPT_HIP_CHECK(hipStreamSynchronize(stream_));  // label is freed on leaving this scope
      col_perm_ = std::move(perm);
    }
  }
  // A^T on the device (compute_transpose_of_problem, problem.cu:472-489): a
  // stable radix sort of the nonzeros by column keeps each column's entries in
  // row order, as the host counting sort of the reference does.
  device_transpose(m_, n_, nnz, A_.ptr.get(), A_.idx.get(), a_orig_.get(), AT_.ptr.get(), AT_.idx.get(),
                   at_orig_.get(), A_.rowid.get(), AT_.rowid.get(), stream_);
  if (col_perm_.get()) {
    // Rows of A sorted by the new labels: the transpose of A^T (whose rows are the
    // relabelled columns, each in row order) lists every row's entries by column.
    device_transpose(n_, m_, nnz, AT_.ptr.get(), AT_.idx.get(), at_orig_.get(), A_.ptr.get(), A_.idx.get(),
                     a_orig_.get(), AT_.rowid.get(), A_.rowid.get(), stream_);
    // The planner reads one column per chunk of a long row: gather those, for
    // every tile capacity the solver or its tuner may plan with, instead of
    // downloading all of A's relabelled indices.
    std::vector<int> pos;
    for (int t : {512, 1024, 2048, A_.tile}) chunk_midpoints(A_.h_ptr.data(), m_, t - kVecPad, pos);
    std::sort(pos.begin(), pos.end());
    pos.erase(std::unique(pos.begin(), pos.end()), pos.end());
    a_cols_.col.assign(pos.size(), 0);
    if (!pos.empty()) {
      DevBuf<int> dpos(pos.size()), dcol(pos.size());
      dpos.upload(pos, stream_);
      hipLaunchKernelGGL(gather_idx_kernel, dim3(grid_for((int)pos.size())), dim3(kBlock), 0, stream_, (int)pos.size(),
                         dpos.get(), A_.idx.get(), dcol.get());
      CUOPT_HIP_CHECK(hipGetLastError());
      dcol.download(a_cols_.col.data(), pos.size(), stream_);
      CUOPT_HIP_CHECK(hipStreamSynchronize(stream_));  // dpos / dcol are freed on leaving this scope
    }
    a_cols_.pos = std::move(pos);
  }
  AT_.h_ptr.resize(n_ + 1);
  AT_.ptr.download(AT_.h_ptr.data(), n_ + 1, stream_);
  CUOPT_HIP_CHECK(hipStreamSynchronize(stream_));  // host sources above are temporaries
  a_src_ = col_perm_.get() ? ColSource{nullptr, &a_cols_} : ColSource{idx_src, nullptr};
  build_plan(A_, stream_, a_src_, st_.long_row_merge);
  build_plan(AT_, stream_, ColSource{}, st_.long_row_merge);
}

template <typename V>
void PdlpSolver<V>::setup_scaling()
{
  const int nnz = A_.nnz;
  const int red_blocks = std::max({A_.plan.red_blocks(), AT_.plan.red_blocks(), ew_grid(std::max(m_, n_)),
                                   ew_grid(std::max(nnz, 1))});
  red_.resize((size_t)red_blocks * kMaxRed);
  scal_.resize(S_COUNT);
  scal_.zero(stream_);
  Dr_.resize(m_);
  Dc_.resize(n_);
  rowv_.resize(m_);
  colv_.resize(n_);
  launch_ew(m_, nullptr, FillF{Dr_.get(), 1.0}, stream_);
  launch_ew(n_, nullptr, FillF{Dc_.get(), 1.0}, stream_);

  // Ruiz infinity-norm equilibration (initial_scaling.cu:181-220).
  for (int it = 0; it < kRuizIters; ++it) {
    launch_tiled(A_, a_orig_.get(), nullptr,
                 ScaleStatEpi<false, true>{A_.rowid.get(), Dr_.get(), Dc_.get(), 1.0, rowv_.get()}, stream_);
    launch_tiled(AT_, at_orig_.get(), nullptr,
                 ScaleStatEpi<true, true>{AT_.rowid.get(), Dr_.get(), Dc_.get(), 1.0, colv_.get()}, stream_);
    launch_ew(m_, nullptr, DivSqrtF{Dr_.get(), rowv_.get()}, stream_);
    launch_ew(n_, nullptr, DivSqrtF{Dc_.get(), colv_.get()}, stream_);
  }
  // Pock-Chambolle, alpha = 1 (initial_scaling.cu:233-364).
  launch_tiled(A_, a_orig_.get(), nullptr,
               ScaleStatEpi<false, false>{A_.rowid.get(), Dr_.get(), Dc_.get(), kPcAlpha, rowv_.get()}, stream_);
  launch_tiled(AT_, at_orig_.get(), nullptr,
               ScaleStatEpi<true, false>{AT_.rowid.get(), Dr_.get(), Dc_.get(), 2.0 - kPcAlpha, colv_.get()},
               stream_);
  launch_ew(m_, nullptr, DivSqrtF{Dr_.get(), rowv_.get()}, stream_);
  launch_ew(n_, nullptr, DivSqrtF{Dc_.get(), colv_.get()}, stream_);

  // Scaled matrix values, FP64 (scale_problem_kernel / scale_transposed_problem_kernel).
  a_s_.resize(nnz + kArrayPad);
  at_s_.resize(nnz + kArrayPad);
  a_s_.zero(stream_);
  at_s_.zero(stream_);
  launch_ew(nnz, nullptr, ScaleValsF{a_orig_.get(), A_.rowid.get(), A_.idx.get(), Dr_.get(), Dc_.get(), a_s_.get()},
            stream_);
  launch_ew(nnz, nullptr,
            ScaleValsF{at_orig_.get(), AT_.rowid.get(), AT_.idx.get(), Dc_.get(), Dr_.get(), at_s_.get()}, stream_);

  // Vectors.
  c_orig_.resize(n_);
  lb_orig_.resize(n_);
  ub_orig_.resize(n_);
  lc_orig_.resize(m_);
  uc_orig_.resize(m_);
  std::vector<double> cmin(lp_.c.begin(), lp_.c.end());
  if (lp_.maximize)
    for (double& v : cmin) v = -v;  // convert_to_maximization_problem (problem_helpers.cuh:132-151)
  upload_columns(cmin.data(), c_orig_);
  upload_columns(lb_.data(), lb_orig_);
  upload_columns(ub_.data(), ub_orig_);
  lc_orig_.upload(lc_, stream_);
  uc_orig_.upload(uc_, stream_);
  cs_.resize(n_);
  lbs_.resize(n_);
  ubs_.resize(n_);
  lcs_.resize(m_);
  ucs_.resize(m_);
  launch_ew(n_, red_.get(),
            ScaleColsF{c_orig_.get(), lb_orig_.get(), ub_orig_.get(), Dc_.get(), cs_.get(), lbs_.get(), ubs_.get()},
            stream_);
  launch_reduce(red_.get(), ew_grid(n_), 2, scal_.get() + S_SETUP0, stream_);
  CUOPT_HIP_CHECK(hipStreamSynchronize(stream_));
  read_scalars();
  const double cs2 = h_scal_[S_SETUP0], c2 = h_scal_[S_SETUP1];
  launch_ew(m_, red_.get(), ScaleRowsF{lc_orig_.get(), uc_orig_.get(), Dr_.get(), lcs_.get(), ucs_.get()}, stream_);
  launch_reduce(red_.get(), ew_grid(m_), 2, scal_.get() + S_SETUP0, stream_);
  read_scalars();
  const double bs2 = h_scal_[S_SETUP0], b2 = h_scal_[S_SETUP1];
  // bound_objective_rescaling (initial_scaling.cu:94-149)
  beta_   = 1.0 / (std::sqrt(bs2) + 1.0);
  gamma_  = 1.0 / (std::sqrt(cs2) + 1.0);
  b_norm_ = std::sqrt(b2);  // compute_sum_bounds on the unscaled problem
  c_norm_ = std::sqrt(c2);

  c_w_.resize(n_);
  bnd_w_.resize(n_);
  cb_w_.resize(m_);
  cb_orig_.resize(m_);
  launch_ew(n_, nullptr,
            FinishColsF<V>{cs_.get(), lbs_.get(), ubs_.get(), beta_, gamma_, true, c_w_.get(), bnd_w_.get()}, stream_);
  launch_ew(m_, nullptr,
            FinishRowsF<V>{lcs_.get(), ucs_.get(), beta_, true, cb_w_.get(), cb_orig_.get(), lc_orig_.get(),
                           uc_orig_.get()},
            stream_);
}

// Largest eigenvalue of the symmetric tridiagonal matrix (a, b) by Sturm-sequence bisection.
double tridiag_max_eig(const std::vector<double>& a, const std::vector<double>& b)
{
  const int k = (int)a.size();
  double lo = 0.0, hi = 0.0;
  for (int i = 0; i < k; ++i) {  // Gershgorin bounds
    const double r = (i > 0 ? std::fabs(b[i - 1]) : 0.0) + (i + 1 < k ? std::fabs(b[i]) : 0.0);
    lo = std::min(lo, a[i] - r);
    hi = std::max(hi, a[i] + r);
  }
  auto count_below = [&](double x) {  // eigenvalues < x
    int c    = 0;
    double d = 1.0;
    for (int i = 0; i < k; ++i) {
      d = a[i] - x - (i > 0 ? b[i - 1] * b[i - 1] / d : 0.0);
      if (d == 0.0) d = -1e-300;
      if (d < 0.0) ++c;
    }
    return c;
  };
  for (int it = 0; it < 200 && hi - lo > 1e-15 * std::max(std::fabs(hi), 1e-300); ++it) {
    const double mid = 0.5 * (lo + hi);
    if (count_below(mid) >= k) hi = mid;  // all eigenvalues below mid
    else lo = mid;
  }
  return hi;
}

// sigma_max^2 of A~ by Lanczos on A~ A~^T from a hashed start vector.  Used by
// FP32-mixed solves in place of the power iteration: both estimate the same
// eigenvalue; Lanczos gets it to near machine precision in tens of matrix
// products instead of hundreds.
template <typename V>
double PdlpSolver<V>::lanczos_sigma_max_sq(int& iters)
{
  constexpr int kGroup = 8;  // steps between host reads of the coefficients
  const int kmax = std::min(m_, 240);
  DevBuf<double> q(m_), qprev(m_), w(m_), zb(m_), atq(n_), ab(2 * (size_t)kmax + 2);
  launch_ew(m_, nullptr, HashStartF{q.get()}, stream_);
  launch_ew(m_, nullptr, FillF{qprev.get(), 0.0}, stream_);
  launch_ew(m_, red_.get(), SumSqF{q.get()}, stream_);
  launch_reduce(red_.get(), ew_grid(m_), 1, scal_.get() + S_SETUP0, stream_);
  read_scalars();
  launch_ew(m_, nullptr, ScaleByF{q.get(), 1.0 / std::sqrt(h_scal_[S_SETUP0])}, stream_);
  std::vector<double> h_ab(2 * (size_t)kmax + 2), alpha, beta;
  double theta = 0.0, prev = -1.0;
  iters = 0;
  while (iters < kmax) {
    const int j0 = iters, j1 = std::min(kmax, iters + kGroup);
    for (int j = j0; j < j1; ++j) {
      launch_tiled(AT_, at_s_.get(), nullptr, PlainEpi<double>{q.get(), atq.get()}, stream_);
      launch_tiled(A_, a_s_.get(), nullptr, PlainEpi<double>{atq.get(), zb.get()}, stream_);
      launch_ew(m_, red_.get(), LanczosAlphaF{zb.get(), q.get(), qprev.get(), ab.get(), j, w.get()}, stream_);
      launch_reduce(red_.get(), ew_grid(m_), 1, ab.get() + 2 * j, stream_);
      launch_ew(m_, red_.get(), LanczosOrthF{q.get(), ab.get(), j, w.get()}, stream_);
      launch_reduce(red_.get(), ew_grid(m_), 1, ab.get() + 2 * j + 1, stream_);
      launch_ew(m_, nullptr, LanczosNextF{q.get(), qprev.get(), w.get(), ab.get(), j}, stream_);
    }
    ab.download(h_ab.data(), 2 * (size_t)j1, stream_);
    CUOPT_HIP_CHECK(hipStreamSynchronize(stream_));
    bool stop = false;
    for (int j = j0; j < j1 && !stop; ++j) {
      alpha.push_back(h_ab[2 * j]);
      iters = j + 1;
      theta = tridiag_max_eig(alpha, beta);
      const double b2 = h_ab[2 * j + 1];
      // Ritz values rise monotonically to sigma_max^2; a relative change of 1e-6
      // leaves an error far inside the 0.2 % margin of eta = 0.998 / sigma_max
      if (j > 2 && std::fabs(theta - prev) <= 1e-6 * theta) stop = true;
      if (!(b2 > 1e-26 * theta * theta)) stop = true;                      // invariant subspace
      prev = theta;
      beta.push_back(std::sqrt(b2));
    }
    if (stop) break;
  }
  return theta;
}

template <typename V>
void PdlpSolver<V>::power_iteration()
{
  // compute_initial_step_size, power-iteration branch (pdlp.cu:2622-2738).
  if constexpr (std::is_same<V, float>::value) {
    sigma_max_sq_ = lanczos_sigma_max_sq(power_iters_);
    eta_          = 0.998 / std::sqrt(sigma_max_sq_);
    return;
  }
  std::vector<double> z(m_);
  std::mt19937 gen(1);
  std::normal_distribution<double> dist(0.0, 1.0);
  for (int i = 0; i < m_; ++i) z[i] = dist(gen);
  DevBuf<double> za(m_), zb(m_), atq(n_);
  DevBuf<PowerState> st(1);
  za.upload(z, stream_);
  launch_ew(m_, red_.get(), SumSqF{za.get()}, stream_);
  launch_reduce(red_.get(), ew_grid(m_), 1, scal_.get() + S_SETUP0, stream_);
  read_scalars();
  PowerState h{};
  h.nrm = std::sqrt(h_scal_[S_SETUP0]);
  CUOPT_HIP_CHECK(hipMemcpyAsync(st.get(), &h, sizeof(h), hipMemcpyHostToDevice, stream_));
  double* zo = za.get();
  double* zn = zb.get();
  auto launch_batch = [&](int batch) {
    for (int b = 0; b < batch; ++b) {
      launch_tiled(AT_, at_s_.get(), nullptr, PowerAtEpi{zo, st.get(), atq.get()}, stream_);
      launch_tiled(A_, a_s_.get(), red_.get(), PowerAEpi{atq.get(), zo, st.get(), zn}, stream_);
      if (m_ <= kPowerSmallRows) {
        hipLaunchKernelGGL(power_step_small, dim3(1), dim3(1024), 0, stream_, red_.get(), A_.plan.red_blocks(),
                           (const double*)zo, (const double*)zn, m_, st.get(), 1e-4);
      } else {
        hipLaunchKernelGGL(power_finalize_s2, dim3(1), dim3(1024), 0, stream_, red_.get(), A_.plan.red_blocks(),
                           st.get());
        launch_ew(m_, red_.get(), PowerResF{zo, zn, st.get()}, stream_);
        hipLaunchKernelGGL(power_finalize_res, dim3(1), dim3(1024), 0, stream_, red_.get(), ew_grid(m_), st.get(),
                           1e-4);
      }
      CUOPT_HIP_CHECK(hipGetLastError());
      std::swap(zo, zn);
    }
  };
  // A full batch of 16 steps (even, so it starts and ends on the same buffer) is
  // captured once and replayed: on small problems its 80 launches cost more than
  // their kernels.
  constexpr int kPowerBatch = 16;
  hipGraphExec_t batch_graph = nullptr;
  struct GraphGuard {
    hipGraphExec_t* g;
    ~GraphGuard()
    {
      if (*g) (void)hipGraphExecDestroy(*g);
    }
  } graph_guard{&batch_graph};
  int launched = 0;
  while (launched < 5000) {
    const int batch = std::min(kPowerBatch, 5000 - launched);
    if (use_graphs_ && batch == kPowerBatch) {
      if (!batch_graph) {
        hipGraph_t g;
        CUOPT_HIP_CHECK(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
        try {
          launch_batch(batch);
        } catch (...) {
          (void)hipStreamEndCapture(stream_, &g);
          if (g) (void)hipGraphDestroy(g);
          throw;
        }
        CUOPT_HIP_CHECK(hipStreamEndCapture(stream_, &g));
        const hipError_t r = hipGraphInstantiate(&batch_graph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        CUOPT_HIP_CHECK(r);
      }
      CUOPT_HIP_CHECK(hipGraphLaunch(batch_graph, stream_));
    } else {
      launch_batch(batch);
    }
    launched += batch;
    CUOPT_HIP_CHECK(hipMemcpyAsync(&h, st.get(), sizeof(h), hipMemcpyDeviceToHost, stream_));
    CUOPT_HIP_CHECK(hipStreamSynchronize(stream_));
    if (h.done != 0.0) break;
  }
  sigma_max_sq_ = h.s2;
  power_iters_  = (int)h.iters;
  eta_          = 0.998 / std::sqrt(sigma_max_sq_);
}

template <typename V>
void PdlpSolver<V>::init_iterates()
{
  for (int b = 0; b < 2; ++b) {
    x_[b].resize(n_);
    y_[b].resize(m_);
  }
  x0_.resize(n_);
  y0_.resize(m_);
  xbar_.resize(n_);
  ybar_.resize(m_);
  xhat_.resize(n_);
  yhat_.resize(m_);
  slack_.resize(n_);
  aty_.resize(n_);
  xu_.resize(n_);
  yu_.resize(m_);
  su_.resize(n_);
  step_sc_.resize(1);
  DevBuf<double> xi, yi;
  if (!st_.initial_primal.empty()) {
    xi.resize(n_);
    upload_columns(st_.initial_primal.data(), xi);
  }
  if (!st_.initial_dual.empty()) {
    yi.resize(m_);
    yi.upload(st_.initial_dual.data(), m_, stream_);
  }
  launch_ew(n_, nullptr,
            InitIterateF<V, true>{xi.get(), Dc_.get(), beta_, true, bnd_w_.get(), x_[0].get(), xhat_.get(), x0_.get()},
            stream_);
  launch_ew(m_, nullptr,
            InitIterateF<V, false>{yi.get(), Dr_.get(), gamma_, true, nullptr, y_[0].get(), yhat_.get(), y0_.get()},
            stream_);
  launch_ew(n_, nullptr, ZeroF<V>{slack_.get()}, stream_);
  CUOPT_HIP_CHECK(hipStreamSynchronize(stream_));  // xi / yi are freed on return
  cur_ = 0;
}

template <typename V>
void PdlpSolver<V>::evaluate_kkt()
{
  const double ib = 1.0 / beta_, ig = 1.0 / gamma_;
  const bool rf   = refine_active_;
  const double* xs = rf ? xstar_.get() : nullptr;
  const double* ys = rf ? ystar_.get() : nullptr;
  const double idp = rf ? 1.0 / refine_.dp : 1.0, idd = rf ? 1.0 / refine_.dd : 1.0;
  // xu_ holds the gathered point until the column pass writes the final x there.
  launch_ew(n_, nullptr, KktPointF<V>{xhat_.get(), Dc_.get(), ib, xs, idp, xu_.get()}, stream_);
  const bool per = st_.per_constraint_residual;
  if (per && rp_res_.size() < (size_t)std::max(m_, 1)) rp_res_.resize(std::max(m_, 1));
  if (per && rd_res_.size() < (size_t)std::max(n_, 1)) rd_res_.resize(std::max(n_, 1));
  launch_tiled(A_, a_orig_.get(), red_.get(),
               KktRowsEpi<V>{xu_.get(), yhat_.get(), Dr_.get(), ig, gamma_, cb_orig_.get(), y0_.get(), yu_.get(), xs,
                             ys, idp, idd, per ? rp_res_.get() : nullptr},
               stream_);
  launch_reduce(red_.get(), A_.plan.red_blocks(), 3, scal_.get() + S_KKT_RP2, stream_);
  launch_tiled(AT_, at_orig_.get(), red_.get(),
               KktColsEpi<V>{yu_.get(), xhat_.get(), slack_.get(), Dc_.get(), ib, beta_, ig, gamma_, c_orig_.get(),
                             x0_.get(), xu_.get(), su_.get(), xs, idp, idd, bnd_w_.get(), lb_orig_.get(),
                             ub_orig_.get(), per ? rd_res_.get() : nullptr},
               stream_);
  launch_reduce(red_.get(), AT_.plan.red_blocks(), 4, scal_.get() + S_KKT_RD2, stream_);
  if (per) {
    const int ab = st_.per_constraint_absolute;
    hipLaunchKernelGGL(linf_rel_kernel, dim3(1), dim3(1024), 0, stream_, rp_res_.get(), cb_orig_.get(), nullptr, m_,
                       st_.rel_primal, ab, scal_.get() + S_LINF_P);
    CUOPT_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(linf_rel_kernel, dim3(1), dim3(1024), 0, stream_, rd_res_.get(), nullptr, c_orig_.get(), n_,
                       st_.rel_dual, ab, scal_.get() + S_LINF_D);
    CUOPT_HIP_CHECK(hipGetLastError());
  }
  read_scalars();
  const double* s = h_scal_;
  l2p_            = std::sqrt(s[S_KKT_RP2]);
  l2d_            = std::sqrt(s[S_KKT_RD2]);
  double pobj     = s[S_KKT_XC];
  double dobj     = s[S_KKT_SX] + s[S_KKT_PSUM];
  if (obj_sf_ != 1.0 || obj_off_ != 0.0) {
    pobj = obj_sf_ * (pobj + obj_off_);
    dobj = obj_sf_ * (dobj + obj_off_);
  }
  pobj_   = pobj;
  dobj_   = dobj;
  gap_    = std::fabs(pobj - dobj);
  absobj_ = std::fabs(pobj) + std::fabs(dobj);
  dp_     = s[S_KKT_DP];
  dd_     = s[S_KKT_DD];
  if (per) {
    linf_p_ = s[S_LINF_P];
    linf_d_ = s[S_LINF_D];
  }
}

template <typename V>
void PdlpSolver<V>::record_best(const Tolerances& t)
{
  const bool feas = primal_feasible(t);
  bool better     = !have_best_;
  if (have_best_) {
    if (feas != best_feasible_) better = feas;
    else if (feas) better = lp_.maximize ? pobj_ > best_pobj_ : pobj_ < best_pobj_;
    else better = l2p_ < best_l2p_;
  }
  if (!better) return;
  if (!have_best_) {
    best_x_.resize(std::max(n_, 1));
    best_y_.resize(std::max(m_, 1));
    best_s_.resize(std::max(n_, 1));
  }
  CUOPT_HIP_CHECK(hipMemcpyAsync(best_x_.get(), xu_.get(), n_ * sizeof(double), hipMemcpyDeviceToDevice, stream_));
  CUOPT_HIP_CHECK(hipMemcpyAsync(best_y_.get(), yu_.get(), m_ * sizeof(double), hipMemcpyDeviceToDevice, stream_));
  CUOPT_HIP_CHECK(hipMemcpyAsync(best_s_.get(), su_.get(), n_ * sizeof(double), hipMemcpyDeviceToDevice, stream_));
  have_best_     = true;
  best_feasible_ = feas;
  best_l2p_      = l2p_;
  best_l2d_      = l2d_;
  best_pobj_     = pobj_;
  best_dobj_     = dobj_;
  best_gap_      = gap_;
  best_absobj_   = absobj_;
}

template <typename V>
void PdlpSolver<V>::restore_best()
{
  if (!have_best_) return;
  CUOPT_HIP_CHECK(hipMemcpyAsync(xu_.get(), best_x_.get(), n_ * sizeof(double), hipMemcpyDeviceToDevice, stream_));
  CUOPT_HIP_CHECK(hipMemcpyAsync(yu_.get(), best_y_.get(), m_ * sizeof(double), hipMemcpyDeviceToDevice, stream_));
  CUOPT_HIP_CHECK(hipMemcpyAsync(su_.get(), best_s_.get(), n_ * sizeof(double), hipMemcpyDeviceToDevice, stream_));
  l2p_    = best_l2p_;
  l2d_    = best_l2d_;
  pobj_   = best_pobj_;
  dobj_   = best_dobj_;
  gap_    = best_gap_;
  absobj_ = best_absobj_;
}

template <typename V>
void PdlpSolver<V>::launch_step(int j, bool major, bool fpe, double* ny_slot, double* fpe_slot)
{
  const int c = cur_, nx = cur_ ^ 1;
#define CUOPT_PRIMAL(MAJ, FP)                                                                               \
  launch_tiled(AT_, at_w_.get(), nullptr,                                                                    \
               PrimalStepEpi<V, MAJ, FP>{y_[c].get(), x_[c].get(), c_w_.get(), bnd_w_.get(), x0_.get(),     \
                                         xbar_.get(), x_[nx].get(), xhat_.get(), slack_.get(), aty_.get(),  \
                                         step_sc_.get(), j, refine_active_},                                \
               stream_)
#define CUOPT_DUAL(MAJ, FP)                                                                                  \
  launch_tiled(A_, a_w_.get(), red_.get(),                                                                    \
               DualStepEpi<V, MAJ, FP>{xbar_.get(), y_[c].get(), cb_w_.get(), y0_.get(), y_[nx].get(),        \
                                       yhat_.get(), ybar_.get(), step_sc_.get(), j,                           \
                                       refine_active_ ? kink_.get() : nullptr},                               \
               stream_)
  if (ev_[0]) CUOPT_HIP_CHECK(hipEventRecord(ev_[0], stream_));
  if (major) {
    if (fpe) CUOPT_PRIMAL(true, true);
    else CUOPT_PRIMAL(true, false);
  } else {
    if (fpe) CUOPT_PRIMAL(false, true);
    else CUOPT_PRIMAL(false, false);
  }
  if (ev_[1]) CUOPT_HIP_CHECK(hipEventRecord(ev_[1], stream_));
  if (major) {
    if (fpe) CUOPT_DUAL(true, true);
    else CUOPT_DUAL(true, false);
  } else {
    if (fpe) CUOPT_DUAL(false, true);
    else CUOPT_DUAL(false, false);
  }
#undef CUOPT_PRIMAL
#undef CUOPT_DUAL
  if (ev_[2]) CUOPT_HIP_CHECK(hipEventRecord(ev_[2], stream_));
  if (fpe) {
    launch_reduce(red_.get(), A_.plan.red_blocks(), 1, ny_slot, stream_);
    launch_tiled(AT_, at_w_.get(), red_.get(), FpeEpi<V>{ybar_.get(), aty_.get(), xbar_.get(), x_[c].get()},
                 stream_);
    launch_reduce(red_.get(), AT_.plan.red_blocks(), 2, fpe_slot, stream_);
  }
  cur_ = nx;
}

template <typename V>
void PdlpSolver<V>::finish(SolveResult& R, int status)
{
  R.status = status;
  R.x.resize(n_);
  R.y.resize(m_);
  R.rc.resize(n_);
  yu_.download(R.y.data(), m_, stream_);
  download_columns(xu_, R.x.data());
  download_columns(su_, R.rc.data());
  if (lp_.maximize) {  // solve.cu:583-588
    for (double& v : R.y) v = -v;
    for (double& v : R.rc) v = -v;
  }
  R.primal_objective              = pobj_;
  R.dual_objective                = dobj_;
  R.info[CUOPT_AMD_INFO_L2_PRIMAL]  = l2p_;
  R.info[CUOPT_AMD_INFO_REL_PRIMAL] = l2p_ / (1.0 + b_norm_);
  R.info[CUOPT_AMD_INFO_L2_DUAL]    = l2d_;
  R.info[CUOPT_AMD_INFO_REL_DUAL]   = l2d_ / (1.0 + c_norm_);
  R.info[CUOPT_AMD_INFO_PRIMAL_OBJ] = pobj_;
  R.info[CUOPT_AMD_INFO_DUAL_OBJ]   = dobj_;
  R.info[CUOPT_AMD_INFO_GAP]        = gap_;
  R.info[CUOPT_AMD_INFO_REL_GAP]    = gap_ / (1.0 + absobj_);
  R.info[CUOPT_AMD_INFO_OMEGA]      = omega_;
  R.solve_time                      = elapsed();
  R.info[CUOPT_AMD_INFO_SOLVE_TIME] = R.solve_time;
}

template <typename V>
SolveResult PdlpSolver<V>::run()
{
  SolveResult R;
  t0_ = clock::now();
  CUOPT_HIP_CHECK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
  CUOPT_HIP_CHECK(hipHostMalloc((void**)&h_scal_, S_COUNT * sizeof(double), hipHostMallocDefault));
  CUOPT_HIP_CHECK(hipHostMalloc((void**)&h_step_, sizeof(StepScalars), hipHostMallocDefault));
  if (const char* g = std::getenv("CUOPT_AMD_GRAPHS")) use_graphs_ = std::atoi(g) != 0;
  obj_sf_  = lp_.maximize ? -1.0 : 1.0;
  obj_off_ = lp_.maximize ? -lp_.offset : lp_.offset;

  auto phase_end = [&](int slot, double since) {
    CUOPT_HIP_CHECK(hipStreamSynchronize(stream_));
    const double now = elapsed();
    R.info[slot]     = now - since;
    return now;
  };
  double tph = elapsed();
  build_structures();
  tph = phase_end(CUOPT_AMD_INFO_STRUCTURE_TIME, tph);
  setup_scaling();
  tph = phase_end(CUOPT_AMD_INFO_SCALING_TIME, tph);
  power_iteration();
  phase_end(CUOPT_AMD_INFO_POWER_TIME, tph);
  // working-precision matrices
  if constexpr (std::is_same<V, double>::value) {
    a_w_ = std::move(a_s_);
    at_w_ = std::move(at_s_);
  } else {
    a_w_.resize(A_.nnz + kArrayPad);
    at_w_.resize(A_.nnz + kArrayPad);
    a_w_.zero(stream_);
    at_w_.zero(stream_);
    launch_ew(A_.nnz, nullptr, ConvertF<V>{a_s_.get(), a_w_.get()}, stream_);
    launch_ew(A_.nnz, nullptr, ConvertF<V>{at_s_.get(), at_w_.get()}, stream_);
  }
  R.info[CUOPT_AMD_INFO_TUNE_TIME] = 0.0;
  if (st_.spmv_autotune) {
    const double t = elapsed();
    SpmvChoice ca, cat;
    const bool timed   = autotune_spmv<V>(A_, a_w_.get(), a_src_, stream_, st_.long_row_merge, false, ca);
    const bool timed_t = autotune_spmv<V>(AT_, at_w_.get(), ColSource{}, stream_, st_.long_row_merge, true, cat);
    CUOPT_HIP_CHECK(hipStreamSynchronize(stream_));
    if (timed || timed_t) R.info[CUOPT_AMD_INFO_TUNE_TIME] = elapsed() - t;
    R.info[CUOPT_AMD_INFO_LANES_BIAS_A]  = ca.bias;
    R.info[CUOPT_AMD_INFO_LANES_BIAS_AT] = cat.bias;
  } else {
    R.info[CUOPT_AMD_INFO_LANES_BIAS_A] = R.info[CUOPT_AMD_INFO_LANES_BIAS_AT] = 0.0;
  }
  R.info[CUOPT_AMD_INFO_TILE_A]   = A_.tile;
  R.info[CUOPT_AMD_INFO_TILE_AT]  = AT_.tile;
  R.info[CUOPT_AMD_INFO_MERGE_A]  = A_.plan.nlong > 0 ? (A_.plan.fused ? 1 : 0) : -1;
  R.info[CUOPT_AMD_INFO_MERGE_AT] = AT_.plan.nlong > 0 ? (AT_.plan.fused ? 1 : 0) : -1;
  // Stable3: initial primal weight 1 (bound_objective_rescaling branch, pdlp.cu:2799-2807).
  omega_ = best_omega_ = 1.0;
  tau_   = eta_ / omega_;
  sigma_ = eta_ * omega_;
  init_iterates();
  if constexpr (!std::is_same<V, double>::value) {
    CUOPT_HIP_CHECK(hipStreamSynchronize(stream_));
    a_s_.release();
    at_s_.release();
  }
  R.info[CUOPT_AMD_INFO_ETA0]         = eta_;
  R.info[CUOPT_AMD_INFO_OMEGA0]       = omega_;
  R.info[CUOPT_AMD_INFO_POWER_ITERS]  = power_iters_;
  R.info[CUOPT_AMD_INFO_SIGMA_MAX_SQ] = sigma_max_sq_;
  R.info[CUOPT_AMD_INFO_BETA]         = beta_;
  R.info[CUOPT_AMD_INFO_GAMMA]        = gamma_;
  R.info[CUOPT_AMD_INFO_PRECISION]    = std::is_same<V, double>::value ? 0 : 1;
  CUOPT_HIP_CHECK(hipStreamSynchronize(stream_));
  setup_time_                       = elapsed();
  R.info[CUOPT_AMD_INFO_SETUP_TIME] = setup_time_;

  const Tolerances user{st_.abs_primal, st_.rel_primal, st_.abs_dual, st_.rel_dual, st_.abs_gap, st_.rel_gap};
  const bool refine = !std::is_same<V, double>::value && st_.refinement != 0 && st_.bench_steps <= 0;
  finish(R, pdhg_loop(R, user, refine));
  R.info[CUOPT_AMD_INFO_STEPS]         = (double)k_done_;
  R.info[CUOPT_AMD_INFO_RESTARTS]      = restarts_;
  R.info[CUOPT_AMD_INFO_CHECKS]        = checks_;
  R.info[CUOPT_AMD_INFO_LOOP_TIME]     = R.solve_time - setup_time_;
  R.info[CUOPT_AMD_INFO_REFINE_ROUNDS] = recentrings_;
  return R;
}

// Recentring at a restart (FP64 iterative refinement of FP32-mixed solves,
// applied at every restart once the relative KKT error is below 1e-2).
// At a restart the PDHG state is the single point (x-hat, y-hat), whose FP64
// values in the original problem the check has just left in xu_ / yu_.  The
// iterates are re-expressed as a correction problem around that point,
//   d = Dp (x - x*) / Dc,   w = Dd (y - y*) / Dr,
// which is the original LP exactly (objective Dd (c - A^T y*), bounds
// Dp ([l, u] - x*), constraint bounds Dp ([l_c, u_c] - A x*), and the kink of
// the dual term at y = 0 moved to w = -Dd y* / Dr), so the FP32 iterates only
// carry the part of the solution not yet resolved.  Dp and Dd put the primal
// and dual residuals of the correction at about one.  The point itself becomes
// d = w = 0: the restart is applied here, and PDHG continues as if restarted
// in the original coordinates (the PID error terms are scale-invariant; the
// primal weight is converted to the new units).
template <typename V>
void PdlpSolver<V>::recentre()
{
  double dp = 1.0 / std::max(l2p_, 1e-14 * (1.0 + b_norm_));
  double dd = 1.0 / std::max(l2d_, 1e-14 * (1.0 + c_norm_));
  dp        = std::max(dp, 1.0);
  dd        = std::max(dd, 1.0);
  if (dp * dd * gap_ > 1.0) {  // also keep the correction's duality gap at about one
    const double f = std::sqrt(1.0 / (dp * dd * gap_));
    dp = std::max(dp * f, 1.0);
    dd = std::max(dd * f, 1.0);
  }
  // Primal weight carried over into the units of the new correction problem:
  // scaled primal iterates were x * beta / Dc before the first round and
  // Dp * (x - x*) / Dc after it (dual: y * gamma / Dr, Dd * (y - y*) / Dr).
  const double w_prev = omega_;
  const double p_prev = refine_.dp > 0.0 ? refine_.dp : beta_;
  const double d_prev = refine_.dd > 0.0 ? refine_.dd : gamma_;
  const double ratio  = (dd / d_prev) / (dp / p_prev);
  if (refine_.dp == 0.0) {
    xstar_.resize(n_);
    ystar_.resize(m_);
    ax_.resize(m_);
    aty64_.resize(n_);
    kink_.resize(m_);
  }
  if (st_.log_to_console && std::getenv("CUOPT_AMD_LOG_REFINEMENT")) {
    std::printf("Recentre: rel KKT %.3e (primal %.3e dual %.3e gap %.3e) -> Dp %.3e Dd %.3e omega %.3e\n",
                rel_kkt_error(), l2p_ / (1.0 + b_norm_), l2d_ / (1.0 + c_norm_), gap_ / (1.0 + absobj_), dp, dd,
                w_prev * ratio);
    std::fflush(stdout);
  }
  refine_.dp = dp;
  refine_.dd = dd;
  launch_ew(n_, nullptr, CaptureColsF{xu_.get(), lb_orig_.get(), ub_orig_.get(), xstar_.get()}, stream_);
  CUOPT_HIP_CHECK(hipMemcpyAsync(ystar_.get(), yu_.get(), (size_t)m_ * sizeof(double), hipMemcpyDeviceToDevice,
                                 stream_));
  launch_tiled(A_, a_orig_.get(), nullptr, PlainEpi<double>{xstar_.get(), ax_.get()}, stream_);
  launch_tiled(AT_, at_orig_.get(), nullptr, PlainEpi<double>{ystar_.get(), aty64_.get()}, stream_);
  launch_ew(n_, nullptr,
            CorrectionColsF<V>{xstar_.get(), aty64_.get(), c_orig_.get(), lb_orig_.get(), ub_orig_.get(), Dc_.get(),
                               dp, dd, c_w_.get(), bnd_w_.get()},
            stream_);
  launch_ew(m_, nullptr,
            CorrectionRowsF<V>{ystar_.get(), ax_.get(), lc_orig_.get(), uc_orig_.get(), Dr_.get(), dp, dd,
                               cb_w_.get(), kink_.get()},
            stream_);
  // Correction problems are solved without the bound / objective rescaling.
  beta_  = 1.0;
  gamma_ = 1.0;
  launch_ew(n_, nullptr,
            InitIterateF<V, true>{nullptr, Dc_.get(), 1.0, false, bnd_w_.get(), x_[0].get(), xhat_.get(), x0_.get()},
            stream_);
  launch_ew(m_, nullptr,
            InitIterateF<V, false>{nullptr, Dr_.get(), 1.0, false, nullptr, y_[0].get(), yhat_.get(), y0_.get()},
            stream_);
  launch_ew(n_, nullptr, ZeroF<V>{slack_.get()}, stream_);
  cur_   = 0;
  omega_      = bound_w(w_prev * ratio);
  best_omega_ = bound_w(best_omega_ * ratio);
  ++recentrings_;
  tau_   = eta_ / omega_;
  sigma_ = eta_ * omega_;
}

template <typename V>
int PdlpSolver<V>::pdhg_loop(SolveResult& R, const Tolerances& tol, bool refine)
{
  const double nan    = std::numeric_limits<double>::quiet_NaN();
  double fpe = nan, fpe0 = nan, last_trial = std::numeric_limits<double>::infinity();
  double err_sum = 0.0, last_err = 0.0, best_gap = std::numeric_limits<double>::infinity();
  long long k        = 0;
  bool restarted_prev = false;   // a restart happened at the previous check
  bool periodic_prev  = false;   // the previous block ran a periodic FPE step
  const long long k0  = k_done_;  // iterations of earlier loops (none: one loop per solve)
  kr_ = 0;
  refine_active_ = false;

  // Benchmark window (cuOptAmdBenchmarkSteps): [bench_w, bench_end) is timed on
  // the production path; the next `ev_iters` iterations run with direct
  // launches bracketed by HIP events for per-kernel times.  Window edges are
  // moved to check iterations, where the host synchronises anyway.
  const bool bench = st_.bench_steps > 0;
  auto round_up_check = [&](long long v) {
    while (!is_check(v)) ++v;
    return v;
  };
  const long long bench_w   = bench ? round_up_check(st_.bench_warmup) : -1;
  const long long bench_end = bench ? round_up_check(bench_w + st_.bench_steps) : -1;
  const long long ev_end    = bench ? round_up_check(bench_end + std::min<long long>(st_.bench_steps, 400)) : -1;
  clock::time_point tb0;
  double window_ms = 0.0;
  int bench_checks = 0;
  std::vector<hipEvent_t> events;  // 3 per evented iteration
  auto done = [&](int status) {
    k_done_ = k0 + k;
    return status;
  };
  auto bench_stop  = [&]() {
    CUOPT_HIP_CHECK(hipStreamSynchronize(stream_));
    double tp = 0.0, td = 0.0;
    const size_t nit = events.size() / 3;
    for (size_t i = 0; i < nit; ++i) {
      float a = 0.f, b = 0.f;
      CUOPT_HIP_CHECK(hipEventElapsedTime(&a, events[3 * i], events[3 * i + 1]));
      CUOPT_HIP_CHECK(hipEventElapsedTime(&b, events[3 * i + 1], events[3 * i + 2]));
      tp += a;
      td += b;
    }
    for (hipEvent_t e : events) (void)hipEventDestroy(e);
    events.clear();
    R.bench[0] = window_ms;
    R.bench[1] = nit ? tp / nit : 0.0;
    R.bench[2] = nit ? td / nit : 0.0;
    R.bench[3] = (double)(bench_end - bench_w);
    R.bench[4] = setup_time_;
    R.bench[5] = bench_checks;
    R.bench[6] = CUOPT_TERIMINATION_STATUS_ITERATION_LIMIT;
    R.bench[7] = std::is_same<V, double>::value ? 0 : 1;
    R.bench[8] = (double)nit;
    return done(CUOPT_TERIMINATION_STATUS_ITERATION_LIMIT);
  };

  while (true) {
    if (bench && k == bench_w) {
      CUOPT_HIP_CHECK(hipStreamSynchronize(stream_));
      tb0 = clock::now();
    }
    if (bench && k == bench_end) {
      CUOPT_HIP_CHECK(hipStreamSynchronize(stream_));
      window_ms = std::chrono::duration<double, std::milli>(clock::now() - tb0).count();
    }
    if (bench && k >= ev_end) return bench_stop();
    if (bench && k >= bench_w && k < bench_end) ++bench_checks;
    // ---- termination check (k is a check iteration) ----
    const bool is_major = ((k % kMajor) == 0 && k > 0) || k <= 0;
    evaluate_kkt();
    ++checks_;
    if (restarted_prev) {
      fpe  = std::sqrt((h_scal_[S_FPE0_NX] * omega_ + h_scal_[S_FPE0_NY] / omega_) + 2.0 * h_scal_[S_FPE0_I] * eta_);
      fpe0 = fpe;
    }
    if (periodic_prev)
      fpe = std::sqrt((h_scal_[S_FPE_NX] * omega_ + h_scal_[S_FPE_NY] / omega_) + 2.0 * h_scal_[S_FPE_I] * eta_);
    const size_t rec = R.trace.size();
    R.trace.insert(R.trace.end(), {(double)(k0 + k), l2p_, l2d_, pobj_, dobj_, gap_, 0.0, omega_, fpe, (double)kr_});
    if (k > 1 && is_optimal(tol)) return done(CUOPT_TERIMINATION_STATUS_OPTIMAL);
    // first_primal_feasible (pdlp.cu:928-983): a primal feasible, not optimal iterate ends the solve
    if (k > 1 && st_.first_primal_feasible && !bench && primal_feasible(tol))
      return done(CUOPT_TERIMINATION_STATUS_PRIMAL_FEASIBLE);
    // pdlp_stop_on_nan (an extension): a NaN residual can never satisfy the test
    // again, so stop with a numerical error.  Off, as in the reference's Stable3
    // loop (pdlp.cu:854-1180), the solve runs on to its limits.
    if (st_.stop_on_nan && (std::isnan(l2p_) || std::isnan(l2d_) || std::isnan(gap_)))
      return done(CUOPT_TERIMINATION_STATUS_NUMERICAL_ERROR);
    if (st_.save_best_primal_so_far && !bench) record_best(tol);  // pdlp.cu:1171-1175
    if (elapsed() >= st_.time_limit) {
      if (st_.save_best_primal_so_far && !bench) restore_best();  // pdlp.cu:341-347
      return done(CUOPT_TERIMINATION_STATUS_TIME_LIMIT);
    }
    if (k0 + k >= st_.iteration_limit) {
      if (st_.save_best_primal_so_far && !bench) restore_best();  // pdlp.cu:394-401
      return done(CUOPT_TERIMINATION_STATUS_ITERATION_LIMIT);
    }

    bool restarted = false;
    if (is_major) {
      bool do_restart = false;
      if (k == kMajor) {
        do_restart = true;
      } else if (k > kMajor) {
        if (fpe <= kSufficient * fpe0)
          do_restart = true;
        else if (fpe <= kNecessary * fpe0 && fpe > last_trial)
          do_restart = true;
        else if ((double)kr_ >= kArtificial * (double)k)
          do_restart = true;
      }
      last_trial = fpe;
      if (do_restart) {
        // cupdlpx_new_primal_weight_computation (pdlp_restart_strategy.cu:757-835)
        const double l2pd  = std::sqrt(dp_), l2dd = std::sqrt(dd_);
        const double rel_d = l2d_ / (1.0 + c_norm_), rel_p = l2p_ / (1.0 + b_norm_);
        const double ratio = (rel_p == 0.0) ? std::numeric_limits<double>::infinity() : rel_d / rel_p;
        if (l2pd > 1e-16 && l2dd > 1e-16 && l2pd < 1e12 && l2dd < 1e12 && ratio > 1e-8 && ratio < 1e8) {
          const double cur = omega_;
          const double err = std::log(l2dd) - std::log(l2pd) - std::log(cur);
          const double s   = err_sum * kISmooth + err;
          err_sum          = s;
          const double de  = err - last_err;
          const double w   = bound_w(std::exp(kKp * err + kKi * s + kKd * de) * cur);
          omega_           = w;
          last_err         = err;
          tau_             = eta_ / w;
          sigma_           = eta_ * w;
        } else {
          omega_   = best_omega_;
          err_sum  = 0.0;
          last_err = 0.0;
          tau_     = eta_ / best_omega_;
          sigma_   = eta_ * best_omega_;
        }
        const double g = std::fabs(std::log10(ratio));
        if (g < best_gap) {
          best_gap    = g;
          best_omega_ = omega_;
        }
        if (refine && (refine_active_ || rel_kkt_error() <= kRecentreFrom)) {
          recentre();  // the restart point becomes the origin of the new coordinates
          refine_active_ = true;
        } else {
          launch_ew(n_, nullptr, RestartCopyF<V>{xhat_.get(), x_[cur_].get(), x0_.get()}, stream_);
          launch_ew(m_, nullptr, RestartCopyF<V>{yhat_.get(), y_[cur_].get(), y0_.get()}, stream_);
        }
        kr_        = 0;
        last_trial = std::numeric_limits<double>::infinity();
        restarted  = true;
        ++restarts_;
      }
    }
    R.trace[rec + 6] = restarted ? 1.0 : 0.0;
    R.trace[rec + 7] = omega_;
    upload_step_scalars();

    // ---- PDHG iterations up to the next check ----
    long long next = k + 1;
    while (!is_check(next)) ++next;
    periodic_prev = false;
    const bool evented = bench && k >= bench_end;
    std::string key;
    if (use_graphs_ && !evented) {
      key = std::to_string(next - k) + (restarted ? "R" : "") + (refine_active_ ? "K:" : ":");
      for (long long s = k; s < next; ++s) {
        const long long s2 = s + 2;
        if (((s + 1) % kMajor == 0) || (s2 % conditional_major(s2)) == 0) key += std::to_string(s - k) + ",";
        if (((s + 1) % kMajor) == 0) key += "p" + std::to_string(s - k) + ",";
      }
    }
    auto launch_block = [&]() {
      for (long long s = k; s < next; ++s) {
        hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
        if (evented) {
          for (hipEvent_t& e : ev) {
            CUOPT_HIP_CHECK(hipEventCreate(&e));
            events.push_back(e);
          }
        }
        ev_ = ev;
        const long long s2      = s + 2;
        const bool should_major = ((s + 1) % kMajor == 0) || (s2 % conditional_major(s2)) == 0;
        const bool periodic     = ((s + 1) % kMajor) == 0;
        const bool post_restart = restarted && s == k;
        const bool fpe_step     = periodic || post_restart;
        double* ny_slot  = scal_.get() + (post_restart ? S_FPE0_NY : S_FPE_NY);
        double* fpe_slot = scal_.get() + (post_restart ? S_FPE0_I : S_FPE_I);
        launch_step((int)(s - k), should_major, fpe_step, ny_slot, fpe_slot);
        ev_ = nullptr;
      }
    };
    for (long long s = k; s < next; ++s)
      if (((s + 1) % kMajor) == 0) periodic_prev = true;
    if (!key.empty()) {
      // Blocks have even length, so the ping-pong buffers are in the same state
      // at every block start and a captured block is valid for every replay.
      auto it = graphs_.find(key);
      if (it == graphs_.end()) {
        const int cur0 = cur_;
        hipGraph_t g;
        CUOPT_HIP_CHECK(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
        launch_block();
        CUOPT_HIP_CHECK(hipStreamEndCapture(stream_, &g));
        hipGraphExec_t ge;
        CUOPT_HIP_CHECK(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
        CUOPT_HIP_CHECK(hipGraphDestroy(g));
        it = graphs_.emplace(key, ge).first;
        cur_ = cur0;
      }
      CUOPT_HIP_CHECK(hipGraphLaunch(it->second, stream_));
    } else {
      launch_block();
    }
    kr_ += next - k;
    k              = next;
    restarted_prev = restarted;
  }
}

// ---------------------------------------------------------------------------
// solve_lp: problem checks and dispatch (solve.cu:1123-1342, PDLP path)
// ---------------------------------------------------------------------------

// Constraint bounds from row types (transform_bounds_functor, problem_helpers.cuh:26-51).
void constraint_bounds(const LpHost& lp, std::vector<double>& lc, std::vector<double>& uc)
{
  const int m = lp.m;
  lc.resize(m);
  uc.resize(m);
  const double inf = std::numeric_limits<double>::infinity();
  for (int i = 0; i < m; ++i) {
    if (lp.ranged) {
      lc[i] = lp.lc[i];
      uc[i] = lp.uc[i];
    } else {
      const char t = lp.sense[i];
      const double b = lp.rhs[i];
      if (t == 'E') {
        lc[i] = b;
        uc[i] = b;
      } else if (t == 'G') {
        lc[i] = b;
        uc[i] = inf;
      } else if (t == 'L') {
        lc[i] = -inf;
        uc[i] = b;
      } else {
        throw validation_error(std::string("invalid constraint sense '") + t + "'");
      }
    }
  }
}

namespace {

// The solve after the crossing-bounds check, on bounds lc/uc/lb/ub of lp.
SolveResult solve_checked(const LpHost& lp, const std::vector<double>& lc, const std::vector<double>& uc,
                          const std::vector<double>& lb, const std::vector<double>& ub, const PdlpSettings& st,
                          std::chrono::steady_clock::time_point t0);

// Presolve (presolve.hpp) around solve_checked: the reference's default
// pipeline presolves LPs before PDLP and postsolves the solution
// (solve.cu:1183-1325).
SolveResult solve_presolved(const LpHost& lp, const std::vector<double>& lc, const std::vector<double>& uc,
                            const std::vector<double>& lb, const std::vector<double>& ub, const PdlpSettings& st,
                            std::chrono::steady_clock::time_point t0)
{
  const Presolved P = presolve(lp, lc, uc, lb, ub);
  SolveResult R;
  auto seconds = [&]() { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); };
  if (P.status != Presolved::REDUCED) {
    R.status = P.status == Presolved::INFEASIBLE ? CUOPT_TERIMINATION_STATUS_INFEASIBLE
                                                 : CUOPT_TERIMINATION_STATUS_UNBOUNDED;
    R.x.assign(lp.n, 0.0);
    R.y.assign(lp.m, 0.0);
    R.rc.assign(lp.n, 0.0);
    R.solve_time = seconds();
    return R;
  }
  if (P.reduced.m == 0) {  // presolve solved the problem
    R.status = CUOPT_TERIMINATION_STATUS_OPTIMAL;
    postsolve(P, lp, R);
    if (!st.dual_postsolve) {
      R.y.clear();
      R.rc.clear();
    }
    R.dual_objective = R.primal_objective;
    R.info[CUOPT_AMD_INFO_STEPS] = 0;
    // Termination information on the original problem: the primal residual of x
    // (convergence_information.cu's r_p and its relative form with the combined
    // bound norm), a zero dual residual (z = c - A^T y exactly) and a zero gap.
    double rp2 = 0.0, bn2 = 0.0;
    for (int i = 0; i < lp.m; ++i) {
      double ax = 0.0;
      for (int k = lp.rowptr[i]; k < lp.rowptr[i + 1]; ++k) ax += lp.vals[k] * R.x[lp.colidx[k]];
      const double d = ax - std::min(std::max(ax, lc[i]), uc[i]);
      rp2 += d * d;
      if (std::isfinite(lc[i]) && lc[i] != uc[i]) bn2 += lc[i] * lc[i];
      if (std::isfinite(uc[i])) bn2 += uc[i] * uc[i];
    }
    R.info[CUOPT_AMD_INFO_L2_PRIMAL]  = std::sqrt(rp2);
    R.info[CUOPT_AMD_INFO_REL_PRIMAL] = std::sqrt(rp2) / (1.0 + std::sqrt(bn2));
    R.info[CUOPT_AMD_INFO_L2_DUAL]    = 0.0;
    R.info[CUOPT_AMD_INFO_REL_DUAL]   = 0.0;
    R.info[CUOPT_AMD_INFO_PRIMAL_OBJ] = R.primal_objective;
    R.info[CUOPT_AMD_INFO_DUAL_OBJ]   = R.dual_objective;
    R.info[CUOPT_AMD_INFO_GAP]        = 0.0;
    R.info[CUOPT_AMD_INFO_REL_GAP]    = 0.0;
    R.solve_time = seconds();
    R.info[CUOPT_AMD_INFO_SOLVE_TIME] = R.solve_time;
    return R;
  }
  PdlpSettings rs = st;
  if (!st.initial_primal.empty()) {
    rs.initial_primal.clear();
    for (int j : P.col_map) rs.initial_primal.push_back(j < (int)st.initial_primal.size() ? st.initial_primal[j] : 0.0);
  }
  if (!st.initial_dual.empty()) {
    rs.initial_dual.clear();
    for (int i : P.row_map) rs.initial_dual.push_back(i < (int)st.initial_dual.size() ? st.initial_dual[i] : 0.0);
  }
  R = solve_checked(P.reduced, P.lc, P.uc, P.lb, P.ub, rs, t0);
  if (R.error == CUOPT_SUCCESS && R.x.size() == (size_t)P.reduced.n && R.y.size() == (size_t)P.reduced.m) {
    postsolve(P, lp, R);
    if (!st.dual_postsolve) {  // no dual solution after presolve (pdlp_test.cu:903-930)
      R.y.clear();
      R.rc.clear();
    }
  } else {
    R.x.assign(lp.n, 0.0);
    R.y.assign(lp.m, 0.0);
    R.rc.assign(lp.n, 0.0);
  }
  R.solve_time = seconds();
  R.info[CUOPT_AMD_INFO_SOLVE_TIME] = R.solve_time;
  return R;
}

// solve_lp without the file parameters.
SolveResult solve_lp_core(const LpHost& lp, const PdlpSettings& st)
{
  const auto t0 = std::chrono::steady_clock::now();
  SolveResult R;
  const int m = lp.m, n = lp.n;
  std::vector<double> lc, uc;
  constraint_bounds(lp, lc, uc);
  std::vector<double> lb(lp.lb), ub(lp.ub);
  // has_crossing_bounds (problem_checking.cu:316-339)
  bool crossing = false;
  for (int j = 0; j < n && !crossing; ++j) crossing = !(ub[j] >= lb[j]);
  if (lp.ranged)
    for (int i = 0; i < m && !crossing; ++i) crossing = !(uc[i] >= lc[i]);
  auto seconds = [&]() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  };
  if (crossing) {
    R.status     = CUOPT_TERIMINATION_STATUS_INFEASIBLE;
    R.solve_time = seconds();
    R.x.assign(n, 0.0);
    R.y.assign(m, 0.0);
    R.rc.assign(n, 0.0);
    return R;
  }
  // presolve: Default, Papilo and PSLP all select this engine's presolve; Off skips it
  if (st.presolve != CUOPT_PRESOLVE_OFF && st.bench_steps <= 0) return solve_presolved(lp, lc, uc, lb, ub, st, t0);
  return solve_checked(lp, lc, uc, lb, ub, st, t0);
}

}  // namespace

SolveResult solve_lp(const LpHost& lp, const PdlpSettings& st)
{
  // The user problem is written before the solve (solve.cu:1281-1284); a file
  // that cannot be written is a validation error, as in the reference's writer
  // (mps_writer.cpp:32).
  if (!st.user_problem_file.empty()) {
    try {
      write_mps(lp, st.user_problem_file);
    } catch (const mps_file_error& e) {
      throw validation_error(e.what());
    }
  }
  SolveResult R = solve_lp_core(lp, st);
  if (!st.solution_file.empty()) write_sol_file(st.solution_file, lp, R);  // solve.cu:1327-1330
  return R;
}

namespace {

SolveResult solve_checked(const LpHost& lp, const std::vector<double>& lc, const std::vector<double>& uc,
                          const std::vector<double>& lb, const std::vector<double>& ub, const PdlpSettings& st,
                          std::chrono::steady_clock::time_point t0)
{
  SolveResult R;
  const int m = lp.m, n = lp.n;
  auto seconds = [&]() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  };
  if (m == 0) {
    R.solve_time = seconds();
    R.x.assign(n, 0.0);
    R.rc.assign(n, 0.0);
    if (st.method == CUOPT_METHOD_PDLP) {  // run_pdlp_solver (solve.cu:555-572)
      R.status = CUOPT_TERIMINATION_STATUS_NUMERICAL_ERROR;
      return R;
    }
    // No constraints: each variable sits at the bound its cost prefers, so the
    // LP is solved in closed form (the reference's concurrent method hands this
    // case to dual simplex, solve.cu:936-1103).
    double obj = lp.offset;
    for (int j = 0; j < n; ++j) {
      const double cj = lp.maximize ? -lp.c[j] : lp.c[j];
      double v        = 0.0;
      if (cj > 0.0) v = lb[j];
      else if (cj < 0.0) v = ub[j];
      else v = std::min(std::max(0.0, lb[j]), ub[j]);
      if (!std::isfinite(v)) {
        R.status = CUOPT_TERIMINATION_STATUS_UNBOUNDED;
        return R;
      }
      R.x[j] = v;
      R.rc[j] = lp.c[j];
      obj += lp.c[j] * v;
    }
    R.status           = CUOPT_TERIMINATION_STATUS_OPTIMAL;
    R.primal_objective = obj;
    R.dual_objective   = obj;
    return R;
  }
  if (st.precision == CUOPT_AMD_PRECISION_FP32_MIXED) {
    PdlpSolver<float> S(lp, lc, uc, lb, ub, st);
    return S.run();
  }
  PdlpSolver<double> S(lp, lc, uc, lb, ub, st);
  return S.run();
}

}  // namespace

// ---------------------------------------------------------------------------
// Test entry point: y = A x
// ---------------------------------------------------------------------------

template <typename V>
static void spmv_impl(int m, int n, const int* rowptr, const int* colidx, const double* vals, const double* x,
                      double* y)
{
  hipStream_t s;
  CUOPT_HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  try {
    MatStruct M;
    M.nrows = m;
    M.ncols = n;
    M.nnz   = rowptr[m];
    M.h_ptr.assign(rowptr, rowptr + m + 1);
    M.ptr.resize(m + 1);
    M.ptr.upload(M.h_ptr, s);
    M.idx.resize(M.nnz + kArrayPad);
    M.idx.zero(s);
    M.idx.upload(colidx, M.nnz, s);
    build_plan(M, s, ColSource{colidx, nullptr});  // XCD-aware order (sums are order independent)
    std::vector<V> hv(vals, vals + M.nnz), hx(x, x + n);
    DevBuf<V> dv(M.nnz + kArrayPad), dx(std::max(n, 1));
    dv.zero(s);
    DevBuf<double> dy(std::max(m, 1));
    dv.upload(hv.data(), M.nnz, s);
    dx.upload(hx.data(), n, s);
    launch_tiled(M, dv.get(), nullptr, PlainEpi<V>{dx.get(), dy.get()}, s);
    dy.download(y, m, s);
    CUOPT_HIP_CHECK(hipStreamSynchronize(s));
  } catch (...) {
    (void)hipStreamDestroy(s);
    throw;
  }
  CUOPT_HIP_CHECK(hipStreamDestroy(s));
}

void tiled_spmv_host(int m, int n, const int* rowptr, const int* colidx, const double* vals, const double* x,
                     double* y, int precision)
{
  if (precision == CUOPT_AMD_PRECISION_FP32_MIXED)
    spmv_impl<float>(m, n, rowptr, colidx, vals, x, y);
  else
    spmv_impl<double>(m, n, rowptr, colidx, vals, x, y);
}

// ---- the SpMV entry point that allocates no device memory -----------------------
//
// Everything the tiled SpMV keeps on the device besides nothing of the caller's
// (the staged CSR arrays, x, y, the tile plan and the long-row partials) lives in
// this fixed array, one copy per GPU, held until the process ends.  Host code
// reaches it only through the address hipGetSymbolAddress returns for the current
// device; the array's name itself is not a device address in host code.
static __device__ __align__(256) unsigned char scratch[kSpmvScratchBytes];

namespace {

std::mutex g_scratch_mu;
std::vector<unsigned char*> g_scratch_ptr;  // per device

unsigned char* spmv_scratch_for_current_device()
{
  int dev = 0;
  CUOPT_HIP_CHECK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if ((int)g_scratch_ptr.size() <= dev) g_scratch_ptr.resize(dev + 1, nullptr);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    CUOPT_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = static_cast<unsigned char*>(p);
  }
  return g_scratch_ptr[dev];
}

template <typename V>
void spmv_noalloc_impl(int m, int n, const int* rowptr, const int* colidx, const double* vals, const double* x,
                       double* y)
{
  const int nnz  = rowptr[m];
  const int tile = 1024;
  TilePlan p = plan_tiles(rowptr, m, tile - kVecPad, kMaxRowsPerTile, kBlock, colidx, n);
  // Layout of the staged data, each region 256-byte aligned.
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t o = off;
    off += (bytes + 255) & ~(size_t)255;
    return o;
  };
  const size_t ntiles = p.tiles.size(), nlong = p.longrows.size();
  const size_t o_ptr = take(sizeof(int) * ((size_t)m + 1)), o_idx = take(sizeof(int) * ((size_t)nnz + kArrayPad));
  const size_t o_val = take(sizeof(V) * ((size_t)nnz + kArrayPad)), o_x = take(sizeof(V) * std::max(n, 1));
  const size_t o_y = take(sizeof(double) * std::max(m, 1)), o_tiles = take(sizeof(Tile) * std::max<size_t>(ntiles, 1));
  const size_t o_meta = take(sizeof(int) * std::max<size_t>(ntiles, 1));
  const size_t o_long = take(sizeof(LongRow) * std::max<size_t>(nlong, 1));
  const size_t o_part = take(sizeof(uint64_t) * std::max<size_t>(ntiles, 1));
  if (off > kSpmvScratchBytes)
    throw validation_error("cuOptAmdSpmvNoAlloc: the problem needs " + std::to_string(off) +
                           " bytes of device scratch, more than its fixed " + std::to_string(kSpmvScratchBytes));
  unsigned char* base = spmv_scratch_for_current_device();
  int* d_ptr = reinterpret_cast<int*>(base + o_ptr);
  int* d_idx = reinterpret_cast<int*>(base + o_idx);
  V* d_val   = reinterpret_cast<V*>(base + o_val);
  V* d_x     = reinterpret_cast<V*>(base + o_x);
  double* d_y       = reinterpret_cast<double*>(base + o_y);
  Tile* d_tiles     = reinterpret_cast<Tile*>(base + o_tiles);
  int* d_meta       = reinterpret_cast<int*>(base + o_meta);
  LongRow* d_long   = reinterpret_cast<LongRow*>(base + o_long);
  uint64_t* d_part  = reinterpret_cast<uint64_t*>(base + o_part);
  // The per-thread default stream: nothing is created.  The array keeps the last
  // call's contents, so the padding the tiles read past the end is cleared here.
  hipStream_t s = hipStreamPerThread;
  std::vector<V> hv(vals, vals + nnz), hx(x, x + n);
  CUOPT_HIP_CHECK(hipMemcpyAsync(d_ptr, rowptr, sizeof(int) * ((size_t)m + 1), hipMemcpyHostToDevice, s));
  CUOPT_HIP_CHECK(hipMemsetAsync(d_idx + nnz, 0, sizeof(int) * kArrayPad, s));
  CUOPT_HIP_CHECK(hipMemsetAsync(d_val + nnz, 0, sizeof(V) * kArrayPad, s));
  if (nnz) {
    CUOPT_HIP_CHECK(hipMemcpyAsync(d_idx, colidx, sizeof(int) * (size_t)nnz, hipMemcpyHostToDevice, s));
    CUOPT_HIP_CHECK(hipMemcpyAsync(d_val, hv.data(), sizeof(V) * (size_t)nnz, hipMemcpyHostToDevice, s));
  }
  if (n) CUOPT_HIP_CHECK(hipMemcpyAsync(d_x, hx.data(), sizeof(V) * (size_t)n, hipMemcpyHostToDevice, s));
  if (ntiles) {
    CUOPT_HIP_CHECK(hipMemcpyAsync(d_tiles, p.tiles.data(), sizeof(Tile) * ntiles, hipMemcpyHostToDevice, s));
    CUOPT_HIP_CHECK(hipMemcpyAsync(d_meta, p.meta.data(), sizeof(int) * ntiles, hipMemcpyHostToDevice, s));
  }
  if (nlong)
    CUOPT_HIP_CHECK(hipMemcpyAsync(d_long, p.longrows.data(), sizeof(LongRow) * nlong, hipMemcpyHostToDevice, s));
  launch_tiled_raw(tile, d_ptr, d_idx, (const V*)d_val, PlanDev{d_tiles, d_meta, d_long, (int)ntiles}, (int)nlong,
                   d_part, nullptr, PlainEpi<V>{d_x, d_y}, s);
  if (m) CUOPT_HIP_CHECK(hipMemcpyAsync(y, d_y, sizeof(double) * (size_t)m, hipMemcpyDeviceToHost, s));
  CUOPT_HIP_CHECK(hipStreamSynchronize(s));
}

}  // namespace

void tiled_spmv_host_noalloc(int m, int n, const int* rowptr, const int* colidx, const double* vals, const double* x,
                             double* y, int precision)
{
  if (precision == CUOPT_AMD_PRECISION_FP32_MIXED)
    spmv_noalloc_impl<float>(m, n, rowptr, colidx, vals, x, y);
  else
    spmv_noalloc_impl<double>(m, n, rowptr, colidx, vals, x, y);
}

int device_count()
{
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess) return 0;
  return c;
}

}  // namespace cuopt_amd
