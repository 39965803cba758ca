// unrolled.hip -- differentiable unrolled PDHG over a batch of objectives (gfx950).
//
// Forward: K steps of the PDHG iteration of the LP path (projections of
// pdhg.cu:431-545, reflection x-bar = 2 x+ - x) for B right-hand sides at once,
// with a ridge term lam/2 ||x||^2 whose proximal step divides by 1 + tau lam:
//
//   x+ = clamp((x - tau (c - A^T y)) / (1 + tau lam), l, u)
//   y+ = sigma (t - clamp(t, -u_c, -l_c)),   t = y / sigma - A (2 x+ - x)
//
// tau = sigma = eta = 0.998 / sigma_max(A) (the power iteration of
// compute_initial_step_size, pdlp.cu:2622-2738, on the unscaled matrix), no
// restarts: x_K is a piecewise-linear function of (c, l_c, u_c).
//
// Backward: the vector-Jacobian product runs the steps in reverse with the
// clamp states of x+ and y+ (transposed SpMV plus projection masks).  States
// are checkpointed every `every` steps during the forward pass; each segment is
// recomputed from its checkpoint before its reverse sweep.  The c and (l_c, u_c)
// gradients are accumulated in FP64 (in FP32 runs per segment in FP32, then
// added into the FP64 totals).
//
// Layout: batch-interleaved vectors, v[i * B + b].  SpMM kernels give a
// wavefront 1, 2 or 4 consecutive matrix rows and one lane to each batch member,
// so every nonzero gathers B contiguous values of the operand (one 256-byte FP32
// line for B = 64) while the rows' value / index stream is read once for the
// whole batch.
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <type_traits>
#include <vector>

#include "batch_size.hpp"
#include "common.hpp"
#include "pdlp_kernels.hpp"
#include "unrolled.hpp"

namespace cuopt_amd {

namespace {

constexpr int kUB = 256;              // 4 wavefronts per workgroup
constexpr int kWavesPerWg = kUB / kWave;

// Rows per wavefront of the SpMM kernels: each wavefront takes R consecutive
// rows, so that one load of their indices and values across its lanes, and up
// to 16 operand gathers, are in flight at once.  Measured at 20 000 x 50 000 with
// 400 k nonzeros and B = 64 (forward / backward pass of 2 000 steps): R = 1 for
// both matrices 78.5 / 188.0 ms, 2 for both 75.2 / 180.3 ms, 4 for both 79.2 /
// 185.7 ms.  The gathers, not the index loads, bound these kernels: a 256-byte
// operand row per nonzero from a vector larger than one XCD's L2.  At least
// 4 096 wavefronts are kept to fill the chip.  CUOPT_AMD_UNR_RPW_A / _T (read
// when a handle is created) force R for A / A^T.
inline int rows_per_wave(int rows, long long nnz)
{
  if (rows <= 0) return 1;
  const double avg = (double)nnz / rows;
  int R = avg <= 4.0 ? 4 : avg <= 24.0 ? 2 : 1;
  while (R > 1 && rows / R < 4096) R /= 2;
  return R;
}

inline dim3 rows_grid(int rows, int B, int R)
{
  const int per_wg = kWavesPerWg * R;
  return dim3((std::max(rows, 1) + per_wg - 1) / per_wg, (B + kWave - 1) / kWave);
}

// Calls f(std::integral_constant<int, R>) for the supported rows per wavefront.
template <class F>
void with_rows(int R, F&& f)
{
  switch (R) {
    case 4: f(std::integral_constant<int, 4>{}); break;
    case 2: f(std::integral_constant<int, 2>{}); break;
    default: f(std::integral_constant<int, 1>{}); break;
  }
}

// First row of this wavefront (wave-uniform), its row count and this lane's
// batch member.  Lanes past B keep b = 0 (an in-bounds address) and bok false:
// they take part in the wavefront's shared loads but store nothing.
template <int R>
__device__ inline bool rows_lane(int rows, int B, int& r0, int& nr, int& b, bool& bok)
{
  r0 = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * kWavesPerWg + threadIdx.x / kWave) * R);
  if (r0 >= rows) return false;
  nr        = min(R, rows - r0);
  const int bl = (int)(blockIdx.y * kWave + (threadIdx.x & (kWave - 1)));
  bok       = bl < B;
  b         = bok ? bl : 0;
  return true;
}

__device__ inline int lane_value(int v, int l) { return __builtin_amdgcn_readlane(v, l); }
__device__ inline float lane_value(float v, int l)
{
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ inline double lane_value(double v, int l)
{
  const unsigned long long u = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, l);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), l);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Dot products of rows r0 .. r0 + nr - 1 with the batch-interleaved operand
// (lane b reads vec[col * B + b]) into out[0 .. nr - 1].  The rows' nonzeros are
// one contiguous run [ptr[r0], ptr[r0 + nr]); the wavefront loads it 64 entries
// at a time across its lanes, broadcasts each index and value with readlane,
// and issues 16 (then 4, then 1) gathers before adding their products.  Each
// row's products are added in increasing k, as a row-at-a-time loop would.
template <typename V, int R>
__device__ inline void rows_dot(const int* __restrict__ ptr, const int* __restrict__ idx,
                                const V* __restrict__ val, const V* __restrict__ vec, int r0, int nr, int b, int B,
                                typename AccFor<V>::type (&out)[R])
{
  using Acc = typename AccFor<V>::type;
  constexpr int G = 16;
  const int lane = (int)(threadIdx.x & (kWave - 1));
  const int pl   = ptr[r0 + min(lane, nr)];  // lane i < nr + 1 holds ptr[r0 + i]
  const int pend = lane_value(pl, nr);
  int cur = 0, bound = lane_value(pl, 1);  // row being summed and its end
#pragma unroll
  for (int i = 0; i < R; ++i) out[i] = Acc::zero();
  Acc a = Acc::zero();
  // A select per row keeps out[] in registers (a store at a run-time index
  // would put it in scratch memory).
  auto close = [&]() {
#pragma unroll
    for (int i = 0; i < R; ++i) out[i] = i == cur ? a : out[i];
  };
  // Before the product at position k: close the rows that end at k (empty ones included).
  auto advance = [&](int k) {
    while (k == bound) {
      close();
      a = Acc::zero();
      ++cur;
      bound = lane_value(pl, min(cur + 1, nr));
    }
  };
  for (int base = lane_value(pl, 0); base < pend; base += kWave) {
    const int cnt = min(kWave, pend - base);
    int ci = 0;
    V vi   = V(0);
    if (lane < cnt) {
      ci = idx[base + lane];
      vi = val[base + lane];
    }
    int t = 0;
    for (; t + G <= cnt; t += G) {
      V x[G];
#pragma unroll
      for (int q = 0; q < G; ++q) x[q] = vec[(size_t)lane_value(ci, t + q) * B + b];
#pragma unroll
      for (int q = 0; q < G; ++q) {
        advance(base + t + q);
        a.add(Acc::prod(lane_value(vi, t + q), x[q]));
      }
    }
    for (; t + 4 <= cnt; t += 4) {
      V x[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) x[q] = vec[(size_t)lane_value(ci, t + q) * B + b];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        advance(base + t + q);
        a.add(Acc::prod(lane_value(vi, t + q), x[q]));
      }
    }
    for (; t < cnt; ++t) {
      const V x = vec[(size_t)lane_value(ci, t) * B + b];
      advance(base + t);
      a.add(Acc::prod(lane_value(vi, t), x));
    }
  }
  close();
}

// ---- forward ----------------------------------------------------------------

template <typename V, int R>
__global__ __launch_bounds__(kUB) void unr_primal(const int* __restrict__ tptr, const int* __restrict__ tidx,
                                                  const V* __restrict__ tval, int n, int B, const V* __restrict__ y,
                                                  const V* __restrict__ x, const V* __restrict__ c,
                                                  const V* __restrict__ lb, const V* __restrict__ ub, V tau,
                                                  V inv_prox, V* __restrict__ xn, V* __restrict__ xb)
{
  int j0, nr, b;
  bool bok;
  if (!rows_lane<R>(n, B, j0, nr, b, bok)) return;
  // The rows' own operands are loaded before their gathers, so their latency
  // overlaps the gathers'.
  V xr[R], cr[R], lo[R], hi[R];
#pragma unroll
  for (int i = 0; i < R; ++i) {
    const int j    = j0 + min(i, nr - 1);
    const size_t o = (size_t)j * B + b;
    xr[i] = x[o];
    cr[i] = c[o];
    lo[i] = lb[j];
    hi[i] = ub[j];
  }
  typename AccFor<V>::type aty[R];
  rows_dot<V, R>(tptr, tidx, tval, y, j0, nr, b, B, aty);
  if (!bok) return;
#pragma unroll
  for (int i = 0; i < R; ++i) {
    if (i < nr) {
      const size_t o = (size_t)(j0 + i) * B + b;
      const V v      = (xr[i] - tau * aty[i].rsub(cr[i])) * inv_prox;
      const V xnew   = mx(mn(v, hi[i]), lo[i]);
      xn[o]          = xnew;
      xb[o]          = V(2) * xnew - xr[i];
    }
  }
}

template <typename V, int R>
__global__ __launch_bounds__(kUB) void unr_dual(const int* __restrict__ aptr, const int* __restrict__ aidx,
                                                const V* __restrict__ aval, int m, int B, const V* __restrict__ xb,
                                                const V* __restrict__ y, const V* __restrict__ lc,
                                                const V* __restrict__ uc, int ldb, V sigma, V* __restrict__ yn)
{
  int i0, nr, b;
  bool bok;
  if (!rows_lane<R>(m, B, i0, nr, b, bok)) return;
  V yo[R], lo[R], hi[R];  // loaded before the gathers
#pragma unroll
  for (int k = 0; k < R; ++k) {
    const int i    = i0 + min(k, nr - 1);
    const size_t q = ldb ? (size_t)i * ldb + b : (size_t)i;  // per-batch bounds, or one per row
    yo[k] = y[(size_t)i * B + b];
    lo[k] = lc[q];
    hi[k] = uc[q];
  }
  typename AccFor<V>::type ax[R];
  rows_dot<V, R>(aptr, aidx, aval, xb, i0, nr, b, B, ax);
  if (!bok) return;
#pragma unroll
  for (int k = 0; k < R; ++k) {
    if (k < nr) {
      const V t    = ax[k].rsub(yo[k] / sigma);
      const V proj = mx(-hi[k], mn(t, -lo[k]));
      yn[(size_t)(i0 + k) * B + b] = (t - proj) * sigma;
    }
  }
}

// ---- backward ---------------------------------------------------------------

// gt = sigma (1 - inside) ay for the step that produced y+ (its clamp states),
// and the FP64 (l_c, u_c) gradient terms.
template <typename V>
__device__ inline V dual_adjoint(V ay, V yplus, V l, V u, V sigma, double& dl, double& du)
{
  const bool eq = l == u;
  if (yplus == V(0) && !eq) {
    dl = du = 0.0;
    return V(0);
  }
  const double g = (double)sigma * (double)ay;
  if (yplus >= V(0)) {  // also y+ == 0 on an equality row: either side, both are b
    dl = g;
    du = 0.0;
  } else {
    dl = 0.0;
    du = g;
  }
  return sigma * ay;
}

// Per-step accumulators of the c and (l_c, u_c) gradients: FP64 in FP64 runs; in
// FP32 runs an FP32 partial per checkpoint segment (at most `every` terms), which
// unr_flush adds into the FP64 totals at the segment's end.  The adjoint kernels
// read and write their accumulators every step, and in FP64 those were about
// half of their traffic.
template <typename V>
struct GradAcc {
  using type = double;
};
template <>
struct GradAcc<float> {
  using type = float;
};

// acc += part; part = 0 (a kernel, not a memset, so that segment graphs may hold it).
__global__ __launch_bounds__(kUB) void unr_flush(size_t n, float* __restrict__ part, double* __restrict__ acc)
{
  for (size_t i = blockIdx.x * (size_t)kUB + threadIdx.x; i < n; i += (size_t)gridDim.x * kUB) {
    acc[i] += (double)part[i];
    part[i] = 0.f;
  }
}

// Start of the reverse sweep: gt for the last step from the upstream y-gradient.
template <typename V>
__global__ __launch_bounds__(kUB) void unr_adj_start(int m, int B, const V* ay,  // may alias gt
                                                     const V* __restrict__ yplus, const V* __restrict__ lc,
                                                     const V* __restrict__ uc, int ldb, V sigma, V* gt,
                                                     typename GradAcc<V>::type* __restrict__ glc,
                                                     typename GradAcc<V>::type* __restrict__ guc)
{
  using GA = typename GradAcc<V>::type;
  const size_t tot = (size_t)m * B;
  for (size_t o = blockIdx.x * (size_t)kUB + threadIdx.x; o < tot; o += (size_t)gridDim.x * kUB) {
    const int i = (int)(o / B), b = (int)(o % B);
    const size_t q = ldb ? (size_t)i * ldb + b : (size_t)i;  // per-batch bounds, or one per row
    double dl, du;
    gt[o] = dual_adjoint(ay ? ay[o] : V(0), yplus[o], lc[q], uc[q], sigma, dl, du);
    glc[o] += (GA)dl;
    guc[o] += (GA)du;
  }
}

// Rows j of A^T: gxb = -A^T gt; gw = inner(x+) (ax + 2 gxb) / (1 + tau lam);
// ax <- gw - gxb (adjoint of x before the step); gc_acc += gw (FP64).
template <typename V, int R>
__global__ __launch_bounds__(kUB) void unr_adj_primal(const int* __restrict__ tptr, const int* __restrict__ tidx,
                                                      const V* __restrict__ tval, int n, int B,
                                                      const V* __restrict__ gt, const V* __restrict__ xplus,
                                                      const V* __restrict__ lb, const V* __restrict__ ub, V inv_prox,
                                                      V* __restrict__ ax, V* __restrict__ gw,
                                                      typename GradAcc<V>::type* __restrict__ gc_acc)
{
  using GA = typename GradAcc<V>::type;
  int j0, nr, b;
  bool bok;
  if (!rows_lane<R>(n, B, j0, nr, b, bok)) return;
  // Element operands first (gt, which the rows gather, is none of them).
  V axo[R];
  bool in[R];
  GA gc[R];
#pragma unroll
  for (int i = 0; i < R; ++i) {
    const int j    = j0 + min(i, nr - 1);
    const size_t o = (size_t)j * B + b;
    const V xp     = xplus[o];
    axo[i]         = ax[o];
    in[i]          = xp > lb[j] && xp < ub[j];
    gc[i]          = gc_acc[o];
  }
  typename AccFor<V>::type s[R];
  rows_dot<V, R>(tptr, tidx, tval, gt, j0, nr, b, B, s);
  if (!bok) return;
#pragma unroll
  for (int i = 0; i < R; ++i) {
    if (i < nr) {
      const size_t o = (size_t)(j0 + i) * B + b;
      const V gxb    = -s[i].value();
      const V w      = in[i] ? (axo[i] + V(2) * gxb) * inv_prox : V(0);
      gw[o]          = w;
      ax[o]          = w - gxb;
      gc_acc[o]      = gc[i] + (GA)w;
    }
  }
}

// Rows i of A: ay_prev = gt / sigma + tau A gw; then, if there is an earlier
// step, its gt from ay_prev and the clamp states of its y+ (= y before this step).
// gt is read and rewritten in place, element by element.
template <typename V, int R>
__global__ __launch_bounds__(kUB) void unr_adj_dual(const int* __restrict__ aptr, const int* __restrict__ aidx,
                                                    const V* __restrict__ aval, int m, int B,
                                                    const V* __restrict__ gw, V tau, V sigma,
                                                    const V* __restrict__ yprev, const V* __restrict__ lc,
                                                    const V* __restrict__ uc, int ldb, V* __restrict__ gt,
                                                    typename GradAcc<V>::type* __restrict__ glc,
                                                    typename GradAcc<V>::type* __restrict__ guc)
{
  using GA = typename GradAcc<V>::type;
  int i0, nr, b;
  bool bok;
  if (!rows_lane<R>(m, B, i0, nr, b, bok)) return;
  // Element operands first (gw, which the rows gather, is none of them).
  V gto[R], yp[R], lo[R], hi[R];
  GA glo[R], guo[R];
#pragma unroll
  for (int k = 0; k < R; ++k) {
    const int i    = i0 + min(k, nr - 1);
    const size_t o = (size_t)i * B + b;
    const size_t q = ldb ? (size_t)i * ldb + b : (size_t)i;  // per-batch bounds, or one per row
    gto[k] = gt[o];
    yp[k]  = yprev[o];
    lo[k]  = lc[q];
    hi[k]  = uc[q];
    glo[k] = glc[o];
    guo[k] = guc[o];
  }
  typename AccFor<V>::type s[R];
  rows_dot<V, R>(aptr, aidx, aval, gw, i0, nr, b, B, s);
  if (!bok) return;
#pragma unroll
  for (int k = 0; k < R; ++k) {
    if (k < nr) {
      const size_t o = (size_t)(i0 + k) * B + b;
      const V ay     = gto[k] / sigma + tau * s[k].value();  // adjoint of y before this step
      double dl, du;
      gt[o]  = dual_adjoint(ay, yp[k], lo[k], hi[k], sigma, dl, du);
      glc[o] = glo[k] + (GA)dl;
      guc[o] = guo[k] + (GA)du;
    }
  }
}

template <typename V>
__global__ __launch_bounds__(kUB) void unr_convert_in(size_t n, const double* __restrict__ in, V* __restrict__ out)
{
  for (size_t i = blockIdx.x * (size_t)kUB + threadIdx.x; i < n; i += (size_t)gridDim.x * kUB) out[i] = (V)in[i];
}
template <typename V>
__global__ __launch_bounds__(kUB) void unr_convert_out(size_t n, const V* __restrict__ in, double scale,
                                                       double* __restrict__ out)
{
  for (size_t i = blockIdx.x * (size_t)kUB + threadIdx.x; i < n; i += (size_t)gridDim.x * kUB)
    out[i] = scale * (double)in[i];
}
template <typename V>
__global__ __launch_bounds__(kUB) void unr_init_x(int n, int B, const V* __restrict__ lb, const V* __restrict__ ub,
                                                  V* __restrict__ x)
{
  const size_t tot = (size_t)n * B;
  for (size_t o = blockIdx.x * (size_t)kUB + threadIdx.x; o < tot; o += (size_t)gridDim.x * kUB) {
    const int j = (int)(o / B);
    x[o]        = mn(mx(V(0), lb[j]), ub[j]);
  }
}
__global__ __launch_bounds__(kUB) void unr_zero64(size_t n, double* __restrict__ v)
{
  for (size_t i = blockIdx.x * (size_t)kUB + threadIdx.x; i < n; i += (size_t)gridDim.x * kUB) v[i] = 0.0;
}
template <typename V>
__global__ __launch_bounds__(kUB) void unr_zero(size_t n, V* __restrict__ v)
{
  for (size_t i = blockIdx.x * (size_t)kUB + threadIdx.x; i < n; i += (size_t)gridDim.x * kUB) v[i] = V(0);
}
// d = c / ||z|| style power step helpers (B = 1): z <- A A^T z, then norms on the host.
__global__ __launch_bounds__(kUB) void unr_sumsq(int n, const double* __restrict__ v, double* __restrict__ out)
{
  __shared__ double lds[kUB];
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += kUB) s += v[i] * v[i];
  lds[threadIdx.x] = s;
  __syncthreads();
  for (int off = kUB / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) lds[threadIdx.x] += lds[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = lds[0];
}
__global__ __launch_bounds__(kUB) void unr_spmv64(const int* __restrict__ ptr, const int* __restrict__ idx,
                                                  const double* __restrict__ val, int rows, const double* __restrict__ x,
                                                  double scale, double* __restrict__ y)
{
  const int r = blockIdx.x * kUB + threadIdx.x;
  if (r >= rows) return;
  double s = 0.0;
  for (int k = ptr[r]; k < ptr[r + 1]; ++k) s += val[k] * (x[idx[k]] * scale);
  y[r] = s;
}

// Y = A X, batch-interleaved (the SpMM timed by the batch-size search).
template <typename V, int R>
__global__ __launch_bounds__(kUB) void unr_spmm(const int* __restrict__ ptr, const int* __restrict__ idx,
                                                const V* __restrict__ val, int rows, int B, const V* __restrict__ x,
                                                V* __restrict__ y)
{
  int r0, nr, b;
  bool bok;
  if (!rows_lane<R>(rows, B, r0, nr, b, bok)) return;
  typename AccFor<V>::type s[R];
  rows_dot<V, R>(ptr, idx, val, x, r0, nr, b, B, s);
  if (!bok) return;
#pragma unroll
  for (int i = 0; i < R; ++i) {
    if (i < nr) {
      y[(size_t)(r0 + i) * B + b] = s[i].value();
    }
  }
}

inline int ew_blocks(size_t n) { return (int)std::max<size_t>(1, std::min<size_t>((n + kUB - 1) / kUB, 4096)); }

// Checkpoint slots addressed by a device-side segment counter, so that one
// captured segment graph serves every segment: store copies src into slot
// *seg of base, load copies slot *seg of base into dst.
template <typename V>
__global__ __launch_bounds__(kUB) void unr_ckpt_store(size_t len, const V* __restrict__ src, V* __restrict__ base,
                                                      const int* __restrict__ seg)
{
  V* dst = base + (size_t)(*seg) * len;
  for (size_t i = blockIdx.x * (size_t)kUB + threadIdx.x; i < len; i += (size_t)gridDim.x * kUB) dst[i] = src[i];
}
template <typename V>
__global__ __launch_bounds__(kUB) void unr_ckpt_load(size_t len, const V* __restrict__ base,
                                                     const int* __restrict__ seg, V* __restrict__ dst)
{
  const V* src = base + (size_t)(*seg) * len;
  for (size_t i = blockIdx.x * (size_t)kUB + threadIdx.x; i < len; i += (size_t)gridDim.x * kUB) dst[i] = src[i];
}
// seg <- set ? value : seg + value (one lane; a vector store)
__global__ void unr_seg_update(int* seg, int value, int set)
{
  if (threadIdx.x == 0 && blockIdx.x == 0) *seg = set ? value : *seg + value;
}

struct UnrolledImpl {
  int m = 0, n = 0, precision = 1;
  int rpw_a = 1, rpw_t = 1;  // rows per wavefront of the SpMM kernels over A and A^T
  double lam = 0.0, eta = 1.0;
  // structure (device)
  DevBuf<int> aptr, aidx, tptr, tidx;
  DevBuf<double> aval64, tval64;
  DevBuf<float> aval32, tval32;
  DevBuf<double> lb64, ub64;
  DevBuf<float> lb32, ub32;
  // last forward
  int B = 0, K = 0, every = 1, ldb = 0;
  DevBuf<unsigned char> work;  // typed views below
  size_t ckpt_off = 0, seg_off = 0;
  DevBuf<int> seg;  // checkpoint segment counter read by the segment graphs
  // Captured segment graphs: a full forward segment starting from either ping-pong
  // buffer, and a full backward segment (checkpoint load, recompute, reverse sweep)
  // with an earlier segment before it.  Valid while graph_key matches.
  hipGraphExec_t fwd_graph[2] = {nullptr, nullptr};
  hipGraphExec_t bwd_graph    = nullptr;
  std::vector<long long> graph_key;
  hipStream_t cap = nullptr;  // private capture stream (default priority)
  void drop_graphs()
  {
    for (hipGraphExec_t* g : {&fwd_graph[0], &fwd_graph[1], &bwd_graph}) {
      if (*g) (void)hipGraphExecDestroy(*g);
      *g = nullptr;
    }
    graph_key.clear();
  }
  ~UnrolledImpl()
  {
    drop_graphs();
    if (cap) (void)hipStreamDestroy(cap);
  }
};

// Graph replay of full segments unless CUOPT_AMD_GRAPHS=0 (direct launches of
// the same kernels, for tests).
inline bool unrolled_graphs_enabled()
{
  const char* g = std::getenv("CUOPT_AMD_GRAPHS");
  return !(g && std::atoi(g) == 0);
}

// Records what f launches on the handle's capture stream and instantiates it.
template <class F>
hipGraphExec_t capture_graph(UnrolledImpl& U, F f)
{
  if (!U.cap) CUOPT_HIP_CHECK(hipStreamCreateWithFlags(&U.cap, hipStreamNonBlocking));
  hipGraph_t g;
  CUOPT_HIP_CHECK(hipStreamBeginCapture(U.cap, hipStreamCaptureModeThreadLocal));
  try {
    f(U.cap);
  } catch (...) {
    (void)hipStreamEndCapture(U.cap, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  CUOPT_HIP_CHECK(hipStreamEndCapture(U.cap, &g));
  hipGraphExec_t e = nullptr;
  const hipError_t r = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  CUOPT_HIP_CHECK(r);
  return e;
}

template <typename V>
struct Views {
  const int *aptr, *aidx, *tptr, *tidx;
  const V *aval, *tval, *lb, *ub;
};

template <typename V>
Views<V> views(UnrolledImpl& U);
template <>
Views<double> views<double>(UnrolledImpl& U)
{
  return {U.aptr.get(), U.aidx.get(), U.tptr.get(), U.tidx.get(), U.aval64.get(), U.tval64.get(), U.lb64.get(),
          U.ub64.get()};
}
template <>
Views<float> views<float>(UnrolledImpl& U)
{
  return {U.aptr.get(), U.aidx.get(), U.tptr.get(), U.tidx.get(), U.aval32.get(), U.tval32.get(), U.lb32.get(),
          U.ub32.get()};
}

// Working storage of one forward/backward pair, carved from U.work:
//   c, lc, uc (working precision), x[2], y[2], xb, checkpoints (x, y) per segment,
//   segment states x+/y+ for every step of one segment, adjoints.
template <typename V>
struct Work {
  V *c, *lc, *uc, *x[2], *y[2], *xb, *ck_x, *ck_y, *sx, *sy, *ax, *gt, *gw;
  double *gc, *glc, *guc;
  // What the adjoint kernels accumulate into: FP32 segment partials in FP32 runs,
  // gc / glc / guc themselves in FP64 runs.
  typename GradAcc<V>::type *pc, *plc, *puc;
};

template <typename V>
Work<V> carve(UnrolledImpl& U, int nseg, bool allocate)
{
  const size_t nB = (size_t)U.n * U.B, mB = (size_t)U.m * U.B, lcn = U.ldb ? mB : (size_t)U.m;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t o = off;
    off += (bytes + 255) & ~(size_t)255;
    return o;
  };
  const size_t o_c = take(nB * sizeof(V)), o_lc = take(lcn * sizeof(V)), o_uc = take(lcn * sizeof(V));
  const size_t o_x0 = take(nB * sizeof(V)), o_x1 = take(nB * sizeof(V)), o_y0 = take(mB * sizeof(V)),
               o_y1 = take(mB * sizeof(V)), o_xb = take(nB * sizeof(V));
  const size_t o_ckx = take((size_t)(nseg + 1) * nB * sizeof(V)), o_cky = take((size_t)(nseg + 1) * mB * sizeof(V));
  const size_t o_sx = take((size_t)(U.every + 1) * nB * sizeof(V)), o_sy = take((size_t)(U.every + 1) * mB * sizeof(V));
  const size_t o_ax = take(nB * sizeof(V)), o_gt = take(mB * sizeof(V)), o_gw = take(nB * sizeof(V));
  const size_t o_gc = take(nB * sizeof(double)), o_glc = take(mB * sizeof(double)), o_guc = take(mB * sizeof(double));
  constexpr bool kPart = std::is_same<V, float>::value;
  const size_t o_pc = take(kPart ? nB * sizeof(float) : 0), o_plc = take(kPart ? mB * sizeof(float) : 0),
               o_puc = take(kPart ? mB * sizeof(float) : 0);
  if (allocate && U.work.size() < off) U.work.resize(off);  // kept across calls: no free / malloc per forward
  unsigned char* p = U.work.get();
  auto at = [&](size_t o) { return reinterpret_cast<V*>(p + o); };
  auto atd = [&](size_t o) { return reinterpret_cast<double*>(p + o); };
  using GA  = typename GradAcc<V>::type;
  auto atg = [&](size_t o, size_t whole) { return reinterpret_cast<GA*>(p + (kPart ? o : whole)); };
  return Work<V>{at(o_c),   at(o_lc),  at(o_uc),  {at(o_x0), at(o_x1)},
                 {at(o_y0), at(o_y1)}, at(o_xb), at(o_ckx), at(o_cky),
                 at(o_sx),  at(o_sy),  at(o_ax),  at(o_gt),
                 at(o_gw),  atd(o_gc), atd(o_glc), atd(o_guc),
                 atg(o_pc, o_gc), atg(o_plc, o_glc), atg(o_puc, o_guc)};
}

// One PDHG step (x, y) -> (xn, yn) through the scratch xb.
template <typename V>
void step(const Views<V>& M, UnrolledImpl& U, const Work<V>& W, const V* x, const V* y, V* xn, V* yn, hipStream_t s)
{
  const V tau = (V)U.eta, sigma = (V)U.eta, inv_prox = (V)(1.0 / (1.0 + U.eta * U.lam));
  with_rows(U.rpw_t, [&](auto RC) {
    constexpr int R = decltype(RC)::value;
    hipLaunchKernelGGL((unr_primal<V, R>), rows_grid(U.n, U.B, R), dim3(kUB), 0, s, M.tptr, M.tidx, M.tval, U.n,
                       U.B, y, x, (const V*)W.c, M.lb, M.ub, tau, inv_prox, xn, W.xb);
  });
  with_rows(U.rpw_a, [&](auto RC) {
    constexpr int R = decltype(RC)::value;
    hipLaunchKernelGGL((unr_dual<V, R>), rows_grid(U.m, U.B, R), dim3(kUB), 0, s, M.aptr, M.aidx, M.aval, U.m, U.B,
                       (const V*)W.xb, y, (const V*)W.lc, (const V*)W.uc, U.ldb, sigma, yn);
  });
}

template <typename V>
void forward_impl(UnrolledImpl& U, const double* c, const double* lc, const double* uc, double* x_out, double* y_out,
                  hipStream_t s)
{
  const int nseg = (U.K + U.every - 1) / U.every;
  Work<V> W      = carve<V>(U, nseg, true);
  Views<V> M     = views<V>(U);
  const size_t nB = (size_t)U.n * U.B, mB = (size_t)U.m * U.B, lcn = U.ldb ? mB : (size_t)U.m;
  if (!U.seg.get()) U.seg.resize(1);
  // The captured graphs hold the work buffer's addresses and the shapes.
  const std::vector<long long> key{(long long)(uintptr_t)U.work.get(), U.B, U.every, U.ldb, (long long)sizeof(V)};
  if (key != U.graph_key) {
    U.drop_graphs();
    U.graph_key = key;
  }
  hipLaunchKernelGGL(unr_convert_in<V>, dim3(ew_blocks(nB)), dim3(kUB), 0, s, nB, c, W.c);
  hipLaunchKernelGGL(unr_convert_in<V>, dim3(ew_blocks(lcn)), dim3(kUB), 0, s, lcn, lc, W.lc);
  hipLaunchKernelGGL(unr_convert_in<V>, dim3(ew_blocks(lcn)), dim3(kUB), 0, s, lcn, uc, W.uc);
  hipLaunchKernelGGL(unr_init_x<V>, dim3(ew_blocks(nB)), dim3(kUB), 0, s, U.n, U.B, M.lb, M.ub, W.x[0]);
  hipLaunchKernelGGL(unr_zero<V>, dim3(ew_blocks(mB)), dim3(kUB), 0, s, mB, W.y[0]);
  hipLaunchKernelGGL(unr_seg_update, dim3(1), dim3(kWave), 0, s, U.seg.get(), 0, 1);
  // One segment from buffer p: checkpoint (x, y) into slot *seg, advance the
  // counter, then L steps.  Full segments replay a captured copy of this.
  auto segment = [&](int p, int L, hipStream_t q) {
    hipLaunchKernelGGL(unr_ckpt_store<V>, dim3(ew_blocks(nB)), dim3(kUB), 0, q, nB, (const V*)W.x[p], W.ck_x,
                       (const int*)U.seg.get());
    hipLaunchKernelGGL(unr_ckpt_store<V>, dim3(ew_blocks(mB)), dim3(kUB), 0, q, mB, (const V*)W.y[p], W.ck_y,
                       (const int*)U.seg.get());
    hipLaunchKernelGGL(unr_seg_update, dim3(1), dim3(kWave), 0, q, U.seg.get(), 1, 0);
    int c = p;
    for (int t = 0; t < L; ++t, c ^= 1) step<V>(M, U, W, W.x[c], W.y[c], W.x[c ^ 1], W.y[c ^ 1], q);
  };
  const bool graphs = unrolled_graphs_enabled() && nseg > 1;
  int cur = 0;
  for (int sg = 0; sg < nseg; ++sg) {
    const int L = std::min(U.K, (sg + 1) * U.every) - sg * U.every;
    if (graphs && L == U.every) {
      if (!U.fwd_graph[cur]) U.fwd_graph[cur] = capture_graph(U, [&](hipStream_t q) { segment(cur, L, q); });
      CUOPT_HIP_CHECK(hipGraphLaunch(U.fwd_graph[cur], s));
    } else {
      segment(cur, L, s);
    }
    cur ^= (L & 1);
  }
  CUOPT_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(unr_convert_out<V>, dim3(ew_blocks(nB)), dim3(kUB), 0, s, nB, W.x[cur], 1.0, x_out);
  if (y_out) hipLaunchKernelGGL(unr_convert_out<V>, dim3(ew_blocks(mB)), dim3(kUB), 0, s, mB, W.y[cur], 1.0, y_out);
  // final state kept for the backward pass (segment states are recomputed)
  CUOPT_HIP_CHECK(hipGetLastError());
}

template <typename V>
void backward_impl(UnrolledImpl& U, const double* gx, const double* gy, double* gc, double* glc, double* guc,
                   hipStream_t s)
{
  const int nseg  = (U.K + U.every - 1) / U.every;
  Work<V> W       = carve<V>(U, nseg, false);
  Views<V> M      = views<V>(U);
  const size_t nB = (size_t)U.n * U.B, mB = (size_t)U.m * U.B;
  const V tau = (V)U.eta, sigma = (V)U.eta, inv_prox = (V)(1.0 / (1.0 + U.eta * U.lam));
  hipLaunchKernelGGL(unr_zero64, dim3(ew_blocks(nB)), dim3(kUB), 0, s, nB, W.gc);
  hipLaunchKernelGGL(unr_zero64, dim3(ew_blocks(mB)), dim3(kUB), 0, s, mB, W.glc);
  hipLaunchKernelGGL(unr_zero64, dim3(ew_blocks(mB)), dim3(kUB), 0, s, mB, W.guc);
  constexpr bool kPart = std::is_same<V, float>::value;
  if constexpr (kPart) {
    hipLaunchKernelGGL(unr_zero<float>, dim3(ew_blocks(nB)), dim3(kUB), 0, s, nB, W.pc);
    hipLaunchKernelGGL(unr_zero<float>, dim3(ew_blocks(mB)), dim3(kUB), 0, s, mB, W.plc);
    hipLaunchKernelGGL(unr_zero<float>, dim3(ew_blocks(mB)), dim3(kUB), 0, s, mB, W.puc);
  }
  hipLaunchKernelGGL(unr_convert_in<V>, dim3(ew_blocks(nB)), dim3(kUB), 0, s, nB, gx, W.ax);
  // y-gradient: start value of ay (zero when absent), staged in gw's slot of size m*B? use gt as ay first.
  if (gy) hipLaunchKernelGGL(unr_convert_in<V>, dim3(ew_blocks(mB)), dim3(kUB), 0, s, mB, gy, W.gt);
  // The segments run last to first; the counter names the checkpoint slot.
  hipLaunchKernelGGL(unr_seg_update, dim3(1), dim3(kWave), 0, s, U.seg.get(), nseg - 1, 1);
  // One segment of L steps: load its checkpoint, recompute states x_{k0..k0+L},
  // y_{k0..k0+L} into the segment store (slot t holds step k0 + t), run the
  // reverse sweep, step the counter back.  last: the segment holding the final
  // step (gt starts from the upstream y-gradient); first: segment 0, whose first
  // step starts from the constant y = 0 and propagates into no earlier step.
  auto segment = [&](int L, bool last, bool first, hipStream_t q) {
    hipLaunchKernelGGL(unr_ckpt_load<V>, dim3(ew_blocks(nB)), dim3(kUB), 0, q, nB, (const V*)W.ck_x,
                       (const int*)U.seg.get(), W.sx);
    hipLaunchKernelGGL(unr_ckpt_load<V>, dim3(ew_blocks(mB)), dim3(kUB), 0, q, mB, (const V*)W.ck_y,
                       (const int*)U.seg.get(), W.sy);
    for (int t = 0; t < L; ++t)
      step<V>(M, U, W, W.sx + (size_t)t * nB, W.sy + (size_t)t * mB, W.sx + (size_t)(t + 1) * nB,
              W.sy + (size_t)(t + 1) * mB, q);
    for (int t = L - 1; t >= 0; --t) {
      const V* xplus = W.sx + (size_t)(t + 1) * nB;
      const V* yplus = W.sy + (size_t)(t + 1) * mB;
      if (last && t == L - 1)
        hipLaunchKernelGGL(unr_adj_start<V>, dim3(ew_blocks(mB)), dim3(kUB), 0, q, U.m, U.B,
                           gy ? (const V*)W.gt : (const V*)nullptr, yplus, (const V*)W.lc, (const V*)W.uc, U.ldb,
                           sigma, W.gt, W.plc, W.puc);
      with_rows(U.rpw_t, [&](auto RC) {
        constexpr int R = decltype(RC)::value;
        hipLaunchKernelGGL((unr_adj_primal<V, R>), rows_grid(U.n, U.B, R), dim3(kUB), 0, q, M.tptr, M.tidx, M.tval,
                           U.n, U.B, (const V*)W.gt, xplus, M.lb, M.ub, inv_prox, W.ax, W.gw, W.pc);
      });
      if (!first || t > 0)
        with_rows(U.rpw_a, [&](auto RC) {
          constexpr int R = decltype(RC)::value;
          hipLaunchKernelGGL((unr_adj_dual<V, R>), rows_grid(U.m, U.B, R), dim3(kUB), 0, q, M.aptr, M.aidx, M.aval,
                             U.m, U.B, (const V*)W.gw, tau, sigma, (const V*)(W.sy + (size_t)t * mB),
                             (const V*)W.lc, (const V*)W.uc, U.ldb, W.gt, W.plc, W.puc);
        });
    }
    if constexpr (kPart) {  // the segment's FP32 partials into the FP64 totals
      hipLaunchKernelGGL(unr_flush, dim3(ew_blocks(nB)), dim3(kUB), 0, q, nB, W.pc, W.gc);
      hipLaunchKernelGGL(unr_flush, dim3(ew_blocks(mB)), dim3(kUB), 0, q, mB, W.plc, W.glc);
      hipLaunchKernelGGL(unr_flush, dim3(ew_blocks(mB)), dim3(kUB), 0, q, mB, W.puc, W.guc);
    }
    hipLaunchKernelGGL(unr_seg_update, dim3(1), dim3(kWave), 0, q, U.seg.get(), -1, 0);
  };
  const bool graphs = unrolled_graphs_enabled();
  for (int sg = nseg - 1; sg >= 0; --sg) {
    const int L = std::min(U.K, (sg + 1) * U.every) - sg * U.every;
    const bool last = sg == nseg - 1, first = sg == 0;
    if (graphs && !last && !first) {  // middle segments are always full
      if (!U.bwd_graph) U.bwd_graph = capture_graph(U, [&](hipStream_t q) { segment(L, false, false, q); });
      CUOPT_HIP_CHECK(hipGraphLaunch(U.bwd_graph, s));
    } else {
      segment(L, last, first, s);
    }
  }
  CUOPT_HIP_CHECK(hipGetLastError());
  // gc = -tau / (1 + tau lam) * sum of gw: the kernel accumulated gw, already divided by (1 + tau lam).
  hipLaunchKernelGGL(unr_convert_out<double>, dim3(ew_blocks(nB)), dim3(kUB), 0, s, nB, (const double*)W.gc,
                     -U.eta, gc);
  CUOPT_HIP_CHECK(hipMemcpyAsync(glc, W.glc, mB * sizeof(double), hipMemcpyDeviceToDevice, s));
  CUOPT_HIP_CHECK(hipMemcpyAsync(guc, W.guc, mB * sizeof(double), hipMemcpyDeviceToDevice, s));
  CUOPT_HIP_CHECK(hipGetLastError());
}

}  // namespace

// ---- C++ entry points used by the C ABI ---------------------------------------------

UnrolledHandle* unrolled_create(int m, int n, const int* rowptr, const int* colidx, const double* vals,
                                const double* lb, const double* ub, double lam, int precision)
{
  auto* U = new UnrolledImpl();
  try {
    U->m = m;
    U->n = n;
    U->lam = lam;
    U->precision = precision;
    const int nnz = rowptr[m];
    U->rpw_a      = rows_per_wave(m, nnz);
    U->rpw_t      = rows_per_wave(n, nnz);
    if (const char* e = std::getenv("CUOPT_AMD_UNR_RPW_A")) U->rpw_a = std::atoi(e);  // 1, 2 or 4
    if (const char* e = std::getenv("CUOPT_AMD_UNR_RPW_T")) U->rpw_t = std::atoi(e);
    // A^T by a stable counting sort
    std::vector<int> tp(n + 1, 0), ti(nnz);
    std::vector<double> tv(nnz);
    for (int k = 0; k < nnz; ++k) tp[colidx[k] + 1]++;
    for (int j = 0; j < n; ++j) tp[j + 1] += tp[j];
    {
      std::vector<int> cur(tp.begin(), tp.end() - 1);
      for (int i = 0; i < m; ++i)
        for (int k = rowptr[i]; k < rowptr[i + 1]; ++k) {
          const int d = cur[colidx[k]]++;
          ti[d]       = i;
          tv[d]       = vals[k];
        }
    }
    hipStream_t s = nullptr;
    U->aptr.resize(m + 1);
    U->aptr.upload(rowptr, m + 1, s);
    U->aidx.resize(std::max(nnz, 1));
    U->aidx.upload(colidx, nnz, s);
    U->tptr.resize(n + 1);
    U->tptr.upload(tp.data(), n + 1, s);
    U->tidx.resize(std::max(nnz, 1));
    U->tidx.upload(ti.data(), nnz, s);
    U->aval64.resize(std::max(nnz, 1));
    U->aval64.upload(vals, nnz, s);
    U->tval64.resize(std::max(nnz, 1));
    U->tval64.upload(tv.data(), nnz, s);
    U->lb64.resize(std::max(n, 1));
    U->lb64.upload(lb, n, s);
    U->ub64.resize(std::max(n, 1));
    U->ub64.upload(ub, n, s);
    std::vector<float> f, g, l, u;  // host sources of the async uploads, alive until the final sync
    if (precision == CUOPT_AMD_PRECISION_FP32_MIXED) {
      f.assign(vals, vals + nnz);
      g.assign(tv.begin(), tv.end());
      l.assign(lb, lb + n);
      u.assign(ub, ub + n);
      U->aval32.resize(std::max(nnz, 1));
      U->aval32.upload(f, s);
      U->tval32.resize(std::max(nnz, 1));
      U->tval32.upload(g, s);
      U->lb32.resize(std::max(n, 1));
      U->lb32.upload(l, s);
      U->ub32.resize(std::max(n, 1));
      U->ub32.upload(u, s);
    }
    // power iteration on A A^T (FP64): z <- A (A^T z / ||z||), sigma_max^2 = ||z_new|| at convergence
    DevBuf<double> z(std::max(m, 1)), w(std::max(n, 1)), nrm(1);
    std::vector<double> z0(m);
    for (int i = 0; i < m; ++i) z0[i] = 1.0 + 0.5 * std::sin(0.7 * i + 0.3);  // deterministic start
    z.upload(z0, s);
    double norm2 = 0.0, prev = 0.0, s2 = 0.0;
    hipLaunchKernelGGL(unr_sumsq, dim3(1), dim3(kUB), 0, s, m, z.get(), nrm.get());
    nrm.download(&norm2, 1, s);
    CUOPT_HIP_CHECK(hipStreamSynchronize(s));
    for (int it = 0; it < 2000 && m > 0 && n > 0; ++it) {
      const double inv = 1.0 / std::sqrt(norm2);
      hipLaunchKernelGGL(unr_spmv64, dim3((n + kUB - 1) / kUB), dim3(kUB), 0, s, U->tptr.get(), U->tidx.get(),
                         U->tval64.get(), n, z.get(), inv, w.get());
      hipLaunchKernelGGL(unr_spmv64, dim3((m + kUB - 1) / kUB), dim3(kUB), 0, s, U->aptr.get(), U->aidx.get(),
                         U->aval64.get(), m, w.get(), 1.0, z.get());
      hipLaunchKernelGGL(unr_sumsq, dim3(1), dim3(kUB), 0, s, m, z.get(), nrm.get());
      nrm.download(&norm2, 1, s);
      CUOPT_HIP_CHECK(hipStreamSynchronize(s));
      s2 = std::sqrt(norm2);  // ||A A^T q|| -> sigma_max^2
      if (it > 10 && std::fabs(s2 - prev) <= 1e-10 * s2) break;
      prev = s2;
    }
    U->eta = s2 > 0.0 ? 0.998 / std::sqrt(s2) : 1.0;
    CUOPT_HIP_CHECK(hipStreamSynchronize(s));
  } catch (...) {
    delete U;
    throw;
  }
  return reinterpret_cast<UnrolledHandle*>(U);
}

void unrolled_destroy(UnrolledHandle* h) { delete reinterpret_cast<UnrolledImpl*>(h); }

double unrolled_step_size(UnrolledHandle* h) { return reinterpret_cast<UnrolledImpl*>(h)->eta; }

void unrolled_forward(UnrolledHandle* h, int B, int K, int every, const double* c, const double* lc, const double* uc,
                      int bounds_per_batch, double* x, double* y, void* stream)
{
  auto* U = reinterpret_cast<UnrolledImpl*>(h);
  if (B <= 0 || K < 0 || every <= 0) throw validation_error("batch and checkpoint interval must be positive");
  U->B     = B;
  U->K     = K;
  U->every = every;
  U->ldb   = bounds_per_batch ? B : 0;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (U->precision == CUOPT_AMD_PRECISION_FP32_MIXED)
    forward_impl<float>(*U, c, lc, uc, x, y, s);
  else
    forward_impl<double>(*U, c, lc, uc, x, y, s);
}

void unrolled_backward(UnrolledHandle* h, const double* gx, const double* gy, double* gc, double* glc, double* guc,
                       void* stream)
{
  auto* U = reinterpret_cast<UnrolledImpl*>(h);
  if (U->B <= 0) throw validation_error("backward called before forward");
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (U->K == 0) {
    const size_t nB = (size_t)U->n * U->B, mB = (size_t)U->m * U->B;
    hipLaunchKernelGGL(unr_zero64, dim3(ew_blocks(nB)), dim3(kUB), 0, s, nB, gc);
    hipLaunchKernelGGL(unr_zero64, dim3(ew_blocks(mB)), dim3(kUB), 0, s, mB, glc);
    hipLaunchKernelGGL(unr_zero64, dim3(ew_blocks(mB)), dim3(kUB), 0, s, mB, guc);
    return;
  }
  if (U->precision == CUOPT_AMD_PRECISION_FP32_MIXED)
    backward_impl<float>(*U, gx, gy, gc, glc, guc, s);
  else
    backward_impl<double>(*U, gx, gy, gc, glc, guc, s);
}

template <typename V>
int optimal_batch_size_impl(UnrolledImpl& U, int max_batch)
{
  const Views<V> M = views<V>(U);
  hipStream_t s;
  CUOPT_HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  hipEvent_t e0 = nullptr, e1 = nullptr;
  DevBuf<V> x, y;
  auto cleanup = [&]() {
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(s);
  };
  try {
    CUOPT_HIP_CHECK(hipEventCreate(&e0));
    CUOPT_HIP_CHECK(hipEventCreate(&e1));
    auto cost = [&](int B) {
      const size_t nB = (size_t)std::max(U.n, 1) * B, mB = (size_t)std::max(U.m, 1) * B;
      if (x.size() < nB) x.resize(nB);
      if (y.size() < mB) y.resize(mB);
      hipLaunchKernelGGL(unr_zero<V>, dim3(ew_blocks(nB)), dim3(kUB), 0, s, nB, x.get());
      hipLaunchKernelGGL(unr_zero<V>, dim3(ew_blocks(mB)), dim3(kUB), 0, s, mB, y.get());
      auto run = [&]() {
        if (U.m > 0)
          with_rows(U.rpw_a, [&](auto RC) {
            constexpr int R = decltype(RC)::value;
            hipLaunchKernelGGL((unr_spmm<V, R>), rows_grid(U.m, B, R), dim3(kUB), 0, s, M.aptr, M.aidx, M.aval, U.m,
                               B, (const V*)x.get(), y.get());
          });
        if (U.n > 0)
          with_rows(U.rpw_t, [&](auto RC) {
            constexpr int R = decltype(RC)::value;
            hipLaunchKernelGGL((unr_spmm<V, R>), rows_grid(U.n, B, R), dim3(kUB), 0, s, M.tptr, M.tidx, M.tval, U.n,
                               B, (const V*)y.get(), x.get());
          });
      };
      run();
      CUOPT_HIP_CHECK(hipEventRecord(e0, s));
      constexpr int kRuns = 5;
      for (int i = 0; i < kRuns; ++i) run();
      CUOPT_HIP_CHECK(hipEventRecord(e1, s));
      CUOPT_HIP_CHECK(hipGetLastError());
      CUOPT_HIP_CHECK(hipEventSynchronize(e1));
      float ms = 0.f;
      CUOPT_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
      return (double)ms / kRuns / B;
    };
    const int best = search_batch_size(max_batch, cost);
    cleanup();
    return best;
  } catch (...) {
    cleanup();
    throw;
  }
}

int unrolled_optimal_batch_size(UnrolledHandle* h, int max_batch)
{
  auto& U = *reinterpret_cast<UnrolledImpl*>(h);
  if (U.precision == CUOPT_AMD_PRECISION_FP32_MIXED) return optimal_batch_size_impl<float>(U, max_batch);
  return optimal_batch_size_impl<double>(U, max_batch);
}

}  // namespace cuopt_amd
