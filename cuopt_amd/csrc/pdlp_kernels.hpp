// pdlp_kernels.hpp -- fused epilogues and elementwise kernels of the PDLP hot loop.
//
// V is the iterate precision (double, or float for the FP32-mixed path).  The
// step epilogues restate, per row, the reference's projection functors
// (cpp/src/pdlp/pdhg.cu:431-545) fused with the Halpern update
// (cpp/src/pdlp/pdlp.cu:2505-2563) and, on the steps that need them, the
// fixed-point-error terms (cpp/src/pdlp/pdlp.cu:1748-1855).  Comparisons are
// written as the std::min / std::max of the FP64 oracle so both pick the same
// operand.
#pragma once

#include <type_traits>

#include "spmv_tiles.hpp"

namespace cuopt_amd {

template <typename V>
struct Vec2;
template <>
struct Vec2<double> {
  using type = double2;
};
template <>
struct Vec2<float> {
  using type = float2;
};

// std::min / std::max as the oracle evaluates them.
template <typename T>
__device__ inline T mn(T a, T b)
{
  return (b < a) ? b : a;
}
template <typename T>
__device__ inline T mx(T a, T b)
{
  return (a < b) ? b : a;
}
__device__ inline double fin0(double v) { return isfinite(v) ? v : 0.0; }

// Deterministic Lanczos start vector in [-1, 1): an integer hash of the index.
__device__ inline double hashed_start(int i)
{
  uint32_t h = (uint32_t)i * 2654435761u + 0x9e3779b9u;
  h ^= h >> 16;
  h *= 0x85ebca6bu;
  h ^= h >> 13;
  h *= 0xc2b2ae35u;
  h ^= h >> 16;
  return (double)h * (2.0 / 4294967296.0) - 1.0;
}

template <typename V>
struct AccFor;
template <>
struct AccFor<double> {
  using type = AccSum64;
};
template <>
struct AccFor<float> {
  using type = AccComp32;
};

// Scalars read by the step kernels from device memory (set between blocks).
struct StepScalars {
  double tau;        // primal step size eta / omega
  double sigma;      // dual step size eta * omega
  long long kr_base; // iterations since the last restart at the block start
  long long pad;
};

// Primal weight range of FP32-mixed solves.  In the normalised units of the
// scaled problem (bound and objective rescaling put ||b||, ||c|| near one and
// start at omega = 1) a weight outside [1e-3, 1e3] makes one of the step sizes
// so small that its FP32 updates fall below the rounding of the iterate, which
// then freezes while the PID keeps moving the weight away.  PDHG converges for
// any fixed weight, so the bound only limits the adaptive heuristic.  FP64
// solves use the reference's unbounded weight.
constexpr double kOmegaMin32 = 1e-3, kOmegaMax32 = 1e3;
template <typename V>
__host__ __device__ inline double bound_primal_weight(double w)
{
  if constexpr (std::is_same<V, float>::value) return w < kOmegaMin32 ? kOmegaMin32 : (w > kOmegaMax32 ? kOmegaMax32 : w);
  return w;
}

template <typename V>
__device__ inline V halpern_weight(const StepScalars* sc, int j)
{
  const long long kr = sc->kr_base + j;
  return (V)((double)(kr + 1) / (double)(kr + 2));
}

// ---- primal step over rows of A~^T (columns j): AtY, projection, reflection, Halpern.
template <typename V, bool MAJOR, bool FPE>
struct PrimalStepEpi {
  using Mat = V;
  using Acc = typename AccFor<V>::type;
  using V2  = typename Vec2<V>::type;
  static constexpr int NRED = 0;
  const V* y;  // gathered
  const V* x;
  const V* c;
  const V2* bnd;
  const V* x0;
  V* xbar;
  V* xnext;
  V* xhat;
  V* slack;
  V* aty_out;
  const StepScalars* sc;
  int j;
  // Recentred (refined) solves: store only the part (x-hat - x) / tau of the slack; the
  // termination check adds the FP64 reduced cost of the corrected point.
  bool small_slack;
  __device__ bool skip() const { return false; }
  struct Pre {
    V x, c, x0;
    V2 b;
  };
  __device__ typename Acc::Elem product(V v, int col, int) const { return Acc::prod(v, y[col]); }
  __device__ Pre load(int r) const { return Pre{x[r], c[r], x0[r], bnd[r]}; }
  __device__ void row(int r, const Acc& aty, const Pre& p, double (&)[kMaxRed]) const
  {
    const V tau     = (V)sc->tau;
    const V w       = halpern_weight<V>(sc, j);
    const V next    = p.x - tau * aty.rsub(p.c);
    const V clamped = mx(mn(next, p.b.y), p.b.x);
    const V xb      = V(2) * clamped - p.x;
    xbar[r]         = xb;
    if constexpr (MAJOR) {
      xhat[r]  = clamped;
      slack[r] = small_slack ? (clamped - p.x) / tau : (clamped - next) / tau;
    }
    if constexpr (FPE) aty_out[r] = aty.value();
    xnext[r] = w * xb + (V(1) - w) * p.x0;
  }
};

// FP32 step kernels (measured at 500 k x 1 M, HIP events, one box, with four
// preloaded rows; the primal step now preloads two):
//  - up to 1 024 slots per tile the next tile's stream goes through LDS, built
//    for at least six waves per SIMD (80 VGPRs then, no spills): primal step
//    25.0 -> 24.5 us, dual step 35.7 -> 34.2 us.  The primal step with two
//    preloaded rows is built for eight (64 VGPRs, 12 bytes per lane spilled):
//    22.8-23.0 -> 22.1-22.4 us tuned, 23.6 -> 23.0 us untuned;
//  - 2 048-slot tiles would drop to four waves by LDS with it: there both steps
//    load each tile's stream at the top of the tile.  The primal step is built
//    for eight waves (64 VGPRs, no spills): on the banded 4 M x 8 M LP 187.9-190.0
//    -> 174.3-174.6 us against the next stream held in registers at six waves.
//    The dual step's best plan (2 048 slots with in-launch merge) measured 32.2 us
//    against 34.3 with the LDS stream at four waves; built for eight waves with two
//    preloaded rows it took the banded dual step 167.3-167.9 -> 151.5-151.8 us but
//    the 500 k x 1 M one 32.5-32.6 -> 33.2-33.4 us, so it keeps six waves.
template <bool MAJOR, bool FPE, int TILE>
struct TilePipelining<PrimalStepEpi<float, MAJOR, FPE>, TILE> {
  static constexpr bool prefetch   = false;
  static constexpr bool lds_stream = TILE <= 1024;
  static constexpr int min_waves   = 8;
  static constexpr int pre_rows    = kPreRows;
};

// FP64 primal step: at 2 048 slots it loads each tile's stream at the top of the
// tile, built for six waves (77 VGPRs, no spills), instead of holding the next
// stream in registers at four: banded 4 M x 8 M 295-297 -> 268-269 us.  The FP64
// dual step keeps the register stream: built the same way it took the banded dual
// step 261.5-262.5 -> 225.8 us but the 500 k x 1 M one (its tuned plan, 2 048
// slots with in-launch merge) 39.4-39.5 -> 40.2 us.
template <bool MAJOR, bool FPE, int TILE>
struct TilePipelining<PrimalStepEpi<double, MAJOR, FPE>, TILE> {
  static constexpr bool prefetch   = TILE <= 1024;
  static constexpr bool lds_stream = false;
  static constexpr int min_waves   = TILE <= 1024 ? 1 : 6;
  static constexpr int pre_rows    = kPreRows;
};

// ---- dual step over rows of A~ (rows i): A x-bar, projection, reflection, Halpern.
template <typename V, bool MAJOR, bool FPE>
struct DualStepEpi {
  using Mat = V;
  using Acc = typename AccFor<V>::type;
  using V2  = typename Vec2<V>::type;
  static constexpr int NRED = FPE ? 1 : 0;
  const V* xbar;  // gathered
  const V* y;
  const V2* cb;  // (lower, upper) of the scaled constraints
  const V* y0;
  V* ynext;
  V* yhat;
  V* ybar;
  const StepScalars* sc;
  int j;
  // Recentred (refined) solves only (else null): per-row position of the kink of the
  // dual objective term, which sits at y = 0 in the original problem and at
  // -Dd * y* / Dr in the coordinates of a correction problem.
  const V* kink;
  __device__ bool skip() const { return false; }
  struct Pre {
    V y, y0, k;
    V2 b;
  };
  __device__ typename Acc::Elem product(V v, int col, int) const { return Acc::prod(v, xbar[col]); }
  __device__ Pre load(int r) const { return Pre{y[r], y0[r], kink ? kink[r] : V(0), cb[r]}; }
  __device__ void row(int r, const Acc& ax, const Pre& p, double (&red)[kMaxRed]) const
  {
    const V sigma = (V)sc->sigma;
    const V w     = halpern_weight<V>(sc, j);
    V nd;
    if (kink) {
      // The projection of the original problem moved to the kink k, written as
      // clamp(k, s + sigma l, s + sigma u) with s = y - sigma A x-bar: equal to
      // k + sigma (t - clamp(t, -u, -l)) for t = (y - k) / sigma - A x-bar, but
      // without forming y - k, whose rounding would undo the refinement.
      const V s = sigma * ax.rsub(p.y / sigma);
      nd        = mx(mn(p.k, s + sigma * p.b.y), s + sigma * p.b.x);
    } else {
      const V t    = ax.rsub(p.y / sigma);
      const V proj = mx(-p.b.y, mn(t, -p.b.x));
      nd           = (t - proj) * sigma;
    }
    const V yb    = V(2) * nd - p.y;
    if constexpr (MAJOR) yhat[r] = nd;
    if constexpr (FPE) {
      ybar[r]        = yb;
      const double d = (double)yb - (double)p.y;
      red[0] += d * d;
    }
    ynext[r] = w * yb + (V(1) - w) * p.y0;
  }
};

// The FP32 dual step preloads four rows per lane group: with two (the default,
// which took the 500 k x 1 M primal steps 23.7-24.3 -> 22.8 us in FP32 and
// 32.7 -> 31.0 us in FP64) its tuned plan measured 32.5-32.7 against 32.2 us,
// and two rows plus the next stream in registers at 2 048 slots (80 VGPRs, six
// waves) 32.5-32.6 us.
template <bool MAJOR, bool FPE, int TILE>
struct TilePipelining<DualStepEpi<float, MAJOR, FPE>, TILE> {
  static constexpr bool prefetch   = false;
  static constexpr bool lds_stream = TILE <= 1024;
  static constexpr int min_waves   = TILE <= 1024 ? 6 : 1;
  static constexpr int pre_rows    = 4;
};

// ---- fixed-point error: A~^T y-bar against the stored A~^T y, over columns j.
// red[0] = (A~^T ybar - A~^T y) . (xbar - x), red[1] = ||xbar - x||^2.
template <typename V>
struct FpeEpi {
  using Mat = V;
  using Acc = typename AccFor<V>::type;
  static constexpr int NRED = 2;
  const V* ybar;  // gathered
  const V* aty;
  const V* xbar;
  const V* x;
  __device__ bool skip() const { return false; }
  struct Pre {
    V aty, xbar, x;
  };
  __device__ typename Acc::Elem product(V v, int col, int) const { return Acc::prod(v, ybar[col]); }
  __device__ Pre load(int r) const { return Pre{aty[r], xbar[r], x[r]}; }
  __device__ void row(int, const Acc& atyb, const Pre& p, double (&red)[kMaxRed]) const
  {
    const double dx = (double)p.xbar - (double)p.x;
    const double t  = (double)atyb.value() - (double)p.aty;
    red[0] += t * dx;
    red[1] += dx * dx;
  }
};

// Unscaled primal point x = Dc x-hat / beta (recentred: x* + that / dp), one
// FP64 value per column, gathered by the row pass of the KKT evaluation: one
// random 8-byte read per nonzero instead of three (x-hat, Dc, x*).  The same
// expression, evaluated in the same order, as the column pass's own x.
template <typename V>
struct KktPointF {
  static constexpr int NRED = 0;
  const V* xhat;
  const double* Dc;
  double inv_beta;
  const double* xstar;
  double inv_dp;
  double* xu;
  __device__ bool skip() const { return false; }
  __device__ void operator()(int j, double (&)[kMaxRed]) const
  {
    const double xs = ((double)xhat[j] * Dc[j]) * inv_beta;
    xu[j]           = xstar ? xstar[j] + xs * inv_dp : xs;
  }
};

// ---- termination evaluation on the unscaled problem (FP64 matrix).
// Rows i of A: gathers unscaled x-hat on the fly, owns y-hat.
// red: 0 ||Ax - clamp(Ax)||^2, 1 sum of primal slack terms, 2 ||y0 - yhat||^2 (rescaled)
template <typename V>
struct KktRowsEpi {
  using Mat = double;
  using Acc = AccSum64;
  static constexpr int NRED = 3;
  const double* xu;  // gathered: the unscaled primal point (KktPointF)
  V* yhat;
  const double* Dr;
  double inv_gamma, gamma;
  const double2* cb_orig;  // unscaled constraint bounds
  const V* y0;
  double* yu_out;
  // Recentred (refined) solves (else null): the iterates are corrections around the
  // FP64 point (xstar, ystar) in units 1/inv_dp, 1/inv_dd; the residuals are
  // those of the original problem at the corrected point, and the iterates are
  // not rewritten.
  const double* xstar;
  const double* ystar;
  double inv_dp, inv_dd;
  double* rp_out = nullptr;  // per_constraint_residual: the row's primal residual (else null)
  __device__ bool skip() const { return false; }
  __device__ double product(double v, int col, int) const { return v * xu[col]; }
  struct Pre {
    double dr;
    double2 b;
    V yhat, y0;
  };
  __device__ Pre load(int r) const { return Pre{Dr[r], cb_orig[r], yhat[r], y0[r]}; }
  __device__ void row(int r, const Acc& axs, const Pre& p, double (&red)[kMaxRed]) const
  {
    const double ax = axs.value();
    const double dr = p.dr;
    double yu       = ((double)p.yhat * dr) * inv_gamma;
    if (ystar) yu = ystar[r] + yu * inv_dd;
    const double2 b = p.b;
    const double cl = mx(b.x, mn(ax, b.y));
    const double rp = ax - cl;
    red[0] += rp * rp;
    if (rp_out) rp_out[r] = rp;
    red[1] += mx(yu, 0.0) * fin0(b.x) + mn(yu, 0.0) * fin0(b.y);
    yu_out[r]     = yu;
    if (ystar) {
      const double d = (double)p.y0 - (double)p.yhat;
      red[2] += d * d;
      return;
    }
    double ys     = dr != 0.0 ? yu / dr : yu;
    ys            = ys * gamma;
    const V ynew  = (V)ys;
    yhat[r]       = ynew;
    const double d = (double)p.y0 - 1.0 * (double)ynew;
    red[2] += d * d;
  }
};

// Rows j of A^T: gathers unscaled y (written by KktRowsEpi), owns x-hat and slack.
// red: 0 ||c - A^T y - s||^2, 1 x.c, 2 s.x, 3 ||x0 - xhat||^2 (rescaled)
template <typename V>
struct KktColsEpi {
  using Mat = double;
  using Acc = AccSum64;
  static constexpr int NRED = 4;
  const double* yu;  // gathered
  V* xhat;
  V* slack;
  const double* Dc;
  double inv_beta, beta, inv_gamma, gamma;
  const double* c_orig;
  const V* x0;
  double* xu_out;
  double* su_out;
  const double* xstar;  // recentred solves, as in KktRowsEpi (else null)
  double inv_dp, inv_dd;
  // Recentred solves: scaled bounds of the correction problem.  A variable
  // whose x-hat sits on one of them has the slack "FP64 reduced cost of the
  // corrected point + stored small part"; any other variable has zero slack.
  const typename Vec2<V>::type* bnd;
  const double* lb_orig;  // recentred solves: a variable on its correction bound
  const double* ub_orig;  // is put exactly on the original bound
  double* rd_out = nullptr;  // per_constraint_residual: the column's dual residual (else null)
  __device__ bool skip() const { return false; }
  struct Pre {
    double dc, c;
    V xhat, slack, x0;
    typename Vec2<V>::type b;
  };
  __device__ double product(double v, int col, int) const { return v * yu[col]; }
  __device__ Pre load(int r) const
  {
    typename Vec2<V>::type bb{};
    if (xstar) bb = bnd[r];
    return Pre{Dc[r], c_orig[r], xhat[r], slack[r], x0[r], bb};
  }
  __device__ void row(int r, const Acc& atys, const Pre& p, double (&red)[kMaxRed]) const
  {
    const double aty = atys.value();
    const double dc = p.dc;
    double xu       = ((double)p.xhat * dc) * inv_beta;
    double su       = dc != 0.0 ? (double)p.slack / dc : (double)p.slack;
    su              = su * inv_gamma;
    const double cj = p.c;
    const double g  = cj - aty;
    if (xstar) {
      const bool lo = p.xhat == p.b.x, hi = p.xhat == p.b.y;
      xu            = lo ? lb_orig[r] : (hi ? ub_orig[r] : xstar[r] + xu * inv_dp);
      su            = (lo || hi) ? g + su * inv_dd : 0.0;
    }
    const double rd = g - su;
    red[0] += rd * rd;
    if (rd_out) rd_out[r] = rd;
    red[1] += xu * cj;
    red[2] += su * xu;
    xu_out[r] = xu;
    su_out[r] = su;
    if (xstar) {
      const double d = (double)p.x0 - (double)p.xhat;
      red[3] += d * d;
      return;
    }
    double xs = dc != 0.0 ? xu / dc : xu;
    xs        = xs * beta;
    double ss = su * dc;
    ss        = ss * gamma;
    const V xnew = (V)xs;
    xhat[r]      = xnew;
    slack[r]     = (V)ss;
    const double d = (double)p.x0 - 1.0 * (double)xnew;
    red[3] += d * d;
  }
};

// ---- power iteration on A~ A~^T (FP64).  State in PowerState (device).
struct PowerState {
  double nrm;   // ||z|| of the current z
  double s2;    // sigma_max^2 estimate (q . A A^T q)
  double res;   // ||A A^T q - s2 q||
  double done;  // 1 once res < tol
  double iters;
  double pad[3];
};

struct PowerAtEpi {  // atq = A^T (z / ||z||)
  using Mat = double;
  using Acc = AccSum64;
  static constexpr int NRED = 0;
  const double* z;
  const PowerState* st;
  double* atq;
  __device__ bool skip() const { return st->done != 0.0; }
  struct Pre {};
  __device__ double product(double v, int col, int) const { return v * (z[col] / st->nrm); }
  __device__ Pre load(int) const { return Pre{}; }
  __device__ void row(int r, const Acc& s, const Pre&, double (&)[kMaxRed]) const { atq[r] = s.value(); }
};

struct PowerAEpi {  // znew = A atq, red[0] = q . znew
  using Mat = double;
  using Acc = AccSum64;
  static constexpr int NRED = 1;
  const double* atq;
  const double* zold;
  const PowerState* st;
  double* znew;
  __device__ bool skip() const { return st->done != 0.0; }
  struct Pre {
    double zold;
  };
  __device__ double product(double v, int col, int) const { return v * atq[col]; }
  __device__ Pre load(int r) const { return Pre{zold[r]}; }
  __device__ void row(int r, const Acc& sum, const Pre& p, double (&red)[kMaxRed]) const
  {
    const double s = sum.value();
    znew[r]        = s;
    const double q = p.zold / st->nrm;
    red[0] += q * s;
  }
};

// ---- Ruiz / Pock-Chambolle row statistics of |(a * D_r) * D_c|.
// For A the gathered index is the column and rowid[k] the row; for A^T the
// roles swap (rows of A^T are columns of A), selected by TRANSPOSED.
template <bool TRANSPOSED, bool MAXNORM>
struct ScaleStatEpi {
  using Mat = double;
  using Acc = typename std::conditional<MAXNORM, AccMax64, AccSum64>::type;
  static constexpr int NRED = 0;
  const int* rowid;
  const double* Dr;
  const double* Dc;
  double alpha_pow;  // exponent applied to |a~| for sums (1 skips pow)
  double* out;
  __device__ bool skip() const { return false; }
  __device__ double product(double v, int col, int k) const
  {
    const int rr = rowid[k];
    const double dr = TRANSPOSED ? Dr[col] : Dr[rr];
    const double dc = TRANSPOSED ? Dc[rr] : Dc[col];
    const double a  = fabs((v * dr) * dc);
    if constexpr (MAXNORM) return a;
    return alpha_pow == 1.0 ? a : pow(a, alpha_pow);
  }
  struct Pre {};
  __device__ Pre load(int) const { return Pre{}; }
  __device__ void row(int r, const Acc& s, const Pre&, double (&)[kMaxRed]) const { out[r] = s.value(); }
};

// ---- plain y = A x (FP64 or compensated FP32) for the test entry point.
template <typename V>
struct PlainEpi {
  using Mat = V;
  using Acc = typename AccFor<V>::type;
  static constexpr int NRED = 0;
  const V* x;
  double* y;
  __device__ bool skip() const { return false; }
  struct Pre {};
  __device__ typename Acc::Elem product(V v, int col, int) const { return Acc::prod(v, x[col]); }
  __device__ Pre load(int) const { return Pre{}; }
  __device__ void row(int r, const Acc& s, const Pre&, double (&)[kMaxRed]) const { y[r] = (double)s.value(); }
};

}  // namespace cuopt_amd
