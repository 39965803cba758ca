// presolve.cpp -- see presolve.hpp.
#include "presolve.hpp"

#include <algorithm>
#include <cmath>

namespace cuopt_amd {

Presolved presolve(const LpHost& lp, const std::vector<double>& lc0, const std::vector<double>& uc0,
                   const std::vector<double>& lb0, const std::vector<double>& ub0)
{
  Presolved P;
  const int m = lp.m, n = lp.n;
  const int nnz = m > 0 ? lp.rowptr[m] : 0;
  std::vector<double> lc(lc0), uc(uc0), lb(lb0), ub(ub0);
  // Columns of A (CSC) for the substitution of fixed columns.
  std::vector<int> cptr(n + 1, 0), crow(nnz);
  std::vector<double> cval(nnz);
  for (int k = 0; k < nnz; ++k) cptr[lp.colidx[k] + 1]++;
  for (int j = 0; j < n; ++j) cptr[j + 1] += cptr[j];
  {
    std::vector<int> cur(cptr.begin(), cptr.end() - 1);
    for (int i = 0; i < m; ++i)
      for (int k = lp.rowptr[i]; k < lp.rowptr[i + 1]; ++k) {
        const int d = cur[lp.colidx[k]]++;
        crow[d]     = i;
        cval[d]     = lp.vals[k];
      }
  }
  std::vector<char> row_on(m, 1), col_on(n, 1);
  std::vector<int> rcount(m), ccount(n);
  for (int i = 0; i < m; ++i) rcount[i] = lp.rowptr[i + 1] - lp.rowptr[i];
  for (int j = 0; j < n; ++j) ccount[j] = cptr[j + 1] - cptr[j];
  P.x_removed.assign(n, 0.0);
  P.lb_from.assign(n, -1);
  P.ub_from.assign(n, -1);
  double offset = lp.offset;
  std::vector<double> cc(lp.c.begin(), lp.c.begin() + n);  // costs (user sense) after substitutions
  auto cost = [&](int j) { return lp.maximize ? -cc[j] : cc[j]; };  // minimisation sense

  std::vector<int> rq, cq;  // rows / columns to look at again
  for (int i = 0; i < m; ++i)
    if (rcount[i] <= 1) rq.push_back(i);
  for (int j = 0; j < n; ++j)
    if (ccount[j] == 0 || lb[j] == ub[j]) cq.push_back(j);

  auto remove_col = [&](int j, double v) {  // x_j = v everywhere
    col_on[j]      = 0;
    P.x_removed[j] = v;
    ++P.removed_cols;
    offset += cc[j] * v;
    for (int k = cptr[j]; k < cptr[j + 1]; ++k) {
      const int i = crow[k];
      if (!row_on[i]) continue;
      const double d = cval[k] * v;
      if (std::isfinite(lc[i])) lc[i] -= d;
      if (std::isfinite(uc[i])) uc[i] -= d;
      if (--rcount[i] <= 1) rq.push_back(i);
    }
  };
  // A column whose remaining entries all sit in removed rows is empty: recount
  // over the active rows until no such column is left.
  auto recount_columns = [&]() {
    std::vector<int> live(n, 0);
    for (int i = 0; i < m; ++i)
      if (row_on[i])
        for (int k = lp.rowptr[i]; k < lp.rowptr[i + 1]; ++k) ++live[lp.colidx[k]];
    for (int j = 0; j < n; ++j)
      if (col_on[j] && live[j] == 0 && ccount[j] != 0) {
        ccount[j] = 0;
        cq.push_back(j);
      }
    return !cq.empty();
  };
  // Redundant rows: a row whose activity range over the current column bounds lies
  // inside [lc, uc] can never bind, so it is dropped (its dual stays 0).  Dual
  // fixing: a column whose cost does not rise when it decreases, and whose every
  // entry only loosens its row when it decreases (a > 0 in a row without a lower
  // bound, a < 0 in a row without an upper bound), takes its lower bound; the
  // mirror case takes its upper bound.  Its reduced cost c - A^T y then has the
  // sign of its bound by the signs of those rows' duals, so postsolve needs nothing
  // more than the recomputed c - A^T y.
  auto redundant_rows = [&]() {
    bool changed = false;
    for (int i = 0; i < m; ++i) {
      if (!row_on[i] || rcount[i] <= 1) continue;
      double lo = 0.0, hi = 0.0;
      for (int k = lp.rowptr[i]; k < lp.rowptr[i + 1]; ++k) {
        const int j    = lp.colidx[k];
        const double a = lp.vals[k];
        if (!col_on[j] || a == 0.0) continue;
        lo += a > 0.0 ? a * lb[j] : a * ub[j];
        hi += a > 0.0 ? a * ub[j] : a * lb[j];
      }
      if (!(lo >= lc[i] && hi <= uc[i])) continue;  // NaN (inf - inf) never qualifies
      row_on[i] = 0;
      ++P.removed_rows;
      changed = true;
      for (int k = lp.rowptr[i]; k < lp.rowptr[i + 1]; ++k) {
        const int j = lp.colidx[k];
        if (col_on[j] && --ccount[j] == 0) cq.push_back(j);
      }
    }
    return changed;
  };
  auto dual_fixing = [&]() {
    bool changed = false;
    for (int j = 0; j < n; ++j) {
      if (!col_on[j] || ccount[j] == 0 || lb[j] == ub[j]) continue;
      bool down_ok = true, up_ok = true;  // decreasing / increasing x_j loosens every row
      for (int k = cptr[j]; k < cptr[j + 1] && (down_ok || up_ok); ++k) {
        const int i = crow[k];
        const double a = cval[k];
        if (!row_on[i] || a == 0.0) continue;
        const bool no_lo = !std::isfinite(lc[i]) && lc[i] < 0, no_up = !std::isfinite(uc[i]) && uc[i] > 0;
        if (a > 0.0) {
          down_ok = down_ok && no_lo;
          up_ok   = up_ok && no_up;
        } else {
          down_ok = down_ok && no_up;
          up_ok   = up_ok && no_lo;
        }
      }
      const double cj = cost(j);
      double v        = 0.0;
      if (down_ok && cj >= 0.0 && std::isfinite(lb[j])) v = lb[j];
      else if (up_ok && cj <= 0.0 && std::isfinite(ub[j])) v = ub[j];
      else continue;  // (toward an infinite bound the cost need not be bounded: left to the solver)
      lb[j] = ub[j] = v;
      cq.push_back(j);
      changed = true;
    }
    return changed;
  };
  // Free column singletons (see presolve.hpp).
  auto free_column_singletons = [&]() {
    bool changed = false;
    for (int j = 0; j < n; ++j) {
      if (!col_on[j] || ccount[j] != 1) continue;
      int i    = -1;
      double a = 0.0;
      bool two = false;
      for (int k = cptr[j]; k < cptr[j + 1]; ++k) {
        if (!row_on[crow[k]]) continue;
        if (i < 0 || crow[k] == i) {
          i = crow[k];
          a += cval[k];  // a repeated column index sums
        } else {
          two = true;
        }
      }
      if (two || i < 0 || a == 0.0) continue;
      double lo = 0.0, hi = 0.0;  // activity range of the row's other entries
      for (int k = lp.rowptr[i]; k < lp.rowptr[i + 1]; ++k) {
        const int jj   = lp.colidx[k];
        const double v = lp.vals[k];
        if (!col_on[jj] || jj == j || v == 0.0) continue;
        lo += v > 0.0 ? v * lb[jj] : v * ub[jj];
        hi += v > 0.0 ? v * ub[jj] : v * lb[jj];
      }
      bool freecol = !std::isfinite(lb[j]) && lb[j] < 0 && !std::isfinite(ub[j]) && ub[j] > 0;
      if (!freecol) {  // implied free: x_j = (t - rest) / a over t in [lc, uc] stays in its bounds
        const double xlo = a > 0.0 ? (lc[i] - hi) / a : (uc[i] - lo) / a;
        const double xhi = a > 0.0 ? (uc[i] - lo) / a : (lc[i] - hi) / a;
        freecol = xlo >= lb[j] && xhi <= ub[j];  // NaN (inf - inf) never qualifies
      }
      if (!freecol) continue;
      const double cj   = cc[j];
      const double cmin = lp.maximize ? -cj : cj;
      double b          = lc[i];
      if (lc[i] != uc[i]) {
        if (cmin == 0.0) {
          b = std::nan("");
        } else {
          // min (c_j / a)(t - rest): t at the lower bound when c_j / a > 0
          b = cmin / a > 0.0 ? lc[i] : uc[i];
          if (!std::isfinite(b)) continue;  // unbounded direction or infeasible rest: left to the solver
        }
      }
      if (std::isfinite(b)) offset += cj * b / a;
      for (int k = lp.rowptr[i]; k < lp.rowptr[i + 1]; ++k) {
        const int jj = lp.colidx[k];
        if (!col_on[jj] || jj == j) continue;
        cc[jj] -= cj * lp.vals[k] / a;
        if (--ccount[jj] == 0) cq.push_back(jj);
      }
      // The row's bounds in terms of the original row: fixed columns were
      // subtracted from lc / uc when removed, and postsolve sums every entry.
      double shift = 0.0;
      for (int k = lp.rowptr[i]; k < lp.rowptr[i + 1]; ++k)
        if (!col_on[lp.colidx[k]]) shift += lp.vals[k] * P.x_removed[lp.colidx[k]];
      P.ops.push_back({'c', (int)P.col_singletons.size()});
      P.col_singletons.push_back({i, j, a, cj, b + shift, lc[i] + shift, uc[i] + shift});
      row_on[i] = 0;
      col_on[j] = 0;
      ccount[j] = 0;
      ++P.removed_rows;
      ++P.removed_cols;
      changed = true;
    }
    return changed;
  };
  while (!rq.empty() || !cq.empty() || recount_columns() || redundant_rows() || dual_fixing() ||
         free_column_singletons()) {
    while (!rq.empty()) {
      const int i = rq.back();
      rq.pop_back();
      if (!row_on[i] || rcount[i] > 1) continue;
      if (rcount[i] == 0) {  // empty row
        const double tol = 1e-9 * (1.0 + std::max(std::fabs(std::isfinite(lc[i]) ? lc[i] : 0.0),
                                                  std::fabs(std::isfinite(uc[i]) ? uc[i] : 0.0)));
        if (lc[i] > tol || uc[i] < -tol) {
          P.status = Presolved::INFEASIBLE;
          return P;
        }
        row_on[i] = 0;
        ++P.removed_rows;
        continue;
      }
      int j = -1;  // singleton row: its one active column
      double a = 0.0;
      for (int k = lp.rowptr[i]; k < lp.rowptr[i + 1]; ++k)
        if (col_on[lp.colidx[k]]) {
          if (j == lp.colidx[k]) {
            a += lp.vals[k];  // a repeated column index sums
          } else {
            j = lp.colidx[k];
            a = lp.vals[k];
          }
        }
      if (j < 0 || a == 0.0) {  // only zero coefficients left: an empty row in effect
        for (int k = lp.rowptr[i]; k < lp.rowptr[i + 1]; ++k) {
          const int jj = lp.colidx[k];
          if (col_on[jj] && --ccount[jj] == 0) cq.push_back(jj);
        }
        rcount[i] = 0;
        rq.push_back(i);
        continue;
      }
      const int s    = (int)P.singletons.size();
      P.ops.push_back({'s', s});
      P.singletons.push_back({i, j, a, cc[j]});
      const double lo = a > 0 ? lc[i] / a : uc[i] / a;
      const double hi = a > 0 ? uc[i] / a : lc[i] / a;
      if (lo > lb[j]) {
        lb[j]        = lo;
        P.lb_from[j] = s;
      }
      if (hi < ub[j]) {
        ub[j]        = hi;
        P.ub_from[j] = s;
      }
      row_on[i] = 0;
      ++P.removed_rows;
      if (lb[j] > ub[j]) {
        const double tol = 1e-9 * (1.0 + std::fabs(lb[j]) + std::fabs(ub[j]));
        if (lb[j] - ub[j] > tol) {
          P.status = Presolved::INFEASIBLE;
          return P;
        }
        lb[j] = ub[j] = 0.5 * (lb[j] + ub[j]);
      }
      if (--ccount[j] == 0 || lb[j] == ub[j]) cq.push_back(j);
    }
    while (!cq.empty()) {
      const int j = cq.back();
      cq.pop_back();
      if (!col_on[j]) continue;
      if (lb[j] == ub[j]) {  // fixed
        if (!std::isfinite(lb[j])) {  // fixed at an infinite value: no finite point exists
          P.status = Presolved::INFEASIBLE;
          return P;
        }
        remove_col(j, lb[j]);
        break;  // its rows may have become singletons
      }
      if (ccount[j] != 0) continue;
      const double cj = cost(j);  // empty column: the bound its cost prefers
      double v        = 0.0;
      if (cj > 0.0) v = lb[j];
      else if (cj < 0.0) v = ub[j];
      else v = std::min(std::max(0.0, lb[j]), ub[j]);
      if (!std::isfinite(v)) {
        P.status = Presolved::UNBOUNDED;
        return P;
      }
      remove_col(j, v);
    }
  }

  // A row whose activity range over the (tightened) column bounds misses
  // [lc, uc] cannot hold: the problem is infeasible.
  for (int i = 0; i < m; ++i) {
    if (!row_on[i]) continue;
    double lo = 0.0, hi = 0.0;  // -inf / +inf or finite: zero coefficients are skipped
    for (int k = lp.rowptr[i]; k < lp.rowptr[i + 1]; ++k) {
      const int j    = lp.colidx[k];
      const double a = lp.vals[k];
      if (!col_on[j] || a == 0.0) continue;
      lo += a > 0.0 ? a * lb[j] : a * ub[j];
      hi += a > 0.0 ? a * ub[j] : a * lb[j];
    }
    auto fin = [](double v) { return std::isfinite(v) ? std::fabs(v) : 0.0; };
    const double tol = 1e-9 * (1.0 + std::max(fin(lc[i]), fin(uc[i])) + std::max(fin(lo), fin(hi)));
    if (lo > uc[i] + tol || hi < lc[i] - tol) {
      P.status = Presolved::INFEASIBLE;
      return P;
    }
  }

  // The reduced problem (ranged form, continuous variables).
  std::vector<int> new_col(n, -1);
  for (int j = 0; j < n; ++j)
    if (col_on[j]) {
      new_col[j] = (int)P.col_map.size();
      P.col_map.push_back(j);
    }
  LpHost& R = P.reduced;
  R.maximize = lp.maximize;
  R.offset   = offset;
  R.ranged   = true;
  R.n        = (int)P.col_map.size();
  R.rowptr.push_back(0);
  for (int i = 0; i < m; ++i) {
    if (!row_on[i]) continue;
    P.row_map.push_back(i);
    for (int k = lp.rowptr[i]; k < lp.rowptr[i + 1]; ++k) {
      const int j = lp.colidx[k];
      if (!col_on[j]) continue;
      R.colidx.push_back(new_col[j]);
      R.vals.push_back(lp.vals[k]);
    }
    R.rowptr.push_back((int)R.colidx.size());
    P.lc.push_back(lc[i]);
    P.uc.push_back(uc[i]);
  }
  R.m = (int)P.row_map.size();
  for (int j : P.col_map) {
    R.c.push_back(cc[j]);
    P.lb.push_back(lb[j]);
    P.ub.push_back(ub[j]);
  }
  R.lc = P.lc;
  R.uc = P.uc;
  R.lb = P.lb;
  R.ub = P.ub;
  R.var_types.assign(R.n, CUOPT_CONTINUOUS);
  return P;
}

void postsolve(const Presolved& P, const LpHost& lp, SolveResult& R)
{
  const int m = lp.m, n = lp.n;
  std::vector<double> x(P.x_removed), y(m, 0.0);
  for (size_t j = 0; j < P.col_map.size() && j < R.x.size(); ++j) x[P.col_map[j]] = R.x[j];
  for (size_t i = 0; i < P.row_map.size() && i < R.y.size(); ++i) y[P.row_map[i]] = R.y[i];
  // Column singletons' values, latest first: each depends only on columns still
  // in the problem after it, whose values are known by then.
  for (int q = (int)P.ops.size() - 1; q >= 0; --q) {
    if (P.ops[q].first != 'c') continue;
    const auto& C = P.col_singletons[P.ops[q].second];
    double rest   = 0.0;
    for (int k = lp.rowptr[C.row]; k < lp.rowptr[C.row + 1]; ++k)
      if (lp.colidx[k] != C.col) rest += lp.vals[k] * x[lp.colidx[k]];
    const double t = std::isnan(C.b) ? std::min(std::max(rest, C.lo), C.hi) : C.b;
    x[C.col]       = (t - rest) / C.a;
  }
  // z = c - A^T y over the original rows (removed rows still at y = 0), then the
  // removed rows' duals, latest first.
  std::vector<double> z(lp.c.begin(), lp.c.begin() + n);
  for (int i = 0; i < m; ++i)
    if (y[i] != 0.0)
      for (int k = lp.rowptr[i]; k < lp.rowptr[i + 1]; ++k) z[lp.colidx[k]] -= lp.vals[k] * y[i];
  auto add_dual = [&](int row, double dy) {
    y[row] += dy;
    for (int k = lp.rowptr[row]; k < lp.rowptr[row + 1]; ++k) z[lp.colidx[k]] -= lp.vals[k] * dy;
  };
  for (int q = (int)P.ops.size() - 1; q >= 0; --q) {
    if (P.ops[q].first == 'c') {  // x_j's reduced cost c_j - a y_i is 0
      const auto& C = P.col_singletons[P.ops[q].second];
      add_dual(C.row, C.cj / C.a);
      continue;
    }
    // A singleton row whose bound holds x_j takes x_j's reduced cost in the
    // problem as it stood then: z_j less the cost moved onto x_j before it.
    const int s      = P.ops[q].second;
    const auto& S    = P.singletons[s];
    const int j      = S.col;
    const bool at_lo = P.lb_from[j] == s, at_hi = P.ub_from[j] == s;
    if (!at_lo && !at_hi) continue;
    const double zt = z[j] - (lp.c[j] - S.cj);
    const double zj = lp.maximize ? -zt : zt;  // minimisation sense
    // at a lower bound the reduced cost is >= 0, at an upper bound <= 0
    const bool use = (at_lo && zj > 0.0) || (at_hi && zj < 0.0);
    if (use) add_dual(S.row, zt / S.a);
  }
  // the reduced costs of the final duals, recomputed in one pass
  z.assign(lp.c.begin(), lp.c.begin() + n);
  for (int i = 0; i < m; ++i)
    if (y[i] != 0.0)
      for (int k = lp.rowptr[i]; k < lp.rowptr[i + 1]; ++k) z[lp.colidx[k]] -= lp.vals[k] * y[i];
  double obj = lp.offset;
  for (int j = 0; j < n; ++j) obj += lp.c[j] * x[j];
  R.x  = std::move(x);
  R.y  = std::move(y);
  R.rc = std::move(z);
  R.primal_objective = obj;
}

}  // namespace cuopt_amd
