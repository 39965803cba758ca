// presolve.hpp -- LP presolve and postsolve on the host.
//
// The reference runs a third-party presolver (PSLP by default, solve.cu:1183-1325)
// before PDLP and undoes it on the solution, duals included (dual_postsolve).
// This engine's presolve applies the reductions below until none applies; they
// are exact (no tolerance changes the problem), and each one's dual effect is
// undone in postsolve:
//   - an empty row: dropped after checking lc <= 0 <= uc (else infeasible); y = 0;
//   - a fixed column (lb == ub): substituted into its rows and the objective offset;
//   - a singleton row a x_j in [lc, uc]: turned into bounds on x_j; in postsolve
//     its y takes the reduced cost of x_j when x_j sits on a bound it supplied;
//   - an empty column: set to the bound its cost prefers (unbounded if that
//     bound is infinite and the cost nonzero), reduced cost = cost;
//   - a redundant row (activity range over the column bounds inside [lc, uc]):
//     dropped; y = 0;
//   - dual fixing: a column that can move toward one of its finite bounds without
//     raising the cost or tightening any row is fixed there; its reduced cost
//     c - A^T y has the right sign by the signs of its rows' duals;
//   - a free column singleton: a column whose only entry a is in row i and whose
//     bounds are free, or implied by row i over the other columns' bounds, is
//     solved from the row and substituted: its cost moves onto the row's other
//     columns (c_k -= c_j a_ik / a) and the offset, and the row goes.  An
//     inequality row settles at the bound its cost prefers (left to the solver if
//     that bound is infinite); with zero cost any point of the row will do.  In
//     postsolve x_j = (b - sum a_ik x_k) / a and y_i = c_j / a, with c_j the cost
//     when it was removed, so x_j's reduced cost is 0.
// Afterwards a row whose activity range over the column bounds misses its
// [lc, uc] (beyond 1e-9 relative) reports the problem infeasible.
// Everything runs in FP64 in the user's objective sense; the reduced problem
// is handed to the solver in ranged form.
#pragma once

#include <vector>

#include "lp_types.hpp"

namespace cuopt_amd {

struct Presolved {
  enum Status { REDUCED, INFEASIBLE, UNBOUNDED } status = REDUCED;
  // Costs below are in the user's objective sense, as they stood when the
  // reduction was made (free column singletons move cost between columns).
  LpHost reduced;                   // ranged form; lc / uc below
  std::vector<double> lc, uc, lb, ub;  // bounds of the reduced problem
  std::vector<int> row_map;         // reduced row -> original row
  std::vector<int> col_map;         // reduced column -> original column
  std::vector<double> x_removed;    // original column -> its value when removed
  struct Singleton {
    int row, col;
    double a;
    double cj;  // the column's cost at that time
  };
  std::vector<Singleton> singletons;  // in removal order
  struct ColSingleton {
    int row, col;
    double a;
    double cj;     // the column's cost at that time: the row's dual is cj / a
    double b;      // the row's activity, original row (NaN: zero cost, any point of [lo, hi])
    double lo, hi; // the row's bounds then, original row (columns fixed before it included)
  };
  std::vector<ColSingleton> col_singletons;  // in removal order
  std::vector<std::pair<char, int>> ops;     // removal order: ('s', singleton) or ('c', column singleton)
  std::vector<int> lb_from, ub_from;  // original column -> singleton that set its bound (-1: none)
  int removed_rows = 0, removed_cols = 0;
};

// lc / uc: the constraint bounds of lp (constraint_bounds); lb / ub its variable bounds.
Presolved presolve(const LpHost& lp, const std::vector<double>& lc, const std::vector<double>& uc,
                   const std::vector<double>& lb, const std::vector<double>& ub);

// Maps a solution of P.reduced back to lp: x, y and the reduced costs c - A^T y
// of the original problem; the primal objective is recomputed on lp.
void postsolve(const Presolved& P, const LpHost& lp, SolveResult& R);

}  // namespace cuopt_amd
