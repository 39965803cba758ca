// mps_io.hpp -- MPS reader and writer behind cuOptReadProblem / cuOptWriteProblem.
//
// Reading follows the semantics of the reference parser in the free format that
// cuOptReadProblem selects (cpp/libmps_parser/src/mps_parser.cpp, cuopt_c.cpp:103):
//   - whitespace-separated records, which also covers fixed-format files whose
//     names hold no blanks;
//   - comment lines start with '*' or '$', and a data field starting with '$'
//     comments out the rest of its line;
//   - OBJSENSE and OBJNAME as one-line or two-line sections; OBJNAME must come
//     before ROWS; without it the first N row is the objective and the entries of
//     further N rows are dropped;
//   - ROWS, COLUMNS, RHS and an objective are required; a column's entries must
//     be contiguous; row names are unique;
//   - the right-hand side of the objective row is the negated objective offset;
//   - integer markers in COLUMNS; an integer column without a BOUNDS entry is
//     binary;
//   - RANGES by row type (E: sign picks [b, b+|R|] or [b-|R|, b]; L: [b-|R|, b];
//     G: [b, b+|R|]);
//   - BOUNDS types UP LO FX FR MI PL BV LI UI with the CPLEX rule that a negative
//     UP / UI on a variable without an earlier bound sets its lower bound to -inf
//     (mps_parser.cpp:1370-1440).  Default variable bounds are [0, +inf).
// One deliberate difference: a row named by OBJNAME keeps its COLUMNS entries
// (the reference files every N row met after OBJNAME among the ignored ones,
// its own included, mps_parser.cpp:851-856 and 995, which drops them).
#pragma once

#include <stdexcept>
#include <string>

#include "lp_types.hpp"

namespace cuopt_amd {

struct mps_file_error : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct mps_parse_error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// Problems without RANGES keep the row-type form (sense / rhs); with RANGES the
// ranged form (lc / uc) is what the solver uses, and sense / rhs stay readable.
LpHost read_mps(const std::string& path);

// Free-format MPS; values printed to round-trip.  Rows, columns, the objective and
// the problem keep their names when every name of the kind can be read back
// (non-empty, no blanks, unique, not a '$' comment, not RHS / RNG / BND / MARKER);
// otherwise the kind gets generated names R<i> / C<j> / OBJ.
void write_mps(const LpHost& lp, const std::string& path);

// The solution file of the solution_file parameter (solve.cu:1327-1330,
// solution_writer.cu:16-40): "# Status: Optimal" or "Primal Feasible" (any other
// status is written as "Infeasible"), then, unless infeasible, "# Objective
// value: <primal objective>" and one "<name> <value>" line per named variable,
// 18 significant digits.  A file that cannot be opened is reported on stderr and
// skipped, as the reference does.
void write_sol_file(const std::string& path, const LpHost& lp, const SolveResult& R);

}  // namespace cuopt_amd
