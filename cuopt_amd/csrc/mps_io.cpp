// mps_io.cpp -- see mps_io.hpp.
#include "mps_io.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <limits>
#include <sstream>
#include <unordered_map>
#include <unordered_set>
#include <vector>

namespace cuopt_amd {

namespace {

constexpr double kInf = std::numeric_limits<double>::infinity();

std::vector<std::string> tokens(const std::string& line)
{
  std::vector<std::string> t;
  std::istringstream ss(line);
  std::string w;
  while (ss >> w) t.push_back(w);
  return t;
}

std::string upper(std::string s)
{
  for (char& ch : s) ch = (char)std::toupper((unsigned char)ch);
  return s;
}

double number(const std::string& s, const std::string& line)
{
  char* end = nullptr;
  const double v = std::strtod(s.c_str(), &end);
  if (end == s.c_str() || *end != '\0') throw mps_parse_error("bad number '" + s + "' in line: " + line);
  return v;
}

}  // namespace

LpHost read_mps(const std::string& path)
{
  std::ifstream in(path);
  if (!in) throw mps_file_error("Error opening MPS file: " + path);

  enum Section { NONE, NAME, OBJSENSE, OBJNAME, ROWS, COLUMNS, RHS, RANGES, BOUNDS };
  Section sec   = NONE;
  bool maximize = false, integer_block = false;
  bool seen_rows = false, seen_columns = false, seen_rhs = false;
  std::string obj_row;                           // the objective: the first N row, or the OBJNAME row
  std::unordered_set<std::string> ignored_rows;  // further N rows: their entries are dropped
  std::vector<std::string> row_names;
  std::vector<char> row_type;
  std::unordered_map<std::string, int> row_id, col_id;
  std::vector<std::string> col_names;
  std::vector<char> vtype;
  struct Entry {
    int row, col;
    double val;
  };
  std::vector<Entry> entries;
  std::string problem_name;
  std::vector<double> cost, lb, ub;
  std::vector<char> has_bound;
  std::unordered_map<int, double> rhs, range;
  double offset = 0.0;

  auto col_of = [&](const std::string& name) {
    auto it = col_id.find(name);
    if (it != col_id.end()) return it->second;
    const int j = (int)col_names.size();
    col_id.emplace(name, j);
    col_names.push_back(name);
    vtype.push_back(integer_block ? 'I' : 'C');
    cost.push_back(0.0);
    lb.push_back(0.0);
    ub.push_back(kInf);
    has_bound.push_back(0);
    return j;
  };
  auto row_of = [&](const std::string& name, const std::string& line) {
    auto it = row_id.find(name);
    if (it == row_id.end()) throw mps_parse_error("unknown row '" + name + "' in line: " + line);
    return it->second;
  };

  auto set_objname = [&](const std::string& name) {
    // OBJNAME must come before ROWS (mps_parser.cpp:1274-1278)
    if (!obj_row.empty()) throw mps_parse_error("OBJNAME section should appear before ROWS section");
    obj_row = name;
  };
  std::string line;
  int last_col = -1;
  while (std::getline(in, line)) {
    if (!line.empty() && line.back() == '\r') line.pop_back();
    auto t = tokens(line);
    // comment lines start with '*' or '$' (mps_parser.cpp:594-596)
    if (t.empty() || line[0] == '*' || line[0] == '$') continue;
    if (!std::isspace((unsigned char)line[0])) {  // section header
      const std::string h = upper(t[0]);
      if (h == "NAME") {
        sec = NAME;
        if (t.size() > 1) problem_name = t[1];
      } else if (h == "OBJSENSE" || h == "OBJSENCE") {
        sec = OBJSENSE;
        if (t.size() > 1) maximize = upper(t[1]).rfind("MAX", 0) == 0;
      } else if (h == "OBJNAME") {
        sec = OBJNAME;
        if (t.size() > 1) set_objname(t[1]);
      } else if (h == "ROWS") {
        sec       = ROWS;
        seen_rows = true;
      } else if (h == "COLUMNS") {
        sec          = COLUMNS;
        seen_columns = true;
      } else if (h == "RHS") {
        sec      = RHS;
        seen_rhs = true;
      } else if (h == "RANGES") {
        sec = RANGES;
      } else if (h == "BOUNDS") {
        sec = BOUNDS;
      } else if (h == "ENDATA") {
        break;
      } else if (h == "QUADOBJ" || h == "QMATRIX" || h == "QSECTION" || h == "QCMATRIX") {
        throw mps_parse_error("quadratic sections are not supported by the LP solver: " + h);
      } else if (sec == OBJSENSE && (h.rfind("MAX", 0) == 0 || h.rfind("MIN", 0) == 0)) {
        maximize = h.rfind("MAX", 0) == 0;
      } else {
        throw mps_parse_error("unknown section: " + line);
      }
      continue;
    }
    // In a data line a field starting with '$' begins a comment that runs to
    // the end of the line (mps_parser.cpp:978, 1082, 1154, 1196).
    for (size_t p = 1; p < t.size(); ++p)
      if (t[p][0] == '$') {
        t.resize(p);
        break;
      }
    switch (sec) {
      case OBJSENSE: maximize = upper(t[0]).rfind("MAX", 0) == 0; break;
      case OBJNAME: set_objname(t[0]); break;
      case ROWS: {
        if (t.size() < 2) throw mps_parse_error("bad ROWS line: " + line);
        const std::string ty = upper(t[0]);
        if (ty == "N") {
          // The first N row is the objective unless OBJNAME named one; every later N
          // row is ignored, the one OBJNAME names included, so that row's COLUMNS
          // entries are dropped and only its RHS (the offset) is read
          // (mps_parser.cpp:851-856, 995, 1095).
          if (obj_row.empty()) obj_row = t[1];
          else ignored_rows.insert(t[1]);
          break;
        }
        if (ty != "E" && ty != "L" && ty != "G") throw mps_parse_error("bad row type in line: " + line);
        if (row_id.count(t[1])) throw mps_parse_error("duplicate row '" + t[1] + "'");
        row_id.emplace(t[1], (int)row_names.size());
        row_names.push_back(t[1]);
        row_type.push_back(ty[0]);
        break;
      }
      case COLUMNS: {
        if (t.size() >= 3 && (t[1] == "'MARKER'" || upper(t[1]) == "MARKER")) {
          const std::string mk = upper(t[2]);
          if (mk.find("INTORG") != std::string::npos) integer_block = true;
          if (mk.find("INTEND") != std::string::npos) integer_block = false;
          break;
        }
        if (t.size() % 2 == 0) throw mps_parse_error("bad COLUMNS line: " + line);
        // a column's entries must be contiguous (mps_parser.cpp:911-918)
        if (last_col < 0 || col_names[last_col] != t[0]) {
          if (col_id.count(t[0]))
            throw mps_parse_error("All rows for the column (" + t[0] + ") should occur contiguously! line=" + line);
          last_col = col_of(t[0]);
        }
        const int j = last_col;
        for (size_t p = 1; p + 1 < t.size(); p += 2) {
          const double v = number(t[p + 1], line);
          if (ignored_rows.count(t[p])) continue;
          if (t[p] == obj_row) {
            cost[j] = v;
          } else {
            auto it = row_id.find(t[p]);
            if (it == row_id.end()) {
              throw mps_parse_error("unknown row '" + t[p] + "' in line: " + line);
            }
            entries.push_back(Entry{it->second, j, v});
          }
        }
        break;
      }
      case RHS:
      case RANGES: {
        // Optional set name.  RHS: there is none when the first field is a row
        // or the objective (mps_parser.cpp:1052-1056); RANGES: an odd field count.
        size_t first = (t.size() % 2 == 1) ? 1 : 0;
        if (sec == RHS) first = (t[0] == obj_row || row_id.count(t[0])) ? 0 : 1;
        if (t.size() - first == 1 || (t.size() - first) % 2 != 0)
          throw mps_parse_error("bad RHS/RANGES line: " + line);
        for (size_t p = first; p + 1 < t.size(); p += 2) {
          const double v = number(t[p + 1], line);
          // RHS of the objective row is the negated offset; any other row must be a
          // constraint, an ignored N row included (mps_parser.cpp:1095-1107)
          if (sec == RHS && t[p] == obj_row) {
            offset = -v;
          } else if (sec == RHS) {
            rhs[row_of(t[p], line)] = v;
          } else {
            range[row_of(t[p], line)] = v;
          }
        }
        break;
      }
      case BOUNDS: {
        const std::string bt = upper(t[0]);
        const bool valueless = bt == "FR" || bt == "MI" || bt == "PL" || bt == "BV";
        // records: type [set] column [value]; there is no set name when the
        // second field is a column (mps_parser.cpp:1147-1151)
        if (t.size() < 2) throw mps_parse_error("bad BOUNDS line: " + line);
        const size_t ci = (t.size() >= 3 && !col_id.count(t[1])) ? 2 : 1;
        const std::string name = t[ci];
        double v               = 0.0;
        if (!valueless) {
          if (t.size() < ci + 2) throw mps_parse_error("bad BOUNDS line: " + line);
          v = number(t[ci + 1], line);
        }
        const int j       = col_of(name);
        const bool first  = !has_bound[j];
        if (bt == "LO") {
          lb[j] = v;
        } else if (bt == "UP") {
          ub[j] = v;
          if (first && v < 0.0) lb[j] = -kInf;
        } else if (bt == "FX") {
          lb[j] = v;
          ub[j] = v;
        } else if (bt == "FR") {
          lb[j] = -kInf;
          ub[j] = kInf;
        } else if (bt == "MI") {
          lb[j] = -kInf;
        } else if (bt == "PL") {
          ub[j] = kInf;
        } else if (bt == "BV") {
          lb[j]    = 0.0;
          ub[j]    = 1.0;
          vtype[j] = 'I';
        } else if (bt == "LI") {
          if (first) ub[j] = kInf;
          lb[j]    = v;
          vtype[j] = 'I';
        } else if (bt == "UI") {
          ub[j] = v;
          if (first && v < 0.0) lb[j] = -kInf;
          vtype[j] = 'I';
        } else if (bt == "SC") {
          throw mps_parse_error("semi-continuous bounds are not supported: " + line);
        } else {
          throw mps_parse_error("invalid bound type in line: " + line);
        }
        has_bound[j] = 1;
        break;
      }
      case NAME: break;
      default: throw mps_parse_error("data line outside a section: " + line);
    }
  }

  // mps_parser.cpp:772-780
  if (obj_row.empty()) throw mps_parse_error("No objective found!");
  if (!seen_rows) throw mps_parse_error("ROWS section is missing");
  if (!seen_columns) throw mps_parse_error("COLUMNS section is missing");
  if (!seen_rhs) throw mps_parse_error("RHS section is missing");
  // an integer column without a BOUNDS entry is binary (mps_parser.cpp:801-807)
  for (size_t j = 0; j < col_names.size(); ++j)
    if (vtype[j] == 'I' && !has_bound[j]) {
      lb[j] = 0.0;
      ub[j] = 1.0;
    }

  LpHost lp;
  lp.m        = (int)row_names.size();
  lp.n        = (int)col_names.size();
  lp.row_names      = row_names;
  lp.col_names      = col_names;
  lp.objective_name = obj_row;
  lp.problem_name   = problem_name;
  lp.maximize = maximize;
  lp.offset   = offset;
  lp.c        = cost;
  lp.lb       = lb;
  lp.ub       = ub;
  lp.var_types = vtype;
  // CSR, entries of a row in column order (duplicates kept as given)
  std::stable_sort(entries.begin(), entries.end(),
                   [](const Entry& a, const Entry& b) { return a.row != b.row ? a.row < b.row : a.col < b.col; });
  lp.rowptr.assign(lp.m + 1, 0);
  for (const Entry& e : entries) lp.rowptr[e.row + 1]++;
  for (int i = 0; i < lp.m; ++i) lp.rowptr[i + 1] += lp.rowptr[i];
  lp.colidx.reserve(entries.size());
  lp.vals.reserve(entries.size());
  for (const Entry& e : entries) {
    lp.colidx.push_back(e.col);
    lp.vals.push_back(e.val);
  }
  lp.sense = row_type;
  lp.rhs.assign(lp.m, 0.0);
  for (const auto& kv : rhs) lp.rhs[kv.first] = kv.second;
  // Both row forms, as the reference parser's data model carries them
  // (mps_parser.cpp:348-412): sense / rhs and the [lower, upper] pairs with RANGES
  // applied.  The pairs are the ones the solver reads.
  lp.ranged = true;
  {
    lp.lc.resize(lp.m);
    lp.uc.resize(lp.m);
    for (int i = 0; i < lp.m; ++i) {
      const double b = lp.rhs[i];
      double lo = b, hi = b;
      if (row_type[i] == 'L') lo = -kInf;
      if (row_type[i] == 'G') hi = kInf;
      auto it = range.find(i);
      if (it != range.end()) {
        const double r = it->second;
        if (row_type[i] == 'E') {
          if (r > 0) hi = b + std::fabs(r);
          else lo = b - std::fabs(r);
        } else if (row_type[i] == 'L') {
          lo = b - std::fabs(r);
        } else {
          hi = b + std::fabs(r);
        }
      }
      lp.lc[i] = lo;
      lp.uc[i] = hi;
    }
  }
  return lp;
}

void write_mps(const LpHost& lp, const std::string& path)
{
  FILE* f = std::fopen(path.c_str(), "w");
  if (!f) throw mps_file_error("Error opening MPS file for writing: " + path);
  auto num = [](double v) {
    char b[64];
    std::snprintf(b, sizeof(b), "%.17g", v);
    return std::string(b);
  };
  // Row bounds in the ranged form, whatever form the problem was created in.
  std::vector<double> lo(lp.m), hi(lp.m);
  for (int i = 0; i < lp.m; ++i) {
    if (lp.ranged) {
      lo[i] = lp.lc[i];
      hi[i] = lp.uc[i];
    } else {
      const double b = lp.rhs[i];
      lo[i]          = lp.sense[i] == 'L' ? -kInf : b;
      hi[i]          = lp.sense[i] == 'G' ? kInf : b;
    }
  }
  // The problem's names when they can be read back (non-empty, no blanks, unique,
  // not a '$' comment, not a set name this writer uses); R<i> / C<j> / OBJ otherwise.
  const std::unordered_set<std::string> reserved{"RHS", "RNG", "BND", "MARKER", "'MARKER'"};
  auto readable = [&](const std::string& s) {
    if (s.empty() || s[0] == '$' || reserved.count(s)) return false;
    for (char ch : s)
      if (std::isspace((unsigned char)ch)) return false;
    return true;
  };
  auto pick = [&](const std::vector<std::string>& v, int count) {
    if ((int)v.size() != count) return false;
    std::unordered_set<std::string> seen;
    for (const auto& s : v)
      if (!readable(s) || !seen.insert(s).second) return false;
    return true;
  };
  const bool rows_readable = pick(lp.row_names, lp.m), cols_readable = pick(lp.col_names, lp.n);
  auto is_row = [&](const std::string& s) {
    return std::find(lp.row_names.begin(), lp.row_names.end(), s) != lp.row_names.end();
  };
  std::string obj = readable(lp.objective_name) ? lp.objective_name : "OBJ";
  if (rows_readable && is_row(obj)) obj = "OBJ";         // the objective row needs a name of its own
  const bool use_rows = rows_readable && !is_row(obj);  // a row named OBJ: generated row names
  auto rname = [&](int i) { return use_rows ? lp.row_names[i] : "R" + std::to_string(i); };
  auto cname = [&](int j) { return cols_readable ? lp.col_names[j] : "C" + std::to_string(j); };
  std::fprintf(f, "NAME %s\n", readable(lp.problem_name) ? lp.problem_name.c_str() : "cuopt_amd");
  if (lp.maximize) std::fprintf(f, "OBJSENSE\n    MAX\n");
  std::fprintf(f, "ROWS\n N %s\n", obj.c_str());
  std::vector<char> ty(lp.m);
  for (int i = 0; i < lp.m; ++i) {
    if (lo[i] == hi[i]) ty[i] = 'E';
    else if (std::isinf(lo[i]) && lo[i] < 0) ty[i] = 'L';
    else ty[i] = 'G';  // finite lower bound; a finite upper bound becomes a range
    std::fprintf(f, " %c %s\n", ty[i], rname(i).c_str());
  }
  // columns: transpose the CSR rows
  std::vector<int> cptr(lp.n + 1, 0);
  const int nnz = lp.rowptr[lp.m];
  for (int k = 0; k < nnz; ++k) cptr[lp.colidx[k] + 1]++;
  for (int j = 0; j < lp.n; ++j) cptr[j + 1] += cptr[j];
  std::vector<int> crow(nnz);
  std::vector<double> cval(nnz);
  {
    std::vector<int> cur(cptr.begin(), cptr.end() - 1);
    for (int i = 0; i < lp.m; ++i)
      for (int k = lp.rowptr[i]; k < lp.rowptr[i + 1]; ++k) {
        const int d = cur[lp.colidx[k]]++;
        crow[d]     = i;
        cval[d]     = lp.vals[k];
      }
  }
  std::fprintf(f, "COLUMNS\n");
  bool in_int = false;
  for (int j = 0; j < lp.n; ++j) {
    const bool is_int = !lp.var_types.empty() && lp.var_types[j] == 'I';
    if (is_int != in_int) {
      std::fprintf(f, "    MARKER 'MARKER' '%s'\n", is_int ? "INTORG" : "INTEND");
      in_int = is_int;
    }
    const std::string cj = cname(j);
    std::fprintf(f, "    %s %s %s\n", cj.c_str(), obj.c_str(), num(lp.c[j]).c_str());
    for (int k = cptr[j]; k < cptr[j + 1]; ++k)
      std::fprintf(f, "    %s %s %s\n", cj.c_str(), rname(crow[k]).c_str(), num(cval[k]).c_str());
  }
  if (in_int) std::fprintf(f, "    MARKER 'MARKER' 'INTEND'\n");
  std::fprintf(f, "RHS\n");
  if (lp.offset != 0.0) std::fprintf(f, "    RHS %s %s\n", obj.c_str(), num(-lp.offset).c_str());
  for (int i = 0; i < lp.m; ++i) {
    const double b = ty[i] == 'L' ? hi[i] : lo[i];
    if (b != 0.0) std::fprintf(f, "    RHS %s %s\n", rname(i).c_str(), num(b).c_str());
  }
  bool any_range = false;
  for (int i = 0; i < lp.m; ++i)
    if (ty[i] == 'G' && std::isfinite(hi[i])) {
      if (!any_range) std::fprintf(f, "RANGES\n");
      any_range = true;
      std::fprintf(f, "    RNG %s %s\n", rname(i).c_str(), num(hi[i] - lo[i]).c_str());
    }
  std::fprintf(f, "BOUNDS\n");
  for (int j = 0; j < lp.n; ++j) {
    const double l = lp.lb[j], u = lp.ub[j];
    const std::string cj = cname(j);
    const char* c = cj.c_str();
    if (l == u) {
      std::fprintf(f, " FX BND %s %s\n", c, num(l).c_str());
      continue;
    }
    if (std::isinf(l) && std::isinf(u)) {
      std::fprintf(f, " FR BND %s\n", c);
      continue;
    }
    bool written = false;
    if (std::isinf(l)) {
      std::fprintf(f, " MI BND %s\n", c);
      written = true;
    } else if (l != 0.0) {
      std::fprintf(f, " LO BND %s %s\n", c, num(l).c_str());
      written = true;
    }
    if (std::isfinite(u)) {
      std::fprintf(f, " UP BND %s %s\n", c, num(u).c_str());
      written = true;
    }
    // an integer column without a BOUNDS entry would read back as binary
    if (!written && !lp.var_types.empty() && lp.var_types[j] == 'I') std::fprintf(f, " PL BND %s\n", c);
  }
  std::fprintf(f, "ENDATA\n");
  const bool ok = std::ferror(f) == 0;
  std::fclose(f);
  if (!ok) throw mps_file_error("Error writing MPS file: " + path);
}

void write_sol_file(const std::string& path, const LpHost& lp, const SolveResult& R)
{
  FILE* f = std::fopen(path.c_str(), "w");
  if (!f) {
    std::fprintf(stderr, "Could not open file: %s for solution output\n", path.c_str());
    return;
  }
  const char* status = R.status == CUOPT_TERIMINATION_STATUS_OPTIMAL           ? "Optimal"
                       : R.status == CUOPT_TERIMINATION_STATUS_PRIMAL_FEASIBLE ? "Primal Feasible"
                                                                               : "Infeasible";
  std::fprintf(f, "# Status: %s\n", status);
  if (std::strcmp(status, "Infeasible") != 0) {
    std::fprintf(f, "# Objective value: %.18g\n", R.primal_objective);
    for (size_t j = 0; j < lp.col_names.size() && j < R.x.size(); ++j)
      std::fprintf(f, "%s %.18g\n", lp.col_names[j].c_str(), R.x[j]);
  }
  std::fclose(f);
}

}  // namespace cuopt_amd
