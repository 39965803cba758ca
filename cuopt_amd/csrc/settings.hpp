// settings.hpp -- string-keyed solver parameter registry behind cuOptSetParameter & co.
//
// Same names, value ranges, defaults and error behaviour as the reference's
// registry (cpp/src/math_optimization/solver_settings.cu:55-330): an unknown
// name or an out-of-range value throws std::invalid_argument, which the C ABI
// maps to CUOPT_INVALID_ARGUMENT.  Parameters that only steer the MIP or
// barrier solvers are accepted and stored so that existing callers keep
// working, but the LP PDLP path does not read them.
#pragma once

#include <string>
#include <vector>

#include "lp_types.hpp"

namespace cuopt_amd {

class SolverSettingsRegistry {
 public:
  SolverSettingsRegistry();
  SolverSettingsRegistry(const SolverSettingsRegistry&)            = delete;
  SolverSettingsRegistry& operator=(const SolverSettingsRegistry&) = delete;

  void set_from_string(const std::string& name, const std::string& value);
  void set_int(const std::string& name, int value);
  void set_float(const std::string& name, double value);
  void set_bool(const std::string& name, bool value);
  int get_int(const std::string& name) const;
  double get_float(const std::string& name) const;
  bool get_bool(const std::string& name) const;
  std::string get_as_string(const std::string& name) const;

  PdlpSettings pdlp;

 private:
  template <typename T>
  struct Entry {
    std::string name;
    T* ptr;
    T lo, hi, def;
  };
  struct BoolEntry {
    std::string name;
    bool* ptr;
    bool def;
  };
  struct StrEntry {
    std::string name;
    std::string* ptr;
  };
  std::vector<Entry<int>> ints_;
  std::vector<Entry<double>> floats_;
  std::vector<BoolEntry> bools_;
  std::vector<StrEntry> strs_;
  // storage of parameters the LP path accepts but does not use
  std::vector<int> int_store_;
  std::vector<double> float_store_;
  std::vector<char> bool_store_;
  std::vector<std::string> str_store_;
};

}  // namespace cuopt_amd
