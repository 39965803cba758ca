// batch_size.cpp -- see batch_size.hpp.
#include "batch_size.hpp"

#include <algorithm>

namespace cuopt_amd {

int search_batch_size(int max_batch, const std::function<double(int)>& cost)
{
  if (max_batch <= 1) return 1;
  constexpr int kSteps = 4;  // the first step is the choice of direction
  int start            = 1;
  while (start * 2 <= std::min(128, max_batch)) start *= 2;
  const int lo = std::max(1, start / 2), hi = std::min(2 * start, max_batch);
  const double c_start = cost(start), c_lo = cost(lo), c_hi = cost(hi);
  int steps = 1;
  if (c_lo < c_start) {  // smaller batches are cheaper per member: keep halving
    if (lo == 1) return 1;
    int best = lo, b = lo;
    double c_best = c_lo;
    do {
      b              = std::max(1, b / 2);
      const double c = cost(b);
      if (!(c < c_best)) break;
      if (b == 1) return 1;
      ++steps;
      best   = b;
      c_best = c;
    } while (steps < kSteps);
    const int mid = std::max(1, (2 * b + b) / 2);
    return cost(mid) < c_best ? mid : best;
  }
  if (c_hi < c_start) {  // larger batches are cheaper per member: keep doubling
    if (hi == max_batch) return hi;
    int best = hi, b = hi;
    double c_best = c_hi;
    do {
      b              = std::min(2 * b, max_batch);
      const double c = cost(b);
      if (!(c < c_best)) break;
      if (b == max_batch) return b;
      ++steps;
      best   = b;
      c_best = c;
    } while (steps < kSteps);
    const int mid = std::min(std::max(1, (b / 2 + b) / 2), max_batch);
    return cost(mid) < c_best ? mid : best;
  }
  return start;
}

}  // namespace cuopt_amd
