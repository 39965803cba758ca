// unrolled.hpp -- differentiable unrolled PDHG (forward with checkpoints, VJP).
#pragma once

#include "lp_types.hpp"

namespace cuopt_amd {

struct UnrolledHandle;

// Host CSR (m x n) and variable bounds; ridge lam >= 0; precision selects FP32
// or FP64 iterates.  Computes the step size from a power iteration on A A^T.
UnrolledHandle* unrolled_create(int m, int n, const int* rowptr, const int* colidx, const double* vals,
                                const double* lb, const double* ub, double lam, int precision);
void unrolled_destroy(UnrolledHandle* h);
double unrolled_step_size(UnrolledHandle* h);
// Device pointers, batch-interleaved (v[i * B + b]); lc / uc are m x B when
// bounds_per_batch, else m.  y may be null.  Work is queued on `stream`.
void unrolled_forward(UnrolledHandle* h, int B, int K, int every, const double* c, const double* lc, const double* uc,
                      int bounds_per_batch, double* x, double* y, void* stream);
// VJP of the last forward: gx (n x B), gy (m x B or null) -> gc (n x B), glc, guc (m x B).
void unrolled_backward(UnrolledHandle* h, const double* gx, const double* gy, double* gc, double* glc, double* guc,
                       void* stream);
// Batch size in [1, max_batch] with the least SpMM time per batch member
// (A X plus A^T Y on the handle's matrices, search_batch_size), timed on the
// device: one untimed run, then the mean of five.
int unrolled_optimal_batch_size(UnrolledHandle* h, int max_batch);

}  // namespace cuopt_amd
