// batch_size.hpp -- choice of the batch size of a same-matrix batch (SpMM).
#pragma once

#include <functional>

namespace cuopt_amd {

// The search of the reference's batch-size handler
// (optimal_batch_size_handler.cu:196-435): start at the largest power of two
// <= min(128, max_batch), compare with half and double that size, walk in the
// better direction by factors of two while the cost per batch member keeps
// falling (at most four evaluations after the first three), then try the size
// halfway back to the last one evaluated.  `cost(b)` is the time of the batch's
// SpMMs at batch size b divided by b.  Returns a size in [1, max_batch].
int search_batch_size(int max_batch, const std::function<double(int)>& cost);

}  // namespace cuopt_amd
