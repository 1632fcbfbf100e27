// Batched dense inversion on the matrix cores (dense.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

namespace cfdg {

// Largest order inverted on the device (two LDS panels of order x 16 doubles per workgroup).
constexpr int kDenseMaxOrder = 512;

// Inverts the row-major order[i] x order[i] host matrices mats[i] into out[i] on the device, one
// workgroup per matrix (block Gauss-Jordan without pivoting, FP64 MFMA). Returns per matrix 1 when
// the device produced the inverse, 0 when the caller must use the host path (order above
// kDenseMaxOrder, or a zero or non-finite pivot).
std::vector<int> dense_inverse_batch(const std::vector<int>& order, const std::vector<const double*>& mats,
                                     const std::vector<double*>& out, hipStream_t s);

} // namespace cfdg
