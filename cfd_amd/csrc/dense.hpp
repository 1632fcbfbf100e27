// Batched dense inversion on the matrix cores (dense.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

namespace cfdg {

// Largest order inverted by one workgroup (two LDS panels of order x 16 doubles per workgroup;
// above it the whole device takes one matrix, which was faster already at order 488: 1 against
// 6 ms), and the largest inverted on the device at all (two launches per 16-row block step).
constexpr int kDenseMaxOrder = 256;
constexpr int kDenseWideMaxOrder = 2048;
// Dense tails of the FP32 cycle up to this order are taken whatever the finest level's size.
constexpr int kTailOrder = 512;

// Inverts the row-major order[i] x order[i] host matrices mats[i] into out[i] on the device
// (block Gauss-Jordan without pivoting, FP64 MFMA): one workgroup per matrix up to kDenseMaxOrder,
// the whole device per matrix above. Returns per matrix 1 when the device produced the inverse,
// 0 when the caller must use the host path (order above kDenseWideMaxOrder, or a zero or
// non-finite pivot).
std::vector<int> dense_inverse_batch(const std::vector<int>& order, const std::vector<const double*>& mats,
                                     const std::vector<double*>& out, hipStream_t s);

} // namespace cfdg
