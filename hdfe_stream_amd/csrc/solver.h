// Internal interface between the pipeline sources and the solver (pcg.hip).
#pragma once
#include <stdint.h>

#include <vector>

namespace hdfe {
// Create the solver's device matrix from a host int32 indptr (L+1) and
// indices / values / inverse diagonal on the host or the device.
// With own_* given (device buffers from dev_malloc holding the row pointers,
// indices and values, all work on them finished), the matrix takes them over
// instead of copying.
int matrix_create(int64_t L, int64_t nnz, const std::vector<int> &ip32, const int *indices,
                  const double *data, const double *dinv, bool on_device, void **out,
                  int *own_indptr = nullptr, int *own_indices = nullptr,
                  double *own_data = nullptr);
}  // namespace hdfe
