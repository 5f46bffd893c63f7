// Internal interface between the pipeline sources and the solver (pcg.hip).
#pragma once
#include <stdint.h>

#include <vector>

namespace hdfe {
// Create the solver's device matrix from a host int32 indptr (L+1) and
// indices / values / inverse diagonal on the host or the device.
int matrix_create(int64_t L, int64_t nnz, const std::vector<int> &ip32, const int *indices,
                  const double *data, const double *dinv, bool on_device, void **out);
}  // namespace hdfe
