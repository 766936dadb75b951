// The allocation-free SpMV (noalloc.hip): y = M v on caller-owned device arrays with a file-scope scratch array.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace tpdhg {

// Largest nonzero count one call covers (the scratch array's capacity).
int64_t spmv_noalloc_max_nnz();
// Enqueues y = M v on stream st (ptr, idx: device int32 CSR of `rows` rows and `nnz` nonzeros; val, v, y: device
// float or double per precision_bits). Asynchronous; allocates nothing. Throws on bad sizes or launch errors.
void spmv_noalloc(int precision_bits, int rows, int64_t nnz, const int* ptr, const int* idx, const void* val,
                  const void* v, void* y, hipStream_t st);

}  // namespace tpdhg
