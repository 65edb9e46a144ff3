// batch_svm.hip -- the batch kernels of hinge-SVM dual instances, in a
// translation unit of their own (it compiles in parallel with batch.hip).
#include <hip/hip_runtime.h>

#include "batch_kernels.h"

namespace acdc {
const void *batch_svm_kernel(bool fp32, int lanes, bool async, int64_t maxlen) {
  return batch_kernel_for<true>(fp32, lanes, async, maxlen);
}
}  // namespace acdc
