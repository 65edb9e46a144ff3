// batch_ws.hip -- the batch kernels of screened Lasso instances with
// asynchronous workers (working sets, Anderson extrapolation, support Newton
// steps), in a translation unit of their own (it compiles in parallel with
// batch.hip).
#include <hip/hip_runtime.h>

#include "batch_kernels.h"

namespace acdc {
const void *batch_ws_kernel(bool fp32, int lanes, int64_t maxlen) {
  return batch_kernel_for<false, true>(fp32, lanes, true, maxlen);
}
}  // namespace acdc
