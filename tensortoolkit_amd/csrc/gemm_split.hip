// Split-precision (BF16-sliced, FP64-emulating) grouped GEMM.
#include "qlt_internal.h"

namespace qlt {

qlt_status split_workspace_bytes(const qlt_plan *plan, int32_t dtype, size_t *bytes) {
  (void)plan;
  (void)dtype;
  *bytes = 0;
  return QLT_OK;
}

qlt_status launch_split(const qlt_plan *, int32_t, const void *, const void *, void *, const double *,
                        const double *, hipStream_t) {
  return fail(QLT_ERR_UNSUPPORTED, "split-BF16 mode is not built yet");
}

}  // namespace qlt
