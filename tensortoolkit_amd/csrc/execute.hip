// qlt_contract_execute / qlt_plan_prepare: validation and dispatch to the
// native or split-precision grouped kernels.

#include <string>

#include "qlt_internal.h"

namespace qlt {
qlt_status launch_native(const qlt_plan *plan, int32_t dtype, const void *a, const void *b, void *c,
                         const double *alpha, const double *beta, hipStream_t st);
qlt_status split_workspace_bytes(const qlt_plan *plan, int32_t dtype, int32_t mode, size_t *bytes);
qlt_status launch_split(const qlt_plan *plan, int32_t dtype, int32_t mode, const void *a, const void *b, void *c,
                        const double *alpha, const double *beta, hipStream_t st);
qlt_status split_prepare_metadata(qlt_plan *plan, int32_t dtype, int32_t mode);
bool split_profitable(const qlt_plan *plan, int32_t dtype, int32_t mode);
}  // namespace qlt

using namespace qlt;

extern "C" {

qlt_status qlt_plan_prepare(qlt_plan *plan, int32_t dtype, int32_t mode) {
  QLT_REQUIRE_INIT();
  QLT_CHECK_ARG(plan != nullptr, "qlt_plan_prepare: null plan");
  QLT_CHECK_ARG(valid_dtype(dtype), "qlt_plan_prepare: bad dtype");
  QLT_CHECK_ARG(valid_mode(mode), "qlt_plan_prepare: bad mode");
  // single-precision element types always run natively (the split modes emulate FP64)
  if (mode == QLT_MODE_NATIVE || plan->n_tiles == 0 || dtype_is_single(dtype)) return QLT_OK;
  if (!split_profitable(plan, dtype, mode)) return QLT_OK;
  if (plan->split_dtype == dtype && plan->split_mode == mode) return QLT_OK;
  size_t need = 0;
  QLT_TRY(split_workspace_bytes(plan, dtype, mode, &need));
  // An earlier split execute of this plan may still be reading the workspace
  // on a non-blocking stream: wait for the device before it is freed or its
  // records are rewritten.
  if (plan->d_work != nullptr) QLT_HIP(hipDeviceSynchronize());
  if (need > plan->work_bytes) {
    if (plan->d_work != nullptr) QLT_HIP(hipFree(plan->d_work));
    plan->d_work = nullptr;
    plan->work_bytes = 0;
    plan->split_dtype = -1;
    plan->split_mode = -1;
    QLT_HIP(hipMalloc(&plan->d_work, need));
    plan->work_bytes = need;
  }
  QLT_TRY(split_prepare_metadata(plan, dtype, mode));
  return QLT_OK;
}

qlt_status qlt_contract_execute(const qlt_plan *plan, int32_t dtype, int32_t mode, const void *a,
                                const void *b, void *c, const double *alpha, const double *beta,
                                void *stream) {
  QLT_REQUIRE_INIT();
  QLT_CHECK_ARG(plan != nullptr, "qlt_contract_execute: null plan");
  QLT_CHECK_ARG(valid_dtype(dtype), "qlt_contract_execute: bad dtype");
  QLT_CHECK_ARG(valid_mode(mode), "qlt_contract_execute: bad mode");
  QLT_CHECK_ARG(alpha != nullptr && beta != nullptr, "qlt_contract_execute: null alpha/beta");
  QLT_CHECK_ARG(plan->device == current_device(),
                "qlt_contract_execute: plan was built on device " + std::to_string(plan->device));
  if (plan->n_tiles == 0 && plan->n_ptiles == 0) return QLT_OK;
  QLT_CHECK_ARG(a != nullptr && b != nullptr && c != nullptr, "qlt_contract_execute: null data buffer");
  hipStream_t st = as_stream(stream);
  if (mode == QLT_MODE_NATIVE || dtype_is_single(dtype)) return launch_native(plan, dtype, a, b, c, alpha, beta, st);
  if (split_profitable(plan, dtype, mode) && (plan->split_dtype != dtype || plan->split_mode != mode)) {
    QLT_CHECK_ARG(!stream_capturing(st),
                  "qlt_contract_execute: split workspace not prepared; call qlt_plan_prepare before capture");
    QLT_TRY(qlt_plan_prepare(const_cast<qlt_plan *>(plan), dtype, mode));
  }
  return launch_split(plan, dtype, mode, a, b, c, alpha, beta, st);
}

}  // extern "C"
