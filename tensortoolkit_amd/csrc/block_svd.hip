// Batched FP64 block SVD / QR.
#include "qlt_internal.h"

extern "C" {

qlt_status qlt_block_svd_f64(int64_t, const int64_t *, const int64_t *, const double *, const int64_t *,
                             double *, const int64_t *, double *, const int64_t *, double *, const int64_t *,
                             void *) {
  return qlt::fail(QLT_ERR_UNSUPPORTED, "qlt_block_svd_f64 is not built yet");
}

qlt_status qlt_block_qr_f64(int64_t, const int64_t *, const int64_t *, const double *, const int64_t *,
                            double *, const int64_t *, double *, const int64_t *, void *) {
  return qlt::fail(QLT_ERR_UNSUPPORTED, "qlt_block_qr_f64 is not built yet");
}

}  // extern "C"
