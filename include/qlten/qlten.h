// Umbrella header of the qlten host API over libqlten_amd (the reference's
// entry header is include/qlten/qlten.h).
#ifndef QLTEN_AMD_QLTEN_H
#define QLTEN_AMD_QLTEN_H

#include "qlten/framework/backend.h"
#include "qlten/framework/value_t.h"
#include "qlten/qltensor/index.h"
#include "qlten/qltensor/qltensor.h"
#include "qlten/qltensor/special_qn.h"
#include "qlten/tensor_manipulation/basic_operations.h"
#include "qlten/tensor_manipulation/contract_contiguous_axes.h"
#include "qlten/tensor_manipulation/ten_ctrct.h"
#include "qlten/tensor_manipulation/ten_decomp/ten_qr.h"
#include "qlten/tensor_manipulation/ten_decomp/ten_svd.h"
#include "qlten/tensor_manipulation/tensor_op_cost.h"

#endif
