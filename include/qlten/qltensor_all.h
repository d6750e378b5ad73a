// Umbrella header (reference name qltensor_all.h) of the qlten host API over libqlten_amd (the reference's
// entry header is include/qlten/qlten.h).
#ifndef QLTEN_AMD_QLTENSOR_ALL_H
#define QLTEN_AMD_QLTENSOR_ALL_H

#include "qlten/framework/backend.h"
#include "qlten/framework/value_t.h"
#include "qlten/qltensor/index.h"
#include "qlten/qltensor/qltensor.h"
#include "qlten/qltensor/special_qn.h"
#include "qlten/tensor_manipulation/basic_operations.h"
#include "qlten/tensor_manipulation/contract_contiguous_axes.h"
#include "qlten/tensor_manipulation/ten_ctrct.h"

#endif
