// Elementwise tensor operations used around the contraction path:
// ToComplex (basic_operations.h:98-130, VectorRealToCplxKernel
// blas_level1.h:825-843), Dag (basic_operations.h:35-60: inverse every
// index, conjugate the data), LinearCombine (ten_linear_combine.h:186) and
// the 2-norm inner product.
#ifndef QLTEN_AMD_TENSOR_MANIPULATION_BASIC_OPERATIONS_H
#define QLTEN_AMD_TENSOR_MANIPULATION_BASIC_OPERATIONS_H

#include <stdexcept>
#include <type_traits>
#include <vector>

#include "qlten/qltensor/qltensor.h"

namespace qlten {

namespace detail {
// The same block layout with the data converted element by element
// (real -> complex or a precision change).
template <typename To, typename From, typename QNT>
QLTensor<To, QNT> ConvertElems(const QLTensor<From, QNT> &t) {
  QLTensor<To, QNT> out(t.GetIndexes());
  if (t.IsDefault()) return out;
  const auto &src = t.GetBlkSparDataTen();
  auto &dst = out.GetBlkSparDataTen();
  std::vector<CoorsT> coors;
  for (const auto &kv : src.GetBlkIdxDataBlkMap()) coors.push_back(kv.second.blk_coors);
  if (out.IsScalar()) dst.RefreshOffsets();
  else dst.DataBlksInsert(coors);
  if (src.GetActualRawDataSize() == 0) return out;
  dst.Allocate();
  QltCheck(qlt_convert(QLTensor<From, QNT>::DType(), QLTensor<To, QNT>::DType(),
                       static_cast<int64_t>(src.GetActualRawDataSize()), src.GetActualRawDataPtr(),
                       dst.GetActualRawDataPtr(), hp_numeric::LibStream()));
  return out;
}
}  // namespace detail

template <typename QNT>
QLTensor<QLTEN_Complex, QNT> ToComplex(const QLTensor<QLTEN_Double, QNT> &real_t) {
  return detail::ConvertElems<QLTEN_Complex>(real_t);
}
template <typename QNT>
QLTensor<QLTEN_ComplexFloat, QNT> ToComplex(const QLTensor<QLTEN_Float, QNT> &real_t) {
  return detail::ConvertElems<QLTEN_ComplexFloat>(real_t);
}

template <typename ElemT, typename QNT>
QLTensor<ElemT, QNT> Dag(const QLTensor<ElemT, QNT> &t) {
  IndexVec<QNT> inv;
  for (const auto &ix : t.GetIndexes()) inv.push_back(InverseIndex(ix));
  QLTensor<ElemT, QNT> shell(inv);
  if (t.IsDefault()) return shell;
  // same block layout and raw data under the inverted indexes, conjugated
  const auto &src = t.GetBlkSparDataTen();
  auto &dst = shell.GetBlkSparDataTen();
  std::vector<CoorsT> coors;
  for (const auto &kv : src.GetBlkIdxDataBlkMap()) coors.push_back(kv.second.blk_coors);
  if (shell.IsScalar()) dst.RefreshOffsets();
  else dst.DataBlksInsert(coors);
  dst.Allocate();
  if (src.GetActualRawDataSize()) {
    hp_numeric::SyncLibStream();
    QltCheck(qlt_memcpy_d2d(dst.GetActualRawDataPtr(), src.GetActualRawDataPtr(),
                            src.GetActualRawDataSize() * sizeof(ElemT)));
    if constexpr (std::is_same<ElemT, QLTEN_Complex>::value)
      QltCheck(qlt_conjugate(static_cast<int64_t>(src.GetActualRawDataSize()),
                             reinterpret_cast<double *>(dst.GetActualRawDataPtr()), hp_numeric::LibStream()));
    else if constexpr (std::is_same<ElemT, QLTEN_ComplexFloat>::value)
      QltCheck(qlt_conjugate_c64(static_cast<int64_t>(src.GetActualRawDataSize()),
                                 reinterpret_cast<float *>(dst.GetActualRawDataPtr()), hp_numeric::LibStream()));
  }
  return shell;
}

// res = sum_i coefs[i] * tens[i]  (LinearCombine, ten_linear_combine.h:186;
// the block map of the result is the union of the terms').
template <typename ElemT, typename QNT>
void LinearCombine(const std::vector<ElemT> &coefs, const std::vector<QLTensor<ElemT, QNT> *> &tens,
                   const ElemT beta, QLTensor<ElemT, QNT> *res) {
  if (coefs.size() != tens.size()) throw std::invalid_argument("LinearCombine: size mismatch");
  QLTensor<ElemT, QNT> acc;
  if (!res->IsDefault() && beta != ElemT(0)) acc = (*res) * beta;
  for (size_t i = 0; i < tens.size(); ++i) {
    if (tens[i]->IsDefault()) continue;
    acc = QLTensor<ElemT, QNT>::AxpbyUnion(ElemT(1), acc, coefs[i], *tens[i]);
  }
  *res = std::move(acc);
}

// <a|b> = sum conj(a_i) b_i over equal block layouts.
template <typename ElemT, typename QNT>
ElemT InnerProduct(const QLTensor<ElemT, QNT> &a, const QLTensor<ElemT, QNT> &b) {
  const auto &sa = a.GetBlkSparDataTen(), &sb = b.GetBlkSparDataTen();
  if (sa.TopologySignature() != sb.TopologySignature())
    throw std::invalid_argument("InnerProduct: tensors have different block layouts");
  double out[2] = {0.0, 0.0};
  if (sa.GetActualRawDataSize())
    QltCheck(qlt_dot(QLTensor<ElemT, QNT>::DType(), static_cast<int64_t>(sa.GetActualRawDataSize()),
                     sa.GetActualRawDataPtr(), sb.GetActualRawDataPtr(), out, hp_numeric::LibStream()));
  using R = typename RealTypeTrait<ElemT>::type;
  if constexpr (IsComplex<ElemT>()) return ElemT(static_cast<R>(out[0]), static_cast<R>(out[1]));
  else return static_cast<R>(out[0]);
}

}  // namespace qlten

#endif
