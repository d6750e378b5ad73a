// Element and coordinate types of the qlten host API.
// Same names as the reference (include/qlten/framework/value_t.h:37-84):
// QLTEN_Double / QLTEN_Complex / QLTEN_Float / QLTEN_ComplexFloat element
// types, CoorsT / ShapeT vectors.
#ifndef QLTEN_AMD_FRAMEWORK_VALUE_T_H
#define QLTEN_AMD_FRAMEWORK_VALUE_T_H

#include <complex>
#include <cstddef>
#include <type_traits>
#include <vector>

namespace qlten {

using QLTEN_Double = double;
using QLTEN_Complex = std::complex<double>;
using QLTEN_Float = float;
using QLTEN_ComplexFloat = std::complex<float>;
using CoorsT = std::vector<size_t>;
using ShapeT = std::vector<size_t>;

template <typename T>
struct RealTypeTrait {
  using type = T;
};
template <>
struct RealTypeTrait<QLTEN_Complex> {
  using type = double;
};
template <>
struct RealTypeTrait<QLTEN_ComplexFloat> {
  using type = float;
};

template <typename ElemT>
constexpr bool IsComplex() {
  return std::is_same<ElemT, QLTEN_Complex>::value || std::is_same<ElemT, QLTEN_ComplexFloat>::value;
}
// Stored in single precision (QLTEN_Float, QLTEN_ComplexFloat).
template <typename ElemT>
constexpr bool IsSingle() {
  return std::is_same<ElemT, QLTEN_Float>::value || std::is_same<ElemT, QLTEN_ComplexFloat>::value;
}

}  // namespace qlten

#endif
