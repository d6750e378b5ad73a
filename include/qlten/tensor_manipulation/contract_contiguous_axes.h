// Contiguous-axes contraction (tensor_manipulation/contract_contiguous_axes.h:
// 849-1071): contracted axes (start + i) % rank, saved axes cyclic from the
// end of the contracted run ((end + i) % rank), result indexes A.saved ++
// B.saved.  The ASide / BSide template arguments are layout hints in the
// reference (which side of each operand the contracted run sits on); here
// the gather prologue handles every layout, so they select nothing.
// Accumulate form: C = beta * C + alpha * A.B into an existing output
// topology (contract_contiguous_axes.h:664-782,953-1041).
#ifndef QLTEN_AMD_TENSOR_MANIPULATION_CONTRACT_CONTIGUOUS_AXES_H
#define QLTEN_AMD_TENSOR_MANIPULATION_CONTRACT_CONTIGUOUS_AXES_H

#include <stdexcept>
#include <string>
#include <vector>

#include "qlten/tensor_manipulation/ten_ctrct.h"

namespace qlten {

enum class CtrctSide { Head, Tail };

struct ContiguousContractStats {
  size_t transpose_prepare_calls = 0;
  size_t transpose_prepare_bytes = 0;
  size_t raw_data_contract_tasks = 0;
  size_t gemm_calls = 0;
  size_t accumulate_calls = 0;
  size_t accumulate_gemm_calls = 0;
  size_t output_tensor_rebuilds = 0;
  size_t temporary_output_bytes_avoided = 0;
  size_t output_topology_expansions = 0;
  size_t output_expand_copy_bytes = 0;
  size_t output_expand_new_blocks = 0;
  size_t output_untouched_scale_bytes = 0;
};

namespace detail {

class ContractAccumulateLayoutMismatch : public std::invalid_argument {
 public:
  explicit ContractAccumulateLayoutMismatch(const std::string &m) : std::invalid_argument(m) {}
};

inline void ContiguousAxes(size_t rank, size_t start, size_t size, std::vector<size_t> &ctrct,
                           std::vector<size_t> &saved) {
  ctrct.clear();
  saved.clear();
  for (size_t i = 0; i < size; ++i) ctrct.push_back((start + i) % rank);
  const size_t end = (start + size) % rank;
  for (size_t i = 0; i < rank - size; ++i) saved.push_back((end + i) % rank);
}

}  // namespace detail

template <typename TenElemT, typename QNT, CtrctSide ASide = CtrctSide::Tail, CtrctSide BSide = CtrctSide::Head>
void ContractContiguousAxes(const QLTensor<TenElemT, QNT> &pa, const QLTensor<TenElemT, QNT> &pb,
                            const size_t a_ctrct_axes_start, const size_t b_ctrct_axes_start,
                            const size_t ctrct_axes_size, QLTensor<TenElemT, QNT> &pc,
                            ContiguousContractStats *stats = nullptr) {
  std::vector<size_t> a_ctrct, a_saved, b_ctrct, b_saved;
  detail::ContiguousAxes(pa.Rank(), a_ctrct_axes_start, ctrct_axes_size, a_ctrct, a_saved);
  detail::ContiguousAxes(pb.Rank(), b_ctrct_axes_start, ctrct_axes_size, b_ctrct, b_saved);
  detail::ContractWithAxes(pa, pb, a_ctrct, b_ctrct, a_saved, b_saved, &pc);
  if (stats) {
    ++stats->output_tensor_rebuilds;
    size_t n = 0;
    if (!pa.IsDefault() && !pb.IsDefault())
      n = detail::GetPlan(pa.GetBlkSparDataTen(), pb.GetBlkSparDataTen(), a_ctrct, b_ctrct, a_saved, b_saved)->n_tasks;
    stats->raw_data_contract_tasks = n;
    stats->gemm_calls = n;
  }
}

template <typename TenElemT, typename QNT>
void ContractTailHeadContiguous(const QLTensor<TenElemT, QNT> &pa, const QLTensor<TenElemT, QNT> &pb,
                                const size_t a_start, const size_t b_start, const size_t size,
                                QLTensor<TenElemT, QNT> &pc, ContiguousContractStats *stats = nullptr) {
  ContractContiguousAxes<TenElemT, QNT, CtrctSide::Tail, CtrctSide::Head>(pa, pb, a_start, b_start, size, pc,
                                                                          stats);
}

// Legacy form Contract<T, Q, a_ctrct_tail, b_ctrct_head>(A, B, a0, b0, n, C)
// (contract_contiguous_axes.h:1057-1071).
template <typename TenElemT, typename QNT, bool a_ctrct_tail = true, bool b_ctrct_head = true>
void Contract(const QLTensor<TenElemT, QNT> &pa, const QLTensor<TenElemT, QNT> &pb, const size_t a_start,
              const size_t b_start, const size_t size, QLTensor<TenElemT, QNT> &pc) {
  constexpr CtrctSide AS = a_ctrct_tail ? CtrctSide::Tail : CtrctSide::Head;
  constexpr CtrctSide BS = b_ctrct_head ? CtrctSide::Head : CtrctSide::Tail;
  ContractContiguousAxes<TenElemT, QNT, AS, BS>(pa, pb, a_start, b_start, size, pc);
}

// C = beta * C + alpha * (A . B) into C's existing topology; blocks of C the
// product does not touch are scaled by beta; blocks the product creates are
// added (unless allow_expansion is false, which throws
// ContractAccumulateLayoutMismatch).  A default C requires beta == 0.
template <typename TenElemT, typename QNT>
void ContractTailHeadContiguousAccumulate(const QLTensor<TenElemT, QNT> &pa, const QLTensor<TenElemT, QNT> &pb,
                                          const size_t a_start, const size_t b_start, const size_t size,
                                          const TenElemT alpha, const TenElemT beta, QLTensor<TenElemT, QNT> &pc,
                                          ContiguousContractStats *stats = nullptr,
                                          const bool allow_expansion = true) {
  if (&pc == &pa || &pc == &pb)
    throw std::invalid_argument(
        "ContractTailHeadContiguousAccumulate does not support aliasing between output and input tensors.");
  std::vector<size_t> a_ctrct, a_saved, b_ctrct, b_saved;
  detail::ContiguousAxes(pa.Rank(), a_start, size, a_ctrct, a_saved);
  detail::ContiguousAxes(pb.Rank(), b_start, size, b_ctrct, b_saved);
  if (stats) ++stats->accumulate_calls;
  if (pc.IsDefault()) {
    if (beta != TenElemT(0))
      throw std::invalid_argument("ContractTailHeadContiguousAccumulate requires beta == 0 when output is default.");
    detail::ContractWithAxes(pa, pb, a_ctrct, b_ctrct, a_saved, b_saved, &pc);
    if (alpha != TenElemT(1)) pc *= alpha;
    if (stats) ++stats->output_tensor_rebuilds;
    return;
  }
  IndexVec<QNT> expect;
  for (size_t ax : a_saved) expect.push_back(pa.GetIndexes()[ax]);
  for (size_t ax : b_saved) expect.push_back(pb.GetIndexes()[ax]);
  if (pc.GetIndexes() != expect)
    throw detail::ContractAccumulateLayoutMismatch(
        "ContractTailHeadContiguousAccumulate output indexes are not compatible with the contraction result.");
  auto plan = detail::GetPlan(pa.GetBlkSparDataTen(), pb.GetBlkSparDataTen(), a_ctrct, b_ctrct, a_saved, b_saved);
  // make sure every block the product touches exists in C
  const auto &cmap = pc.GetBlkSparDataTen().GetBlkIdxDataBlkMap();
  bool expand = false;
  for (size_t i = 0; i < plan->c_idx.size(); ++i) {
    auto it = cmap.find(plan->c_idx[i]);
    if (it == cmap.end()) {
      expand = true;
    } else if (it->second.shape != plan->c_dims[i]) {
      throw detail::ContractAccumulateLayoutMismatch(
          "ContractTailHeadContiguousAccumulate output block shape is not compatible with the contraction result.");
    }
  }
  if (expand) {
    if (!allow_expansion)
      throw detail::ContractAccumulateLayoutMismatch(
          "ContractTailHeadContiguousAccumulate output block topology requires expansion.");
    QLTensor<TenElemT, QNT> grown(pc.GetIndexes());
    std::vector<CoorsT> coors;
    std::map<size_t, CoorsT> all;
    for (const auto &kv : cmap) all[kv.first] = kv.second.blk_coors;
    for (size_t i = 0; i < plan->c_idx.size(); ++i) all[plan->c_idx[i]] = plan->c_coords[i];
    for (const auto &kv : all) coors.push_back(kv.second);
    grown.GetBlkSparDataTen().DataBlksInsert(coors);
    grown.GetBlkSparDataTen().Allocate(true);
    pc = QLTensor<TenElemT, QNT>::AxpbyUnion(TenElemT(1), grown, TenElemT(1), pc);
    if (stats) ++stats->output_topology_expansions;
  }
  if (plan->n_tasks == 0) {
    if (beta != TenElemT(1)) pc *= beta;
    return;
  }
  // untouched blocks: scale by beta
  auto &cb = pc.GetBlkSparDataTen();
  const auto &cm = cb.GetBlkIdxDataBlkMap();
  std::vector<int64_t> offs, lens, tgt;
  size_t j = 0;
  for (const auto &kv : cm) {
    while (j < plan->c_idx.size() && plan->c_idx[j] < kv.first) ++j;
    const bool touched = (j < plan->c_idx.size() && plan->c_idx[j] == kv.first);
    if (touched) {
      tgt.push_back(static_cast<int64_t>(kv.second.data_offset));
    } else if (beta != TenElemT(1)) {
      offs.push_back(static_cast<int64_t>(kv.second.data_offset));
      lens.push_back(static_cast<int64_t>(kv.second.size));
    }
  }
  const std::complex<double> bz(beta), az(alpha);
  const double b2[2] = {bz.real(), bz.imag()}, a2[2] = {az.real(), az.imag()};
  if (!offs.empty())
    QltCheck(qlt_scale_segments(QLTensor<TenElemT, QNT>::DType(), static_cast<int64_t>(offs.size()), offs.data(),
                                lens.data(), b2, cb.GetActualRawDataPtr(), hp_numeric::LibStream()));
  // touched blocks: C_blk = alpha * sum A'B' + beta * C_blk, at C's own offsets
  QltCheck(qlt_plan_set_c_offsets(plan->plan, tgt.data()));
  QltCheck(qlt_contract_execute(plan->plan, QLTensor<TenElemT, QNT>::DType(), hp_numeric::ContractModeFlag(),
                                pa.GetBlkSparDataTen().GetActualRawDataPtr(),
                                pb.GetBlkSparDataTen().GetActualRawDataPtr(), cb.GetActualRawDataPtr(), a2, b2,
                                hp_numeric::LibStream()));
  // restore the plan's own offsets for later plain contractions
  std::vector<int64_t> own(plan->c_offsets.begin(), plan->c_offsets.end());
  QltCheck(qlt_plan_set_c_offsets(plan->plan, own.data()));
  if (stats) stats->accumulate_gemm_calls = plan->n_tasks;
}

template <typename TenElemT, typename QNT>
bool TryContractTailHeadContiguousAccumulate(const QLTensor<TenElemT, QNT> &pa, const QLTensor<TenElemT, QNT> &pb,
                                             const size_t a_start, const size_t b_start, const size_t size,
                                             const TenElemT alpha, const TenElemT beta, QLTensor<TenElemT, QNT> &pc,
                                             ContiguousContractStats *stats = nullptr) {
  try {
    ContractTailHeadContiguousAccumulate(pa, pb, a_start, b_start, size, alpha, beta, pc, stats, false);
    return true;
  } catch (const detail::ContractAccumulateLayoutMismatch &) {
    return false;
  }
}

}  // namespace qlten

#endif
