// Cost estimate of a contraction, read off the device plan that Contract runs
// for the same operands (the plan is built once and cached, so a later
// Contract reuses it).
//
// Interface of the reference's estimator: EstimateContractCost with its
// TensorOpCost, BlockSparseTensorLayout and BlockSparseTensorBlock records and
// MakeBlockSparseTensorLayout (tensor_manipulation/tensor_op_cost.h:36-150,
// 434-530).  The counts follow the reference's schedule, task by task
// (m, k, n, beta):
//   flops       += c m k n, c = 2 (real) or 8 (complex);
//   read_bytes  += (m k + k n) elements, plus m n when the task reads the
//                  output block (beta != 0, i.e. not the block's first task);
//   write_bytes += m n elements.
// An operand whose axes are not already in (saved, contracted) order (A) or
// (contracted, saved) order (B) is transposed into a temporary in that
// schedule: the elements of its distinct touched blocks are counted once in
// temp_peak_bytes, read_bytes and write_bytes.  These are logical bytes of the
// reference's schedule; the device kernels fold the transposes into their
// operand gathers.
#ifndef QLTEN_AMD_TENSOR_MANIPULATION_TENSOR_OP_COST_H
#define QLTEN_AMD_TENSOR_MANIPULATION_TENSOR_OP_COST_H

#include <map>
#include <stdexcept>
#include <type_traits>
#include <unordered_set>
#include <utility>
#include <vector>

#include "qlten/tensor_manipulation/ten_ctrct.h"

namespace qlten {

// One stored block of a block-sparse tensor, metadata only.
struct BlockSparseTensorBlock {
  size_t block_index = 0;
  CoorsT block_coords;
  ShapeT shape;
  size_t data_offset = 0;
  size_t element_count = 0;
};

inline bool operator==(const BlockSparseTensorBlock &x, const BlockSparseTensorBlock &y) {
  return x.block_index == y.block_index && x.block_coords == y.block_coords && x.shape == y.shape &&
         x.data_offset == y.data_offset && x.element_count == y.element_count;
}

// Topology of a block-sparse tensor (indexes, blocks, offsets) without data.
template <typename QNT>
struct BlockSparseTensorLayout {
  bool is_default = true;
  IndexVec<QNT> indexes;
  std::map<size_t, BlockSparseTensorBlock> blocks;
  size_t scalar_raw_elem_count = 0;

  bool IsDefault() const { return is_default; }
  bool IsScalar() const { return !is_default && indexes.empty(); }
  size_t Rank() const { return indexes.size(); }
  size_t BlockCount() const { return blocks.size(); }
  size_t RawElementCount() const {
    if (IsScalar()) return scalar_raw_elem_count == 0 ? 1 : scalar_raw_elem_count;
    size_t n = 0;
    for (const auto &kv : blocks) n += kv.second.element_count;
    return n;
  }
};

template <typename QNT>
using TensorLayout = BlockSparseTensorLayout<QNT>;

template <typename QNT>
bool operator==(const BlockSparseTensorLayout<QNT> &x, const BlockSparseTensorLayout<QNT> &y) {
  return x.is_default == y.is_default && x.indexes == y.indexes && x.blocks == y.blocks &&
         x.scalar_raw_elem_count == y.scalar_raw_elem_count;
}

// Estimated work of one tensor operation (logical bytes of the schedule, not
// measured memory traffic; temp_peak_bytes is the peak temporary workspace).
template <typename QNT>
struct TensorOpCost {
  double flops = 0.0;
  size_t gemm_count = 0;
  size_t executable_block_pair_count = 0;
  size_t candidate_block_pair_count = 0;
  size_t output_block_count = 0;
  size_t output_raw_elem_count = 0;
  size_t read_bytes = 0;
  size_t write_bytes = 0;
  size_t temp_peak_bytes = 0;
  BlockSparseTensorLayout<QNT> output_layout;
};

template <typename ElemT, typename QNT>
BlockSparseTensorLayout<QNT> MakeBlockSparseTensorLayout(const QLTensor<ElemT, QNT> &t) {
  BlockSparseTensorLayout<QNT> lay;
  lay.is_default = t.IsDefault();
  if (lay.is_default) return lay;
  lay.indexes = t.GetIndexes();
  if (t.IsScalar()) {
    lay.scalar_raw_elem_count = 1;
    return lay;
  }
  for (const auto &kv : t.GetBlkSparDataTen().GetBlkIdxDataBlkMap()) {
    BlockSparseTensorBlock b;
    b.block_index = kv.first;
    b.block_coords = kv.second.blk_coors;
    b.shape = kv.second.shape;
    b.data_offset = kv.second.data_offset;
    b.element_count = kv.second.size;
    lay.blocks.emplace(kv.first, std::move(b));
  }
  return lay;
}

// A tensor with the layout's indexes and blocks and no raw data, so that the
// estimate of one contraction can feed the next (tensor_op_cost.h:159-182).
template <typename ElemT, typename QNT>
QLTensor<ElemT, QNT> MakeTensorShellFromLayout(const BlockSparseTensorLayout<QNT> &layout) {
  if (layout.IsDefault()) return QLTensor<ElemT, QNT>();
  QLTensor<ElemT, QNT> shell(layout.indexes);
  if (layout.IsScalar() || layout.blocks.empty()) return shell;
  std::vector<CoorsT> coords;
  coords.reserve(layout.blocks.size());
  for (const auto &kv : layout.blocks) coords.push_back(kv.second.block_coords);
  shell.GetBlkSparDataTen().DataBlksInsert(coords);
  return shell;
}

namespace detail {

inline bool IsIdentityOrder(const std::vector<size_t> &first, const std::vector<size_t> &second) {
  size_t want = 0;
  for (const auto *v : {&first, &second})
    for (size_t ax : *v)
      if (ax != want++) return false;
  return true;
}

template <typename ElemT, typename QNT>
void CheckCostInputs(const QLTensor<ElemT, QNT> &a, const QLTensor<ElemT, QNT> &b,
                     const std::vector<std::vector<size_t>> &axes_set) {
  if (a.IsDefault() || b.IsDefault()) throw std::invalid_argument("EstimateContractCost: default input tensor");
  if (a.IsScalar() || b.IsScalar()) throw std::invalid_argument("EstimateContractCost: scalar input tensor");
  if (axes_set.size() != 2 || axes_set[0].size() != axes_set[1].size())
    throw std::invalid_argument("EstimateContractCost: axes_set needs two lists of equal length");
  std::vector<char> used_a(a.Rank(), 0), used_b(b.Rank(), 0);
  for (size_t i = 0; i < axes_set[0].size(); ++i) {
    const size_t x = axes_set[0][i], y = axes_set[1][i];
    if (x >= a.Rank() || y >= b.Rank()) throw std::invalid_argument("EstimateContractCost: axis out of range");
    if (used_a[x] || used_b[y]) throw std::invalid_argument("EstimateContractCost: axis contracted twice");
    used_a[x] = used_b[y] = 1;
    if (a.GetIndex(x) != InverseIndex(b.GetIndex(y)))
      throw std::invalid_argument("EstimateContractCost: index mismatch at contracted axis " + std::to_string(i));
  }
}

// Elements of the distinct blocks of `t` named by `blk_of(task)` over the tasks.
template <typename ElemT, typename QNT, typename F>
size_t TouchedElements(const QLTensor<ElemT, QNT> &t, const std::vector<qlt_task> &tasks, F blk_of) {
  const auto &blocks = t.GetBlkSparDataTen().GetBlkIdxDataBlkMap();
  std::unordered_set<size_t> seen;
  size_t n = 0;
  for (const auto &tk : tasks) {
    const size_t idx = static_cast<size_t>(blk_of(tk));
    if (seen.insert(idx).second) n += blocks.at(idx).size;
  }
  return n;
}

}  // namespace detail

// EstimateContractCost(A, B, {{a_axes}, {b_axes}}) -- tensor_op_cost.h:434-496.
template <typename ElemT, typename QNT>
TensorOpCost<QNT> EstimateContractCost(const QLTensor<ElemT, QNT> &a, const QLTensor<ElemT, QNT> &b,
                                       const std::vector<std::vector<size_t>> &axes_set) {
  detail::CheckCostInputs(a, b, axes_set);
  const auto &a_ax = axes_set[0], &b_ax = axes_set[1];
  const auto a_saved = detail::SavedAxes(a.Rank(), a_ax);
  const auto b_saved = detail::SavedAxes(b.Rank(), b_ax);
  const auto plan = detail::GetPlan(a.GetBlkSparDataTen(), b.GetBlkSparDataTen(), a_ax, b_ax, a_saved, b_saved);
  std::vector<qlt_task> tasks(plan->n_tasks);
  if (!tasks.empty()) QltCheck(qlt_plan_get_tasks(plan->plan, tasks.data()));

  TensorOpCost<QNT> cost;
  const size_t esz = sizeof(ElemT);
  const double per_mkn = std::is_same<ElemT, QLTEN_Complex>::value ? 8.0 : 2.0;
  cost.candidate_block_pair_count =
      a.GetBlkSparDataTen().GetBlkIdxDataBlkMap().size() * b.GetBlkSparDataTen().GetBlkIdxDataBlkMap().size();
  cost.executable_block_pair_count = tasks.size();
  for (const auto &tk : tasks) {
    const size_t m = static_cast<size_t>(tk.m), k = static_cast<size_t>(tk.k), n = static_cast<size_t>(tk.n);
    cost.flops += per_mkn * static_cast<double>(m) * static_cast<double>(k) * static_cast<double>(n);
    cost.gemm_count += 1;
    cost.read_bytes += (m * k + k * n + (tk.beta != 0.0 ? m * n : 0)) * esz;
    cost.write_bytes += m * n * esz;
  }

  // output layout: A's saved indexes then B's, blocks as the plan places them
  auto &lay = cost.output_layout;
  lay.is_default = false;
  for (size_t ax : a_saved) lay.indexes.push_back(a.GetIndex(ax));
  for (size_t ax : b_saved) lay.indexes.push_back(b.GetIndex(ax));
  if (lay.indexes.empty()) {
    lay.scalar_raw_elem_count = 1;
  } else {
    for (size_t i = 0; i < plan->c_idx.size(); ++i) {
      BlockSparseTensorBlock blk;
      blk.block_index = plan->c_idx[i];
      blk.block_coords = plan->c_coords[i];
      blk.shape = plan->c_dims[i];
      blk.data_offset = plan->c_offsets[i];
      blk.element_count = 1;
      for (size_t d : blk.shape) blk.element_count *= d;
      lay.blocks.emplace(blk.block_index, std::move(blk));
    }
  }
  cost.output_block_count = lay.BlockCount();
  cost.output_raw_elem_count = lay.RawElementCount();

  size_t temp = 0;
  if (!detail::IsIdentityOrder(a_saved, a_ax))
    temp += detail::TouchedElements(a, tasks, [](const qlt_task &t) { return t.a_blk_idx; }) * esz;
  if (!detail::IsIdentityOrder(b_ax, b_saved))
    temp += detail::TouchedElements(b, tasks, [](const qlt_task &t) { return t.b_blk_idx; }) * esz;
  cost.temp_peak_bytes = temp;
  cost.read_bytes += temp;
  cost.write_bytes += temp;
  return cost;
}

// Pair form: {(a_axis, b_axis), ...} -- tensor_op_cost.h:498-510.
template <typename ElemT, typename QNT>
TensorOpCost<QNT> EstimateContractCost(const QLTensor<ElemT, QNT> &a, const QLTensor<ElemT, QNT> &b,
                                       const std::vector<std::pair<size_t, size_t>> &contract_axes) {
  std::vector<std::vector<size_t>> axes_set(2);
  for (const auto &p : contract_axes) {
    axes_set[0].push_back(p.first);
    axes_set[1].push_back(p.second);
  }
  return EstimateContractCost(a, b, axes_set);
}

}  // namespace qlten

#endif
