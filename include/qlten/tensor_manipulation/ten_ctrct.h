// qlten::Contract on the device.
//
// Same signatures and result index order as the reference
// (tensor_manipulation/ten_ctrct.h:145-362): C's indexes are A's saved
// indexes followed by B's saved indexes, saved axes ascending.  The plan
// (DataBlkGenForTenCtrct, data_blk_operations.h:411-578) is built by the
// library's device kernels and cached by the topologies of A and B and the
// axes, so a repeated contraction pattern (a Lanczos / DMRG sweep) reuses
// it; execution is one grouped launch (CtrctTwoBSDTAndAssignIn,
// global_operations.h:895-992, folded transposes).  Contract behaves as if
// synchronous: every later host read synchronises the library stream.
#ifndef QLTEN_AMD_TENSOR_MANIPULATION_TEN_CTRCT_H
#define QLTEN_AMD_TENSOR_MANIPULATION_TEN_CTRCT_H

#include <algorithm>
#include <cassert>
#include <list>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "qlten/qltensor/qltensor.h"
#include "qlten/tensor_manipulation/basic_operations.h"

namespace qlten {

namespace detail {

// A device plan plus the output block map it produces.
struct ContractPlan {
  qlt_plan *plan = nullptr;
  std::vector<size_t> c_idx;
  std::vector<CoorsT> c_coords;
  std::vector<ShapeT> c_dims;
  std::vector<size_t> c_offsets;
  size_t c_raw_size = 0;
  size_t n_tasks = 0;
  bool c_scalar = false;
  ContractPlan() = default;
  ContractPlan(const ContractPlan &) = delete;
  ContractPlan &operator=(const ContractPlan &) = delete;
  ~ContractPlan() {
    if (plan) qlt_plan_destroy(plan);
  }
};

// LRU cache of plans keyed by (A topology, B topology, axes).  The key is
// compared in full, so a hit always has identical block layouts.
class PlanCache {
 public:
  static PlanCache &Instance() {
    static PlanCache c;
    return c;
  }
  size_t capacity = 256;
  size_t hits = 0, misses = 0;

  std::shared_ptr<ContractPlan> Find(const std::vector<int64_t> &key) {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = map_.find(Hash(key));
    if (it == map_.end()) return nullptr;
    for (auto li = it->second.begin(); li != it->second.end(); ++li) {
      if ((*li)->key == key) {
        ++hits;
        (*li)->stamp = ++clock_;
        return (*li)->plan;
      }
    }
    return nullptr;
  }
  void Insert(const std::vector<int64_t> &key, std::shared_ptr<ContractPlan> plan) {
    std::lock_guard<std::mutex> lk(mu_);
    ++misses;
    auto e = std::make_shared<Entry>();
    e->key = key;
    e->plan = std::move(plan);
    e->stamp = ++clock_;
    map_[Hash(key)].push_back(e);
    ++size_;
    while (size_ > capacity) EvictOldest();
  }
  void Clear() {
    std::lock_guard<std::mutex> lk(mu_);
    map_.clear();
    size_ = 0;
  }

 private:
  struct Entry {
    std::vector<int64_t> key;
    std::shared_ptr<ContractPlan> plan;
    uint64_t stamp = 0;
  };
  static size_t Hash(const std::vector<int64_t> &k) {
    uint64_t h = 1469598103934665603ull;
    for (int64_t v : k) {
      h ^= static_cast<uint64_t>(v);
      h *= 1099511628211ull;
    }
    return static_cast<size_t>(h);
  }
  void EvictOldest() {
    uint64_t best = ~0ull;
    size_t best_h = 0;
    for (auto &kv : map_)
      for (auto &e : kv.second)
        if (e->stamp < best) {
          best = e->stamp;
          best_h = kv.first;
        }
    auto &lst = map_[best_h];
    lst.remove_if([&](const std::shared_ptr<Entry> &e) { return e->stamp == best; });
    if (lst.empty()) map_.erase(best_h);
    --size_;
  }
  std::mutex mu_;
  std::unordered_map<size_t, std::list<std::shared_ptr<Entry>>> map_;
  size_t size_ = 0;
  uint64_t clock_ = 0;
};

inline std::vector<size_t> SavedAxes(size_t rank, const std::vector<size_t> &ctrct) {
  std::vector<size_t> s;
  for (size_t i = 0; i < rank; ++i)
    if (std::find(ctrct.begin(), ctrct.end(), i) == ctrct.end()) s.push_back(i);
  return s;
}

template <typename ElemT, typename QNT>
std::shared_ptr<ContractPlan> GetPlan(const BlockSparseDataTensor<ElemT, QNT> &a,
                                      const BlockSparseDataTensor<ElemT, QNT> &b,
                                      const std::vector<size_t> &a_ctrct, const std::vector<size_t> &b_ctrct,
                                      const std::vector<size_t> &a_saved, const std::vector<size_t> &b_saved) {
  std::vector<int64_t> key = a.TopologySignature();
  key.push_back(-1);
  const auto kb = b.TopologySignature();
  key.insert(key.end(), kb.begin(), kb.end());
  for (const auto *v : {&a_ctrct, &b_ctrct, &a_saved, &b_saved}) {
    key.push_back(-2);
    for (size_t x : *v) key.push_back(static_cast<int64_t>(x));
  }
  auto &cache = PlanCache::Instance();
  if (auto hit = cache.Find(key)) return hit;

  auto da = a.MakeDesc();
  auto db = b.MakeDesc();
  std::vector<int32_t> ac(a_ctrct.begin(), a_ctrct.end()), bc(b_ctrct.begin(), b_ctrct.end());
  std::vector<int32_t> as(a_saved.begin(), a_saved.end()), bs(b_saved.begin(), b_saved.end());
  auto p = std::make_shared<ContractPlan>();
  QltCheck(qlt_plan_build(&da->d, &db->d, static_cast<int32_t>(ac.size()), ac.data(), bc.data(),
                          static_cast<int32_t>(as.size()), as.data(), static_cast<int32_t>(bs.size()), bs.data(),
                          hp_numeric::LibStream(), &p->plan));
  qlt_plan_info info;
  QltCheck(qlt_plan_get_info(p->plan, &info));
  p->c_raw_size = static_cast<size_t>(info.c_raw_size);
  p->n_tasks = static_cast<size_t>(info.n_tasks);
  p->c_scalar = info.c_is_scalar != 0;
  const size_t n = static_cast<size_t>(info.n_c_blks), r = static_cast<size_t>(info.c_rank);
  std::vector<int64_t> idx(n), coords(n * r), dims(n * r), offs(n);
  if (n) QltCheck(qlt_plan_get_c_blocks(p->plan, idx.data(), coords.data(), dims.data(), offs.data()));
  for (size_t i = 0; i < n; ++i) {
    p->c_idx.push_back(static_cast<size_t>(idx[i]));
    p->c_coords.emplace_back(coords.begin() + i * r, coords.begin() + (i + 1) * r);
    p->c_dims.emplace_back(dims.begin() + i * r, dims.begin() + (i + 1) * r);
    p->c_offsets.push_back(static_cast<size_t>(offs[i]));
  }
  cache.Insert(key, p);
  return p;
}

// Contract with explicit contracted and saved axes (shared by the general and
// the contiguous-axes paths).  pc receives the result (its previous content
// is discarded).
template <typename ElemT, typename QNT>
void ContractWithAxes(const QLTensor<ElemT, QNT> &a, const QLTensor<ElemT, QNT> &b,
                      const std::vector<size_t> &a_ctrct, const std::vector<size_t> &b_ctrct,
                      const std::vector<size_t> &a_saved, const std::vector<size_t> &b_saved,
                      QLTensor<ElemT, QNT> *pc) {
  if (a_ctrct.size() != b_ctrct.size()) throw std::invalid_argument("Contract: axis lists differ in length");
  for (size_t i = 0; i < a_ctrct.size(); ++i)
    if (a.GetIndexes()[a_ctrct[i]] != InverseIndex(b.GetIndexes()[b_ctrct[i]]))
      throw std::invalid_argument("Contract: index mismatch at contracted axis " + std::to_string(i));
  IndexVec<QNT> c_idxs;
  for (size_t ax : a_saved) c_idxs.push_back(a.GetIndexes()[ax]);
  for (size_t ax : b_saved) c_idxs.push_back(b.GetIndexes()[ax]);
  *pc = QLTensor<ElemT, QNT>(std::move(c_idxs));
  if (a.IsDefault() || b.IsDefault()) return;
  auto plan = GetPlan(a.GetBlkSparDataTen(), b.GetBlkSparDataTen(), a_ctrct, b_ctrct, a_saved, b_saved);
  auto &c = pc->GetBlkSparDataTen();
  c.SetBlocks(plan->c_idx, plan->c_coords, plan->c_dims, plan->c_offsets, plan->c_raw_size);
  if (plan->n_tasks == 0) return;   // no matching blocks: an empty result (global_operations.h:904)
  c.Allocate();
  const double one[2] = {1.0, 0.0}, zero[2] = {0.0, 0.0};
  QltCheck(qlt_contract_execute(plan->plan, QLTensor<ElemT, QNT>::DType(), hp_numeric::ContractModeFlag(),
                                a.GetBlkSparDataTen().GetActualRawDataPtr(),
                                b.GetBlkSparDataTen().GetActualRawDataPtr(), c.GetActualRawDataPtr(), one, zero,
                                hp_numeric::LibStream()));
}

}  // namespace detail

// Contract(A, B, {{a_axes}, {b_axes}}, C) -- ten_ctrct.h:276-290.
template <typename TenElemT, typename QNT>
void Contract(const QLTensor<TenElemT, QNT> *pa, const QLTensor<TenElemT, QNT> *pb,
              const std::vector<std::vector<size_t>> &axes_set, QLTensor<TenElemT, QNT> *pc) {
  assert(pc->IsDefault());   // only an empty tensor takes the result (ten_ctrct.h:184)
  const auto a_saved = detail::SavedAxes(pa->Rank(), axes_set[0]);
  const auto b_saved = detail::SavedAxes(pb->Rank(), axes_set[1]);
  detail::ContractWithAxes(*pa, *pb, axes_set[0], axes_set[1], a_saved, b_saved, pc);
}

// Mixed real / complex overloads (ten_ctrct.h:292-350): promote the real side.
template <typename QNT>
void Contract(const QLTensor<QLTEN_Double, QNT> *pa, const QLTensor<QLTEN_Complex, QNT> *pb,
              const std::vector<std::vector<size_t>> &axes_set, QLTensor<QLTEN_Complex, QNT> *pc) {
  auto ca = ToComplex(*pa);
  Contract(&ca, pb, axes_set, pc);
}

template <typename QNT>
void Contract(const QLTensor<QLTEN_Complex, QNT> *pa, const QLTensor<QLTEN_Double, QNT> *pb,
              const std::vector<std::vector<size_t>> &axes_set, QLTensor<QLTEN_Complex, QNT> *pc) {
  auto cb = ToComplex(*pb);
  Contract(pa, &cb, axes_set, pc);
}

template <typename QNT>
void Contract(const QLTensor<QLTEN_Float, QNT> *pa, const QLTensor<QLTEN_ComplexFloat, QNT> *pb,
              const std::vector<std::vector<size_t>> &axes_set, QLTensor<QLTEN_ComplexFloat, QNT> *pc) {
  auto ca = ToComplex(*pa);
  Contract(&ca, pb, axes_set, pc);
}

template <typename QNT>
void Contract(const QLTensor<QLTEN_ComplexFloat, QNT> *pa, const QLTensor<QLTEN_Float, QNT> *pb,
              const std::vector<std::vector<size_t>> &axes_set, QLTensor<QLTEN_ComplexFloat, QNT> *pc) {
  auto cb = ToComplex(*pb);
  Contract(pa, &cb, axes_set, pc);
}

// Contract(A, axes_a, B, axes_b, C) -- ten_ctrct.h:352-362.
template <typename TenElemT1, typename TenElemT2, typename TenElemT3, typename QNT>
void Contract(const QLTensor<TenElemT1, QNT> *pa, const std::vector<size_t> &axes_a,
              const QLTensor<TenElemT2, QNT> *pb, const std::vector<size_t> &axes_b,
              QLTensor<TenElemT3, QNT> *pc) {
  std::vector<std::vector<size_t>> axes_set = {axes_a, axes_b};
  Contract(pa, pb, axes_set, pc);
}

}  // namespace qlten

#endif
