// QLTensor<ElemT, QNT> and its BlockSparseDataTensor.
//
// Same interface and layout semantics as the reference's QLTensor
// (qltensor/qltensor.h:61-660, qltensor/qltensor_impl.h) and
// BlockSparseDataTensor (qltensor/blk_spar_data_ten/blk_spar_data_ten.h:53-595):
// a block map ordered by block index (row-major over the sector counts of
// the indexes), every block row-major over its degeneracy shape, blocks
// contiguous in one raw buffer in ascending block index.  The raw buffer
// lives in device memory (as in the reference's GPU build,
// blk_spar_data_ten.h:481-487); element access copies through the host.
#ifndef QLTEN_AMD_QLTENSOR_QLTENSOR_H
#define QLTEN_AMD_QLTENSOR_QLTENSOR_H

#include <cassert>
#include <cmath>
#include <cstdint>
#include <ios>
#include <istream>
#include <ostream>
#include <complex>
#include <map>
#include <memory>
#include <numeric>
#include <stdexcept>
#include <unordered_map>
#include <utility>
#include <vector>

#include "qlten/framework/backend.h"
#include "qlten/framework/value_t.h"
#include "qlten/qltensor/index.h"

namespace qlten {

// Global generator state of Random (the reference keeps a global
// std::mt19937_64 seeded 0, utility/random.h:20-45).  Each Random call draws
// the next raw-size doubles of a SplitMix64 counter stream.
struct RandomState {
  uint64_t seed = 0;
  int64_t counter = 0;
};
inline RandomState &GlobalRandomState() {
  static RandomState s;
  return s;
}
inline void SetRandomSeed(const uint64_t seed) {
  GlobalRandomState().seed = seed;
  GlobalRandomState().counter = 0;
}

template <typename QNT>
struct DataBlk {
  DataBlk() = default;
  DataBlk(const CoorsT &coors, const ShapeT &shp) : blk_coors(coors), shape(shp) {
    size = std::accumulate(shape.begin(), shape.end(), size_t(1), std::multiplies<size_t>());
  }
  CoorsT blk_coors;
  ShapeT shape;
  size_t size = 0;
  size_t data_offset = 0;
};

template <typename ElemT, typename QNT>
class BlockSparseDataTensor {
 public:
  using BlkIdxDataBlkMap = std::map<size_t, DataBlk<QNT>>;

  explicit BlockSparseDataTensor(const IndexVec<QNT> *pidx) : pqlten_indexes(pidx) {
    ten_rank = pidx->size();
    blk_shape = CalcQNSctNumOfIdxs(*pidx);
    blk_multi_dim_offsets_.assign(ten_rank, 1);
    for (int i = static_cast<int>(ten_rank) - 2; i >= 0; --i)
      blk_multi_dim_offsets_[i] = blk_multi_dim_offsets_[i + 1] * blk_shape[i + 1];
  }

  // deep copy (device data included)
  BlockSparseDataTensor(const BlockSparseDataTensor &o)
      : ten_rank(o.ten_rank), blk_shape(o.blk_shape), pqlten_indexes(o.pqlten_indexes),
        blk_multi_dim_offsets_(o.blk_multi_dim_offsets_), blk_idx_data_blk_map_(o.blk_idx_data_blk_map_),
        raw_data_size_(o.raw_data_size_) {
    if (o.raw_.bytes()) {
      raw_.Allocate(o.raw_.bytes());
      hp_numeric::SyncLibStream();
      QltCheck(qlt_memcpy_d2d(raw_.get(), o.raw_.get(), o.raw_.bytes()));
    }
  }
  BlockSparseDataTensor &operator=(const BlockSparseDataTensor &) = delete;

  size_t ten_rank = 0;
  ShapeT blk_shape;
  const IndexVec<QNT> *pqlten_indexes = nullptr;

  const BlkIdxDataBlkMap &GetBlkIdxDataBlkMap() const { return blk_idx_data_blk_map_; }
  size_t BlkCoorsToBlkIdx(const CoorsT &coors) const {
    size_t idx = 0;
    for (size_t i = 0; i < coors.size(); ++i) idx += coors[i] * blk_multi_dim_offsets_[i];
    return idx;
  }
  bool IsScalar() const { return ten_rank == 0; }
  size_t GetActualRawDataSize() const { return raw_.bytes() / sizeof(ElemT); }
  size_t RawDataSize() const { return raw_data_size_; }
  const ElemT *GetActualRawDataPtr() const { return static_cast<const ElemT *>(raw_.get()); }
  ElemT *GetActualRawDataPtr() { return static_cast<ElemT *>(raw_.get()); }

  void Clear() {
    blk_idx_data_blk_map_.clear();
    raw_.Release();
    raw_data_size_ = 0;
  }

  // Insert blocks (metadata only) and lay out offsets by prefix sum in
  // ascending block index (DataBlksInsert, data_blk_operations.h:182-208).
  void DataBlksInsert(const std::vector<CoorsT> &coors_list) {
    for (const auto &c : coors_list) {
      ShapeT shp(ten_rank);
      for (size_t i = 0; i < ten_rank; ++i) shp[i] = (*pqlten_indexes)[i].GetQNSct(c[i]).dim();
      blk_idx_data_blk_map_[BlkCoorsToBlkIdx(c)] = DataBlk<QNT>(c, shp);
    }
    RefreshOffsets();
  }

  // Install a block layout produced by a contraction plan (already in
  // ascending block index with prefix-sum offsets).
  void SetBlocks(const std::vector<size_t> &idx, const std::vector<CoorsT> &coors, const std::vector<ShapeT> &dims,
                 const std::vector<size_t> &offsets, size_t raw_size) {
    blk_idx_data_blk_map_.clear();
    for (size_t i = 0; i < idx.size(); ++i) {
      DataBlk<QNT> b(coors[i], dims[i]);
      b.data_offset = offsets[i];
      blk_idx_data_blk_map_.emplace(idx[i], std::move(b));
    }
    raw_data_size_ = raw_size;
  }

  void RefreshOffsets() {
    raw_data_size_ = 0;
    for (auto &kv : blk_idx_data_blk_map_) {
      kv.second.data_offset = raw_data_size_;
      raw_data_size_ += kv.second.size;
    }
    if (IsScalar()) raw_data_size_ = 1;
  }

  // Allocate the raw buffer (not initialised, as the reference's Allocate()).
  void Allocate(bool zero = false) {
    raw_.Allocate(raw_data_size_ * sizeof(ElemT));
    if (zero && raw_data_size_) QltCheck(qlt_memset_zero(raw_.get(), raw_.bytes(), hp_numeric::LibStream()));
  }

  ElemT ElemGet(const CoorsT &blk_coors, const CoorsT &data_coors) const {
    if (IsScalar()) {
      if (raw_.bytes() < sizeof(ElemT)) return ElemT(0);
      return ReadElem(0);
    }
    auto it = blk_idx_data_blk_map_.find(BlkCoorsToBlkIdx(blk_coors));
    if (it == blk_idx_data_blk_map_.end() || raw_.bytes() == 0) return ElemT(0);
    return ReadElem(it->second.data_offset + InBlkOffset(it->second, data_coors));
  }

  void ElemSet(const CoorsT &blk_coors, const CoorsT &data_coors, const ElemT v) {
    if (IsScalar()) {
      if (raw_data_size_ == 0) raw_data_size_ = 1;
      if (raw_.bytes() == 0) Allocate(true);
      WriteElem(0, v);
      return;
    }
    const size_t idx = BlkCoorsToBlkIdx(blk_coors);
    if (blk_idx_data_blk_map_.find(idx) == blk_idx_data_blk_map_.end()) InsertZeroBlock(blk_coors);
    const auto &blk = blk_idx_data_blk_map_.at(idx);
    WriteElem(blk.data_offset + InBlkOffset(blk, data_coors), v);
  }

  // Host copy of the raw data (test / IO helper).
  std::vector<ElemT> RawDataToHost() const {
    std::vector<ElemT> out(raw_.bytes() / sizeof(ElemT));
    if (!out.empty()) {
      hp_numeric::SyncLibStream();
      QltCheck(qlt_memcpy_d2h(out.data(), raw_.get(), raw_.bytes()));
    }
    return out;
  }
  void RawDataFromHost(const std::vector<ElemT> &v) {
    if (v.size() != raw_data_size_) throw std::invalid_argument("RawDataFromHost: size mismatch");
    if (raw_.bytes() != v.size() * sizeof(ElemT)) Allocate();
    if (!v.empty()) {
      hp_numeric::SyncLibStream();
      QltCheck(qlt_memcpy_h2d(raw_.get(), v.data(), raw_.bytes()));
    }
  }

  // ---- on-disk format (BlockSparseDataTensor::StreamRead/StreamWrite,
  // blk_spar_data_ten.h:790-830; RawDataRead_/RawDataWrite_,
  // raw_data_operations.h:582-633) ----
  // Text: the block count, then every block's coordinates one per line in
  // ascending block index; then one line break and the raw data as binary
  // ElemT (host byte order), then a line break.  A scalar with no data writes
  // one zero element.
  void StreamWrite(std::ostream &os) const {
    os << blk_idx_data_blk_map_.size() << "\n";
    for (const auto &kv : blk_idx_data_blk_map_)
      for (size_t c : kv.second.blk_coors) os << c << "\n";
    if (IsScalar() && GetActualRawDataSize() == 0) {
      const ElemT zero{};
      os.write(reinterpret_cast<const char *>(&zero), sizeof(ElemT));
      os << "\n";
    } else {
      const std::vector<ElemT> h = RawDataToHost();
      os.write(reinterpret_cast<const char *>(h.data()), static_cast<std::streamsize>(h.size() * sizeof(ElemT)));
      os << std::endl;
    }
  }
  void StreamRead(std::istream &is) {
    size_t n = 0;
    is >> n;
    std::vector<CoorsT> coors(n, CoorsT(ten_rank));
    for (auto &c : coors)
      for (size_t j = 0; j < ten_rank; ++j) is >> c[j];
    blk_idx_data_blk_map_.clear();
    DataBlksInsert(coors);
    Allocate();
    is.seekg(1, std::ios::cur);   // the line break before the binary data
    std::vector<ElemT> h(raw_data_size_);
    is.read(reinterpret_cast<char *>(h.data()), static_cast<std::streamsize>(h.size() * sizeof(ElemT)));
    if (!is) throw std::ios_base::failure("BlockSparseDataTensor::StreamRead: truncated raw data");
    RawDataFromHost(h);
  }

  // Descriptor arrays handed to qlt_plan_build.
  struct Desc {
    std::vector<int64_t> shape, coords, dims, offsets;
    qlt_bsdt_desc d{};
  };
  std::unique_ptr<Desc> MakeDesc() const {
    auto out = std::make_unique<Desc>();
    for (size_t s : blk_shape) out->shape.push_back(static_cast<int64_t>(s));
    for (const auto &kv : blk_idx_data_blk_map_) {
      for (size_t c : kv.second.blk_coors) out->coords.push_back(static_cast<int64_t>(c));
      for (size_t d : kv.second.shape) out->dims.push_back(static_cast<int64_t>(d));
      out->offsets.push_back(static_cast<int64_t>(kv.second.data_offset));
    }
    out->d.rank = static_cast<int32_t>(ten_rank);
    out->d.n_blks = static_cast<int64_t>(blk_idx_data_blk_map_.size());
    out->d.blk_shape = out->shape.data();
    out->d.blk_coords = out->coords.data();
    out->d.blk_dims = out->dims.data();
    out->d.blk_offsets = out->offsets.data();
    return out;
  }

  // Topology signature: sector counts plus every block's index, shape and
  // offset.  Two tensors with equal signatures produce the same plan.
  std::vector<int64_t> TopologySignature() const {
    std::vector<int64_t> sig;
    sig.push_back(static_cast<int64_t>(ten_rank));
    for (size_t s : blk_shape) sig.push_back(static_cast<int64_t>(s));
    for (const auto &kv : blk_idx_data_blk_map_) {
      sig.push_back(static_cast<int64_t>(kv.first));
      for (size_t d : kv.second.shape) sig.push_back(static_cast<int64_t>(d));
      sig.push_back(static_cast<int64_t>(kv.second.data_offset));
    }
    return sig;
  }

 private:
  size_t InBlkOffset(const DataBlk<QNT> &blk, const CoorsT &data_coors) const {
    size_t off = 0;
    for (size_t i = 0; i < data_coors.size(); ++i) off = off * blk.shape[i] + data_coors[i];
    return off;
  }
  ElemT ReadElem(size_t off) const {
    ElemT v;
    hp_numeric::SyncLibStream();
    QltCheck(qlt_memcpy_d2h(&v, static_cast<const ElemT *>(raw_.get()) + off, sizeof(ElemT)));
    return v;
  }
  void WriteElem(size_t off, const ElemT v) {
    hp_numeric::SyncLibStream();
    QltCheck(qlt_memcpy_h2d(static_cast<ElemT *>(raw_.get()) + off, &v, sizeof(ElemT)));
  }
  void InsertZeroBlock(const CoorsT &coors) {
    std::vector<ElemT> old = RawDataToHost();
    std::map<size_t, size_t> old_off;
    for (const auto &kv : blk_idx_data_blk_map_) old_off[kv.first] = kv.second.data_offset;
    DataBlksInsert({coors});
    std::vector<ElemT> fresh(raw_data_size_, ElemT(0));
    for (const auto &kv : blk_idx_data_blk_map_) {
      auto it = old_off.find(kv.first);
      if (it == old_off.end() || old.empty()) continue;
      std::copy(old.begin() + it->second, old.begin() + it->second + kv.second.size,
                fresh.begin() + kv.second.data_offset);
    }
    Allocate();
    RawDataFromHost(fresh);
  }

  std::vector<size_t> blk_multi_dim_offsets_;
  BlkIdxDataBlkMap blk_idx_data_blk_map_;
  DeviceArray raw_;
  size_t raw_data_size_ = 0;
};

template <typename ElemT, typename QNT>
class QLTensor {
 public:
  using BSDT = BlockSparseDataTensor<ElemT, QNT>;

  QLTensor() = default;
  explicit QLTensor(const IndexVec<QNT> &indexes) : indexes_(indexes) { Init(); }
  explicit QLTensor(IndexVec<QNT> &&indexes) : indexes_(std::move(indexes)) { Init(); }

  QLTensor(const QLTensor &o) : rank_(o.rank_), shape_(o.shape_), size_(o.size_), indexes_(o.indexes_) {
    if (o.bsdt_) {
      bsdt_ = std::make_unique<BSDT>(*o.bsdt_);
      bsdt_->pqlten_indexes = &indexes_;
    }
  }
  QLTensor &operator=(const QLTensor &o) {
    if (this == &o) return *this;
    rank_ = o.rank_;
    shape_ = o.shape_;
    size_ = o.size_;
    indexes_ = o.indexes_;
    bsdt_.reset();
    if (o.bsdt_) {
      bsdt_ = std::make_unique<BSDT>(*o.bsdt_);
      bsdt_->pqlten_indexes = &indexes_;
    }
    return *this;
  }
  QLTensor(QLTensor &&o) noexcept
      : rank_(o.rank_), shape_(std::move(o.shape_)), size_(o.size_), indexes_(std::move(o.indexes_)),
        bsdt_(std::move(o.bsdt_)) {
    if (bsdt_) bsdt_->pqlten_indexes = &indexes_;
    o.rank_ = 0;
    o.size_ = 0;
  }
  QLTensor &operator=(QLTensor &&o) noexcept {
    rank_ = o.rank_;
    shape_ = std::move(o.shape_);
    size_ = o.size_;
    indexes_ = std::move(o.indexes_);
    bsdt_ = std::move(o.bsdt_);
    if (bsdt_) bsdt_->pqlten_indexes = &indexes_;
    o.rank_ = 0;
    o.size_ = 0;
    return *this;
  }

  size_t Rank() const { return rank_; }
  const ShapeT &GetShape() const { return shape_; }
  size_t size() const { return size_; }
  const IndexVec<QNT> &GetIndexes() const { return indexes_; }
  const Index<QNT> &GetIndex(const size_t i) const { return indexes_[i]; }
  bool IsDefault() const { return size_ == 0; }
  bool IsScalar() const { return rank_ == 0 && size_ == 1; }
  // True once raw data is allocated (qltensor_impl.h:1071-1073); a tensor
  // with blocks but no raw data is a layout shell.
  bool HasActualData() const { return !IsDefault() && bsdt_ && bsdt_->GetActualRawDataSize() != 0; }
  BSDT &GetBlkSparDataTen() { return *bsdt_; }
  const BSDT &GetBlkSparDataTen() const { return *bsdt_; }

  // Fill every allowed block (CalcDiv == div) with uniform [0, 1) values
  // (QLTensor::Random, qltensor_impl.h:374-403; blocks in row-major order).
  void Random(const QNT &div) {
    AllowedBlocks(div);
    const int64_t n = static_cast<int64_t>(bsdt_->RawDataSize()) * (IsComplex<ElemT>() ? 2 : 1);
    auto &st = GlobalRandomState();
    if constexpr (IsSingle<ElemT>())
      QltCheck(qlt_fill_uniform_f32(bsdt_->GetActualRawDataPtr(), n, st.seed, st.counter, 0.0, 1.0,
                                    hp_numeric::LibStream()));
    else
      QltCheck(qlt_fill_uniform(bsdt_->GetActualRawDataPtr(), n, st.seed, st.counter, 0.0, 1.0,
                                hp_numeric::LibStream()));
    st.counter += n;
  }

  // Fill every allowed block with `value` (QLTensor::Fill, qltensor_impl.h:412-441).
  void Fill(const QNT &div, const ElemT &value) {
    AllowedBlocks(div);
    std::vector<ElemT> v(bsdt_->RawDataSize(), value);
    bsdt_->RawDataFromHost(v);
  }

  ElemT GetElem(const std::vector<size_t> &coors) const {
    assert(coors.size() == rank_);
    if (IsDefault()) return ElemT(0);
    auto bd = BlkAndDataCoors(coors);
    return bsdt_->ElemGet(bd.first, bd.second);
  }
  void SetElem(const std::vector<size_t> &coors, const ElemT v) {
    assert(coors.size() == rank_);
    auto bd = BlkAndDataCoors(coors);
    bsdt_->ElemSet(bd.first, bd.second, v);
  }

  QNT Div() const {
    if (IsScalar() || IsDefault() || bsdt_->GetBlkIdxDataBlkMap().empty()) return QNT::Zero();
    return CalcDiv(indexes_, bsdt_->GetBlkIdxDataBlkMap().begin()->second.blk_coors);
  }

  using RealT = typename RealTypeTrait<ElemT>::type;
  // sqrt(sum |x|^2) over the raw data (accumulated in FP64).
  RealT Get2Norm() const {
    if (IsDefault() || bsdt_->GetActualRawDataSize() == 0) return RealT(0);
    double out[2];
    const void *p = bsdt_->GetActualRawDataPtr();
    QltCheck(qlt_dot(DType(), static_cast<int64_t>(bsdt_->GetActualRawDataSize()), p, p, out,
                     hp_numeric::LibStream()));
    return static_cast<RealT>(std::sqrt(out[0]));
  }
  RealT Normalize() {
    const RealT norm = Get2Norm();
    if (norm > RealT(0)) *this *= ElemT(RealT(1) / norm);
    return norm;
  }

  QLTensor &operator*=(const ElemT s) {
    if (IsDefault() || bsdt_->GetActualRawDataSize() == 0) return *this;
    const double a[2] = {std::real(std::complex<double>(s)), std::imag(std::complex<double>(s))};
    const double zero[2] = {0.0, 0.0};
    void *p = bsdt_->GetActualRawDataPtr();
    QltCheck(qlt_axpby(DType(), static_cast<int64_t>(bsdt_->GetActualRawDataSize()), a, p, zero, p,
                       hp_numeric::LibStream()));
    return *this;
  }
  QLTensor operator*(const ElemT s) const {
    QLTensor r(*this);
    r *= s;
    return r;
  }
  QLTensor operator-() const { return (*this) * ElemT(-1.0); }

  // Block-union sum (LinearCombine semantics for two terms).
  QLTensor operator+(const QLTensor &rhs) const { return AxpbyUnion(ElemT(1.0), *this, ElemT(1.0), rhs); }
  QLTensor operator-(const QLTensor &rhs) const { return AxpbyUnion(ElemT(1.0), *this, ElemT(-1.0), rhs); }
  QLTensor &operator+=(const QLTensor &rhs) { return *this = *this + rhs; }

  static constexpr int32_t DType() {
    return IsSingle<ElemT>() ? (IsComplex<ElemT>() ? QLT_C64 : QLT_F32) : (IsComplex<ElemT>() ? QLT_C128 : QLT_F64);
  }

  // a * x + b * y over the union of the two block maps.
  static QLTensor AxpbyUnion(const ElemT a, const QLTensor &x, const ElemT b, const QLTensor &y) {
    if (x.IsDefault()) return y * b;
    if (y.IsDefault()) return x * a;
    if (x.indexes_ != y.indexes_) throw std::invalid_argument("QLTensor +/-: index mismatch");
    QLTensor out(x.indexes_);
    std::vector<CoorsT> coors;
    std::map<size_t, CoorsT> all;
    for (const auto &kv : x.bsdt_->GetBlkIdxDataBlkMap()) all[kv.first] = kv.second.blk_coors;
    for (const auto &kv : y.bsdt_->GetBlkIdxDataBlkMap()) all[kv.first] = kv.second.blk_coors;
    for (const auto &kv : all) coors.push_back(kv.second);
    if (out.IsScalar()) {
      out.bsdt_->RefreshOffsets();
    } else {
      out.bsdt_->DataBlksInsert(coors);
    }
    out.bsdt_->Allocate(true);
    AddScaledBlocks(out, a, x);
    AddScaledBlocks(out, b, y);
    return out;
  }

  // ---- on-disk format (QLTensor::StreamRead/StreamWrite, qltensor_impl.h:798-830) ----
  // The rank, every index, then the block-sparse data tensor.  Only a default
  // tensor reads.
  void StreamRead(std::istream &is) {
    if (!IsDefault()) throw std::invalid_argument("QLTensor::StreamRead: only a default tensor can read");
    size_t rank = 0;
    is >> rank;
    indexes_ = IndexVec<QNT>(rank);
    for (auto &ix : indexes_) ix.StreamRead(is);
    bsdt_.reset();
    Init();
    if (!bsdt_) throw std::ios_base::failure("QLTensor::StreamRead: the stream holds a tensor of size 0");
    bsdt_->StreamRead(is);
  }
  void StreamWrite(std::ostream &os) const {
    if (IsDefault()) throw std::invalid_argument("QLTensor::StreamWrite: default tensor");
    os << rank_ << "\n";
    for (const auto &ix : indexes_) {
      ix.StreamWrite(os);
      if (!os) throw std::ios_base::failure("Failed to write streamable object.");
    }
    bsdt_->StreamWrite(os);
    if (!os) throw std::ios_base::failure("Failed to write streamable object.");
  }

  // Equality as the reference's (qltensor_impl.h:353-366, blk_spar_data_ten.h
  // operator==): both default, or equal indexes, the same block indices and
  // every raw element within 1e-15 (kDoubleEpsilon, framework/consts.h:19).
  bool operator==(const QLTensor &rhs) const {
    if (IsDefault() || rhs.IsDefault()) return IsDefault() && rhs.IsDefault();
    if (indexes_ != rhs.indexes_) return false;
    const auto &lm = bsdt_->GetBlkIdxDataBlkMap(), &rm = rhs.bsdt_->GetBlkIdxDataBlkMap();
    if (!IsScalar()) {
      if (lm.size() != rm.size()) return false;
      for (auto l = lm.begin(), r = rm.begin(); l != lm.end(); ++l, ++r)
        if (l->first != r->first) return false;
    }
    const std::vector<ElemT> a = bsdt_->RawDataToHost(), b = rhs.bsdt_->RawDataToHost();
    if (IsScalar()) {
      const ElemT x = a.empty() ? ElemT{} : a[0], y = b.empty() ? ElemT{} : b[0];
      return std::abs(x - y) < 1.0e-15;
    }
    if (a.size() != b.size()) return false;
    for (size_t i = 0; i < a.size(); ++i)
      if (!(std::abs(a[i] - b[i]) < 1.0e-15)) return false;
    return true;
  }
  bool operator!=(const QLTensor &rhs) const { return !(*this == rhs); }

 private:
  void Init() {
    rank_ = indexes_.size();
    shape_.clear();
    size_ = 1;
    for (const auto &ix : indexes_) {
      shape_.push_back(ix.dim());
      size_ *= ix.dim();
    }
    if (!IsDefault()) bsdt_ = std::make_unique<BSDT>(&indexes_);
  }

  // Enumerate the allowed blocks: every choice of the first rank-1 sector
  // coordinates in row-major order, then the sectors of the last index that
  // close the divergence, in ascending order (the same blocks and order as
  // scanning all coordinates, qltensor_impl.h:374-403).
  void AllowedBlocks(const QNT &div) {
    assert(!IsDefault());
    bsdt_->Clear();
    if (IsScalar()) {
      bsdt_->RefreshOffsets();
      bsdt_->Allocate();
      return;
    }
    std::vector<CoorsT> coors;
    const auto &last = indexes_.back();
    std::unordered_map<size_t, std::vector<size_t>> by_hash;   // contribution hash -> sectors
    for (size_t s = 0; s < last.GetQNSctNum(); ++s) {
      QNT q = last.GetQNSct(s).GetQn();
      if (last.GetDir() == IN) q = -q;
      else if (last.GetDir() == NDIR) q = QNT::Zero();
      by_hash[q.Hash()].push_back(s);
    }
    CoorsT head(rank_ - 1, 0);
    const size_t n_head = std::accumulate(bsdt_->blk_shape.begin(), bsdt_->blk_shape.end() - 1, size_t(1),
                                          std::multiplies<size_t>());
    for (size_t h = 0; h < n_head; ++h) {
      size_t rem = h;
      for (int i = static_cast<int>(rank_) - 2; i >= 0; --i) {
        head[i] = rem % bsdt_->blk_shape[i];
        rem /= bsdt_->blk_shape[i];
      }
      QNT part = QNT::Zero();
      for (size_t i = 0; i + 1 < rank_; ++i) {
        const QNT &q = indexes_[i].GetQNSct(head[i]).GetQn();
        if (indexes_[i].GetDir() == IN) part += -q;
        else if (indexes_[i].GetDir() == OUT) part += q;
      }
      const QNT need = div - part;
      auto it = by_hash.find(need.Hash());
      if (it == by_hash.end()) continue;
      for (size_t s : it->second) {
        CoorsT c(head);
        c.push_back(s);
        coors.push_back(std::move(c));
      }
    }
    bsdt_->DataBlksInsert(coors);
    bsdt_->Allocate();
  }

  std::pair<CoorsT, CoorsT> BlkAndDataCoors(const std::vector<size_t> &coors) const {
    CoorsT blk(rank_), data(rank_);
    for (size_t i = 0; i < rank_; ++i) {
      auto p = indexes_[i].CoorToBlkCoorDataCoor(coors[i]);
      blk[i] = p.first;
      data[i] = p.second;
    }
    return {blk, data};
  }

  static void AddScaledBlocks(QLTensor &out, const ElemT s, const QLTensor &x) {
    const auto &om = out.bsdt_->GetBlkIdxDataBlkMap();
    const double sc[2] = {std::real(std::complex<double>(s)), std::imag(std::complex<double>(s))};
    const double one[2] = {1.0, 0.0};
    if (out.IsScalar()) {
      if (x.bsdt_->GetActualRawDataSize())
        QltCheck(qlt_axpby(DType(), 1, sc, x.bsdt_->GetActualRawDataPtr(), one, out.bsdt_->GetActualRawDataPtr(),
                           hp_numeric::LibStream()));
      return;
    }
    const size_t esz = sizeof(ElemT);
    for (const auto &kv : x.bsdt_->GetBlkIdxDataBlkMap()) {
      const auto &dst = om.at(kv.first);
      const char *src = reinterpret_cast<const char *>(x.bsdt_->GetActualRawDataPtr()) + kv.second.data_offset * esz;
      char *d = reinterpret_cast<char *>(out.bsdt_->GetActualRawDataPtr()) + dst.data_offset * esz;
      QltCheck(qlt_axpby(DType(), static_cast<int64_t>(kv.second.size), sc, src, one, d, hp_numeric::LibStream()));
    }
  }

  size_t rank_ = 0;
  ShapeT shape_;
  size_t size_ = 0;
  IndexVec<QNT> indexes_;
  std::unique_ptr<BSDT> bsdt_;
};

// Stream operators of a tensor (qltensor.h:702-712); writing raises
// std::ios_base::failure when the stream fails, as the reference's block
// data tensor and indexes do through Streamable (streamable.h:58-64).
template <typename ElemT, typename QNT>
std::istream &operator>>(std::istream &is, QLTensor<ElemT, QNT> &t) {
  t.StreamRead(is);
  return is;
}
template <typename ElemT, typename QNT>
std::ostream &operator<<(std::ostream &os, const QLTensor<ElemT, QNT> &t) {
  t.StreamWrite(os);
  return os;
}

}  // namespace qlten

#endif
