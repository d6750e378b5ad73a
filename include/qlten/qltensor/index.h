// QNSector, Index and the divergence of a block.
// API as the reference's QNSector (qltensor/qnsct.h:35-139) and Index
// (qltensor/index.h:46-285): sectors of (quantum number, degeneracy), a
// direction IN = -1 / OUT = 1, equality by hash, InverseIndex, CalcDiv.
#ifndef QLTEN_AMD_QLTENSOR_INDEX_H
#define QLTEN_AMD_QLTENSOR_INDEX_H

#include <cassert>
#include <functional>
#include <ios>
#include <iostream>
#include <utility>
#include <vector>

#include "qlten/framework/value_t.h"
#include "qlten/qltensor/special_qn.h"

namespace qlten {

enum TenIndexDirType { NDIR = 0, IN = -1, OUT = 1 };

// Hash of a vector of hashable items (framework/vec_hash.h:22-34).
template <typename T>
size_t VecHasher(const std::vector<T> &vec) {
  constexpr size_t kP1 = 11400714785074694791ull, kP2 = 14029467366897019727ull,
                   kP5 = 2870177450012600261ull;
  size_t h = kP5;
  for (const auto &item : vec) {
    h += item.Hash() * kP2;
    h = (h << 31) | (h >> 33);
    h *= kP1;
  }
  return h + (vec.size() ^ kP5);
}

template <typename QNT>
class QNSector {
 public:
  QNSector() : QNSector(QNT(), 0) {}
  QNSector(const QNT &qn, const size_t dgnc) : qn_(qn), dgnc_(dgnc), hash_(qn.Hash() ^ dgnc) {}

  const QNT &GetQn() const { return qn_; }
  size_t GetDegeneracy() const { return dgnc_; }
  size_t dim() const { return dgnc_ * qn_.dim(); }
  size_t Hash() const { return hash_; }
  bool operator==(const QNSector &rhs) const { return hash_ == rhs.hash_; }
  bool operator!=(const QNSector &rhs) const { return !(*this == rhs); }

  // Stream format (qnsct.h:106-115): the quantum number, the degeneracy and
  // the stored hash.
  void StreamRead(std::istream &is) { is >> qn_ >> dgnc_ >> hash_; }
  void StreamWrite(std::ostream &os) const {
    os << qn_;
    os << dgnc_ << "\n";
    os << hash_ << "\n";
  }

 private:
  QNT qn_;
  size_t dgnc_;
  size_t hash_;
};

template <typename QNT>
using QNSectorVec = std::vector<QNSector<QNT>>;

template <typename QNT>
class Index {
 public:
  Index() : Index({}, TenIndexDirType::NDIR) {}
  Index(const QNSectorVec<QNT> &qnscts, const TenIndexDirType dir) : qnscts_(qnscts), dir_(dir) {
    dim_ = 0;
    for (const auto &s : qnscts_) dim_ += s.dim();
    hash_ = CalcHash();
  }

  size_t dim() const { return dim_; }
  TenIndexDirType GetDir() const { return dir_; }
  size_t GetQNSctNum() const { return qnscts_.size(); }
  const QNSector<QNT> &GetQNSct(const size_t i) const { return qnscts_[i]; }
  const QNSectorVec<QNT> &GetQNScts() const { return qnscts_; }

  // (sector coordinate, coordinate inside the sector) of an actual coordinate
  std::pair<size_t, size_t> CoorToBlkCoorDataCoor(size_t coor) const {
    assert(coor < dim_);
    for (size_t b = 0; b < qnscts_.size(); ++b) {
      const size_t d = qnscts_[b].dim();
      if (coor < d) return {b, coor};
      coor -= d;
    }
    return {0, 0};
  }

  void Inverse() {
    if (dir_ == IN) dir_ = OUT;
    else if (dir_ == OUT) dir_ = IN;
    hash_ = CalcHash();
  }

  size_t Hash() const { return hash_; }
  bool operator==(const Index &rhs) const { return hash_ == rhs.hash_; }
  bool operator!=(const Index &rhs) const { return !(*this == rhs); }

  // Stream format (index.h:170-188): sector count, the sectors, the direction
  // as an int, the dimension and the stored hash.
  void StreamRead(std::istream &is) {
    size_t n = 0;
    is >> n;
    qnscts_ = QNSectorVec<QNT>(n);
    for (auto &s : qnscts_) s.StreamRead(is);
    int dir = 0;
    is >> dir;
    dir_ = static_cast<TenIndexDirType>(dir);
    is >> dim_ >> hash_;
  }
  void StreamWrite(std::ostream &os) const {
    os << qnscts_.size() << "\n";
    for (const auto &s : qnscts_) s.StreamWrite(os);
    os << static_cast<int>(dir_) << "\n";
    os << dim_ << "\n";
    os << hash_ << "\n";
  }

 private:
  size_t CalcHash() const { return VecHasher(qnscts_) ^ std::hash<int>()(static_cast<int>(dir_)); }
  QNSectorVec<QNT> qnscts_;
  TenIndexDirType dir_;
  size_t dim_ = 0;
  size_t hash_ = 0;
};

// Stream operators; a write that leaves the stream failed throws, as the
// reference's Streamable operator<< does (framework/bases/streamable.h:58-64).
template <typename QNT>
std::istream &operator>>(std::istream &is, QNSector<QNT> &s) {
  s.StreamRead(is);
  return is;
}
template <typename QNT>
std::ostream &operator<<(std::ostream &os, const QNSector<QNT> &s) {
  s.StreamWrite(os);
  if (!os) throw std::ios_base::failure("Failed to write streamable object.");
  return os;
}
template <typename QNT>
std::istream &operator>>(std::istream &is, Index<QNT> &ix) {
  ix.StreamRead(is);
  return is;
}
template <typename QNT>
std::ostream &operator<<(std::ostream &os, const Index<QNT> &ix) {
  ix.StreamWrite(os);
  if (!os) throw std::ios_base::failure("Failed to write streamable object.");
  return os;
}

template <typename IndexT>
IndexT InverseIndex(const IndexT &idx) {
  IndexT inv(idx);
  inv.Inverse();
  return inv;
}

template <typename QNT>
using IndexVec = std::vector<Index<QNT>>;

template <typename IndexVecT>
std::vector<size_t> CalcQNSctNumOfIdxs(const IndexVecT &indexes) {
  std::vector<size_t> n;
  for (const auto &ix : indexes) n.push_back(ix.GetQNSctNum());
  return n;
}

// Divergence of a block: sum over indexes of +qn (OUT) or -qn (IN).
template <typename QNT>
QNT CalcDiv(const IndexVec<QNT> &indexes, const CoorsT &blk_coors) {
  assert(indexes.size() == blk_coors.size());
  QNT div = QNT::Zero();
  for (size_t i = 0; i < indexes.size(); ++i) {
    const QNT &q = indexes[i].GetQNSct(blk_coors[i]).GetQn();
    if (indexes[i].GetDir() == IN) div += -q;
    else if (indexes[i].GetDir() == OUT) div += q;
  }
  return div;
}

}  // namespace qlten

#endif
