// Block matrices of a tensor decomposition: the tensor's blocks grouped by
// (left flux, right flux) into dense matrices (GenIdxTenDecompDataBlkMats,
// tensor_manipulation/ten_decomp/ten_decomp_basic.h:68-106; TenDecompDataBlkMat
// and InsertDataBlkMat, qltensor/blk_spar_data_ten/data_blk_mat.h:39-131), and
// their assembly on the device with one batched copy
// (RawDataGenDenseDataBlkMat_, raw_data_operations.h:550-575).
#ifndef QLTEN_AMD_TENSOR_MANIPULATION_TEN_DECOMP_TEN_DECOMP_BASIC_H
#define QLTEN_AMD_TENSOR_MANIPULATION_TEN_DECOMP_TEN_DECOMP_BASIC_H

#include <array>
#include <cstdint>
#include <tuple>
#include <vector>

#include "qlten/qltensor/qltensor.h"

namespace qlten {
namespace detail {

// One dense block matrix: row sectors (left-part coordinates, row offset,
// rows) and column sectors in first-seen order, and its member blocks.
template <typename QNT>
struct DecompBlkMat {
  QNT lqnflux, rqnflux, mid_qn;
  size_t rows = 0, cols = 0;
  std::vector<std::tuple<CoorsT, size_t, size_t>> row_scts, col_scts;
  std::vector<std::tuple<size_t, size_t, size_t>> elems;   // (row sector, col sector, raw-data offset)

  static size_t FindOrAdd(std::vector<std::tuple<CoorsT, size_t, size_t>> &scts, const CoorsT &c, size_t dim,
                          size_t &total) {
    for (size_t i = 0; i < scts.size(); ++i)
      if (std::get<0>(scts[i]) == c) return i;
    scts.emplace_back(c, total, dim);
    total += dim;
    return scts.size() - 1;
  }
};

// Left flux of a block: sum over the first `ldims` legs of +qn (OUT) or -qn
// (IN) (CalcDataBlkMatLeftQNFlux, ten_decomp_basic.h:40-66).
template <typename QNT>
QNT LeftFlux(const IndexVec<QNT> &indexes, const CoorsT &coors, size_t ldims) {
  QNT f = QNT::Zero();
  for (size_t i = 0; i < ldims; ++i) {
    const QNT &q = indexes[i].GetQNSct(coors[i]).GetQn();
    if (indexes[i].GetDir() == TenIndexDirType::OUT) f += q;
    else if (indexes[i].GetDir() == TenIndexDirType::IN) f += -q;
  }
  return f;
}

template <typename ElemT, typename QNT>
std::vector<DecompBlkMat<QNT>> GenDecompBlkMats(const QLTensor<ElemT, QNT> &t, size_t ldims, const QNT &lqndiv) {
  std::vector<DecompBlkMat<QNT>> mats;
  const QNT div = t.Div();
  for (const auto &kv : t.GetBlkSparDataTen().GetBlkIdxDataBlkMap()) {
    const auto &blk = kv.second;
    const QNT lf = LeftFlux(t.GetIndexes(), blk.blk_coors, ldims);
    const QNT rf = div - lf;
    size_t m = 1, n = 1;
    for (size_t i = 0; i < blk.shape.size(); ++i) (i < ldims ? m : n) *= blk.shape[i];
    DecompBlkMat<QNT> *bm = nullptr;
    for (auto &x : mats)
      if (x.lqnflux == lf && x.rqnflux == rf) {
        bm = &x;
        break;
      }
    if (bm == nullptr) {
      mats.emplace_back();
      bm = &mats.back();
      bm->lqnflux = lf;
      bm->rqnflux = rf;
      bm->mid_qn = lqndiv - lf;
    }
    CoorsT lp(blk.blk_coors.begin(), blk.blk_coors.begin() + ldims);
    CoorsT rp(blk.blk_coors.begin() + ldims, blk.blk_coors.end());
    const size_t i = DecompBlkMat<QNT>::FindOrAdd(bm->row_scts, lp, m, bm->rows);
    const size_t j = DecompBlkMat<QNT>::FindOrAdd(bm->col_scts, rp, n, bm->cols);
    bm->elems.emplace_back(i, j, blk.data_offset);
  }
  return mats;
}

// Batched 2-D copy (qlt_copy_blocks_2d) of rectangles given as
// (rows, cols, src_off, src_ld, dst_off, dst_ld).
inline void Copy2D(int32_t dtype, const std::vector<std::array<int64_t, 6>> &rects, const void *src, void *dst) {
  if (rects.empty()) return;
  std::vector<int64_t> c[6];
  for (const auto &r : rects)
    for (int k = 0; k < 6; ++k) c[k].push_back(r[k]);
  hp_numeric::SyncLibStream();
  QltCheck(qlt_copy_blocks_2d(dtype, static_cast<int64_t>(rects.size()), c[0].data(), c[1].data(), src, c[2].data(),
                              c[3].data(), dst, c[4].data(), c[5].data(), hp_numeric::LibStream()));
}

// Dense row-major block matrices, concatenated on the device; `offs`
// receives each matrix's element offset.
template <typename ElemT, typename QNT>
DeviceArray AssembleBlkMats(const QLTensor<ElemT, QNT> &t, const std::vector<DecompBlkMat<QNT>> &mats,
                            std::vector<int64_t> &offs) {
  int64_t total = 0;
  offs.clear();
  for (const auto &bm : mats) {
    offs.push_back(total);
    total += static_cast<int64_t>(bm.rows * bm.cols);
  }
  DeviceArray a(static_cast<size_t>(std::max<int64_t>(total, 1)) * sizeof(ElemT));
  QltCheck(qlt_memset_zero(a.get(), a.bytes(), hp_numeric::LibStream()));
  std::vector<std::array<int64_t, 6>> rects;
  for (size_t mi = 0; mi < mats.size(); ++mi) {
    const auto &bm = mats[mi];
    for (const auto &e : bm.elems) {
      const auto &rs = bm.row_scts[std::get<0>(e)];
      const auto &cs = bm.col_scts[std::get<1>(e)];
      const int64_t m = static_cast<int64_t>(std::get<2>(rs)), n = static_cast<int64_t>(std::get<2>(cs));
      rects.push_back({m, n, static_cast<int64_t>(std::get<2>(e)), n,
                       offs[mi] + static_cast<int64_t>(std::get<1>(rs) * bm.cols + std::get<1>(cs)),
                       static_cast<int64_t>(bm.cols)});
    }
  }
  Copy2D(QLTensor<ElemT, QNT>::DType(), rects, t.GetBlkSparDataTen().GetActualRawDataPtr(), a.get());
  return a;
}

// Installs `coors` (any order) as the tensor's blocks and allocates its raw
// data; returns each input block's raw-data offset.
template <typename ElemT, typename QNT>
std::vector<size_t> InsertBlocks(QLTensor<ElemT, QNT> &t, const std::vector<CoorsT> &coors, bool zero) {
  auto &bsdt = t.GetBlkSparDataTen();
  bsdt.DataBlksInsert(coors);
  bsdt.Allocate(zero);
  std::vector<size_t> offs;
  for (const auto &c : coors) offs.push_back(bsdt.GetBlkIdxDataBlkMap().at(bsdt.BlkCoorsToBlkIdx(c)).data_offset);
  return offs;
}

}  // namespace detail
}  // namespace qlten

#endif
