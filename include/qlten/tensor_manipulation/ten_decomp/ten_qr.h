// Block QR of a QLTensor (QR / TensorQRExecutor,
// tensor_manipulation/ten_decomp/ten_qr.h:56-293): block matrices grouped by
// flux, decomposed in one batched call (qlt_block_qr), Q with the new OUT
// middle index of one sector of dim min(rows, cols) per block matrix, R with
// the IN middle index (CreatQRResTens_ / FillQRResTens_, ten_qr.h:186-289).
#ifndef QLTEN_AMD_TENSOR_MANIPULATION_TEN_DECOMP_TEN_QR_H
#define QLTEN_AMD_TENSOR_MANIPULATION_TEN_DECOMP_TEN_QR_H

#include <algorithm>
#include <array>
#include <cassert>
#include <tuple>
#include <vector>

#include "qlten/tensor_manipulation/ten_decomp/ten_decomp_basic.h"

namespace qlten {

template <typename TenElemT, typename QNT>
void QR(const QLTensor<TenElemT, QNT> *pt, const size_t ldims, const QNT &lqndiv, QLTensor<TenElemT, QNT> *pq,
        QLTensor<TenElemT, QNT> *pr) {
  using Ten = QLTensor<TenElemT, QNT>;
  assert(pt->Rank() >= 2 && ldims >= 1 && ldims < pt->Rank());
  assert(pq->IsDefault() && pr->IsDefault());
  const auto mats = detail::GenDecompBlkMats(*pt, ldims, lqndiv);
  std::vector<int64_t> a_offs;
  DeviceArray a = detail::AssembleBlkMats(*pt, mats, a_offs);
  const size_t nm = mats.size();
  std::vector<int64_t> rows(nm), cols(nm), k(nm), q_offs(nm), r_offs(nm);
  int64_t nq = 0, nr = 0;
  QNSectorVec<QNT> mid;
  for (size_t i = 0; i < nm; ++i) {
    rows[i] = static_cast<int64_t>(mats[i].rows);
    cols[i] = static_cast<int64_t>(mats[i].cols);
    k[i] = std::min(rows[i], cols[i]);
    q_offs[i] = nq;
    r_offs[i] = nr;
    nq += rows[i] * k[i];
    nr += k[i] * cols[i];
    mid.push_back(QNSector<QNT>(mats[i].mid_qn, static_cast<size_t>(k[i])));
  }
  DeviceArray q(static_cast<size_t>(nq) * sizeof(TenElemT)), r(static_cast<size_t>(nr) * sizeof(TenElemT));
  if (nm) {
    hp_numeric::SyncLibStream();
    QltCheck(qlt_block_qr(Ten::DType(), static_cast<int64_t>(nm), rows.data(), cols.data(), a.get(), a_offs.data(),
                          q.get(), q_offs.data(), r.get(), r_offs.data(), hp_numeric::LibStream()));
  }
  const Index<QNT> mid_out(mid, TenIndexDirType::OUT);
  const Index<QNT> mid_in = InverseIndex(mid_out);
  const auto &tidx = pt->GetIndexes();
  IndexVec<QNT> q_idx(tidx.begin(), tidx.begin() + ldims);
  q_idx.push_back(mid_out);
  IndexVec<QNT> r_idx{mid_in};
  r_idx.insert(r_idx.end(), tidx.begin() + ldims, tidx.end());
  *pq = Ten(std::move(q_idx));
  *pr = Ten(std::move(r_idx));
  if (nm == 0) return;
  std::vector<CoorsT> q_coors, r_coors;
  std::vector<std::array<int64_t, 4>> q_src, r_src;
  for (size_t i = 0; i < nm; ++i) {
    for (const auto &rs : mats[i].row_scts) {
      CoorsT c = std::get<0>(rs);
      c.push_back(i);
      q_coors.push_back(c);
      q_src.push_back({static_cast<int64_t>(std::get<2>(rs)), k[i],
                       q_offs[i] + static_cast<int64_t>(std::get<1>(rs)) * k[i], k[i]});
    }
    for (const auto &cs : mats[i].col_scts) {
      CoorsT c{i};
      c.insert(c.end(), std::get<0>(cs).begin(), std::get<0>(cs).end());
      r_coors.push_back(c);
      r_src.push_back({k[i], static_cast<int64_t>(std::get<2>(cs)),
                       r_offs[i] + static_cast<int64_t>(std::get<1>(cs)), cols[i]});
    }
  }
  const auto q_dst = detail::InsertBlocks(*pq, q_coors, false);
  const auto r_dst = detail::InsertBlocks(*pr, r_coors, false);
  std::vector<std::array<int64_t, 6>> rects;
  for (size_t b = 0; b < q_coors.size(); ++b)
    rects.push_back({q_src[b][0], q_src[b][1], q_src[b][2], q_src[b][3], static_cast<int64_t>(q_dst[b]), q_src[b][1]});
  detail::Copy2D(Ten::DType(), rects, q.get(), pq->GetBlkSparDataTen().GetActualRawDataPtr());
  rects.clear();
  for (size_t b = 0; b < r_coors.size(); ++b)
    rects.push_back({r_src[b][0], r_src[b][1], r_src[b][2], r_src[b][3], static_cast<int64_t>(r_dst[b]), r_src[b][1]});
  detail::Copy2D(Ten::DType(), rects, r.get(), pr->GetBlkSparDataTen().GetActualRawDataPtr());
}

}  // namespace qlten

#endif
