// Truncated block SVD of a QLTensor (SVD / TensorSVDExecutor,
// tensor_manipulation/ten_decomp/ten_svd.h:84-263): block matrices grouped by
// flux and assembled on the device, decomposed in one batched call
// (qlt_block_svd), singular values truncated globally on the host
// (CalcTruncedSVInfo_, ten_svd.h:281-372), and U, S, Vt assembled with a new
// middle index (CreateSVDResTens_ / FillSVDResTens_, ten_svd.h:455-581).
#ifndef QLTEN_AMD_TENSOR_MANIPULATION_TEN_DECOMP_TEN_SVD_H
#define QLTEN_AMD_TENSOR_MANIPULATION_TEN_DECOMP_TEN_SVD_H

#include <algorithm>
#include <array>
#include <cassert>
#include <stdexcept>
#include <tuple>
#include <vector>

#include "qlten/tensor_manipulation/ten_decomp/ten_decomp_basic.h"

namespace qlten {

// No singular value survives (ten_svd.h:196-203).
class EmptySVDResultError : public std::runtime_error {
 public:
  using std::runtime_error::runtime_error;
};

namespace detail {

// CalcTruncedSVInfo_ (ten_svd.h:281-372): singular values != 0 sorted
// descending; keep at least Dmin; grow while the kept squared weight does not
// exceed (1 - trunc_err) of the total and the count stays <= Dmax.  Returns D;
// `kept[i]` is the count kept from block matrix i (a prefix of its descending
// values) and `err` = 1 - kept weight / total weight, clamped at 0.
inline size_t TruncateSV(const std::vector<std::vector<double>> &svals, double trunc_err, size_t Dmin, size_t Dmax,
                         std::vector<size_t> &kept, double &err) {
  std::vector<std::pair<double, size_t>> all;
  for (size_t i = 0; i < svals.size(); ++i)
    for (double s : svals[i])
      if (s != 0.0) all.emplace_back(s, i);
  kept.assign(svals.size(), 0);
  if (all.size() <= Dmin) {
    for (const auto &x : all) ++kept[x.second];
    err = 0.0;
    return all.size();
  }
  std::stable_sort(all.begin(), all.end(), [](const auto &a, const auto &b) { return a.first > b.first; });
  double total = 0.0;
  for (const auto &x : all) total += x.first * x.first;
  const double target = total * (1.0 - trunc_err);
  size_t d = Dmin;
  double ksum = 0.0;
  for (size_t i = 0; i < d; ++i) ksum += all[i].first * all[i].first;
  while (true) {
    if (ksum > target) break;
    const size_t nxt = d + 1;
    if (nxt > all.size()) break;
    const double nsum = ksum + all[nxt - 1].first * all[nxt - 1].first;
    if (nxt > Dmax) break;
    d = nxt;
    ksum = nsum;
  }
  for (size_t i = 0; i < d; ++i) ++kept[all[i].second];
  const double e = 1.0 - ksum / total;
  err = e > 0.0 ? e : 0.0;
  return d;
}

}  // namespace detail

// T = U S Vt over the bipartition after `ldims` legs; U carries the new OUT
// middle index (divergence lqndiv), S is (IN, OUT) diagonal, Vt starts with
// the IN middle index.  `trunc_err` uses the squared-norm convention.
template <typename TenElemT, typename QNT>
void SVD(const QLTensor<TenElemT, QNT> *pt, const size_t ldims, const QNT &lqndiv,
         const typename RealTypeTrait<TenElemT>::type trunc_err, const size_t Dmin, const size_t Dmax,
         QLTensor<TenElemT, QNT> *pu, QLTensor<typename RealTypeTrait<TenElemT>::type, QNT> *ps,
         QLTensor<TenElemT, QNT> *pvt, typename RealTypeTrait<TenElemT>::type *pactual_trunc_err, size_t *pD) {
  using Ten = QLTensor<TenElemT, QNT>;
  assert(pt->Rank() >= 2 && ldims >= 1 && ldims < pt->Rank());
  assert(pu->IsDefault() && ps->IsDefault() && pvt->IsDefault());
  assert(trunc_err >= 0 && Dmin <= Dmax);
  const auto mats = detail::GenDecompBlkMats(*pt, ldims, lqndiv);
  if (mats.empty()) throw EmptySVDResultError("SVD failed: empty result (the tensor has no blocks)");
  std::vector<int64_t> a_offs;
  DeviceArray a = detail::AssembleBlkMats(*pt, mats, a_offs);
  const size_t nm = mats.size();
  std::vector<int64_t> rows(nm), cols(nm), r(nm), u_offs(nm), s_offs(nm), vt_offs(nm);
  int64_t nu = 0, ns = 0, nvt = 0;
  for (size_t i = 0; i < nm; ++i) {
    rows[i] = static_cast<int64_t>(mats[i].rows);
    cols[i] = static_cast<int64_t>(mats[i].cols);
    r[i] = std::min(rows[i], cols[i]);
    u_offs[i] = nu;
    s_offs[i] = ns;
    vt_offs[i] = nvt;
    nu += rows[i] * r[i];
    ns += r[i];
    nvt += r[i] * cols[i];
  }
  DeviceArray u(static_cast<size_t>(nu) * sizeof(TenElemT)), s(static_cast<size_t>(ns) * sizeof(double)),
      vt(static_cast<size_t>(nvt) * sizeof(TenElemT));
  hp_numeric::SyncLibStream();
  QltCheck(qlt_block_svd(Ten::DType(), static_cast<int64_t>(nm), rows.data(), cols.data(), a.get(), a_offs.data(),
                         u.get(), u_offs.data(), static_cast<double *>(s.get()), s_offs.data(), vt.get(),
                         vt_offs.data(), nullptr, hp_numeric::LibStream()));
  std::vector<double> h_s(static_cast<size_t>(ns));
  QltCheck(qlt_memcpy_d2h(h_s.data(), s.get(), h_s.size() * sizeof(double)));
  std::vector<std::vector<double>> svals(nm);
  for (size_t i = 0; i < nm; ++i) svals[i].assign(h_s.begin() + s_offs[i], h_s.begin() + s_offs[i] + r[i]);
  std::vector<size_t> kept;
  double err = 0.0;
  const size_t D = detail::TruncateSV(svals, static_cast<double>(trunc_err), Dmin, Dmax, kept, err);
  if (D == 0)
    throw EmptySVDResultError("SVD failed: empty result. The input tensor has no data blocks "
                              "compatible with the specified quantum number divergence (lqndiv).");

  // middle index: one sector per block matrix with kept values, in matrix order
  QNSectorVec<QNT> mid;
  std::vector<size_t> live;
  for (size_t i = 0; i < nm; ++i)
    if (kept[i] > 0) {
      mid.push_back(QNSector<QNT>(mats[i].mid_qn, kept[i]));
      live.push_back(i);
    }
  const Index<QNT> mid_out(mid, TenIndexDirType::OUT);
  const Index<QNT> mid_in = InverseIndex(mid_out);
  const auto &tidx = pt->GetIndexes();
  IndexVec<QNT> u_idx(tidx.begin(), tidx.begin() + ldims);
  u_idx.push_back(mid_out);
  IndexVec<QNT> vt_idx{mid_in};
  vt_idx.insert(vt_idx.end(), tidx.begin() + ldims, tidx.end());
  *pu = Ten(std::move(u_idx));
  *pvt = Ten(std::move(vt_idx));
  *ps = QLTensor<typename RealTypeTrait<TenElemT>::type, QNT>(IndexVec<QNT>{mid_in, mid_out});

  std::vector<CoorsT> u_coors, vt_coors, s_coors;
  std::vector<std::array<int64_t, 4>> u_src, vt_src;   // rows, cols, src offset, src ld
  for (size_t coor = 0; coor < live.size(); ++coor) {
    const size_t i = live[coor];
    const int64_t k = static_cast<int64_t>(kept[i]);
    for (const auto &rs : mats[i].row_scts) {
      CoorsT c = std::get<0>(rs);
      c.push_back(coor);
      u_coors.push_back(c);
      u_src.push_back({static_cast<int64_t>(std::get<2>(rs)), k,
                       u_offs[i] + static_cast<int64_t>(std::get<1>(rs)) * r[i], r[i]});
    }
    for (const auto &cs : mats[i].col_scts) {
      CoorsT c{coor};
      c.insert(c.end(), std::get<0>(cs).begin(), std::get<0>(cs).end());
      vt_coors.push_back(c);
      vt_src.push_back({k, static_cast<int64_t>(std::get<2>(cs)),
                        vt_offs[i] + static_cast<int64_t>(std::get<1>(cs)), cols[i]});
    }
    s_coors.push_back({coor, coor});
  }
  const auto u_dst = detail::InsertBlocks(*pu, u_coors, false);
  const auto vt_dst = detail::InsertBlocks(*pvt, vt_coors, false);
  std::vector<std::array<int64_t, 6>> rects;
  for (size_t b = 0; b < u_coors.size(); ++b)
    rects.push_back({u_src[b][0], u_src[b][1], u_src[b][2], u_src[b][3], static_cast<int64_t>(u_dst[b]),
                     u_src[b][1]});
  detail::Copy2D(Ten::DType(), rects, u.get(), pu->GetBlkSparDataTen().GetActualRawDataPtr());
  rects.clear();
  for (size_t b = 0; b < vt_coors.size(); ++b)
    rects.push_back({vt_src[b][0], vt_src[b][1], vt_src[b][2], vt_src[b][3], static_cast<int64_t>(vt_dst[b]),
                     vt_src[b][1]});
  detail::Copy2D(Ten::DType(), rects, vt.get(), pvt->GetBlkSparDataTen().GetActualRawDataPtr());

  // S: diagonal blocks holding each matrix's kept singular values
  auto &sb = ps->GetBlkSparDataTen();
  sb.DataBlksInsert(s_coors);
  using SRealT = typename RealTypeTrait<TenElemT>::type;
  std::vector<SRealT> sdata(sb.RawDataSize(), SRealT(0));
  for (size_t coor = 0; coor < live.size(); ++coor) {
    const size_t i = live[coor], k = kept[i];
    const size_t off = sb.GetBlkIdxDataBlkMap().at(sb.BlkCoorsToBlkIdx(s_coors[coor])).data_offset;
    for (size_t j = 0; j < k; ++j) sdata[off + j * k + j] = static_cast<SRealT>(svals[i][j]);
  }
  sb.RawDataFromHost(sdata);
  *pactual_trunc_err = static_cast<typename RealTypeTrait<TenElemT>::type>(err);
  *pD = D;
}

}  // namespace qlten

#endif
