// Tests of the qlten host API (include/qlten/) on the device, written the way
// the reference tests contraction (tests/test_tensor_manipulation/
// test_ten_ctrct.cc): U(1) indexes with sectors {-1, 0, +1} of degeneracy 3
// ("s") or 10 ("l"), Random tensors, each result checked element by element
// against a dense contraction computed on the host; path equivalence of the
// contiguous-axes API against the general Contract; accumulate semantics.
// Exit status 0 means every check passed.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <functional>
#include <sstream>
#include <streambuf>
#include <map>
#include <numeric>
#include <set>
#include <string>
#include <type_traits>
#include <vector>

#include "qlten/qlten.h"

using namespace qlten;
using U1Idx = Index<U1QN>;
using QNSctT = QNSector<U1QN>;
using DTen = QLTensor<QLTEN_Double, U1QN>;
using ZTen = QLTensor<QLTEN_Complex, U1QN>;

static int g_fail = 0;
static int g_checks = 0;
#define CHECK(cond, msg)                                                     \
  do {                                                                       \
    ++g_checks;                                                              \
    if (!(cond)) {                                                           \
      ++g_fail;                                                              \
      std::printf("FAIL %s:%d %s\n", __FILE__, __LINE__, std::string(msg).c_str()); \
    }                                                                        \
  } while (0)

// Dense host copy of a tensor (row-major over the full index dimensions).
template <typename T, typename Q>
std::vector<T> Densify(const QLTensor<T, Q> &t) {
  size_t total = 1;
  for (size_t d : t.GetShape()) total *= d;
  std::vector<T> out(total, T(0));
  if (t.IsDefault()) return out;
  if (t.IsScalar()) {
    auto raw = t.GetBlkSparDataTen().RawDataToHost();
    if (!raw.empty()) out[0] = raw[0];
    return out;
  }
  auto raw = t.GetBlkSparDataTen().RawDataToHost();
  const size_t r = t.Rank();
  std::vector<std::vector<size_t>> soff(r);
  for (size_t i = 0; i < r; ++i) {
    size_t acc = 0;
    for (size_t s = 0; s < t.GetIndex(i).GetQNSctNum(); ++s) {
      soff[i].push_back(acc);
      acc += t.GetIndex(i).GetQNSct(s).dim();
    }
  }
  for (const auto &kv : t.GetBlkSparDataTen().GetBlkIdxDataBlkMap()) {
    const auto &blk = kv.second;
    std::vector<size_t> c(r, 0);
    for (size_t e = 0; e < blk.size; ++e) {
      size_t rem = e, lin = 0;
      for (int i = static_cast<int>(r) - 1; i >= 0; --i) {
        c[i] = rem % blk.shape[i];
        rem /= blk.shape[i];
      }
      for (size_t i = 0; i < r; ++i) lin = lin * t.GetShape()[i] + soff[i][blk.blk_coors[i]] + c[i];
      out[lin] = raw[blk.data_offset + e];
    }
  }
  return out;
}

// Dense contraction of A and B over the given axes; result axes are A's saved
// (in a_saved order) then B's saved (in b_saved order).
template <typename T>
std::vector<T> DenseContract(const std::vector<T> &a, const ShapeT &sa, const std::vector<T> &b, const ShapeT &sb,
                             const std::vector<size_t> &a_ax, const std::vector<size_t> &b_ax,
                             const std::vector<size_t> &a_sv, const std::vector<size_t> &b_sv) {
  size_t m = 1, n = 1, k = 1;
  for (size_t x : a_sv) m *= sa[x];
  for (size_t x : b_sv) n *= sb[x];
  for (size_t x : a_ax) k *= sa[x];
  std::vector<T> c(m * n, T(0));
  auto strides = [](const ShapeT &s) {
    std::vector<size_t> st(s.size(), 1);
    for (int i = static_cast<int>(s.size()) - 2; i >= 0; --i) st[i] = st[i + 1] * s[i + 1];
    return st;
  };
  const auto ast = strides(sa), bst = strides(sb);
  auto offset = [](size_t lin, const std::vector<size_t> &axes, const ShapeT &shape, const std::vector<size_t> &st) {
    size_t off = 0;
    for (int j = static_cast<int>(axes.size()) - 1; j >= 0; --j) {
      off += (lin % shape[axes[j]]) * st[axes[j]];
      lin /= shape[axes[j]];
    }
    return off;
  };
  for (size_t i = 0; i < m; ++i) {
    const size_t ai = offset(i, a_sv, sa, ast);
    for (size_t j = 0; j < n; ++j) {
      const size_t bj = offset(j, b_sv, sb, bst);
      T s = T(0);
      for (size_t p = 0; p < k; ++p) {
        size_t ap = 0, bp = 0, rem = p;
        for (int q = static_cast<int>(a_ax.size()) - 1; q >= 0; --q) {
          const size_t c = rem % sa[a_ax[q]];
          rem /= sa[a_ax[q]];
          ap += c * ast[a_ax[q]];
          bp += c * bst[b_ax[q]];
        }
        s += a[ai + ap] * b[bj + bp];
      }
      c[i * n + j] = s;
    }
  }
  return c;
}

template <typename T>
double RelDiff(const std::vector<T> &x, const std::vector<T> &y) {
  double num = 0, den = 0;
  for (size_t i = 0; i < x.size(); ++i) {
    num += std::norm(std::complex<double>(x[i] - y[i]));
    den += std::norm(std::complex<double>(y[i]));
  }
  return den > 0 ? std::sqrt(num / den) : std::sqrt(num);
}

// S is real; the complex tests contract it after ToComplex.
template <typename T>
QLTensor<T, U1QN> ToComplexIfNeeded(const QLTensor<QLTEN_Double, U1QN> &s) {
  if constexpr (IsComplex<T>()) return ToComplex(s);
  else return s;
}

struct Fixture {
  U1QN qn0 = U1QN(0), qnp1 = U1QN(1), qnm1 = U1QN(-1);
  U1Idx idx_in_s, idx_out_s, idx_in_l, idx_out_l;
  Fixture() {
    QNSectorVec<U1QN> s = {QNSctT(qnm1, 3), QNSctT(qn0, 3), QNSctT(qnp1, 3)};
    QNSectorVec<U1QN> l = {QNSctT(qnm1, 10), QNSctT(qn0, 10), QNSctT(qnp1, 10)};
    idx_in_s = U1Idx(s, IN);
    idx_out_s = U1Idx(s, OUT);
    idx_in_l = U1Idx(l, IN);
    idx_out_l = U1Idx(l, OUT);
  }
};

template <typename T, typename Q>
void CheckContract(const QLTensor<T, Q> &a, const QLTensor<T, Q> &b, const std::vector<size_t> &a_ax,
                   const std::vector<size_t> &b_ax, const std::string &name) {
  QLTensor<T, Q> c;
  Contract(&a, &b, {a_ax, b_ax}, &c);
  std::vector<size_t> a_sv, b_sv;
  for (size_t i = 0; i < a.Rank(); ++i)
    if (std::find(a_ax.begin(), a_ax.end(), i) == a_ax.end()) a_sv.push_back(i);
  for (size_t i = 0; i < b.Rank(); ++i)
    if (std::find(b_ax.begin(), b_ax.end(), i) == b_ax.end()) b_sv.push_back(i);
  const auto ref = DenseContract(Densify(a), a.GetShape(), Densify(b), b.GetShape(), a_ax, b_ax, a_sv, b_sv);
  const auto got = Densify(c);
  CHECK(got.size() == ref.size(), name + ": size");
  const double err = RelDiff(got, ref);
  CHECK(err < 1e-13, name + ": rel err " + std::to_string(err));
  CHECK(c.Rank() == a_sv.size() + b_sv.size(), name + ": rank");
}

// Contiguous-axes contraction in the four (ASide, BSide) forms and the four
// legacy bool forms (RunTestTenMatBasedCtrct, test_ten_ctrct.cc:638-735).
// The benchmark follows the reference test: A is viewed with its axes after
// the contracted run first, then the axes up to the run's end (trans_axes_a,
// :673-680), which puts A's saved axes in the cyclic order
// back + 1, ..., rank - 1, 0, ..., front - 1; likewise B.  The dense
// contraction with those saved orders is computed on the host.
template <typename T>
void CheckMatBasedContract(const QLTensor<T, U1QN> &ta, const std::vector<size_t> &a_ax,
                           const QLTensor<T, U1QN> &tb, const std::vector<size_t> &b_ax, const std::string &name) {
  auto cyclic_saved = [](size_t rank, const std::vector<size_t> &ax) {
    std::vector<size_t> sv;
    for (size_t i = ax.back() + 1; i < rank; ++i) sv.push_back(i);
    for (size_t i = 0; i < ax.front(); ++i) sv.push_back(i);
    return sv;
  };
  const auto a_sv = cyclic_saved(ta.Rank(), a_ax), b_sv = cyclic_saved(tb.Rank(), b_ax);
  const auto ref = DenseContract(Densify(ta), ta.GetShape(), Densify(tb), tb.GetShape(), a_ax, b_ax, a_sv, b_sv);
  IndexVec<U1QN> ref_idx;
  for (size_t x : a_sv) ref_idx.push_back(ta.GetIndex(x));
  for (size_t x : b_sv) ref_idx.push_back(tb.GetIndex(x));
  const size_t a0 = a_ax[0], b0 = b_ax[0], n = a_ax.size();
  std::vector<std::pair<std::string, QLTensor<T, U1QN>>> res(8);
  Contract<T, U1QN, true, true>(ta, tb, a0, b0, n, res[0].second);
  Contract<T, U1QN, true, false>(ta, tb, a0, b0, n, res[1].second);
  Contract<T, U1QN, false, true>(ta, tb, a0, b0, n, res[2].second);
  Contract<T, U1QN, false, false>(ta, tb, a0, b0, n, res[3].second);
  ContractContiguousAxes<T, U1QN, CtrctSide::Tail, CtrctSide::Head>(ta, tb, a0, b0, n, res[4].second);
  ContractContiguousAxes<T, U1QN, CtrctSide::Tail, CtrctSide::Tail>(ta, tb, a0, b0, n, res[5].second);
  ContractContiguousAxes<T, U1QN, CtrctSide::Head, CtrctSide::Head>(ta, tb, a0, b0, n, res[6].second);
  ContractContiguousAxes<T, U1QN, CtrctSide::Head, CtrctSide::Tail>(ta, tb, a0, b0, n, res[7].second);
  const char *forms[8] = {"<true,true>", "<true,false>", "<false,true>", "<false,false>",
                          "<Tail,Head>", "<Tail,Tail>", "<Head,Head>", "<Head,Tail>"};
  for (size_t i = 0; i < res.size(); ++i) {
    const auto &c = res[i].second;
    const std::string nm = name + " " + forms[i];
    CHECK(c.GetIndexes() == ref_idx, nm + ": result indexes in cyclic saved order");
    const auto got = Densify(c);
    CHECK(got.size() == ref.size(), nm + ": size");
    if (got.size() == ref.size()) {
      const double err = RelDiff(got, ref);
      CHECK(err < (IsSingle<T>() ? 1e-5 : 1e-13), nm + ": rel err " + std::to_string(err));
    }
    if (i > 0) {
      CHECK(c.GetBlkSparDataTen().GetActualRawDataSize() == res[0].second.GetBlkSparDataTen().GetActualRawDataSize(),
            nm + ": same block shell");
    }
  }
}

// .qlten round trip through a file (RunTestQLTensorFileIOCase,
// tests/test_qltensor/test_qltensor.cc:1029-1041).
template <typename TenT>
void CheckFileIO(const TenT &t, const std::string &name) {
  const std::string file = "test_host_api_io.qlten";
  {
    std::ofstream out(file, std::ofstream::binary);
    out << t;
  }
  TenT cpy;
  {
    std::ifstream in(file, std::ifstream::binary);
    in >> cpy;
  }
  std::remove(file.c_str());
  CHECK(cpy == t, "file IO round trip: " + name);
}

// A stream buffer that accepts `limit` bytes and then fails
// (test_qltensor.cc:37-65).
class WriteFailAfter : public std::streambuf {
 public:
  explicit WriteFailAfter(std::streamsize limit) : left_(limit) {}

 protected:
  std::streamsize xsputn(const char *, std::streamsize n) override {
    if (left_ <= 0) return 0;
    const std::streamsize w = std::min(n, left_);
    left_ -= w;
    return w;
  }
  int_type overflow(int_type ch) override {
    if (traits_type::eq_int_type(ch, traits_type::eof())) return traits_type::not_eof(ch);
    if (left_ <= 0) return traits_type::eof();
    --left_;
    return traits_type::not_eof(ch);
  }

 private:
  std::streamsize left_;
};

DTen d2_for_fail(const Fixture &f) {
  DTen d2({f.idx_in_s, f.idx_out_s});
  d2.Random(U1QN(0));
  return d2;
}

// EstimateContractCost (tensor_op_cost.h:434-530) against counts derived on
// the host from the operands' block maps alone: blocks meet when their
// contracted sector coordinates agree; the first task of an output block does
// not read it; a transposed operand contributes its distinct touched blocks.
// The output layout must be the one Contract produces.
template <typename T>
void CheckContractCost(const QLTensor<T, U1QN> &a, const QLTensor<T, U1QN> &b, const std::vector<size_t> &a_ax,
                       const std::vector<size_t> &b_ax, bool a_trans, bool b_trans, const std::string &name) {
  const auto cost = EstimateContractCost(a, b, {a_ax, b_ax});
  std::vector<size_t> a_sv, b_sv;
  for (size_t i = 0; i < a.Rank(); ++i)
    if (std::find(a_ax.begin(), a_ax.end(), i) == a_ax.end()) a_sv.push_back(i);
  for (size_t i = 0; i < b.Rank(); ++i)
    if (std::find(b_ax.begin(), b_ax.end(), i) == b_ax.end()) b_sv.push_back(i);
  const auto &am = a.GetBlkSparDataTen().GetBlkIdxDataBlkMap();
  const auto &bm = b.GetBlkSparDataTen().GetBlkIdxDataBlkMap();
  const size_t esz = sizeof(T);
  double flops = 0;
  size_t pairs = 0, rd = 0, wr = 0;
  std::map<CoorsT, size_t> c_blocks;   // output coordinates -> tasks seen
  std::set<size_t> a_used, b_used;
  for (const auto &ka : am)
    for (const auto &kb : bm) {
      bool meet = true;
      for (size_t i = 0; i < a_ax.size(); ++i)
        meet = meet && ka.second.blk_coors[a_ax[i]] == kb.second.blk_coors[b_ax[i]];
      if (!meet) continue;
      size_t m = 1, k = 1, n = 1;
      CoorsT cc;
      for (size_t ax : a_sv) m *= ka.second.shape[ax], cc.push_back(ka.second.blk_coors[ax]);
      for (size_t ax : a_ax) k *= ka.second.shape[ax];
      for (size_t ax : b_sv) n *= kb.second.shape[ax], cc.push_back(kb.second.blk_coors[ax]);
      ++pairs;
      flops += (std::is_same<T, QLTEN_Complex>::value ? 8.0 : 2.0) * double(m) * double(k) * double(n);
      rd += (m * k + k * n) * esz;
      if (c_blocks[cc]++ > 0) rd += m * n * esz;
      wr += m * n * esz;
      a_used.insert(ka.first);
      b_used.insert(kb.first);
    }
  size_t temp = 0;
  if (a_trans)
    for (size_t i : a_used) temp += am.at(i).size * esz;
  if (b_trans)
    for (size_t i : b_used) temp += bm.at(i).size * esz;
  CHECK(cost.flops == flops, name + ": flops");
  CHECK(cost.gemm_count == pairs && cost.executable_block_pair_count == pairs, name + ": tasks");
  CHECK(cost.candidate_block_pair_count == am.size() * bm.size(), name + ": candidate pairs");
  CHECK(cost.output_block_count == c_blocks.size(), name + ": output blocks");
  CHECK(cost.temp_peak_bytes == temp, name + ": temp bytes");
  CHECK(cost.read_bytes == rd + temp, name + ": read bytes " + std::to_string(cost.read_bytes) + " vs " +
                                          std::to_string(rd + temp));
  CHECK(cost.write_bytes == wr + temp, name + ": write bytes");
  QLTensor<T, U1QN> c;
  Contract(&a, &b, {a_ax, b_ax}, &c);
  CHECK(cost.output_layout == MakeBlockSparseTensorLayout(c), name + ": output layout");
  CHECK(cost.output_raw_elem_count == c.GetBlkSparDataTen().RawDataSize(), name + ": output elements");
  std::vector<std::pair<size_t, size_t>> pairs_form;
  for (size_t i = 0; i < a_ax.size(); ++i) pairs_form.emplace_back(a_ax[i], b_ax[i]);
  const auto cost2 = EstimateContractCost(a, b, pairs_form);
  CHECK(cost2.flops == cost.flops && cost2.read_bytes == cost.read_bytes &&
            cost2.output_layout == cost.output_layout, name + ": pair form");
}

void TestContractCost(const Fixture &f) {
  // test_tensor_op_cost.cc:19-42: one (2 x 3) . (3 x 4) block pair
  const U1Idx lf({QNSctT(f.qn0, 2)}, OUT), cl({QNSctT(f.qn0, 3)}, OUT), cr({QNSctT(f.qn0, 3)}, IN),
      rf({QNSctT(f.qn0, 4)}, OUT);
  DTen l({lf, cl}), r({cr, rf});
  l.Fill(f.qn0, 1.0);
  r.Fill(f.qn0, 1.0);
  const auto c = EstimateContractCost(l, r, {{1}, {0}});
  CHECK(c.flops == 48.0, "cost: flops 48");
  CHECK(c.gemm_count == 1 && c.executable_block_pair_count == 1, "cost: one task");
  CHECK(c.output_block_count == 1 && c.output_raw_elem_count == 8, "cost: one 8-element block");
  CHECK(c.read_bytes == 18 * sizeof(QLTEN_Double) && c.write_bytes == 8 * sizeof(QLTEN_Double), "cost: bytes");
  CHECK(c.temp_peak_bytes == 0, "cost: no transpose");
  CHECK(c.output_layout.indexes == (IndexVec<U1QN>{lf, rf}), "cost: output indexes");
  CHECK(c.output_layout.blocks.size() == 1 && c.output_layout.blocks.begin()->second.shape == (ShapeT{2, 4}),
        "cost: output block shape");

  // several sectors, accumulating output blocks, with and without transposes
  DTen a({f.idx_in_s, f.idx_out_l, f.idx_out_s});
  DTen b({f.idx_in_s, f.idx_in_l, f.idx_out_s});
  a.Random(U1QN(0));
  b.Random(U1QN(1));
  CheckContractCost(a, b, {2}, {0}, false, false, "cost rank3 tail-head");
  CheckContractCost(a, b, {1, 2}, {1, 0}, false, true, "cost rank3 two axes, B permuted");
  DTen b2({f.idx_in_l, f.idx_out_s, f.idx_in_s});
  b2.Random(U1QN(0));
  CheckContractCost(a, b2, {1}, {0}, true, false, "cost rank3 middle axis of A");
  ZTen za({f.idx_in_s, f.idx_out_l, f.idx_out_s});
  ZTen zb({f.idx_in_l, f.idx_in_s, f.idx_out_s});
  za.Random(U1QN(0));
  zb.Random(U1QN(-1));
  CheckContractCost(za, zb, {2, 1}, {1, 0}, true, true, "cost complex, both permuted");

  // composition over a layout shell (test_tensor_op_cost.cc:44-80 pattern): the
  // shell keeps the layout, holds no data, and the next estimate on it equals
  // the estimate on the real contraction result
  {
    const auto c1 = EstimateContractCost(a, b, {{2}, {0}});
    const auto shell = MakeTensorShellFromLayout<QLTEN_Double>(c1.output_layout);
    CHECK(MakeBlockSparseTensorLayout(shell) == c1.output_layout, "shell: layout round trip");
    CHECK(!shell.HasActualData(), "shell: no raw data");
    DTen real_c;
    Contract(&a, &b, {{2}, {0}}, &real_c);
    CHECK(real_c.HasActualData(), "shell: the contraction result holds data");
    DTen e({f.idx_in_s, f.idx_out_s});
    e.Random(U1QN(0));
    const auto on_shell = EstimateContractCost(shell, e, {{3}, {0}});
    const auto on_real = EstimateContractCost(real_c, e, {{3}, {0}});
    CHECK(on_shell.flops == on_real.flops && on_shell.flops > 0 && on_shell.read_bytes == on_real.read_bytes &&
              on_shell.write_bytes == on_real.write_bytes && on_shell.output_layout == on_real.output_layout,
          "shell: composed estimate");
    CHECK(MakeTensorShellFromLayout<QLTEN_Double>(BlockSparseTensorLayout<U1QN>()).IsDefault(),
          "shell: default layout gives a default tensor");
  }

  // errors: index mismatch, default operand
  bool threw = false;
  try {
    EstimateContractCost(a, b, {{0}, {0}});
  } catch (const std::invalid_argument &) {
    threw = true;
  }
  CHECK(threw, "cost: index mismatch throws");
  threw = false;
  try {
    EstimateContractCost(DTen(), b, {{0}, {0}});
  } catch (const std::invalid_argument &) {
    threw = true;
  }
  CHECK(threw, "cost: default operand throws");
}

int main() {
  Fixture f;
  SetRandomSeed(7);

  // ---- general Contract, real, rank 1..3, every div (test_ten_ctrct.cc:226-494 pattern) ----
  for (int div = -1; div <= 1; ++div) {
    DTen t1({f.idx_out_s});
    t1.Random(U1QN(div));
    DTen t1d({f.idx_in_s});
    t1d.Random(U1QN(div));
    CheckContract(t1, t1d, {0}, {0}, "rank1 full div " + std::to_string(div));

    DTen a2({f.idx_in_s, f.idx_out_l});
    DTen b2({f.idx_in_l, f.idx_out_s});
    a2.Random(U1QN(div));
    b2.Random(U1QN(0));
    CheckContract(a2, b2, {1}, {0}, "rank2 tail-head");
    DTen b2t({f.idx_out_s, f.idx_in_l});
    b2t.Random(U1QN(div));
    CheckContract(a2, b2t, {1}, {1}, "rank2 tail-tail");

    DTen a3({f.idx_in_s, f.idx_out_l, f.idx_out_s});
    DTen b3({f.idx_in_l, f.idx_in_s, f.idx_out_l});
    a3.Random(U1QN(div));
    b3.Random(U1QN(div));
    CheckContract(a3, b3, {1}, {0}, "rank3 one axis");
    CheckContract(a3, b3, {1, 2}, {0, 1}, "rank3 two axes");
    CheckContract(a3, b3, {2, 1}, {1, 0}, "rank3 two axes reversed");
  }

  // ---- complex and mixed real x complex (ten_ctrct.h:292-350) ----
  {
    ZTen a({f.idx_in_l, f.idx_out_s, f.idx_out_l});
    ZTen b({f.idx_in_l, f.idx_in_s, f.idx_out_s});
    a.Random(U1QN(0));
    b.Random(U1QN(1));
    CheckContract(a, b, {2}, {0}, "complex rank3");
    CheckContract(a, b, {2, 1}, {0, 1}, "complex rank3 two axes");
    DTen ar({f.idx_in_l, f.idx_out_s, f.idx_out_l});
    ar.Random(U1QN(0));
    ZTen c_mixed;
    Contract(&ar, &b, {{2}, {0}}, &c_mixed);
    ZTen c_promoted;
    auto arz = ToComplex(ar);
    Contract(&arz, &b, {{2}, {0}}, &c_promoted);
    CHECK(RelDiff(Densify(c_mixed), Densify(c_promoted)) == 0.0, "mixed real x complex == promoted");
  }

  // ---- t . dagger(t) = sum |t_i|^2 (test_ten_ctrct.cc:187-206) ----
  {
    ZTen t({f.idx_out_l});
    t.Random(U1QN(0));
    ZTen td = Dag(t);
    ZTen s;
    Contract(&t, &td, {{0}, {0}}, &s);
    double ref = 0;
    for (auto v : Densify(t)) ref += std::norm(v);
    const auto got = s.GetElem({});
    CHECK(std::abs(got - QLTEN_Complex(ref, 0.0)) < 1e-12 * ref, "t.dag(t)");
  }

  // ---- contiguous-axes path == general Contract (test_ten_ctrct.cc:638-859) ----
  {
    DTen a({f.idx_in_s, f.idx_out_l, f.idx_out_s, f.idx_in_l});
    DTen b({f.idx_in_s, f.idx_out_l, f.idx_in_l});
    a.Random(U1QN(0));
    b.Random(U1QN(1));
    // A axes 2,3 with B axes 0,1 (Tail/Head): saved A = (0,1), B = (2)
    DTen c_gen, c_cont;
    Contract(&a, &b, {{2, 3}, {0, 1}}, &c_gen);
    ContiguousContractStats stats;
    ContractContiguousAxes<QLTEN_Double, U1QN, CtrctSide::Tail, CtrctSide::Head>(a, b, 2, 0, 2, c_cont, &stats);
    CHECK(c_gen.GetIndexes() == c_cont.GetIndexes(), "contiguous indexes");
    CHECK(RelDiff(Densify(c_cont), Densify(c_gen)) < 1e-13, "contiguous tail-head values");
    CHECK(stats.raw_data_contract_tasks > 0, "stats tasks");
    // wrap-around: A axes (3, 0) contracted -> saved cyclic (1, 2)
    DTen a2({f.idx_out_s, f.idx_out_l, f.idx_in_s, f.idx_in_l});
    a2.Random(U1QN(0));
    DTen b2({f.idx_out_l, f.idx_in_s, f.idx_out_s});
    b2.Random(U1QN(0));
    DTen c_wrap, c_ref;
    ContractContiguousAxes<QLTEN_Double, U1QN, CtrctSide::Head, CtrctSide::Head>(a2, b2, 3, 0, 2, c_wrap);
    Contract(&a2, &b2, {{3, 0}, {0, 1}}, &c_ref);   // general: saved A ascending (1, 2) -- same order here
    CHECK(RelDiff(Densify(c_wrap), Densify(c_ref)) < 1e-13, "contiguous wrap-around");
    // legacy bool form
    DTen c_legacy;
    Contract<QLTEN_Double, U1QN, true, true>(a, b, 2, 0, 2, c_legacy);
    CHECK(RelDiff(Densify(c_legacy), Densify(c_gen)) < 1e-13, "legacy bool Contract");
  }

  // ---- the reference's MatBasedContract cases (test_ten_ctrct.cc:737-859):
  // axes {2}/{0} (cyclic = ascending), {1}/{0} (A saved (2, 0): cyclic order
  // differs from ascending), {1,2}/{0,1}, {0,1,2}/{0,1,2} (scalar), every
  // divergence pair the reference uses; real and complex ----
  {
    const U1QN q0(0), qp(1), qm(-1);
    auto cases = [&](auto tag, const std::string &ty) {
      using T = decltype(tag);
      using Ten = QLTensor<T, U1QN>;
      Ten s3({f.idx_in_s, f.idx_out_s, f.idx_out_s});
      Ten s3b({f.idx_in_s, f.idx_out_s, f.idx_out_s});
      Ten s3c({f.idx_in_s, f.idx_in_s, f.idx_out_s});
      Ten s3d({f.idx_out_s, f.idx_in_s, f.idx_in_s});
      const std::vector<std::pair<U1QN, U1QN>> divs = {{q0, q0}, {qp, q0}, {qp, qm}};
      for (const auto &dv : divs) {
        const std::string d = ty + " div (" + std::to_string(dv.first.GetVal()) + "," +
                              std::to_string(dv.second.GetVal()) + ")";
        s3.Random(dv.first);
        s3b.Random(dv.second);
        CheckMatBasedContract(s3, {2}, s3b, {0}, d + " {2}/{0}");
        CheckMatBasedContract(s3, {1}, s3b, {0}, d + " {1}/{0}");
        s3c.Random(dv.second);
        CheckMatBasedContract(s3, {1, 2}, s3c, {0, 1}, d + " {1,2}/{0,1}");
        s3d.Random(dv.second);
        CheckMatBasedContract(s3, {0, 1, 2}, s3d, {0, 1, 2}, d + " {0,1,2}/{0,1,2}");
      }
      // rank 4 with the run in the middle of A and at the end of B: A saved
      // (3, 0), B saved (0, 1)
      Ten a4({f.idx_in_s, f.idx_out_l, f.idx_out_s, f.idx_in_l});
      Ten b4({f.idx_in_s, f.idx_out_l, f.idx_in_l, f.idx_in_s});
      a4.Random(q0);
      b4.Random(qp);
      CheckMatBasedContract(a4, {1, 2}, b4, {2, 3}, ty + " rank4 {1,2}/{2,3}");
    };
    cases(QLTEN_Double(0), "real");
    cases(QLTEN_Complex(0), "complex");
    // the reference runs the same cases on QLTEN_Float / QLTEN_ComplexFloat
    // tensors (test_ten_ctrct.cc:798-859)
    cases(QLTEN_Float(0), "float");
    cases(QLTEN_ComplexFloat(0), "complex float");
  }

  // ---- single-precision element types (value_t.h:38-41): stored as float,
  // contracted with FP64 arithmetic and rounded once, so the result is the
  // FP64 contraction of the float inputs to within float rounding ----
  {
    auto widen = [](const auto &v) {
      using T = typename std::decay_t<decltype(v)>::value_type;
      using W = std::conditional_t<IsComplex<T>(), QLTEN_Complex, QLTEN_Double>;
      std::vector<W> out(v.size());
      for (size_t i = 0; i < v.size(); ++i) out[i] = W(v[i]);
      return out;
    };
    auto check_single = [&](auto tag, const std::string &name) {
      using T = decltype(tag);
      using Ten = QLTensor<T, U1QN>;
      Ten a({f.idx_in_l, f.idx_out_s, f.idx_out_l});
      Ten b({f.idx_in_l, f.idx_in_s, f.idx_out_s});
      a.Random(U1QN(0));
      b.Random(U1QN(1));
      for (const auto &ax : std::vector<std::pair<std::vector<size_t>, std::vector<size_t>>>{
               {{2}, {0}}, {{2, 1}, {0, 1}}}) {
        Ten c;
        Contract(&a, &b, {ax.first, ax.second}, &c);
        std::vector<size_t> a_sv, b_sv;
        for (size_t i = 0; i < a.Rank(); ++i)
          if (std::find(ax.first.begin(), ax.first.end(), i) == ax.first.end()) a_sv.push_back(i);
        for (size_t i = 0; i < b.Rank(); ++i)
          if (std::find(ax.second.begin(), ax.second.end(), i) == ax.second.end()) b_sv.push_back(i);
        const auto ref = DenseContract(widen(Densify(a)), a.GetShape(), widen(Densify(b)), b.GetShape(), ax.first,
                                       ax.second, a_sv, b_sv);
        const double err = RelDiff(widen(Densify(c)), ref);
        CHECK(err < 1e-7, name + " contraction vs FP64 of the float inputs: rel err " + std::to_string(err));
      }
      const auto n2 = a.Get2Norm();
      double ref2 = 0;
      for (auto v : widen(Densify(a))) ref2 += std::norm(std::complex<double>(v));
      CHECK(std::abs(static_cast<double>(n2) - std::sqrt(ref2)) < 1e-6 * std::sqrt(ref2), name + " Get2Norm");
      Ten a2 = a * T(2);
      CHECK(RelDiff(widen(Densify(a2)), [&] { auto d = widen(Densify(a)); for (auto &v : d) v *= 2.0; return d; }()) ==
                0.0,
            name + " operator*");
    };
    check_single(QLTEN_Float(0), "float");
    check_single(QLTEN_ComplexFloat(0), "complex float");
    // mixed float x complex float (ten_ctrct.h:322-350) equals the promoted contraction
    QLTensor<QLTEN_Float, U1QN> ar({f.idx_in_l, f.idx_out_s, f.idx_out_l});
    QLTensor<QLTEN_ComplexFloat, U1QN> bz({f.idx_in_l, f.idx_in_s, f.idx_out_s});
    ar.Random(U1QN(0));
    bz.Random(U1QN(1));
    QLTensor<QLTEN_ComplexFloat, U1QN> c_mixed, c_promoted, c_mixed2, c_promoted2;
    Contract(&ar, &bz, {{2}, {0}}, &c_mixed);
    auto arz = ToComplex(ar);
    Contract(&arz, &bz, {{2}, {0}}, &c_promoted);
    CHECK(RelDiff(widen(Densify(c_mixed)), widen(Densify(c_promoted))) == 0.0, "mixed float x complex float");
    Contract(&bz, &ar, {{0}, {2}}, &c_mixed2);
    Contract(&bz, &arz, {{0}, {2}}, &c_promoted2);
    CHECK(RelDiff(widen(Densify(c_mixed2)), widen(Densify(c_promoted2))) == 0.0, "mixed complex float x float");
    // t . dag(t) = sum |t|^2 for complex float
    QLTensor<QLTEN_ComplexFloat, U1QN> t({f.idx_out_l});
    t.Random(U1QN(0));
    auto td = Dag(t);
    QLTensor<QLTEN_ComplexFloat, U1QN> sc;
    Contract(&t, &td, {{0}, {0}}, &sc);
    double ref = 0;
    for (auto v : widen(Densify(t))) ref += std::norm(v);
    CHECK(std::abs(std::complex<double>(sc.GetElem({})) - std::complex<double>(ref, 0.0)) < 1e-6 * ref,
          "complex float t.dag(t)");
    // block SVD of a float tensor: the known answer in single precision
    QNSectorVec<U1QN> one = {QNSctT(U1QN(0), 2)};
    QLTensor<QLTEN_Float, U1QN> m({U1Idx(one, IN), U1Idx(one, OUT)});
    m.Fill(U1QN(0), 0.0f);
    m.SetElem({0, 0}, 1.0f);
    m.SetElem({0, 1}, 2.0f);
    m.SetElem({1, 0}, 3.0f);
    m.SetElem({1, 1}, 4.0f);
    QLTensor<QLTEN_Float, U1QN> mu, mvt, ms;
    float merr = -1;
    size_t mD = 0;
    SVD(&m, 1, U1QN(0), 0.0f, 1, 10, &mu, &ms, &mvt, &merr, &mD);
    CHECK(mD == 2, "float svd D");
    CHECK(std::abs(ms.GetElem({0, 0}) - 5.464985704219043f) < 1e-6f * 5.5f, "float svd sigma_0");
    CHECK(std::abs(ms.GetElem({1, 1}) - 0.365966190626257f) < 1e-6f * 5.5f, "float svd sigma_1");
  }

  // ---- accumulate: C = beta C + alpha A.B (contract_contiguous_axes.h:664-782) ----
  {
    DTen a({f.idx_in_l, f.idx_out_l});
    DTen b({f.idx_in_l, f.idx_out_s});
    a.Random(U1QN(0));
    b.Random(U1QN(0));
    DTen c0;
    Contract(&a, &b, {{1}, {0}}, &c0);
    DTen acc(c0);
    ContractTailHeadContiguousAccumulate(a, b, 1, 0, 1, 2.0, 0.5, acc);
    auto expect = Densify(c0);
    for (auto &v : expect) v = 2.5 * v;   // 0.5 * C0 + 2 * A.B with C0 = A.B
    CHECK(RelDiff(Densify(acc), expect) < 1e-13, "accumulate alpha/beta");
    DTen fresh;
    ContractTailHeadContiguousAccumulate(a, b, 1, 0, 1, 3.0, 0.0, fresh);
    auto e3 = Densify(c0);
    for (auto &v : e3) v *= 3.0;
    CHECK(RelDiff(Densify(fresh), e3) < 1e-13, "accumulate into default output");
    bool threw = false;
    try {
      DTen empty;
      ContractTailHeadContiguousAccumulate(a, b, 1, 0, 1, 1.0, 1.0, empty);
    } catch (const std::invalid_argument &) {
      threw = true;
    }
    CHECK(threw, "beta != 0 on a default output throws");
    threw = false;
    try {
      ContractTailHeadContiguousAccumulate(a, b, 1, 0, 1, 1.0, 1.0, a);
    } catch (const std::invalid_argument &) {
      threw = true;
    }
    CHECK(threw, "aliasing throws");
  }

  // ---- accumulate into an output whose layout is not the plan's own, in every
  // arithmetic mode (contract_contiguous_axes.h:389-470, 953-1041): C holds
  // blocks of another divergence between the touched ones, so the touched
  // blocks sit at offsets other than the plan's; untouched blocks are scaled
  // by beta; a missing touched block expands the topology; a plain Contract
  // on the same cached plan afterwards writes at the plan's own offsets ----
  {
    QNSectorVec<U1QN> big = {QNSctT(U1QN(-1), 40), QNSctT(U1QN(0), 50), QNSctT(U1QN(1), 36)};
    const U1Idx bi(big, IN), bo(big, OUT);
    auto run = [&](auto tag, hp_numeric::ContractPrecision prec, auto alpha, auto beta, const std::string &name) {
      using T = decltype(tag);
      using Ten = QLTensor<T, U1QN>;
      Ten a({bi, bo}), b({bi, bo}), x0({bi, bo}), x1({bi, bo});
      a.Random(U1QN(0));
      b.Random(U1QN(0));
      x0.Random(U1QN(0));
      x1.Random(U1QN(1));
      if (prec != hp_numeric::ContractPrecision::kNativeFP64) QltCheck(qlt_set_split_min_mean_k(0.0));
      hp_numeric::SetContractPrecision(prec);
      const auto ab = DenseContract(Densify(a), a.GetShape(), Densify(b), b.GetShape(), {1}, {0}, {0}, {1});
      auto expect_of = [&](const Ten &init) {
        auto e = Densify(init);
        for (size_t i = 0; i < e.size(); ++i) e[i] = beta * e[i] + alpha * ab[i];
        return e;
      };
      // interleaved layout: blocks of divergence 0 and 1
      Ten c = x0 + x1;
      CHECK(c.GetBlkSparDataTen().GetBlkIdxDataBlkMap().size() == 3 + 2, name + ": interleaved layout");
      const auto e1 = expect_of(c);
      ContiguousContractStats st;
      ContractTailHeadContiguousAccumulate(a, b, 1, 0, 1, T(alpha), T(beta), c, &st);
      double err = RelDiff(Densify(c), e1);
      CHECK(err < 1e-12, name + ": accumulate into interleaved layout, rel err " + std::to_string(err));
      CHECK(st.output_topology_expansions == 0, name + ": no expansion needed");
      // a second call on the same target reuses the redirected plan
      const auto e2 = expect_of(c);
      ContractTailHeadContiguousAccumulate(a, b, 1, 0, 1, T(alpha), T(beta), c);
      err = RelDiff(Densify(c), e2);
      CHECK(err < 1e-12, name + ": second accumulate, rel err " + std::to_string(err));
      // expansion: only the untouched blocks exist
      Ten g = x1;
      CHECK(!TryContractTailHeadContiguousAccumulate(a, b, 1, 0, 1, T(alpha), T(beta), g),
            name + ": Try refuses expansion");
      const auto e3 = expect_of(g);
      ContractTailHeadContiguousAccumulate(a, b, 1, 0, 1, T(alpha), T(beta), g, &st);
      err = RelDiff(Densify(g), e3);
      CHECK(err < 1e-12, name + ": accumulate with expansion, rel err " + std::to_string(err));
      CHECK(st.output_topology_expansions == 1, name + ": one expansion");
      // the cached plan, restored, serves a plain Contract
      Ten p;
      Contract(&a, &b, {{1}, {0}}, &p);
      err = RelDiff(Densify(p), ab);
      CHECK(err < 1e-12, name + ": plain Contract after accumulate, rel err " + std::to_string(err));
      hp_numeric::SetContractPrecision(hp_numeric::ContractPrecision::kNativeFP64);
      QltCheck(qlt_set_split_min_mean_k(48.0));
    };
    for (auto prec : {hp_numeric::ContractPrecision::kNativeFP64, hp_numeric::ContractPrecision::kSplitBF16,
                      hp_numeric::ContractPrecision::kSplitInt8}) {
      const std::string m = "accumulate mode " + std::to_string(static_cast<int>(prec));
      run(QLTEN_Double(0), prec, 1.5, -0.5, m + " real");
      run(QLTEN_Complex(0), prec, QLTEN_Complex(0.5, -1.5), QLTEN_Complex(2.0, 0.25), m + " complex");
      run(QLTEN_Complex(0), prec, QLTEN_Complex(2.0, 0.0), QLTEN_Complex(0.5, 0.0), m + " complex, real scalars");
    }
  }

  // ---- plan cache reuse on a repeated pattern ----
  {
    DTen a({f.idx_in_l, f.idx_out_l});
    DTen b({f.idx_in_l, f.idx_out_l});
    a.Random(U1QN(0));
    b.Random(U1QN(0));
    DTen c1, c2;
    const size_t h0 = detail::PlanCache::Instance().hits;
    Contract(&a, &b, {{1}, {0}}, &c1);
    a.Random(U1QN(0));   // same topology, new values
    Contract(&a, &b, {{1}, {0}}, &c2);
    CHECK(detail::PlanCache::Instance().hits > h0, "plan cache hit");
    CheckContract(a, b, {1}, {0}, "after cache hit");
  }

  // ---- split precision modes (BF16 slices, int8 digits) keep 1e-12 ----
  for (auto prec : {hp_numeric::ContractPrecision::kSplitBF16, hp_numeric::ContractPrecision::kSplitInt8}) {
    QNSectorVec<U1QN> big = {QNSctT(U1QN(-1), 70), QNSctT(U1QN(0), 90), QNSctT(U1QN(1), 80)};
    U1Idx bi(big, IN), bo(big, OUT);
    ZTen a({bi, bo});
    ZTen b({bi, bo});
    a.Random(U1QN(0));
    b.Random(U1QN(0));
    ZTen cn, cs;
    Contract(&a, &b, {{1}, {0}}, &cn);
    QltCheck(qlt_set_split_min_mean_k(0.0));
    hp_numeric::SetContractPrecision(prec);
    Contract(&a, &b, {{1}, {0}}, &cs);
    hp_numeric::SetContractPrecision(hp_numeric::ContractPrecision::kNativeFP64);
    QltCheck(qlt_set_split_min_mean_k(48.0));
    const double err = RelDiff(Densify(cs), Densify(cn));
    CHECK(err < 1e-12, "split mode " + std::to_string(static_cast<int>(prec)) + " rel err " + std::to_string(err));
  }

  // ---- element access, norms, linear combination ----
  {
    DTen t({f.idx_in_s, f.idx_out_s});
    t.Random(U1QN(0));
    const double n2 = t.Get2Norm();
    double ref = 0;
    for (double v : Densify(t)) ref += v * v;
    CHECK(std::abs(n2 - std::sqrt(ref)) < 1e-12 * std::sqrt(ref), "Get2Norm");
    t.SetElem({0, 0}, 42.0);
    CHECK(t.GetElem({0, 0}) == 42.0, "SetElem/GetElem");
    CHECK(t.GetElem({0, 3}) == 0.0, "element of an absent block is 0");
    DTen u = t + t * 2.0;
    CHECK(RelDiff(Densify(u), [&] { auto d = Densify(t); for (auto &v : d) v *= 3.0; return d; }()) < 1e-14,
          "operator+ / operator*");
    CHECK(t.Div() == U1QN(0), "Div");
  }

  // ---- block SVD and QR (ten_svd.h / ten_qr.h semantics) ----
  {
    // known answer: tests/package_consumer/smoke_gpu.cu:100-117
    QNSectorVec<U1QN> one = {QNSctT(U1QN(0), 2)};
    DTen t({U1Idx(one, IN), U1Idx(one, OUT)});
    t.Fill(U1QN(0), 0.0);
    t.SetElem({0, 0}, 1.0);
    t.SetElem({0, 1}, 2.0);
    t.SetElem({1, 0}, 3.0);
    t.SetElem({1, 1}, 4.0);
    DTen u, vt;
    QLTensor<QLTEN_Double, U1QN> s;
    double err = -1;
    size_t D = 0;
    SVD(&t, 1, U1QN(0), 0.0, 1, 10, &u, &s, &vt, &err, &D);
    CHECK(D == 2 && err == 0.0, "svd 2x2 D / err");
    CHECK(std::abs(s.GetElem({0, 0}) - 5.464985704219043) < 1e-14 * 5.5, "svd 2x2 sigma_0");
    CHECK(std::abs(s.GetElem({1, 1}) - 0.365966190626257) < 1e-14 * 5.5, "svd 2x2 sigma_1");
  }
  auto svd_roundtrip = [&](auto tag, double trunc, const std::string &name) {
    using T = decltype(tag);
    QLTensor<T, U1QN> t({f.idx_in_l, f.idx_in_s, f.idx_out_s, f.idx_out_l});
    t.Random(U1QN(0));
    QLTensor<T, U1QN> u, vt, us, back;
    QLTensor<QLTEN_Double, U1QN> s;
    double err = -1;
    size_t D = 0;
    SVD(&t, 2, U1QN(0), trunc, 1, 100000, &u, &s, &vt, &err, &D);
    CHECK(D == s.GetShape()[0] && u.GetShape()[2] == D && vt.GetShape()[0] == D, name + ": D");
    QLTensor<T, U1QN> sc = ToComplexIfNeeded<T>(s);
    Contract(&u, &sc, {{2}, {0}}, &us);
    Contract(&us, &vt, {{2}, {0}}, &back);
    const auto dt = Densify(t), db = Densify(back);
    const double rel = RelDiff(db, dt);
    if (trunc == 0.0) {
      CHECK(rel < 1e-12, name + ": U S Vt == T, rel " + std::to_string(rel));
      CHECK(err < 1e-14, name + ": no truncation error");
    } else {
      CHECK(std::abs(rel * rel - err) < 1e-10, name + ": |T - U S Vt|^2 / |T|^2 == err");
    }
    QLTensor<T, U1QN> ud = Dag(u), uu;
    Contract(&ud, &u, {{0, 1}, {0, 1}}, &uu);
    const auto duu = Densify(uu);
    double dev = 0;
    for (size_t i = 0; i < D; ++i)
      for (size_t j = 0; j < D; ++j) dev = std::max(dev, std::abs(duu[i * D + j] - T(i == j ? 1.0 : 0.0)));
    CHECK(dev < 1e-12, name + ": U^dag U == 1");
  };
  svd_roundtrip(QLTEN_Double(0), 0.0, "svd real");
  svd_roundtrip(QLTEN_Complex(0), 0.0, "svd complex");
  svd_roundtrip(QLTEN_Double(0), 1e-2, "svd real truncated");
  {
    ZTen t({f.idx_in_l, f.idx_in_s, f.idx_out_s, f.idx_out_l});
    t.Random(U1QN(0));
    ZTen q, r, back;
    QR(&t, 2, U1QN(0), &q, &r);
    Contract(&q, &r, {{2}, {0}}, &back);
    CHECK(RelDiff(Densify(back), Densify(t)) < 1e-12, "qr: Q R == T");
    ZTen qd = Dag(q), qq;
    Contract(&qd, &q, {{0, 1}, {0, 1}}, &qq);
    const size_t k = q.GetShape()[2];
    const auto dq = Densify(qq);
    double dev = 0;
    for (size_t i = 0; i < k; ++i)
      for (size_t j = 0; j < k; ++j) dev = std::max(dev, std::abs(dq[i * k + j] - QLTEN_Complex(i == j ? 1.0 : 0.0)));
    CHECK(dev < 1e-12, "qr: Q^dag Q == 1");
  }

  // ---- .qlten file format (QLTensor::StreamRead/StreamWrite, qltensor_impl.h:798-830;
  // the cases of tests/test_qltensor/test_qltensor.cc:1043-1073) ----
  {
    const U1QN qn0(0), qnp1(1);
    DTen dscalar{IndexVec<U1QN>{}};
    dscalar.Random(U1QN());
    CheckFileIO(dscalar, "real scalar");
    ZTen zscalar{IndexVec<U1QN>{}};
    zscalar.Random(U1QN());
    CheckFileIO(zscalar, "complex scalar");
    for (const U1QN &q : {qn0, qnp1}) {
      DTen d1({f.idx_out_s});
      d1.Random(q);
      CheckFileIO(d1, "real 1d");
      DTen d2({f.idx_in_s, f.idx_out_s});
      d2.Random(q);
      CheckFileIO(d2, "real 2d");
      DTen d3({f.idx_in_s, f.idx_out_l, f.idx_out_s});
      d3.Random(q);
      CheckFileIO(d3, "real 3d");
      ZTen z3({f.idx_in_s, f.idx_out_l, f.idx_out_s});
      z3.Random(q);
      CheckFileIO(z3, "complex 3d");
    }
    using U11Ten = QLTensor<QLTEN_Complex, U1U1QN>;
    const Index<U1U1QN> u_in({QNSector<U1U1QN>(U1U1QN(0, 0), 2), QNSector<U1U1QN>(U1U1QN(1, -1), 3),
                              QNSector<U1U1QN>(U1U1QN(-1, 1), 1)},
                             TenIndexDirType::IN);
    const Index<U1U1QN> u_out = InverseIndex(u_in);
    U11Ten u3({u_in, u_out, u_out});
    u3.Random(U1U1QN(0, 0));
    CheckFileIO(u3, "complex U1U1 3d");

    // the read tensor is an ordinary tensor: it contracts like the original
    DTen a2({f.idx_in_s, f.idx_out_l}), b2({f.idx_in_l, f.idx_out_s});
    a2.Random(qn0);
    b2.Random(qnp1);
    std::stringstream ss;
    ss << a2;
    DTen a2r;
    ss >> a2r;
    DTen c0, c1;
    Contract(&a2, &b2, {{1}, {0}}, &c0);
    Contract(&a2r, &b2, {{1}, {0}}, &c1);
    CHECK(c0 == c1, "a read tensor contracts as the original");

    // a write into a failing stream throws (test_qltensor.cc:1075-1082)
    WriteFailAfter failing(32);
    std::ostream bad(&failing);
    bool threw = false;
    try {
      bad << d2_for_fail(f);
    } catch (const std::ios_base::failure &) {
      threw = true;
    }
    CHECK(threw, "writing into a failing stream throws std::ios_base::failure");

    // a file for the Python reader / writer cross-check
    if (const char *path = std::getenv("QLTEN_IO_OUT")) {
      std::ofstream out(path, std::ofstream::binary);
      out << u3;
      std::ofstream out2(std::string(path) + ".real", std::ofstream::binary);
      DTen d3({f.idx_in_s, f.idx_out_l, f.idx_out_s});
      d3.Random(qnp1);
      out2 << d3;
    }
  }

  // ---- Z2 and trivial quantum numbers (special_qn/znqn.h, trivial_rep_qn.h) ----
  {
    using Z2Idx = Index<Z2QN>;
    const Z2QN z0(0), z1(1);
    CHECK(z1 + z1 == z0 && -z1 == z1 && -z0 == z0 && z1 - z1 == z0, "Z2 arithmetic");
    Z2QN acc = z1;
    acc += z1;
    CHECK(acc == z0 && acc.GetVal() == 0, "Z2 +=");
    CHECK(ZnQN<3>(1) + ZnQN<3>(2) == ZnQN<3>(0) && -ZnQN<3>(1) == ZnQN<3>(2), "Z3 arithmetic");
    CHECK(z0.Hash() != z1.Hash(), "Z2 hashes differ");
    const Z2Idx zo({QNSector<Z2QN>(z0, 2), QNSector<Z2QN>(z1, 3)}, OUT), zi = InverseIndex(zo);
    const Z2Idx wo({QNSector<Z2QN>(z0, 4), QNSector<Z2QN>(z1, 1)}, OUT), wi = InverseIndex(wo);
    QLTensor<QLTEN_Double, Z2QN> za({zi, wo, zo}), zb({zi, wi, zo});
    za.Random(z0);
    zb.Random(z1);
    CheckContract(za, zb, {2}, {0}, "Z2 rank3 tail-head");
    CheckContract(za, zb, {1, 2}, {1, 0}, "Z2 rank3 two axes");
    QLTensor<QLTEN_Complex, Z2QN> zc({zi, wo}), zd({wi, zo});
    zc.Random(z1);
    zd.Random(z1);
    CheckContract(zc, zd, {1}, {0}, "Z2 complex rank2");
    CheckFileIO(za, "Z2 tensor");

    using Q3 = U1U1ZnQN<2>;
    CHECK(Q3().Hash() == (size_t(1) << 30) + (size_t(1) << 51), "U1U1Z2 zero hash (2^20 + 2^41, rotated by 10)");
    CHECK(Q3(1, -2, 1) + Q3(-1, 2, 1) == Q3() && -Q3(3, 1, 1) == Q3(-3, -1, 1), "U1U1Z2 arithmetic");
    const Index<Q3> qo({QNSector<Q3>(Q3(0, 0, 0), 2), QNSector<Q3>(Q3(1, 0, 1), 3), QNSector<Q3>(Q3(0, 1, 1), 2)},
                       OUT),
        qi = InverseIndex(qo);
    QLTensor<QLTEN_Double, Q3> qa({qi, qo, qo}), qb({qi, qi, qo});
    qa.Random(Q3(1, 0, 1));
    qb.Random(Q3(0, 0, 0));
    CheckContract(qa, qb, {2}, {0}, "U1U1Z2 rank3 tail-head");
    CheckContract(qa, qb, {1, 2}, {1, 0}, "U1U1Z2 rank3 two axes");
    CheckFileIO(qa, "U1U1Z2 tensor");

    using TIdx = Index<TrivialRepQN>;
    const TIdx to({QNSector<TrivialRepQN>(TrivialRepQN(), 5)}, OUT), ti = InverseIndex(to);
    const TIdx so({QNSector<TrivialRepQN>(TrivialRepQN(), 3)}, OUT);
    QLTensor<QLTEN_Double, TrivialRepQN> ta({so, to}), tb({ti, so});
    ta.Random(TrivialRepQN());
    tb.Random(TrivialRepQN());
    CheckContract(ta, tb, {1}, {0}, "trivial QN dense contraction");
  }

  // ---- contraction cost estimate (tensor_op_cost.h) ----
  TestContractCost(f);

  std::printf("%d checks, %d failures\n", g_checks, g_fail);
  return g_fail == 0 ? 0 : 1;
}
