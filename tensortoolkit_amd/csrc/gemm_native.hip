// Grouped block GEMM over a contraction plan, native FP64 / complex128 MFMA.
//
// Replaces the execution loop of CtrctTwoBSDTAndAssignIn
// (include/qlten/qltensor/blk_spar_data_ten/global_operations.h:895-992):
// there, every task lazily transposes its A and B blocks with HPTT/cuTENSOR
// (hp_numeric::TensorTranspose, framework/hp_numeric/ten_trans.h:56-114,245-349)
// and then issues one BLAS call C_blk = A'(m x k) B'(k x n) + beta C_blk
// (hp_numeric::MatMultiply, framework/hp_numeric/blas_level3.h:35-56,798-856).
// Here one launch per tile class covers the whole plan.  Each workgroup owns
// one TM x TN tile (64 or 32 each way) of one output block, or one k-slice of
// it, and walks every task of that block, so the per-block beta chain (0 for
// the first task, 1 after) becomes a single write of the finished tile (or of
// its slices' sum).  Narrow outputs (n <= 16) go to a row-streaming kernel.  The transposes are never materialised: the
// prologue gathers A'(i, p) = A[a_off + row_off(i) + col_off(p)] straight
// from the original block layout through per-tile offset tables, staging the
// k-chunk in LDS, with the thread mapping chosen so that the block's
// innermost memory axis is read contiguously.

#include <cstdio>
#include <memory>
#include <type_traits>

#include "qlt_internal.h"

using namespace qlt;

namespace {

typedef double v4d __attribute__((ext_vector_type(4)));

// Element access of the raw data: E is the stored real type (double, or
// float for QLTEN_Float / QLTEN_ComplexFloat tensors), complex elements are
// interleaved (re, im); arithmetic is FP64 either way.  `p` points to the
// element (its real part).
template <typename E, bool CPLX>
__device__ __forceinline__ double2 ld_elem(const E *p) {
  if constexpr (CPLX) {
    if constexpr (sizeof(E) == 8) {
      return *reinterpret_cast<const double2 *>(p);
    } else {
      const float2 f = *reinterpret_cast<const float2 *>(p);
      return double2{static_cast<double>(f.x), static_cast<double>(f.y)};
    }
  } else {
    return double2{static_cast<double>(*p), 0.0};
  }
}
// The same load with the non-temporal hint, for streams read exactly once
// (FP64 storage; the single-precision types take the plain load).
template <typename E, bool CPLX>
__device__ __forceinline__ double2 ld_elem_stream(const E *p) {
  typedef double v2d __attribute__((ext_vector_type(2)));
  if constexpr (CPLX && sizeof(E) == 8) {
    const v2d v = __builtin_nontemporal_load(reinterpret_cast<const v2d *>(p));
    return double2{v[0], v[1]};
  } else if constexpr (!CPLX && sizeof(E) == 8) {
    return double2{__builtin_nontemporal_load(reinterpret_cast<const double *>(p)), 0.0};
  } else {
    return ld_elem<E, CPLX>(p);
  }
}
// The stored form of one element as it sits in registers between its load and
// its LDS staging: double2 / float2 (complex), double / float (real).  The
// GEMM fetches keep this form and convert to FP64 only when staging, so a
// single-precision load is not followed by a conversion inside its guard,
// which would make every load wait for its data before the next one issues.
template <typename E, bool CPLX>
using raw_t = std::conditional_t<CPLX, std::conditional_t<sizeof(E) == 8, double2, float2>, E>;
template <typename E, bool CPLX>
__device__ __forceinline__ raw_t<E, CPLX> ld_raw(const E *p) {
  return *reinterpret_cast<const raw_t<E, CPLX> *>(p);
}
__device__ __forceinline__ double re_of(double2 v) { return v.x; }
__device__ __forceinline__ double re_of(float2 v) { return static_cast<double>(v.x); }
__device__ __forceinline__ double re_of(double v) { return v; }
__device__ __forceinline__ double re_of(float v) { return static_cast<double>(v); }
__device__ __forceinline__ double im_of(double2 v) { return v.y; }
__device__ __forceinline__ double im_of(float2 v) { return static_cast<double>(v.y); }
__device__ __forceinline__ double im_of(double) { return 0.0; }
__device__ __forceinline__ double im_of(float) { return 0.0; }
template <typename E, bool CPLX>
__device__ __forceinline__ void st_elem(E *p, double re, double im) {
  if constexpr (CPLX) {
    if constexpr (sizeof(E) == 8) *reinterpret_cast<double2 *>(p) = double2{re, im};
    else *reinterpret_cast<float2 *>(p) = float2{static_cast<float>(re), static_cast<float>(im)};
  } else {
    *p = static_cast<E>(re);
  }
}

// Epilogue of one output element: C = alpha * x + beta * old, with `old` the
// element's earlier value, loaded by the caller beforehand (not read when
// beta == 0).  The epilogues load every old value of a thread first, at
// addresses clamped into the block, so the loads are in flight together
// rather than each waited on before its store.
template <typename E, bool CPLX>
__device__ __forceinline__ void st_scaled(E *cp, double xr, double xi, const raw_t<E, CPLX> &old, bool use_beta,
                                          double alpha_r, double alpha_i, double beta_r, double beta_i) {
  if constexpr (CPLX) {
    double yr = alpha_r * xr - alpha_i * xi, yi = alpha_r * xi + alpha_i * xr;
    if (use_beta) {
      const double o_r = re_of(old), o_i = im_of(old);
      yr += beta_r * o_r - beta_i * o_i;
      yi += beta_r * o_i + beta_i * o_r;
    }
    st_elem<E, true>(cp, yr, yi);
  } else {
    double y = alpha_r * xr;
    if (use_beta) y += beta_r * re_of(old);
    st_elem<E, false>(cp, y, 0.0);
  }
}

// Epilogue of an MFMA output tile: element (mt, nt, v) of a thread sits at
// row r0 + 16 mt + 4 v, column c0 + 16 nt.  With beta == 0 the stores go out
// back to back; otherwise every old value is loaded first (addresses clamped
// to the block's first element) and the stores follow.  The two cases are
// separate code: merged, the waits for the loads would also hold back the
// stores of the beta == 0 case, since stores count in the same counter.
// ACC is one 16 x 16 sub-tile's 4 values of a thread (v4d, or double[4]).
template <bool CPLX, int BM, int BN, typename E, typename ACC>
__device__ __forceinline__ void tile_epilogue(const Seg &seg, int r0, int c0, const ACC (&acc)[BM][BN],
                                              const ACC (&acc_i)[BM][BN], E *__restrict__ C, double alpha_r,
                                              double alpha_i, double beta_r, double beta_i) {
  constexpr int P = CPLX ? 2 : 1;
  auto store_all = [&](auto with_beta, const raw_t<E, CPLX>(&old)[BM][BN][4]) {
    constexpr bool WB = decltype(with_beta)::value;
#pragma unroll
    for (int mt = 0; mt < BM; ++mt)
#pragma unroll
      for (int nt = 0; nt < BN; ++nt)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int row = r0 + mt * 16 + 4 * v, col = c0 + nt * 16;
          if (row < seg.m && col < seg.n) {
            const int64_t idx = seg.c_off + static_cast<int64_t>(row) * seg.n + col;
            st_scaled<E, CPLX>(C + P * idx, acc[mt][nt][v], CPLX ? acc_i[mt][nt][v] : 0.0, old[mt][nt][v], WB,
                               alpha_r, alpha_i, beta_r, beta_i);
          }
        }
  };
  if (beta_r != 0.0 || beta_i != 0.0) {
    raw_t<E, CPLX> old[BM][BN][4];
#pragma unroll
    for (int mt = 0; mt < BM; ++mt)
#pragma unroll
      for (int nt = 0; nt < BN; ++nt)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int row = r0 + mt * 16 + 4 * v, col = c0 + nt * 16;
          const bool ok = row < seg.m && col < seg.n;
          old[mt][nt][v] = ld_raw<E, CPLX>(C + P * (seg.c_off + (ok ? static_cast<int64_t>(row) * seg.n + col : 0)));
        }
    store_all(std::true_type{}, old);
  } else {
    const raw_t<E, CPLX> none[BM][BN][4] = {};
    store_all(std::false_type{}, none);
  }
}

constexpr int KC = kKChunk;   // plan step: k-chunk of 16
constexpr double kPanelStreamBytes = 256.0 * 1024 * 1024;   // the MI355X last-level (Infinity) cache
constexpr int kRealUnits = 1; // plan steps staged per pipeline step (FP64); 2 halves the occupancy and measured slower

// Tile order for the XCDs: workgroups are dispatched to the 8 XCDs round
// robin, so workgroup b runs on XCD b % 8.  Mapping b to logical tile
// (r / G) * 8G + (b % 8) * G + r % G, with r = b / 8, gives each XCD runs of G
// consecutive tiles (neighbouring tiles of one output block, which share A
// rows or B columns) while keeping the global longest-first order at a
// granularity of 8G tiles; the tail beyond the last full 8G group maps to
// itself.
constexpr int kXcds = 8;
constexpr int kXcdRun = 4;
__device__ __forceinline__ int xcd_tile(int b, int n) {
  constexpr int span = kXcds * kXcdRun;
  if (b >= (n / span) * span) return b;
  const int r = b / kXcds;
  return (r / kXcdRun) * span + (b % kXcds) * kXcdRun + (r % kXcdRun);
}

// Grouped MFMA kernel, one 256-thread workgroup per TM x TN output tile
// (TM, TN in {32, 64}); the four waves split the tile 2 x 2, each wave
// computing (TM/2) x (TN/2) as BM x BN blocks of v_mfma_f64_16x16x4f64.
// Per-thread gather of the A' (TM x 16) and B' (16 x TN) k-chunk:
//   A' k-contiguous -> p = tid & 15, rows (tid >> 4) + 16 q;
//      otherwise    -> r = tid % TM, cols tid / TM + (256 / TM) q;
//   B' n-contiguous -> c = tid % TN, rows tid / TN + (256 / TN) q;
//      otherwise    -> p = tid & 15, cols (tid >> 4) + 16 q.
template <bool CPLX, int TMv, int TNv, typename E = double>
__global__ __launch_bounds__(256) void k_grouped_gemm_f64(
    const Tile *__restrict__ tiles, const Seg *__restrict__ segs, const ExecTask *__restrict__ exec,
    const BlkGather *__restrict__ ga, const BlkGather *__restrict__ gb, const GatherLayout L,
    const E *__restrict__ A, const E *__restrict__ B, E *__restrict__ C, double *__restrict__ part,
    double alpha_r, double alpha_i, double beta_r, double beta_i) {
  constexpr int P = CPLX ? 2 : 1;
  // k extent per pipeline step: U plan steps of 16 (FP64 stages 32 k per
  // step; complex128, whose LDS tiles are twice as large, stages 16)
  constexpr int U = CPLX ? 1 : kRealUnits;
  constexpr int KCv = KC * U;
  constexpr int QA0 = TMv / 16, QB0 = TNv / 16;        // elements per thread per 16 k
  constexpr int QA = QA0 * U, QB = QB0 * U;            // elements per thread per step
  constexpr int BM = TMv / 32, BN = TNv / 32;          // MFMA blocks per wave
  __shared__ double As[2][P][KCv][TMv + 1];
  __shared__ double Bs[2][P][KCv][TNv + 1];

  const Tile tile = tiles[xcd_tile(static_cast<int>(blockIdx.x), static_cast<int>(gridDim.x))];
  const Seg seg = segs[tile.seg];
  const int i0 = tile.tm * TMv, j0 = tile.tn * TNv;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wm = (w >> 1) * (TMv / 2), wn = (w & 1) * (TNv / 2);
  const int lr = lane & 15, lk = lane >> 4;
  // thread mapping of the chunk being fetched, chosen per task: plain
  // blocks always read along memory order
  bool akf = L.a_kfast != 0, bnf = L.b_nfast != 0;

  v4d acc[BM][BN], acc_i[BM][BN], acc_s[BM][BN];
#pragma unroll
  for (int a = 0; a < BM; ++a)
#pragma unroll
    for (int b = 0; b < BN; ++b) {
      acc[a][b] = v4d{0.0, 0.0, 0.0, 0.0};
      acc_i[a][b] = v4d{0.0, 0.0, 0.0, 0.0};
      acc_s[a][b] = v4d{0.0, 0.0, 0.0, 0.0};
    }

  auto a_r = [&](int q) { return akf ? (tid >> 4) + 16 * (q % QA0) : tid % TMv; };
  auto a_p = [&](int q) { return akf ? (tid & 15) + 16 * (q / QA0) : tid / TMv + (256 / TMv) * q; };
  auto b_c = [&](int q) { return bnf ? tid % TNv : (tid >> 4) + 16 * (q % QB0); };
  auto b_p = [&](int q) { return bnf ? tid / TNv + (256 / TNv) * q : (tid & 15) + 16 * (q / QB0); };

  // Per task: the row offsets of the A' elements and column offsets of the
  // B' elements this thread loads.  A block whose permuted view is plain
  // (row i, col p at i * ncols + p) needs no decomposition at all.
  int32_t a_row[QA], b_col[QB];
  bool a_plain = false, b_plain = false;
  raw_t<E, CPLX> pa[QA], pb[QB];
  // every task of this output block reads plain blocks: no descriptor reads
  const bool seg_plain = (seg.flags & (kSegAPlain | kSegBPlain)) == (kSegAPlain | kSegBPlain);
  auto task_rows = [&](const BlkGather &gA, const BlkGather &gB, const ExecTask &et) {
    a_plain = seg_plain || gA.plain != 0;
    b_plain = seg_plain || gB.plain != 0;
    akf = a_plain || gA.fastcol != 0;
    bnf = b_plain || gB.fastcol != 0;
#pragma unroll
    for (int q = 0; q < QA; ++q) {
      const int r = i0 + a_r(q);
      a_row[q] = (r < seg.m) ? (a_plain ? r * et.k : split_offset(r, gA.rdim, gA.rstr, gA.rmag, gA.rsh, L.a_nr)) : -1;
    }
#pragma unroll
    for (int q = 0; q < QB; ++q) {
      const int c = j0 + b_c(q);
      b_col[q] = (c < seg.n) ? (b_plain ? c : split_offset(c, gB.cdim, gB.cstr, gB.cmag, gB.csh, L.b_nc)) : -1;
    }
  };
  auto fetch = [&](const ExecTask &et, const BlkGather &gA, const BlkGather &gB, int kk, int kend) {
#pragma unroll
    for (int q = 0; q < QA; ++q) {
      const int pk = kk + a_p(q);
      const int32_t ac = (pk < kend) ? (a_plain ? pk : split_offset(pk, gA.cdim, gA.cstr, gA.cmag, gA.csh, L.a_nc)) : -1;
      const bool ok = (a_row[q] >= 0) && (ac >= 0);
      const int64_t e = et.a_off + a_row[q] + ac;
      pa[q] = ok ? ld_raw<E, CPLX>(A + e * P) : raw_t<E, CPLX>{};
    }
#pragma unroll
    for (int q = 0; q < QB; ++q) {
      const int pk = kk + b_p(q);
      const int32_t br = (pk < kend) ? (b_plain ? pk * seg.n : split_offset(pk, gB.rdim, gB.rstr, gB.rmag, gB.rsh, L.b_nr)) : -1;
      const bool ok = (b_col[q] >= 0) && (br >= 0);
      const int64_t e = et.b_off + br + b_col[q];
      pb[q] = ok ? ld_raw<E, CPLX>(B + e * P) : raw_t<E, CPLX>{};
    }
  };
  auto stage = [&](int buf) {
#pragma unroll
    for (int q = 0; q < QA; ++q) {
      As[buf][0][a_p(q)][a_r(q)] = re_of(pa[q]);
      if (CPLX) As[buf][P - 1][a_p(q)][a_r(q)] = im_of(pa[q]);
    }
#pragma unroll
    for (int q = 0; q < QB; ++q) {
      Bs[buf][0][b_p(q)][b_c(q)] = re_of(pb[q]);
      if (CPLX) Bs[buf][P - 1][b_p(q)][b_c(q)] = im_of(pb[q]);
    }
  };

  // ---- MFMAs on one staged step ----
  // Work outside the output block is skipped at the MFMA's own granularity:
  // 16 x 16 sub-blocks of the wave that lie wholly beyond m or n, and k
  // steps of 4 beyond the step's valid extent `kv` (all wave-uniform).
  bool mt_ok[BM], nt_ok[BN];
#pragma unroll
  for (int mt = 0; mt < BM; ++mt) mt_ok[mt] = i0 + wm + mt * 16 < seg.m;
#pragma unroll
  for (int nt = 0; nt < BN; ++nt) nt_ok[nt] = j0 + wn + nt * 16 < seg.n;
  auto compute = [&](int buf, int kv) {
#pragma unroll
    for (int ks = 0; ks < KCv; ks += 4) {
      if (ks >= kv) break;
      double a[BM], b[BN];
#pragma unroll
      for (int mt = 0; mt < BM; ++mt) a[mt] = As[buf][0][ks + lk][wm + mt * 16 + lr];
#pragma unroll
      for (int nt = 0; nt < BN; ++nt) b[nt] = Bs[buf][0][ks + lk][wn + nt * 16 + lr];
#pragma unroll
      for (int mt = 0; mt < BM; ++mt)
#pragma unroll
        for (int nt = 0; nt < BN; ++nt)
          if (mt_ok[mt] && nt_ok[nt])
            acc[mt][nt] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[mt], b[nt], acc[mt][nt], 0, 0, 0);
      if (CPLX) {
        double ai[BM], bi[BN];
#pragma unroll
        for (int mt = 0; mt < BM; ++mt) ai[mt] = As[buf][P - 1][ks + lk][wm + mt * 16 + lr];
#pragma unroll
        for (int nt = 0; nt < BN; ++nt) bi[nt] = Bs[buf][P - 1][ks + lk][wn + nt * 16 + lr];
        // three real products per complex one (Gauss): acc = sum ar*br (above),
        // acc_i = sum ai*bi, acc_s = sum (ar + ai)*(br + bi); after the last
        // step re = acc - acc_i and im = acc_s - acc - acc_i
        double as[BM], bs[BN];
#pragma unroll
        for (int mt = 0; mt < BM; ++mt) as[mt] = a[mt] + ai[mt];
#pragma unroll
        for (int nt = 0; nt < BN; ++nt) bs[nt] = b[nt] + bi[nt];
#pragma unroll
        for (int mt = 0; mt < BM; ++mt)
#pragma unroll
          for (int nt = 0; nt < BN; ++nt) {
            if (!(mt_ok[mt] && nt_ok[nt])) continue;
            acc_i[mt][nt] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[mt], bi[nt], acc_i[mt][nt], 0, 0, 0);
            acc_s[mt][nt] = __builtin_amdgcn_mfma_f64_16x16x4f64(as[mt], bs[nt], acc_s[mt][nt], 0, 0, 0);
          }
      }
    }
  };

  // ---- pipeline over this tile's steps of the (task, k-chunk) sequence ----
  // A step stages up to U plan steps of 16 k, never crossing a task or the
  // end of the tile's slice; k beyond the step's end loads as zero.
  auto span_of = [&](const ExecTask &e, int kk_, int left_) {
    return min(U, min((e.k - kk_ + KC - 1) / KC, left_));
  };

  const bool seg_affine = (seg.flags & (kSegAColAffine | kSegBRowAffine)) == (kSegAColAffine | kSegBRowAffine);
  if (seg_affine && U == 1) {
    // Output block whose every task reads A and B blocks that are affine in
    // k: A' element (r, p) at a_off + roff(r) + p * cs, B' element (p, c) at
    // b_off + p * rs + coff(c) (plain blocks: roff = r*k, cs = 1, rs = n,
    // coff = c).  Each thread keeps one pointer per element it loads, set
    // once per task (the row / column offsets decomposed there) and advanced
    // by 16 k per step; row and column bounds are fixed for the tile, and the
    // k bound only matters on a task's last, partial step.
    const int ap = tid & 15, ar0 = tid >> 4;
    const int bc = tid % TNv, bp0 = tid / TNv;
    constexpr int BPS = 256 / TNv;                     // B' rows between a thread's elements
    bool aok[QA], bok[QB];
    const E *aptr[QA], *bptr[QB];
    int64_t a_adv = 0, b_adv = 0;
#pragma unroll
    for (int q = 0; q < QA; ++q) aok[q] = i0 + ar0 + 16 * q < seg.m;
#pragma unroll
    for (int q = 0; q < QB; ++q) bok[q] = j0 + bc < seg.n;
    auto setup = [&](const ExecTask &e, int kk0) {
      const BlkGather &gA = ga[e.a_blk];
      const BlkGather &gB = gb[e.b_blk];
      const bool pa_ = seg_plain || gA.plain != 0, pb_ = seg_plain || gB.plain != 0;
      const int64_t cs = pa_ ? 1 : gA.cstep;
      const int64_t rs = pb_ ? seg.n : gB.rstep;
      a_adv = KC * cs * P;
      b_adv = KC * rs * P;
#pragma unroll
      for (int q = 0; q < QA; ++q) {
        const int r = aok[q] ? i0 + ar0 + 16 * q : 0;
        const int64_t roff = pa_ ? static_cast<int64_t>(r) * e.k
                                 : split_offset(r, gA.rdim, gA.rstr, gA.rmag, gA.rsh, L.a_nr);
        aptr[q] = A + (e.a_off + roff + (ap + kk0) * cs) * P;
      }
#pragma unroll
      for (int q = 0; q < QB; ++q) {
        const int c = bok[q] ? j0 + bc : 0;
        const int64_t coff = pb_ ? c : split_offset(c, gB.cdim, gB.cstr, gB.cmag, gB.csh, L.b_nc);
        bptr[q] = B + (e.b_off + static_cast<int64_t>(bp0 + BPS * q + kk0) * rs + coff) * P;
      }
    };
    auto pfetch = [&](int kvalid) {
      const bool full = kvalid >= KC;
#pragma unroll
      for (int q = 0; q < QA; ++q) {
        const bool ok = aok[q] && (full || ap < kvalid);
        pa[q] = ok ? ld_raw<E, CPLX>(aptr[q]) : raw_t<E, CPLX>{};
        aptr[q] += a_adv;
      }
#pragma unroll
      for (int q = 0; q < QB; ++q) {
        const bool ok = bok[q] && (full || bp0 + BPS * q < kvalid);
        pb[q] = ok ? ld_raw<E, CPLX>(bptr[q]) : raw_t<E, CPLX>{};
        bptr[q] += b_adv;
      }
    };
    auto pstage = [&](int buf) {
#pragma unroll
      for (int q = 0; q < QA; ++q) {
        As[buf][0][ap][ar0 + 16 * q] = re_of(pa[q]);
        if (CPLX) As[buf][P - 1][ap][ar0 + 16 * q] = im_of(pa[q]);
      }
#pragma unroll
      for (int q = 0; q < QB; ++q) {
        Bs[buf][0][bp0 + BPS * q][bc] = re_of(pb[q]);
        if (CPLX) Bs[buf][P - 1][bp0 + BPS * q][bc] = im_of(pb[q]);
      }
    };
    int t = tile.t_begin, kk = tile.k_begin, left = tile.n_chunks;
    ExecTask et = exec[t];
    setup(et, kk);
    int kv = min(KC, et.k - kk);
    pfetch(kv);
    pstage(0);
    __syncthreads();
    int buf = 0;
    while (left > 0) {
      const int left2 = left - 1;
      int t2 = t, kk2 = kk + KC;
      if (kk2 >= et.k) {
        t2 = t + 1;
        kk2 = 0;
      }
      const bool has_next = left2 > 0;
      ExecTask et2 = et;
      int kv2 = 0;
      if (has_next) {
        if (t2 != t) {
          et2 = exec[t2];
          setup(et2, 0);
        }
        kv2 = min(KC, et2.k - kk2);
        pfetch(kv2);
      }
      compute(buf, kv);
      if (has_next) pstage(buf ^ 1);
      __syncthreads();
      t = t2;
      kk = kk2;
      et = et2;
      kv = kv2;
      left = left2;
      buf ^= 1;
    }
  } else {
    int t = tile.t_begin, kk = tile.k_begin, left = tile.n_chunks;
    ExecTask et = exec[t];
    task_rows(ga[et.a_blk], gb[et.b_blk], et);
    int sp = span_of(et, kk, left);
    fetch(et, ga[et.a_blk], gb[et.b_blk], kk, min(et.k, kk + KC * sp));
    stage(0);
    __syncthreads();
    int buf = 0;
    while (left > 0) {
      const int left2 = left - sp;
      int t2 = t, kk2 = kk + KC * sp;
      if (kk2 >= et.k) {
        t2 = t + 1;
        kk2 = 0;
      }
      const bool has_next = left2 > 0;
      ExecTask et2 = et;
      int sp2 = 0;
      if (has_next) {
        if (t2 != t) {
          et2 = exec[t2];
          task_rows(ga[et2.a_blk], gb[et2.b_blk], et2);
        }
        sp2 = span_of(et2, kk2, left2);
        fetch(et2, ga[et2.a_blk], gb[et2.b_blk], kk2, min(et2.k, kk2 + KC * sp2));
      }
      compute(buf, min(et.k - kk, KC * sp));
      if (has_next) stage(buf ^ 1);
      __syncthreads();
      t = t2;
      kk = kk2;
      et = et2;
      left = left2;
      sp = sp2;
      buf ^= 1;
    }
  }

  // ---- complex: the real and imaginary parts from the three product sums ----
  if (CPLX) {
#pragma unroll
    for (int mt = 0; mt < BM; ++mt)
#pragma unroll
      for (int nt = 0; nt < BN; ++nt) {
        const v4d p1 = acc[mt][nt], p2 = acc_i[mt][nt];
        acc[mt][nt] = p1 - p2;
        acc_i[mt][nt] = acc_s[mt][nt] - p1 - p2;
      }
  }

  // ---- a k-slice stores its unscaled partial sum, in the thread's own order ----
  if (tile.part >= 0) {
    double *pp = part + static_cast<int64_t>(tile.part) * (TMv * TNv * P);
#pragma unroll
    for (int mt = 0; mt < BM; ++mt)
#pragma unroll
      for (int nt = 0; nt < BN; ++nt)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int e = ((mt * BN + nt) * 4 + v) * 256 + tid;
          pp[e] = acc[mt][nt][v];
          if (CPLX) pp[TMv * TNv + e] = acc_i[mt][nt][v];
        }
    return;
  }

  // ---- epilogue: C = alpha * acc + beta * C (C untouched when beta == 0) ----
  // f64 16x16x4 output: row 4*reg + lane/16
  tile_epilogue<CPLX, BM, BN, E>(seg, i0 + wm + lk, j0 + wn + lr, acc, acc_i, C, alpha_r, alpha_i, beta_r, beta_i);
}

// Sum of the k-slices of one output tile, in slice order, then the usual
// epilogue.  The thread-to-element map is the MFMA kernel's, so thread tid
// reads exactly the partial elements it would have held.
template <bool CPLX, int TMv, int TNv, typename E = double>
__global__ __launch_bounds__(256) void k_reduce_slices(const ReduceItem *__restrict__ items,
                                                       const Seg *__restrict__ segs,
                                                       const double *__restrict__ part, E *__restrict__ C,
                                                       double alpha_r, double alpha_i, double beta_r, double beta_i) {
  constexpr int P = CPLX ? 2 : 1;
  constexpr int BM = TMv / 32, BN = TNv / 32;
  const ReduceItem it = items[blockIdx.x];
  const Seg seg = segs[it.seg];
  const int i0 = it.tm * TMv, j0 = it.tn * TNv;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wm = (w >> 1) * (TMv / 2), wn = (w & 1) * (TNv / 2);
  const int lr = lane & 15, lk = lane >> 4;
  // slices outer, elements inner: every slice's loads are independent
  double sr[BM][BN][4], si[BM][BN][4];
#pragma unroll
  for (int mt = 0; mt < BM; ++mt)
#pragma unroll
    for (int nt = 0; nt < BN; ++nt)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        sr[mt][nt][v] = 0.0;
        si[mt][nt][v] = 0.0;
      }
  for (int s = 0; s < it.n_parts; ++s) {
    const double *pp = part + static_cast<int64_t>(it.part0 + s) * (TMv * TNv * P);
#pragma unroll
    for (int mt = 0; mt < BM; ++mt)
#pragma unroll
      for (int nt = 0; nt < BN; ++nt)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int e = ((mt * BN + nt) * 4 + v) * 256 + tid;
          sr[mt][nt][v] += pp[e];
          if (CPLX) si[mt][nt][v] += pp[TMv * TNv + e];
        }
  }
  tile_epilogue<CPLX, BM, BN, E>(seg, i0 + wm + lk, j0 + wn + lr, sr, si, C, alpha_r, alpha_i, beta_r, beta_i);
}

// Row-streaming kernel for narrow outputs (n <= NMAX): one thread per output
// row, RPT rows per thread (rows tid, tid + 256, ...).  A' is read straight
// from global memory -- consecutive rows are consecutive threads, so the
// loads are contiguous whenever A's innermost axis is a saved one, and with
// RPT independent rows each thread keeps several loads in flight.  The
// 16 x n chunk of B' is staged once per workgroup in LDS and every thread
// accumulates its n outputs with FP64 FMAs.  Bound by reading A and writing
// C once.
constexpr int RK = 16;            // k per stage
constexpr int kRowsLoadGroup = 16;   // A loads in flight per thread

template <bool CPLX, int NMAX, int RPT, typename E>
__device__ __forceinline__ void rows_epilogue(const Seg &seg, int r0, int tid, const double (&acc)[RPT][NMAX],
                                              const double (&acc_i)[RPT][NMAX], E *__restrict__ C,
                                              double alpha_r, double alpha_i, double beta_r, double beta_i) {
  constexpr int P = CPLX ? 2 : 1;
  const int n = seg.n;
  // beta == 0 and beta != 0 as separate code, as in tile_epilogue
  auto store_all = [&](auto with_beta, const raw_t<E, CPLX>(&old)[RPT][NMAX]) {
    constexpr bool WB = decltype(with_beta)::value;
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
      const int row = r0 + tid + 256 * q;
      if (row >= seg.m) continue;
      const int64_t base = seg.c_off + static_cast<int64_t>(row) * n;
#pragma unroll
      for (int j = 0; j < NMAX; ++j) {
        if (j < n) {
          st_scaled<E, CPLX>(C + P * (base + j), acc[q][j], CPLX ? acc_i[q][j] : 0.0, old[q][j], WB,
                             alpha_r, alpha_i, beta_r, beta_i);
        }
      }
    }
  };
  if (beta_r != 0.0 || beta_i != 0.0) {
    raw_t<E, CPLX> old[RPT][NMAX];
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
      const int row = r0 + tid + 256 * q;
#pragma unroll
      for (int j = 0; j < NMAX; ++j) {
        const bool ok = row < seg.m && j < n;
        old[q][j] = ld_raw<E, CPLX>(C + P * (seg.c_off + (ok ? static_cast<int64_t>(row) * n + j : 0)));
      }
    }
    store_all(std::true_type{}, old);
  } else {
    const raw_t<E, CPLX> none[RPT][NMAX] = {};
    store_all(std::false_type{}, none);
  }
}

template <bool CPLX, int NMAX, int RPT, typename E = double>
__global__ __launch_bounds__(256) void k_rows_f64(
    const Tile *__restrict__ tiles, const Seg *__restrict__ segs, const ExecTask *__restrict__ exec,
    const BlkGather *__restrict__ ga, const BlkGather *__restrict__ gb, const GatherLayout L,
    const E *__restrict__ A, const E *__restrict__ B, E *__restrict__ C, double *__restrict__ part,
    double alpha_r, double alpha_i, double beta_r, double beta_i) {
  constexpr int P = CPLX ? 2 : 1;
  static_assert(RK * NMAX <= 256, "one B' element per thread");
  static_assert(RK == kKChunk, "row kernel steps are the plan's k-chunks");
  __shared__ double Bs[P][RK][NMAX];
  __shared__ int32_t s_acol[RK], s_brow[RK], s_bcol[NMAX];

  const Tile tile = tiles[xcd_tile(static_cast<int>(blockIdx.x), static_cast<int>(gridDim.x))];
  const Seg seg = segs[tile.seg];
  const int tid = threadIdx.x;
  const int r0 = tile.tm * (256 * RPT);
  const int n = seg.n;

  double acc[RPT][NMAX], acc_i[RPT][NMAX];
#pragma unroll
  for (int q = 0; q < RPT; ++q)
#pragma unroll
    for (int j = 0; j < NMAX; ++j) {
      acc[q][j] = 0.0;
      acc_i[q][j] = 0.0;
    }

  // this tile's steps of the (task, k-chunk) sequence
  int t = tile.t_begin, kk = tile.k_begin;
  ExecTask et{};
  bool a_plain = false, b_plain = false;
  int32_t roff[RPT];
  int cur_task = -1;
  for (int step = 0; step < tile.n_chunks; ++step) {
    if (t != cur_task) {
      cur_task = t;
      et = exec[t];
      const BlkGather &gA = ga[et.a_blk];
      const BlkGather &gB = gb[et.b_blk];
      a_plain = gA.plain != 0;
      b_plain = gB.plain != 0;
#pragma unroll
      for (int q = 0; q < RPT; ++q) {
        const int row = r0 + tid + 256 * q;
        roff[q] = (row < seg.m) ? (a_plain ? row * et.k : split_offset(row, gA.rdim, gA.rstr, gA.rmag, gA.rsh, L.a_nr)) : -1;
      }
      __syncthreads();
      if (tid < NMAX)
        s_bcol[tid] = (tid < n) ? (b_plain ? tid : split_offset(tid, gB.cdim, gB.cstr, gB.cmag, gB.csh, L.b_nc)) : -1;
    }
    const BlkGather &gA = ga[et.a_blk];
    const BlkGather &gB = gb[et.b_blk];
    if (tid < RK) {
      const int p = kk + tid;
      s_acol[tid] = (p < et.k) ? (a_plain ? p : split_offset(p, gA.cdim, gA.cstr, gA.cmag, gA.csh, L.a_nc)) : -1;
    } else if (tid >= 64 && tid < 64 + RK) {
      const int p = kk + tid - 64;
      s_brow[tid - 64] = (p < et.k) ? (b_plain ? p * n : split_offset(p, gB.rdim, gB.rstr, gB.rmag, gB.rsh, L.b_nr)) : -1;
    }
    __syncthreads();
    if (tid < RK * NMAX) {
      const int p = tid / NMAX, j = tid % NMAX;
      const int32_t ro = s_brow[p], co = s_bcol[j];
      const bool ok = (ro >= 0) && (co >= 0);
      const int64_t e = et.b_off + ro + co;
      const double2 v = ok ? ld_elem<E, CPLX>(B + e * P) : double2{0.0, 0.0};
      Bs[0][p][j] = v.x;
      if (CPLX) Bs[P - 1][p][j] = v.y;
    }
    __syncthreads();
    const int kc = min(RK, et.k - kk);
    // The A loads of G consecutive k steps go out together (G * RPT = 16 per
    // thread); a step beyond the chunk or a row beyond the block loads the
    // block's first element and counts zero.
#pragma unroll
    for (int p0 = 0; p0 < RK; p0 += kRowsLoadGroup / RPT) {
      constexpr int G = kRowsLoadGroup / RPT;
      if (p0 < kc) {
        double av[G][RPT], av_i[G][RPT];
#pragma unroll
        for (int g = 0; g < G; ++g) {
          const int32_t co = s_acol[p0 + g];
#pragma unroll
          for (int q = 0; q < RPT; ++q) {
            const bool ok = roff[q] >= 0 && co >= 0;
            const double2 v = ld_elem<E, CPLX>(A + (ok ? et.a_off + roff[q] + co : et.a_off) * P);
            av[g][q] = ok ? v.x : 0.0;
            av_i[g][q] = ok ? v.y : 0.0;
          }
        }
#pragma unroll
        for (int g = 0; g < G; ++g) {
          if (p0 + g >= kc) continue;
          const int p = p0 + g;
#pragma unroll
          for (int j = 0; j < NMAX; ++j) {
            if (j < n) {
              const double b = Bs[0][p][j];
              const double b_im = Bs[P - 1][p][j];
#pragma unroll
              for (int q = 0; q < RPT; ++q) {
                if (CPLX) {
                  acc[q][j] = fma(av[g][q], b, acc[q][j]);
                  acc[q][j] = fma(-av_i[g][q], b_im, acc[q][j]);
                  acc_i[q][j] = fma(av[g][q], b_im, acc_i[q][j]);
                  acc_i[q][j] = fma(av_i[g][q], b, acc_i[q][j]);
                } else {
                  acc[q][j] = fma(av[g][q], b, acc[q][j]);
                }
              }
            }
          }
        }
      }
    }
    __syncthreads();
    kk += RK;
    if (kk >= et.k) {
      ++t;
      kk = 0;
    }
  }

  if (tile.part >= 0) {
    // k-slice: unscaled partial sums in the thread's own order
    double *pp = part + static_cast<int64_t>(tile.part) * (256 * RPT * NMAX * P);
#pragma unroll
    for (int q = 0; q < RPT; ++q)
#pragma unroll
      for (int j = 0; j < NMAX; ++j) {
        const int e = (q * NMAX + j) * 256 + tid;
        pp[e] = acc[q][j];
        if (CPLX) pp[256 * RPT * NMAX + e] = acc_i[q][j];
      }
    return;
  }
  rows_epilogue<CPLX, NMAX, RPT, E>(seg, r0, tid, acc, acc_i, C, alpha_r, alpha_i, beta_r, beta_i);
}

// Sum of the k-slices of one row tile, in slice order, then the epilogue.
template <bool CPLX, int NMAX, int RPT, typename E = double>
__global__ __launch_bounds__(256) void k_reduce_rows(const ReduceItem *__restrict__ items,
                                                     const Seg *__restrict__ segs, const double *__restrict__ part,
                                                     E *__restrict__ C, double alpha_r, double alpha_i,
                                                     double beta_r, double beta_i) {
  constexpr int P = CPLX ? 2 : 1;
  constexpr int SZ = 256 * RPT * NMAX * P;
  const ReduceItem it = items[blockIdx.x];
  const Seg seg = segs[it.seg];
  const int tid = threadIdx.x;
  double acc[RPT][NMAX], acc_i[RPT][NMAX];
#pragma unroll
  for (int q = 0; q < RPT; ++q)
#pragma unroll
    for (int j = 0; j < NMAX; ++j) {
      acc[q][j] = 0.0;
      acc_i[q][j] = 0.0;
    }
  // Each slice's slab holds the full RPT x NMAX layout, so every load is in
  // bounds; columns j >= n are loaded and dropped, which keeps the loads of a
  // slice in flight together.
  for (int s = 0; s < it.n_parts; ++s) {
    const double *pp = part + static_cast<int64_t>(it.part0 + s) * SZ;
#pragma unroll
    for (int q = 0; q < RPT; ++q)
#pragma unroll
      for (int j = 0; j < NMAX; ++j) {
        const int e = (q * NMAX + j) * 256 + tid;
        const bool ok = j < seg.n;
        const double x = pp[e];
        acc[q][j] += ok ? x : 0.0;
        if (CPLX) {
          const double xi = pp[256 * RPT * NMAX + e];
          acc_i[q][j] += ok ? xi : 0.0;
        }
      }
  }
  rows_epilogue<CPLX, NMAX, RPT, E>(seg, it.tm * (256 * RPT), tid, acc, acc_i, C, alpha_r, alpha_i, beta_r, beta_i);
}

// Row panels (PanelTile): one 512-row tile of a panel per workgroup, 2 rows
// per thread.  While each thread loads the A' elements of its rows in every
// column of the panel, the panel's terms (their B' elements, one thread
// each) and output columns are staged in LDS; then each output column is
// summed from registers in the row kernel's order and stored with its alpha
// / beta epilogue.  Replaces, for these blocks, the per-task loop of
// global_operations.h:919-978 (one BLAS call per task).
template <bool CPLX, bool STREAM, typename E = double>
__global__ __launch_bounds__(256) void k_panel_rows(const PanelTile *__restrict__ tiles,
                                                    const PanelOut *__restrict__ outs,
                                                    const PanelTerm *__restrict__ terms, const Seg *__restrict__ segs,
                                                    const E *__restrict__ A, const E *__restrict__ B,
                                                    E *__restrict__ C, double alpha_r, double alpha_i,
                                                    double beta_r, double beta_i) {
  constexpr int P = CPLX ? 2 : 1;
  constexpr int RPT = kPanelRpt;
  static_assert(kPanelRows == 512 && RPT == 2, "pair staging: 4 loads per thread");
  __shared__ double s_br[kPanelMaxTerms], s_bi[kPanelMaxTerms];
  __shared__ double s_pa[2][kPanelRows], s_pai[CPLX ? 2 : 1][kPanelRows];
  __shared__ int64_t s_ocoff[kPanelMaxOuts];
  __shared__ int32_t s_on[kPanelMaxOuts], s_ot0[kPanelMaxOuts], s_omask[kPanelMaxOuts];
  static_assert(kPanelMaxTerms <= 128 && kPanelMaxOuts <= 128, "one staging thread per term and per output");

  const PanelTile &pt = tiles[blockIdx.x];
  const int tid = threadIdx.x;
  const int n_cols = pt.n_cols, n_outs = pt.n_outs, n_terms = pt.n_terms;

  // row offsets in every column: the row's digits over rdim (axes past the
  // saved ones have extent 1 and stride 0), dotted with each column's
  // strides.  Rows past the block read its first row (in bounds); their
  // sums are never stored, and the columns past n_cols are never read.
  int32_t roff[kPanelCols][RPT];
  bool valid[RPT];
#pragma unroll
  for (int q = 0; q < RPT; ++q) {
    const int row = pt.r0 + tid + 256 * q;
    valid[q] = row < pt.m;
    int32_t x = valid[q] ? row : 0;
#pragma unroll
    for (int c = 0; c < kPanelCols; ++c) roff[c][q] = 0;
#pragma unroll
    for (int l = kPanelRowAxes - 1; l >= 0; --l) {
      const int32_t qd = static_cast<int32_t>((static_cast<uint64_t>(x) * pt.rmag[l]) >> pt.rsh[l]);
      const int32_t d = x - qd * pt.rdim[l];
      x = qd;
#pragma unroll
      for (int c = 0; c < kPanelCols; ++c) roff[c][q] += d * pt.cstr[c][l];
    }
  }
  // every A' load is issued before anything waits on memory; each A'
  // element is read once, so when the panels' data exceed the last-level
  // cache (STREAM) the loads carry the non-temporal hint.  A pair
  // of columns that are adjacent elements of one A block (pt.pair1 - 1 and
  // pt.pair1) is loaded as (row, column) = (e / 2, e % 2) over the lanes, so
  // that one instruction reads both columns of 32 consecutive rows, and
  // reaches its rows' threads through LDS.
  const int pc = pt.pair1 - 1;   // first column of the pair, or -1
  double av[kPanelCols][RPT], av_i[kPanelCols][RPT];
  if (pc >= 0) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int rl = (tid >> 1) + 128 * j, cc = tid & 1;
      const int row = pt.r0 + rl;
      int32_t x = row < pt.m ? row : 0, off = 0;
#pragma unroll
      for (int l = kPanelRowAxes - 1; l >= 0; --l) {
        const int32_t qd = static_cast<int32_t>((static_cast<uint64_t>(x) * pt.rmag[l]) >> pt.rsh[l]);
        off += (x - qd * pt.rdim[l]) * pt.cstr[pc][l];
        x = qd;
      }
      const E *ap = A + (pt.cbase[pc] + off + cc) * P;
      const double2 v = STREAM ? ld_elem_stream<E, CPLX>(ap) : ld_elem<E, CPLX>(ap);
      s_pa[cc][rl] = v.x;
      if (CPLX) s_pai[CPLX ? cc : 0][rl] = v.y;
    }
  }
#pragma unroll
  for (int c = 0; c < kPanelCols; ++c)
    if (c < n_cols && (pc < 0 || (c != pc && c != pc + 1))) {
#pragma unroll
      for (int q = 0; q < RPT; ++q) {
        const E *ap = A + (pt.cbase[c] + roff[c][q]) * P;
        const double2 v = STREAM ? ld_elem_stream<E, CPLX>(ap) : ld_elem<E, CPLX>(ap);
        av[c][q] = v.x;
        av_i[c][q] = v.y;
      }
    }
  // stage the terms (B' elements) and the output columns
  if (tid < n_terms) {
    const PanelTerm tm = terms[pt.term0 + tid];
    const double2 v = ld_elem<E, CPLX>(B + tm.b * P);
    s_br[tid] = v.x;
    s_bi[tid] = v.y;
  } else if (tid >= 128 && tid < 128 + n_outs) {
    const int o = tid - 128;
    const PanelOut po = outs[pt.out0 + o];
    const Seg &sg = segs[po.seg];
    s_ocoff[o] = sg.c_off + po.j;
    s_on[o] = sg.n;
    s_ot0[o] = po.term0;
    s_omask[o] = po.mask;
  }
  __syncthreads();
  // the paired columns of this thread's rows
#pragma unroll
  for (int c = 0; c < kPanelCols; ++c)
    if (pc >= 0 && (c == pc || c == pc + 1)) {
      const int cc = c - pc;
#pragma unroll
      for (int q = 0; q < RPT; ++q) {
        av[c][q] = s_pa[cc][tid + 256 * q];
        av_i[c][q] = CPLX ? s_pai[CPLX ? cc : 0][tid + 256 * q] : 0.0;
      }
    }

  // The output loop exists twice, for beta == 0 and beta != 0: with the
  // loads of the old values inside a shared loop, every output column waited
  // for the previous column's stores (stores count in the same counter).
  auto out_loop = [&](auto with_beta) {
    constexpr bool WB = decltype(with_beta)::value;
    for (int o = 0; o < n_outs; ++o) {
      const int64_t coff = s_ocoff[o];
      const int32_t n = __builtin_amdgcn_readfirstlane(s_on[o]);
      const int32_t mask = __builtin_amdgcn_readfirstlane(s_omask[o]);
      int32_t t = __builtin_amdgcn_readfirstlane(s_ot0[o]);
      double y[RPT], y_i[RPT];
#pragma unroll
      for (int q = 0; q < RPT; ++q) {
        y[q] = 0.0;
        y_i[q] = 0.0;
      }
      // the output's terms in ascending column order, which is their (task, k)
      // order: the column loop is static, so no register array is indexed
#pragma unroll
      for (int c = 0; c < kPanelCols; ++c) {
        if (!((mask >> c) & 1)) continue;
        const double b = s_br[t], b_im = s_bi[t];
        ++t;
#pragma unroll
        for (int q = 0; q < RPT; ++q) {
          if (CPLX) {
            y[q] = fma(av[c][q], b, y[q]);
            y[q] = fma(-av_i[c][q], b_im, y[q]);
            y_i[q] = fma(av[c][q], b_im, y_i[q]);
            y_i[q] = fma(av_i[c][q], b, y_i[q]);
          } else {
            y[q] = fma(av[c][q], b, y[q]);
          }
        }
      }
      raw_t<E, CPLX> old[RPT] = {};
      if constexpr (WB) {
#pragma unroll
        for (int q = 0; q < RPT; ++q)
          old[q] = ld_raw<E, CPLX>(C + P * (coff + (valid[q] ? static_cast<int64_t>(pt.r0 + tid + 256 * q) * n : 0)));
      }
#pragma unroll
      for (int q = 0; q < RPT; ++q) {
        if (!valid[q]) continue;
        const int64_t e = coff + static_cast<int64_t>(pt.r0 + tid + 256 * q) * n;
        st_scaled<E, CPLX>(C + P * e, y[q], CPLX ? y_i[q] : 0.0, old[q], WB, alpha_r, alpha_i, beta_r, beta_i);
      }
    }
  };
  if (beta_r != 0.0 || beta_i != 0.0) out_loop(std::true_type{});
  else out_loop(std::false_type{});
}

}  // namespace

namespace qlt {

// Native grouped execution of a plan: MFMA tiles for the first tile class,
// the row-streaming kernel for the second.  `mfma_class` = false skips the
// MFMA class (the split path runs it on its own kernel).
qlt_status launch_native_classes(const qlt_plan *plan, int32_t dtype, const void *a, const void *b, void *c,
                                 const double *alpha, const double *beta, hipStream_t st, bool mfma_class,
                                 bool probe, ClassStreams *shared) {
  const bool cplx = dtype_is_complex(dtype);
  // classes with work run on the class streams when there are two or more
  int n_cls = 0;
  for (int32_t cls = mfma_class ? 0 : kNumMfmaClasses; cls < kNumTileClasses; ++cls)
    n_cls += plan->cls_tiles[cls] > 0 || (cls == kTileRows4 && plan->n_ptiles > 0);
  std::unique_ptr<ClassStreams> own;
  ClassStreams *cs = shared;
  if (cs == nullptr && n_cls >= 2) {
    own = std::make_unique<ClassStreams>(st);
    cs = own.get();
  }
  auto stream_of = [&](int32_t cls) {
    if (cls == kNumTileClasses) cls = kTileRows4;   // the row panels
    return cs != nullptr ? cs->get(cls) : st;
  };
  // the class with the most algorithmic work is the one a kernel probe times
  // (kPanelId: the row panels, on the stream of the n <= 4 row class)
  constexpr int32_t kPanelId = kNumTileClasses;
  int32_t dom = -1;
  if (probe && probe_armed(st)) {
    double best = 0.0;
    for (int32_t cls = mfma_class ? 0 : kNumMfmaClasses; cls < kNumTileClasses; ++cls)
      if (plan->cls_tiles[cls] > 0 && plan->cls_mkn[cls] > best) {
        best = plan->cls_mkn[cls];
        dom = cls;
      }
    if (plan->n_ptiles > 0 && plan->panel_mkn > best) dom = kPanelId;
  }
  auto probe_in = [&](int32_t cls) {
    if (cls == dom) probe_begin(stream_of(cls));
  };
  auto probe_out = [&](int32_t cls) {
    if (cls != dom) return;
    const hipStream_t ps = stream_of(cls);
    static const char *const kNames[kNumTileClasses + 1] = {
        "k_grouped_gemm_f64<64x64>", "k_grouped_gemm_f64<32x64>", "k_grouped_gemm_f64<64x32>",
        "k_grouped_gemm_f64<32x32>", "k_rows_f64<n<=4>",          "k_rows_f64<n<=16>",
        "k_panel_rows"};
    char name[64];
    static const char *const kTypes[4] = {"f64", "c128", "f32", "c64"};
    std::snprintf(name, sizeof(name), "%s %s", kNames[cls], kTypes[dtype & 3]);
    const double mkn = cls == kPanelId ? plan->panel_mkn : plan->cls_mkn[cls];
    const double elems = cls == kPanelId ? plan->panel_elems : plan->cls_elems[cls];
    probe_end(ps, name, (cplx ? 8.0 : 2.0) * mkn, static_cast<double>(dtype_bytes(dtype)) * elems);
  };
  const double ar = alpha[0], ai = cplx ? alpha[1] : 0.0;
  const double br = beta[0], bi = cplx ? beta[1] : 0.0;
  const dim3 block(256);
  // every kernel instance of one element storage type E (double, or float
  // for the single-precision element types) and complexity
  auto run = [&](auto e_tag, auto cplx_tag) {
    using E = decltype(e_tag);
    constexpr bool CX = decltype(cplx_tag)::value;
    const auto *A = static_cast<const E *>(a);
    const auto *B = static_cast<const E *>(b);
    auto *Cp = static_cast<E *>(c);
    // MFMA classes: 64x64, 32x64, 64x32, 32x32 (tiles stored class-major)
    int64_t base = 0, red_base = 0;
    for (int cls = 0; cls < kNumMfmaClasses; ++cls) {
      const int64_t cnt = plan->cls_tiles[cls];
      const int64_t n_red = plan->cls_red[cls];
      const Tile *tp = plan->d_tiles + base;
      const ReduceItem *rp = plan->d_red + red_base;
      base += cnt;
      red_base += n_red;
      if (!mfma_class || cnt == 0) continue;
      const dim3 grid(static_cast<unsigned>(cnt));
      const hipStream_t ls = stream_of(cls);
      // this class's region of the k-slice partial sums
      double *const cls_part = plan->d_part != nullptr ? plan->d_part + plan->cls_part_off[cls] : nullptr;
      // k-sliced tiles: the slices write partial sums, then one reduction
      // workgroup per sliced tile adds them and applies alpha / beta
      auto mfma = [&](auto tm_tag, auto tn_tag) {
        constexpr int TMV = decltype(tm_tag)::value, TNV = decltype(tn_tag)::value;
        probe_in(cls);
        hipLaunchKernelGGL((k_grouped_gemm_f64<CX, TMV, TNV, E>), grid, block, 0, ls, tp, plan->d_segs, plan->d_exec,
                           plan->d_ga, plan->d_gb, plan->layout, A, B, Cp, cls_part, ar, ai, br, bi);
        probe_out(cls);
        if (n_red > 0)
          hipLaunchKernelGGL((k_reduce_slices<CX, TMV, TNV, E>), dim3(static_cast<unsigned>(n_red)), block, 0, ls, rp,
                             plan->d_segs, cls_part, Cp, ar, ai, br, bi);
      };
      using I32 = std::integral_constant<int, 32>;
      using I64 = std::integral_constant<int, 64>;
      if (cls == 0) mfma(I64{}, I64{});
      else if (cls == 1) mfma(I32{}, I64{});
      else if (cls == 2) mfma(I64{}, I32{});
      else mfma(I32{}, I32{});
    }
    // row classes (n <= 4, n <= 16), each followed by its k-slice reduction
    for (int32_t cls = kTileRows4; cls <= kTileRows16; ++cls) {
      const int64_t cnt = plan->cls_tiles[cls];
      const int64_t n_red = plan->cls_red[cls];
      const Tile *tp = plan->d_tiles + base;
      const ReduceItem *rp = plan->d_red + red_base;
      base += cnt;
      red_base += n_red;
      if (cnt == 0) continue;
      const dim3 grid(static_cast<unsigned>(cnt)), rgrid(static_cast<unsigned>(n_red > 0 ? n_red : 1));
      const hipStream_t ls = stream_of(cls);
      // this class's region of the k-slice partial sums
      double *const cls_part = plan->d_part != nullptr ? plan->d_part + plan->cls_part_off[cls] : nullptr;
      auto rows = [&](auto nmax_tag, auto rpt_tag) {
        constexpr int NMAXV = decltype(nmax_tag)::value, RPTV = decltype(rpt_tag)::value;
        probe_in(cls);
        hipLaunchKernelGGL((k_rows_f64<CX, NMAXV, RPTV, E>), grid, block, 0, ls, tp, plan->d_segs, plan->d_exec,
                           plan->d_ga, plan->d_gb, plan->layout, A, B, Cp, cls_part, ar, ai, br, bi);
        probe_out(cls);
        if (n_red > 0)
          hipLaunchKernelGGL((k_reduce_rows<CX, NMAXV, RPTV, E>), rgrid, block, 0, ls, rp, plan->d_segs, cls_part, Cp,
                             ar, ai, br, bi);
      };
      if (cls == kTileRows4) rows(std::integral_constant<int, 4>{}, std::integral_constant<int, 4>{});
      else rows(std::integral_constant<int, 16>{}, std::integral_constant<int, 1>{});
    }
    // row panels, after the n <= 4 row class on its stream
    if (plan->n_ptiles > 0) {
      const hipStream_t ls = stream_of(kPanelId);
      probe_in(kPanelId);
      // non-temporal A' loads only when the panels' A' and C exceed the
      // last-level cache: smaller ones are served from it (measured: a
      // 0.83 GB W contraction 0.169 -> 0.148 ms with the hint, a small one
      // 0.028 -> 0.032 ms)
      const bool stream_a = plan->panel_elems * static_cast<double>(sizeof(E) * (CX ? 2 : 1)) > kPanelStreamBytes;
      if (stream_a)
        hipLaunchKernelGGL((k_panel_rows<CX, true, E>), dim3(static_cast<unsigned>(plan->n_ptiles)), block, 0, ls,
                           plan->d_ptiles, plan->d_pouts, plan->d_pterms, plan->d_segs, A, B, Cp, ar, ai, br, bi);
      else
        hipLaunchKernelGGL((k_panel_rows<CX, false, E>), dim3(static_cast<unsigned>(plan->n_ptiles)), block, 0, ls,
                           plan->d_ptiles, plan->d_pouts, plan->d_pterms, plan->d_segs, A, B, Cp, ar, ai, br, bi);
      probe_out(kPanelId);
    }
  };
  switch (dtype) {
    case QLT_F64: run(double{}, std::false_type{}); break;
    case QLT_C128: run(double{}, std::true_type{}); break;
    case QLT_F32: run(float{}, std::false_type{}); break;
    default: run(float{}, std::true_type{}); break;
  }
  QLT_HIP(hipGetLastError());
  return QLT_OK;
}

qlt_status launch_native(const qlt_plan *plan, int32_t dtype, const void *a, const void *b, void *c,
                         const double *alpha, const double *beta, hipStream_t st) {
  return launch_native_classes(plan, dtype, a, b, c, alpha, beta, st, true, true, nullptr);
}

}  // namespace qlt
