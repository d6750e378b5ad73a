// Split-precision grouped block GEMM: FP64 (and complex128) emulated on the
// BF16 MFMA units, or on the int8 MFMA units (below).
//
// Same contract as the native kernel (it replaces the per-task MatMultiply
// loop of CtrctTwoBSDTAndAssignIn, global_operations.h:895-992); only the
// arithmetic differs.  Two steps per execute:
//
// 1. pack (k_pack_short for blocks with k_pad <= 1024, which are read once;
//    k_pack phases 1 and 2 for longer ones): every block the plan touches is
//    read through the same gather descriptors as the native kernel (so the
//    transposes stay folded), scaled per row of A' and per column of B' by a
//    power of two 2^e with max|x| / 2^e in [1/2, 1], and split by
//    round-to-nearest into S = 5 slices
//        y = x / 2^e,  c1 = rint(y * 2^8),  r1 = y - c1 2^-8,
//        c_s = rint(r_{s-1} * 2^(8 + 9(s-1))),  |c_s| <= 256,
//    so y = 2^-8 * sum_s c_s 2^(-9(s-1)) + r_5 with |r_5| <= 2^-45.  Slice s is stored as the
//    BF16 number c_s 2^(-9(s-1)) (integers |c| <= 256 have at most 8
//    significant bits, so every slice is exact in BF16).
// 2. GEMM: C' = sum over slice pairs with sa + sb <= S + 1 (15 products) of
//    A_sa B_sb on v_mfma_f32_32x32x16_bf16.  Products of level L = sa + sb
//    carry the weight 2^(-9(L-2)) in their BF16 values.  They go to four FP32
//    accumulators {L=2}, {L=3}, {L=4}, {L=5,6}; each of the first three is
//    flushed to FP64 before its partial sum can reach 2^24 ulps of its level
//    (so it stays exact), {L=5,6} together with {L=4}, and the flush sum of
//    the flushed groups is exact in FP64.  The flushed value is scaled by
//    2^(e_row + e_col - 16) with ldexp and added to the FP64 accumulator.
//    Truncating at level 6 leaves a relative error of order 2^-45.
// Complex128 is the 4-real-product form Cr = Ar Br - Ai Bi, Ci = Ar Bi + Ai Br
// on the same slices, with one exponent per row/column over both parts.
//
// The int8 form (QLT_MODE_SPLIT_I8) keeps the scaling and replaces the digits:
// the scaled value y is rounded to the integer Y = rint(y * 2^46) and written
// in balanced base 255 as six digits d_0..d_5 in [-127, 127],
//     Y = sum_s d_s 255^(5-s),
// which negate without overflow (the complex real part needs -Ai).  The
// GEMM runs v_mfma_i32_32x32x32_i8 on the 21 digit pairs with sa + sb <= 5
// (the dropped pairs weigh at most 2^-46 of the row and column scales) and
// sums every level L = sa + sb in its own int32 accumulator, which is exact
// for 8192 k (a complex k adds at most 12 * 127^2 to a level).  A flush
// recombines the six levels in FP64 by Horner's rule in 255, scales by
// 255^5 2^(e_row + e_col - 92) and adds to the FP64 accumulator.  The int8
// MFMA runs at twice the BF16 rate and a digit plane is half the bytes of a
// BF16 plane.

#include <algorithm>
#include <cstring>
#include <memory>
#include <type_traits>
#include <vector>

#include "qlt_internal.h"

using namespace qlt;

namespace {

typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef float v16f __attribute__((ext_vector_type(16)));
typedef uint32_t v4u __attribute__((ext_vector_type(4)));

constexpr int S = 5;          // slices per operand
constexpr int KC = 16;        // k per LDS stage (one MFMA k-step)
constexpr int TM = 64;        // tile rows
constexpr int TN = 64;        // tile cols
constexpr int RPAD = 64;      // packed rows are padded to a multiple of this
constexpr int LDK = KC + 8;   // padded LDS row (bf16)
constexpr int S8 = 6;         // int8 form: base-255 digits per element
constexpr int KC8 = 32;       // int8 form: k per LDS stage (one int8 MFMA k-step)
constexpr int kI8FlushK = 8192;   // int8 form: k between flushes of the int32 levels

// Packed layout of one block (rows R_pad = round_up(R, 64), k_pad = round_up(K, KC)):
//   plane element (part, s, r, p) at
//   (((part * S + s) * (k_pad / KC) + p / KC) * R_pad + r) * KC + p % KC
// with S = 5, KC = 16 and BF16 elements, or S = 6, KC = 32 and int8 digits.
struct PackMeta {
  int64_t plane_off;  // bf16 elements into the plane workspace
  int64_t exp_off;    // int32 elements into the exponent workspace
  int32_t rows;       // R (actual)
  int32_t r_pad;
  int32_t cols;       // K (actual)
  int32_t k_pad;
};

struct PackItem {
  int32_t side;       // 0: A block (rows = saved, cols = contracted); 1: B block transposed
  int32_t blk;        // block position
  int32_t r0;         // first row of this 64-row group
  int32_t k0;         // first k of this chunk (digit phase)
};

__device__ __forceinline__ uint16_t bf16_bits_exact(double v) {
  // v is c * 2^-m with |c| <= 256 an integer: exactly representable in BF16,
  // so truncating the FP32 bit pattern is exact.
  const float f = static_cast<float>(v);
  return static_cast<uint16_t>(__float_as_uint(f) >> 16);
}

// Source view of one pack item: A' (rows = saved, cols = contracted) or B'^T
// (rows = saved = cols of B's descriptor, cols = contracted = rows of B's
// descriptor).  A plain A block has row r, col p at r * cols + p (k
// contiguous); a plain B block seen transposed has row j, col p at
// p * rows + j (rows contiguous).
struct PackView {
  const double *src;
  int64_t base;
  int32_t rows, cols;
  const int32_t *rdim, *rstr, *cdim, *cstr, *rsh, *csh;
  const uint64_t *rmag, *cmag;
  int nr, nc;
  bool is_a, plain;
  bool kfast;   // whether consecutive k are consecutive in memory
  __device__ __forceinline__ int32_t row_off(int r) const {
    return plain ? (is_a ? r * cols : r) : split_offset(r, rdim, rstr, rmag, rsh, nr);
  }
  __device__ __forceinline__ int32_t col_off(int p) const {
    return plain ? (is_a ? p : p * rows) : split_offset(p, cdim, cstr, cmag, csh, nc);
  }
};

__device__ __forceinline__ PackView pack_view(bool is_a, const PackMeta &pm, const BlkGather &g, int64_t base,
                                              const GatherLayout &L, const double *src) {
  PackView v;
  v.src = src;
  v.base = base;
  v.rows = pm.rows;
  v.cols = pm.cols;
  v.rdim = is_a ? g.rdim : g.cdim;
  v.rstr = is_a ? g.rstr : g.cstr;
  v.cdim = is_a ? g.cdim : g.rdim;
  v.cstr = is_a ? g.cstr : g.rstr;
  v.rmag = is_a ? g.rmag : g.cmag;
  v.rsh = is_a ? g.rsh : g.csh;
  v.cmag = is_a ? g.cmag : g.rmag;
  v.csh = is_a ? g.csh : g.rsh;
  v.nr = is_a ? L.a_nr : L.b_nc;
  v.nc = is_a ? L.a_nc : L.b_nr;
  v.is_a = is_a;
  v.plain = g.plain != 0;
  v.kfast = v.plain ? is_a : (is_a ? (g.fastcol != 0) : (g.fastcol == 0));
  return v;
}

// Pack kernels, 256 threads per workgroup, in two launches.  Phase 1: one
// workgroup per (side, block, 64-row group) finds the max |x| of each row
// over all k and writes the row exponents.  Phase 2: one workgroup per
// (side, block, 64-row group, k-chunk of KP) splits its chunk into BF16
// slices or (I8) int8 digits and writes them, so the digit work of a long
// block spreads over many workgroups.
template <bool CPLX, bool I8, int PHASE>
__global__ __launch_bounds__(256) void k_pack(const PackItem *__restrict__ items,
                                              const PackMeta *__restrict__ meta_a,
                                              const PackMeta *__restrict__ meta_b,
                                              const BlkGather *__restrict__ ga,
                                              const BlkGather *__restrict__ gb,
                                              const int64_t *__restrict__ a_offs,
                                              const int64_t *__restrict__ b_offs, const GatherLayout L,
                                              const double *__restrict__ A, const double *__restrict__ B,
                                              void *__restrict__ planes_v, int32_t *__restrict__ exps) {
  constexpr int KP = CPLX ? 32 : 64;   // k per phase-2 chunk
  constexpr int KCF = I8 ? KC8 : KC;   // k per packed chunk
  constexpr int PARTS = CPLX ? 2 : 1;
  __shared__ int32_t s_roff[64];
  __shared__ int32_t s_coff[64];
  __shared__ double s_max[64];
  __shared__ double s_max4[4][64];
  __shared__ double s_tile[PARTS][64][KP + 1];

  const PackItem it = items[blockIdx.x];
  const bool is_a = (it.side == 0);
  const PackMeta pm = is_a ? meta_a[it.blk] : meta_b[it.blk];
  const PackView v = pack_view(is_a, pm, is_a ? ga[it.blk] : gb[it.blk], is_a ? a_offs[it.blk] : b_offs[it.blk], L,
                               is_a ? A : B);
  const double *src = v.src;
  const int64_t base = v.base;
  const bool kfast = v.kfast;
  const int tid = threadIdx.x;
  auto row_off = [&](int r) { return v.row_off(r); };
  auto col_off = [&](int p) { return v.col_off(p); };

  if (tid < 64) {
    const int r = it.r0 + tid;
    s_roff[tid] = (r < pm.rows) ? row_off(r) : -1;
  }
  __syncthreads();

  __shared__ int32_t s_exp[64];
  if constexpr (PHASE == 1) {
  // ---- pass 1: row maxima ----
  // Each thread keeps running maxima of the elements it reads; no atomics.
  // k contiguous: wave w reads 64 consecutive k of rows w, w + 4, ..., w + 60
  // (one row per load, coalesced) and reduces each row's maximum across its
  // lanes at the end.  Rows contiguous: lane l reads row l at k = w, w + 4,
  // ... and the four waves' maxima meet in LDS.
  auto absmax = [&](int64_t e) {
    if (CPLX) {
      const double2 v = reinterpret_cast<const double2 *>(src)[e];
      return fmax(fabs(v.x), fabs(v.y));
    }
    return fabs(src[e]);
  };
  const int lane = tid & 63, wv = tid >> 6;
  double mx[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) mx[q] = 0.0;
  for (int p0 = 0; p0 < pm.cols; p0 += 64) {
    if (tid < 64) {
      const int p = p0 + tid;
      s_coff[tid] = (p < pm.cols) ? col_off(p) : -1;
    }
    __syncthreads();
    // the 16 loads go out together: an element outside the block reads row
    // r0, k p0 (inside it) and is not counted
    const int64_t e0 = base + s_roff[0] + s_coff[0];
    if (kfast) {
      const int32_t co = s_coff[lane];
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int32_t ro = s_roff[wv + 4 * q];
        const bool ok = ro >= 0 && co >= 0;
        const double x = absmax(ok ? base + ro + co : e0);
        mx[q] = fmax(mx[q], ok ? x : 0.0);   // the value is used either way, so the load is not sunk into a branch
      }
    } else {
      const int32_t ro = s_roff[lane];
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int32_t co = s_coff[wv + 4 * q];
        const bool ok = ro >= 0 && co >= 0;
        const double x = absmax(ok ? base + ro + co : e0);
        mx[0] = fmax(mx[0], ok ? x : 0.0);
      }
    }
    __syncthreads();
  }
  if (kfast) {
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      double m = mx[q];
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
      if (lane == 0) s_max[wv + 4 * q] = m;
    }
  } else {
    s_max4[wv][lane] = mx[0];
  }
  __syncthreads();
  if (!kfast && tid < 64)
    s_max[tid] = fmax(fmax(s_max4[0][tid], s_max4[1][tid]), fmax(s_max4[2][tid], s_max4[3][tid]));
  __syncthreads();

  // exponent of each row: max / 2^e in [1/2, 1]
  if (tid < 64) {
    const double m = s_max[tid];
    int e = 0;
    if (m > 0.0) (void)frexp(m, &e);
    const int r = it.r0 + tid;
    if (r < pm.r_pad) exps[pm.exp_off + r] = e;
  }
  return;
  }

  // ---- phase 2: split and write one chunk (including the zero padding) ----
  if (tid < 64) s_exp[tid] = exps[pm.exp_off + it.r0 + tid];   // r0 + 63 < r_pad
  const int nkc = pm.k_pad / KCF;
  {
    const int p0 = it.k0;
    if (tid < KP) {
      const int p = p0 + tid;
      s_coff[tid] = (p < pm.cols) ? col_off(p) : -1;
    }
    __syncthreads();
    // all loads go out before the first LDS store: an element outside the
    // block reads row r0, k p0 (inside it) and is stored as zero
    constexpr int NQ = (64 * KP) / 256;
    using Elem = std::conditional_t<CPLX, double2, double>;
    const int64_t e0 = base + s_roff[0] + s_coff[0];
    Elem xs[NQ];
    bool oks[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int r = kfast ? tid / KP + (256 / KP) * q : tid & 63;
      const int p = kfast ? tid % KP : (tid >> 6) + 4 * q;
      const int32_t ro = s_roff[r], co = s_coff[p];
      oks[q] = (ro >= 0) && (co >= 0);
      xs[q] = reinterpret_cast<const Elem *>(src)[oks[q] ? base + ro + co : e0];
    }
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int r = kfast ? tid / KP + (256 / KP) * q : tid & 63;
      const int p = kfast ? tid % KP : (tid >> 6) + 4 * q;
      if constexpr (CPLX) {
        s_tile[0][r][p] = oks[q] ? xs[q].x : 0.0;
        s_tile[PARTS - 1][r][p] = oks[q] ? xs[q].y : 0.0;
      } else {
        s_tile[0][r][p] = oks[q] ? xs[q] : 0.0;
      }
    }
    __syncthreads();
    // thread -> (row, run of KP/4 consecutive k)
    constexpr int RUN = KP / 4;
    const int r = tid >> 2;
    const int pr = (tid & 3) * RUN;
    const int grow = it.r0 + r;
    if (I8 && grow < pm.r_pad) {
      int8_t *planes = static_cast<int8_t *>(planes_v);
      // x 2^(46 - e) is exact (a power of two) and below 2^46 in magnitude;
      // it is applied by ldexp on each element, since 2^(46 - e) alone
      // overflows for rows whose maximum is below 2^-977.
      // Digits from the least significant up, all exact in FP64: q is the
      // integer nearest to Y / 255 (the product's error, below 2^-14, stays
      // far inside the 1/510 distance of Y / 255 from a half integer), so
      // Y - 255 q is the balanced digit in [-127, 127].
      const int sc = 46 - s_exp[r];
#pragma unroll
      for (int part = 0; part < PARTS; ++part) {
        uint32_t wd[S8][RUN / 4];
#pragma unroll
        for (int sl = 0; sl < S8; ++sl)
#pragma unroll
          for (int v = 0; v < RUN / 4; ++v) wd[sl][v] = 0u;
#pragma unroll
        for (int u = 0; u < RUN; ++u) {
          double Y = rint(ldexp(s_tile[part][r][pr + u], sc));
#pragma unroll
          for (int sl = S8 - 1; sl >= 0; --sl) {
            const double q = rint(Y * (1.0 / 255.0));
            const int d = static_cast<int>(fma(-q, 255.0, Y));
            wd[sl][u >> 2] |= static_cast<uint32_t>(d & 0xff) << (8 * (u & 3));
            Y = q;
          }
        }
        const int p = p0 + pr;
        if (p < pm.k_pad) {
          // RUN (8 or 16) consecutive k never cross a KC8 boundary
#pragma unroll
          for (int sl = 0; sl < S8; ++sl) {
            const int64_t idx =
                pm.plane_off +
                ((static_cast<int64_t>(part * S8 + sl) * nkc + p / KC8) * pm.r_pad + grow) * KC8 + (p % KC8);
            if constexpr (RUN == 16) {
              *reinterpret_cast<v4u *>(planes + idx) = v4u{wd[sl][0], wd[sl][1], wd[sl][2], wd[sl][3]};
            } else {
              *reinterpret_cast<uint2 *>(planes + idx) = uint2{wd[sl][0], wd[sl][1]};
            }
          }
        }
      }
    } else if (!I8 && grow < pm.r_pad) {
      uint16_t *planes = static_cast<uint16_t *>(planes_v);
      const int inv = -s_exp[r];   // ldexp per element: 2^-e overflows for subnormal maxima
#pragma unroll
      for (int part = 0; part < PARTS; ++part) {
        uint16_t out[S][RUN];
#pragma unroll
        for (int u = 0; u < RUN; ++u) {
          double y = ldexp(s_tile[part][r][pr + u], inv);   // exact (power of two)
#pragma unroll
          for (int s = 0; s < S; ++s) {
            const double scale = ldexp(1.0, 8 + 9 * s);
            const double c = rint(y * scale);           // |c| <= 256
            y -= c / scale;                             // exact
            out[s][u] = bf16_bits_exact(c * ldexp(1.0, -9 * s));
          }
        }
        const int p = p0 + pr;
        if (p < pm.k_pad) {
          // RUN consecutive k never cross a KC boundary (RUN divides KC or equals it)
#pragma unroll
          for (int s = 0; s < S; ++s) {
            const int64_t idx =
                pm.plane_off +
                ((static_cast<int64_t>(part * S + s) * nkc + p / KC) * pm.r_pad + grow) * KC + (p % KC);
            uint16_t *dst = planes + idx;
#pragma unroll
            for (int u = 0; u < RUN; u += 8) {
              v4u w;
              w[0] = static_cast<uint32_t>(out[s][u + 0]) | (static_cast<uint32_t>(out[s][u + 1]) << 16);
              w[1] = static_cast<uint32_t>(out[s][u + 2]) | (static_cast<uint32_t>(out[s][u + 3]) << 16);
              w[2] = static_cast<uint32_t>(out[s][u + 4]) | (static_cast<uint32_t>(out[s][u + 5]) << 16);
              w[3] = static_cast<uint32_t>(out[s][u + 6]) | (static_cast<uint32_t>(out[s][u + 7]) << 16);
              *reinterpret_cast<v4u *>(dst + u) = w;
            }
          }
        }
      }
    }
    __syncthreads();
  }
}

// Single-read pack of the blocks whose k_pad is at most kShortK.  One
// workgroup of 256 threads per (side, block, group of RW real rows) stages
// the rows' whole k extent in LDS (the block is read once), then the
// TPR = k_pad / 8 consecutive threads of a row take the row maximum and split
// 8 consecutive k each.  RW = floor(256 / TPR), so at least half and usually
// nearly all lanes carry data; a row may straddle waves.  The zero padding
// of the planes (rows R .. R_pad) and of the exponents is written once, when
// the workspace is prepared, and never touched here.  LDS rows hold one pad
// double per 8 so that the 8-element runs are read without bank conflicts.
constexpr int kShortK = 1024;
template <bool CPLX, bool I8>
__global__ __launch_bounds__(256) void k_pack_short(const PackItem *__restrict__ items,
                                                    const PackMeta *__restrict__ meta_a,
                                                    const PackMeta *__restrict__ meta_b,
                                                    const BlkGather *__restrict__ ga,
                                                    const BlkGather *__restrict__ gb,
                                                    const int64_t *__restrict__ a_offs,
                                                    const int64_t *__restrict__ b_offs, const GatherLayout L,
                                                    const double *__restrict__ A, const double *__restrict__ B,
                                                    void *__restrict__ planes_v, int32_t *__restrict__ exps) {
  constexpr int PARTS = CPLX ? 2 : 1;
  constexpr int RUN = 8;                         // consecutive k per thread
  constexpr int ELEMS = 256 * RUN;               // elements per part per workgroup
  constexpr int KCF = I8 ? KC8 : KC;
  constexpr int kMaxRows = 256 / (KC / RUN);     // k_pad >= KC
  __shared__ int32_t s_roff[kMaxRows];
  __shared__ int32_t s_coff[kShortK];
  __shared__ double s_tile[PARTS][ELEMS + ELEMS / 8 + kMaxRows];
  __shared__ double s_lead[256];

  const PackItem it = items[blockIdx.x];
  const bool is_a = (it.side == 0);
  const PackMeta pm = is_a ? meta_a[it.blk] : meta_b[it.blk];
  const PackView v = pack_view(is_a, pm, is_a ? ga[it.blk] : gb[it.blk], is_a ? a_offs[it.blk] : b_offs[it.blk], L,
                               is_a ? A : B);
  const int KP = pm.k_pad, TPR = KP / RUN, RW = 256 / TPR, lrow = KP + KP / 8 + 1;
  const int nrows = min(RW, pm.rows - it.r0);    // >= 1
  const int tid = threadIdx.x;
  if (tid < nrows) s_roff[tid] = v.row_off(it.r0 + tid);
  for (int p = tid; p < KP; p += 256) s_coff[p] = (p < pm.cols) ? v.col_off(p) : -1;
  __syncthreads();

  // ---- one read of the rows: element f = tid + 256 j of the nrows x KP
  // extent, k-major when consecutive k are consecutive in memory, else
  // row-major; (q, w) = (f / d, f % d) advance by (256 / d, 256 % d).  All
  // eight loads are issued before the first LDS store: an element outside the
  // block loads the block's first element instead and is stored as zero. ----
  {
    const int d = v.kfast ? KP : nrows;
    const int dq = 256 / d, dw = 256 - dq * d;
    int q = tid / d, w = tid - q * d;
    const int n = nrows * KP;
    const int64_t e0 = v.base + s_roff[0] + s_coff[0];   // row r0, k 0: inside the block
    using Elem = std::conditional_t<CPLX, double2, double>;
    Elem xs[RUN];
    int lps[RUN];
    bool oks[RUN];
#pragma unroll
    for (int j = 0; j < RUN; ++j) {
      const bool inside = tid + 256 * j < n;
      const int r = inside ? (v.kfast ? q : w) : 0, p = inside ? (v.kfast ? w : q) : 0;
      const int32_t co = s_coff[p];
      oks[j] = inside && co >= 0;
      lps[j] = inside ? r * lrow + p + (p >> 3) : -1;
      xs[j] = reinterpret_cast<const Elem *>(v.src)[oks[j] ? v.base + s_roff[r] + co : e0];
      q += dq;
      w += dw;
      if (w >= d) {
        w -= d;
        ++q;
      }
    }
#pragma unroll
    for (int j = 0; j < RUN; ++j) {
      if (lps[j] < 0) continue;
      if constexpr (CPLX) {
        s_tile[0][lps[j]] = oks[j] ? xs[j].x : 0.0;
        s_tile[PARTS - 1][lps[j]] = oks[j] ? xs[j].y : 0.0;
      } else {
        s_tile[0][lps[j]] = oks[j] ? xs[j] : 0.0;
      }
    }
  }
  __syncthreads();

  // ---- row maximum and exponent ----
  const int r = tid / TPR, run = tid - r * TPR, p0 = run * RUN;
  const bool active = r < nrows;
  const int lx = (active ? r : 0) * lrow + 9 * run;
  double x[PARTS][RUN];
  double m = 0.0;
#pragma unroll
  for (int part = 0; part < PARTS; ++part)
#pragma unroll
    for (int u = 0; u < RUN; ++u) {
      x[part][u] = active ? s_tile[part][lx + u] : 0.0;
      m = fmax(m, fabs(x[part][u]));
    }
  // Segmented suffix maximum inside the wave: after the step of offset o a
  // thread holds the maximum over its next 2o threads of the same row (and
  // wave), so the first thread of a row in each wave ends with the row's
  // maximum over that wave.
  const int lane = tid & 63;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const double y = __shfl_down(m, o, 64);
    if (run + o < TPR && lane + o < 64) m = fmax(m, y);
  }
  if (run == 0 || lane == 0) s_lead[tid] = m;
  __syncthreads();
  if (!active) return;
  const int first = r * TPR;                     // the row's threads: first .. first + TPR - 1 < 256
  double mx = s_lead[first];
  for (int wv = (first >> 6) + 1; wv <= ((first + TPR - 1) >> 6); ++wv) mx = fmax(mx, s_lead[wv << 6]);
  int ex = 0;
  if (mx > 0.0) (void)frexp(mx, &ex);
  const int grow = it.r0 + r;                    // < rows
  if (run == 0) exps[pm.exp_off + grow] = ex;

  // ---- split the run p0 .. p0 + 7 (inside one packed k-chunk) ----
  const int nkc = KP / KCF;
  if constexpr (I8) {
    int8_t *planes = static_cast<int8_t *>(planes_v);
    const int sc = 46 - ex;   // ldexp per element: 2^(46 - e) overflows below 2^-977
#pragma unroll
    for (int part = 0; part < PARTS; ++part) {
      uint32_t wd[S8][2] = {};
#pragma unroll
      for (int u = 0; u < RUN; ++u) {
        double Y = rint(ldexp(x[part][u], sc));
#pragma unroll
        for (int sl = S8 - 1; sl >= 0; --sl) {
          const double q = rint(Y * (1.0 / 255.0));
          const int d = static_cast<int>(fma(-q, 255.0, Y));
          wd[sl][u >> 2] |= static_cast<uint32_t>(d & 0xff) << (8 * (u & 3));
          Y = q;
        }
      }
#pragma unroll
      for (int sl = 0; sl < S8; ++sl) {
        const int64_t idx =
            pm.plane_off + ((static_cast<int64_t>(part * S8 + sl) * nkc + p0 / KC8) * pm.r_pad + grow) * KC8 + (p0 % KC8);
        *reinterpret_cast<uint2 *>(planes + idx) = uint2{wd[sl][0], wd[sl][1]};
      }
    }
  } else {
    uint16_t *planes = static_cast<uint16_t *>(planes_v);
    const int inv = -ex;   // ldexp per element: 2^-e overflows for subnormal maxima
#pragma unroll
    for (int part = 0; part < PARTS; ++part) {
      uint16_t out[S][RUN];
#pragma unroll
      for (int u = 0; u < RUN; ++u) {
        double y = ldexp(x[part][u], inv);           // exact (power of two)
#pragma unroll
        for (int s = 0; s < S; ++s) {
          const double scale = ldexp(1.0, 8 + 9 * s);
          const double c = rint(y * scale);           // |c| <= 256
          y -= c / scale;                             // exact
          out[s][u] = bf16_bits_exact(c * ldexp(1.0, -9 * s));
        }
      }
#pragma unroll
      for (int s = 0; s < S; ++s) {
        const int64_t idx =
            pm.plane_off + ((static_cast<int64_t>(part * S + s) * nkc + p0 / KC) * pm.r_pad + grow) * KC + (p0 % KC);
        v4u w;
        w[0] = static_cast<uint32_t>(out[s][0]) | (static_cast<uint32_t>(out[s][1]) << 16);
        w[1] = static_cast<uint32_t>(out[s][2]) | (static_cast<uint32_t>(out[s][3]) << 16);
        w[2] = static_cast<uint32_t>(out[s][4]) | (static_cast<uint32_t>(out[s][5]) << 16);
        w[3] = static_cast<uint32_t>(out[s][6]) | (static_cast<uint32_t>(out[s][7]) << 16);
        *reinterpret_cast<v4u *>(planes + idx) = w;
      }
    }
  }
}

__device__ __forceinline__ v8bf neg8(v8bf x) {
  v4u u = __builtin_bit_cast(v4u, x);
  u ^= 0x80008000u;
  return __builtin_bit_cast(v8bf, u);
}

// Self-contained tile and task records of the persistent split kernels, so
// that a tile or task boundary reads one record instead of a chain of
// tile -> segment -> task -> block metadata.
struct SplitTile {
  int64_t c_off;    // output block
  int32_t m, n;
  int32_t i0, j0;   // tile origin in the output block
  int32_t t0, t1;   // the block's tasks
};
struct SplitTask {
  int64_t a_plane, b_plane;   // plane offsets of the task's A and B blocks
  int32_t a_exp, b_exp;       // exponent offsets (a plan's exponents stay below 2^31)
  int32_t a_rpad, b_rpad;     // padded rows
  int32_t nkc;                // packed k-chunks
  int32_t pad;
};

// Record loads whose fields are pinned to scalar registers (the records are
// uniform across the workgroup).
__device__ __forceinline__ int64_t uniform64(int64_t x) {
  const uint64_t u = static_cast<uint64_t>(x);
  const uint32_t lo = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(u));
  const uint32_t hi = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(u >> 32));
  return static_cast<int64_t>((static_cast<uint64_t>(hi) << 32) | lo);
}
__device__ __forceinline__ SplitTile load_tile(const SplitTile *p, int i) {
  SplitTile r = p[i];
  r.c_off = uniform64(r.c_off);
  r.m = __builtin_amdgcn_readfirstlane(r.m);
  r.n = __builtin_amdgcn_readfirstlane(r.n);
  r.i0 = __builtin_amdgcn_readfirstlane(r.i0);
  r.j0 = __builtin_amdgcn_readfirstlane(r.j0);
  r.t0 = __builtin_amdgcn_readfirstlane(r.t0);
  r.t1 = __builtin_amdgcn_readfirstlane(r.t1);
  return r;
}
__device__ __forceinline__ SplitTask load_task(const SplitTask *p, int i) {
  SplitTask r = p[i];
  r.a_plane = uniform64(r.a_plane);
  r.b_plane = uniform64(r.b_plane);
  r.a_exp = __builtin_amdgcn_readfirstlane(r.a_exp);
  r.b_exp = __builtin_amdgcn_readfirstlane(r.b_exp);
  r.a_rpad = __builtin_amdgcn_readfirstlane(r.a_rpad);
  r.b_rpad = __builtin_amdgcn_readfirstlane(r.b_rpad);
  r.nkc = __builtin_amdgcn_readfirstlane(r.nkc);
  return r;
}

// Tile order for the XCDs: workgroup b runs on XCD b % 8.  Mapping b to
// logical tile (r / G) * 8G + (b % 8) * G + r % G, with r = b / 8, hands each
// XCD runs of G consecutive tiles of the (longest-first) tile list, i.e.
// neighbouring 64 x 64 tiles of one output block, which read the same A row
// panel or B column panel of slice planes and so find it in that XCD's L2;
// the tail beyond the last full 8G group maps to itself.
constexpr int kXcds = 8;
constexpr int kSplitXcdRun = 8;
__device__ __forceinline__ int split_xcd_tile(int b, int n) {
  constexpr int span = kXcds * kSplitXcdRun;
  if (b >= (n / span) * span) return b;
  const int r = b / kXcds;
  return (r / kSplitXcdRun) * span + (b % kXcds) * kSplitXcdRun + (r % kSplitXcdRun);
}

// Level group of a slice pair (sa, sb are 0-based, level = sa + sb + 2).
__device__ __forceinline__ constexpr int group_of(int sa, int sb) {
  return (sa + sb >= 3) ? 3 : (sa + sb);
}

// Real split GEMM: one 256-thread workgroup per 64 x 64 output tile, 2 x 2
// waves of 32 x 32, one 32x32x16 MFMA block per slice product.  The (task,
// k-chunk) sequence of the tile's output block is walked as one software
// pipeline: while the MFMAs consume the chunk staged in one LDS buffer, the
// next chunk's planes (possibly of the next task) are already in flight into
// registers, and are written to the other LDS buffer afterwards, so there is
// a single barrier per chunk and the load latency hides behind MFMA work.
// The FP32 level groups are flushed to FP64 every KCH = 64 k (levels 2..4
// stay exact) and at the end of every task (the exponents change).
__global__ __launch_bounds__(256) void k_split_gemm_f64(
    const Tile *__restrict__ tiles, const Seg *__restrict__ segs, const ExecTask *__restrict__ exec,
    const PackMeta *__restrict__ meta_a, const PackMeta *__restrict__ meta_b,
    const uint16_t *__restrict__ planes, const int32_t *__restrict__ exps, double *__restrict__ C,
    double alpha, double beta) {
  constexpr int KCH = 64;
  constexpr int CHUNKS = S * 128;                  // 16-byte pieces per operand per k-chunk
  constexpr int PER_THREAD = (CHUNKS + 255) / 256;
  __shared__ __attribute__((aligned(16))) uint16_t As[2][S][TM][LDK];
  __shared__ __attribute__((aligned(16))) uint16_t Bs[2][S][TN][LDK];
  __shared__ int32_t s_ea[2][TM];

  const Tile tile = tiles[split_xcd_tile(static_cast<int>(blockIdx.x), static_cast<int>(gridDim.x))];
  const Seg seg = segs[tile.seg];
  const int i0 = tile.tm * TM, j0 = tile.tn * TN;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
  const int lrow = lane & 31, lk = (lane >> 5) * 8;
  // a wave whose 32 x 32 block lies wholly outside the output block only
  // helps to stage operands (its MFMAs and flushes are skipped)
  const bool active = (i0 + wm < seg.m) && (j0 + wn < seg.n);

  double acc64[16];
#pragma unroll
  for (int r = 0; r < 16; ++r) acc64[r] = 0.0;
  v16f g[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) g[q] = v16f{};

  v4u pa[PER_THREAD], pb[PER_THREAD];
  int32_t ea_next = 0;
  // issue the global loads of chunk kc of the task described by (ma, mb)
  auto fetch = [&](const PackMeta &ma, const PackMeta &mb, int kc) {
    const int nkc = ma.k_pad / KC;
#pragma unroll
    for (int u = 0; u < PER_THREAD; ++u) {
      const int c = tid + 256 * u;
      if (c < CHUNKS) {
        const int ps = c >> 7, rr = (c & 127) >> 1, half = c & 1;
        const int64_t ia = ma.plane_off + ((static_cast<int64_t>(ps) * nkc + kc) * ma.r_pad + i0 + rr) * KC + half * 8;
        const int64_t ib = mb.plane_off + ((static_cast<int64_t>(ps) * nkc + kc) * mb.r_pad + j0 + rr) * KC + half * 8;
        pa[u] = *reinterpret_cast<const v4u *>(planes + ia);
        pb[u] = *reinterpret_cast<const v4u *>(planes + ib);
      }
    }
  };
  auto stage = [&](int buf) {
#pragma unroll
    for (int u = 0; u < PER_THREAD; ++u) {
      const int c = tid + 256 * u;
      if (c < CHUNKS) {
        const int ps = c >> 7, rr = (c & 127) >> 1, half = c & 1;
        *reinterpret_cast<v4u *>(&As[buf][ps][rr][half * 8]) = pa[u];
        *reinterpret_cast<v4u *>(&Bs[buf][ps][rr][half * 8]) = pb[u];
      }
    }
  };

  // ---- pipeline prologue: chunk 0 of the first task ----
  int t = seg.t0, kc = 0;
  PackMeta ma = meta_a[exec[t].a_blk], mb = meta_b[exec[t].b_blk];
  int nkc = ma.k_pad / KC;
  int32_t eb = exps[mb.exp_off + j0 + wn + lrow];
  fetch(ma, mb, 0);
  if (tid < TM) s_ea[0][tid] = exps[ma.exp_off + i0 + tid];
  stage(0);
  __syncthreads();
  int buf = 0, since_flush = 0;

  while (t < seg.t1) {
    // ---- prefetch the next chunk (same task or the next one) ----
    int t2 = t, kc2 = kc + 1;
    if (kc2 == nkc) {
      t2 = t + 1;
      kc2 = 0;
    }
    const bool has_next = t2 < seg.t1;
    PackMeta ma2 = ma, mb2 = mb;
    int32_t eb2 = eb;
    if (has_next) {
      if (t2 != t) {
        const ExecTask e2 = exec[t2];
        ma2 = meta_a[e2.a_blk];
        mb2 = meta_b[e2.b_blk];
        eb2 = exps[mb2.exp_off + j0 + wn + lrow];
        if (tid < TM) ea_next = exps[ma2.exp_off + i0 + tid];
      }
      fetch(ma2, mb2, kc2);
    }
    // ---- 15 slice products on buffer `buf` ----
    if (active) {
      v8bf bfr[S];
#pragma unroll
      for (int s2 = 0; s2 < S; ++s2) bfr[s2] = *reinterpret_cast<const v8bf *>(&Bs[buf][s2][wn + lrow][lk]);
#pragma unroll
      for (int sa = 0; sa < S; ++sa) {
        const v8bf ar = *reinterpret_cast<const v8bf *>(&As[buf][sa][wm + lrow][lk]);
#pragma unroll
        for (int sb = 0; sb < S - sa; ++sb) {
          const int q = group_of(sa, sb);
          g[q] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ar, bfr[sb], g[q], 0, 0, 0);
        }
      }
    }
    since_flush += KC;
    // ---- flush to FP64 (always at the end of a task: exponents change) ----
    if (active && (since_flush >= KCH || kc == nkc - 1)) {
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int row = wm + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        const int e = s_ea[buf][row] + eb - 16;
        const double sum = static_cast<double>(g[0][r]) + static_cast<double>(g[1][r]) +
                           static_cast<double>(g[2][r]) + static_cast<double>(g[3][r]);
        acc64[r] += ldexp(sum, e);
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) g[q] = v16f{};
      since_flush = 0;
    }
    // ---- stage the prefetched chunk into the other buffer ----
    if (has_next) {
      stage(buf ^ 1);
      if (tid < TM) s_ea[buf ^ 1][tid] = (t2 != t) ? ea_next : s_ea[buf][tid];
    }
    __syncthreads();
    if (t2 != t) {
      ma = ma2;
      mb = mb2;
      nkc = ma.k_pad / KC;
      eb = eb2;
    }
    t = t2;
    kc = kc2;
    buf ^= 1;
  }

  // ---- epilogue ----
  const int col = j0 + wn + lrow;
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int row = i0 + wm + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
    if (row < seg.m && col < seg.n) {
      const int64_t idx = seg.c_off + static_cast<int64_t>(row) * seg.n + col;
      double y = alpha * acc64[r];
      if (beta != 0.0) y += beta * C[idx];
      C[idx] = y;
    }
  }
}

// Complex split GEMM with the output components on separate waves, as a
// persistent kernel: one 512-thread workgroup per CU walks the tiles
// blockIdx.x, blockIdx.x + gridDim.x, ... of the 64 x 64 tile list.  Waves
// 0-3 accumulate the real part (ar*br - ai*bi) and waves 4-7 the imaginary
// part (ar*bi + ai*br) of the same 2 x 2 arrangement of 32 x 32 blocks, each
// over all 15 slice products, with the FP32 level groups of k_split_gemm_f64.
// Two waves per SIMD hide part of the load and barrier latency.  The
// (tile, task, k-chunk) sequence of the workgroup is one software pipeline:
// while the MFMAs consume one chunk from LDS, the next chunk's planes are in
// flight into registers, also across a tile boundary, so a tile's metadata
// and first chunk load behind the previous tile's last chunk.  When a tile
// ends, each wave stores its own output component.  The scalars are real: a
// complex alpha or beta is applied around the kernel (launch_split).
__global__ __launch_bounds__(512) void k_split_gemm_c2(
    const Tile *__restrict__ tiles, int32_t n_tiles, const Seg *__restrict__ segs,
    const ExecTask *__restrict__ exec, const PackMeta *__restrict__ meta_a, const PackMeta *__restrict__ meta_b,
    const uint16_t *__restrict__ planes, const int32_t *__restrict__ exps, double *__restrict__ C, double alpha,
    double beta) {
  constexpr int PARTS = 2;
  constexpr int CHUNKS = PARTS * S * 128;          // 16-byte pieces per operand per k-chunk
  constexpr int NT = 512;
  constexpr int PER_THREAD = (CHUNKS + NT - 1) / NT;
  __shared__ __attribute__((aligned(16))) uint16_t As[2][PARTS][S][TM][LDK];
  __shared__ __attribute__((aligned(16))) uint16_t Bs[2][PARTS][S][TN][LDK];
  __shared__ int32_t s_ea[2][TM];

  const int tid = threadIdx.x, lane = tid & 63, w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int ws = w & 3, comp = w >> 2;             // spatial block, output component
  const int wm = (ws >> 1) * 32, wn = (ws & 1) * 32;
  const int lrow = lane & 31, lk = (lane >> 5) * 8;
  const int G = static_cast<int>(gridDim.x);

  struct TileRef {
    int i0, j0;
    Seg seg;
  };
  auto tile_at = [&](int bidx) __attribute__((always_inline)) {
    TileRef r;
    const Tile tl = tiles[split_xcd_tile(bidx, n_tiles)];
    r.seg = segs[tl.seg];
    r.i0 = tl.tm * TM;
    r.j0 = tl.tn * TN;
    return r;
  };

  double acc64[16];
#pragma unroll
  for (int r = 0; r < 16; ++r) acc64[r] = 0.0;
  v16f g[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) g[q] = v16f{};

  v4u pa[PER_THREAD], pb[PER_THREAD];
  int32_t ea_next = 0;
  // Plane offsets of a thread's pieces within a task's k-chunk 0 (elements;
  // a block's planes stay below 2^31 elements), recomputed when the task
  // changes.  A chunk then is one uniform base pointer per operand plus these
  // offsets, so the per-chunk address work is scalar.
  uint32_t oa[PER_THREAD], ob[PER_THREAD];
  auto task_offsets = [&](const PackMeta &ma, const PackMeta &mb) __attribute__((always_inline)) {
    const int nkc = ma.k_pad / KC;
#pragma unroll
    for (int u = 0; u < PER_THREAD; ++u) {
      const int c = min(tid + NT * u, CHUNKS - 1);
      const int ps = c >> 7, rr = (c & 127) >> 1, half = c & 1;
      oa[u] = static_cast<uint32_t>((ps * nkc * ma.r_pad + rr) * KC + half * 8);
      ob[u] = static_cast<uint32_t>((ps * nkc * mb.r_pad + rr) * KC + half * 8);
    }
  };
  auto fetch = [&](const PackMeta &ma, const PackMeta &mb, int kc, int i0, int j0) __attribute__((always_inline)) {
    const uint16_t *ba = planes + ma.plane_off + (static_cast<int64_t>(kc) * ma.r_pad + i0) * KC;
    const uint16_t *bb = planes + mb.plane_off + (static_cast<int64_t>(kc) * mb.r_pad + j0) * KC;
#pragma unroll
    for (int u = 0; u < PER_THREAD; ++u) {
      if (tid + NT * u < CHUNKS) {
        pa[u] = *reinterpret_cast<const v4u *>(ba + oa[u]);
        pb[u] = *reinterpret_cast<const v4u *>(bb + ob[u]);
      }
    }
  };
  auto stage = [&](int buf) __attribute__((always_inline)) {
#pragma unroll
    for (int u = 0; u < PER_THREAD; ++u) {
      const int c = tid + NT * u;
      if (c < CHUNKS) {
        const int ps = c >> 7, rr = (c & 127) >> 1, half = c & 1;
        *reinterpret_cast<v4u *>(&As[buf][ps / S][ps % S][rr][half * 8]) = pa[u];
        *reinterpret_cast<v4u *>(&Bs[buf][ps / S][ps % S][rr][half * 8]) = pb[u];
      }
    }
  };

  // ---- pipeline prologue: chunk 0 of the first task of the first tile ----
  int bt = static_cast<int>(blockIdx.x);
  TileRef cur = tile_at(bt);
  int t = cur.seg.t0, kc = 0;
  PackMeta ma = meta_a[exec[t].a_blk], mb = meta_b[exec[t].b_blk];
  int nkc = ma.k_pad / KC;
  int32_t eb = exps[mb.exp_off + cur.j0 + wn + lrow];
  task_offsets(ma, mb);
  fetch(ma, mb, 0, cur.i0, cur.j0);
  if (tid < TM) s_ea[0][tid] = exps[ma.exp_off + cur.i0 + tid];
  stage(0);
  __syncthreads();
  int buf = 0;

  // ---- flushes to FP64: each FP32 group at its own exactness limit ----
  // A level-L product is an integer multiple of 2^(-9(L-2)) of magnitude
  // <= 2^16, and one complex k adds 2(L-1) of them to the level's group, so
  // the group stays exact in FP32 for 256 / (2(L-1)) values of k: {L=2}
  // for 128, {L=3} for 64, {L=4} for 42 (flushed every 32).  {L=5,6} is
  // flushed with {L=4}.  Every group is flushed when a task ends, since the
  // exponents change.
  auto flush = [&](auto ng_tag) __attribute__((always_inline)) {
    constexpr int NG = decltype(ng_tag)::value;   // groups 4-NG..3
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int row = wm + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
      const int e = s_ea[buf][row] + eb - 16;
      double sum = static_cast<double>(g[2][r]) + static_cast<double>(g[3][r]);
      if constexpr (NG >= 3) sum += static_cast<double>(g[1][r]);
      if constexpr (NG >= 4) sum += static_cast<double>(g[0][r]);
      acc64[r] += ldexp(sum, e);
    }
    g[2] = v16f{};
    g[3] = v16f{};
    if constexpr (NG >= 3) g[1] = v16f{};
    if constexpr (NG >= 4) g[0] = v16f{};
  };

  while (true) {
    // ---- the next chunk: same task, next task, or the next tile's first ----
    int t2 = t, kc2 = kc + 1;
    bool new_task = false;
    if (kc2 == nkc) {
      t2 = t + 1;
      kc2 = 0;
      new_task = true;
    }
    const bool tile_end = t2 >= cur.seg.t1;
    const int bt2 = tile_end ? bt + G : bt;
    const bool has_next = bt2 < n_tiles;
    TileRef nxt = cur;
    PackMeta ma2 = ma, mb2 = mb;
    int32_t eb2 = eb;
    if (has_next) {
      if (tile_end) {
        nxt = tile_at(bt2);
        t2 = nxt.seg.t0;
      }
      if (new_task) {
        const ExecTask e2 = exec[t2];
        ma2 = meta_a[e2.a_blk];
        mb2 = meta_b[e2.b_blk];
        eb2 = exps[mb2.exp_off + nxt.j0 + wn + lrow];
        if (tid < TM) ea_next = exps[ma2.exp_off + nxt.i0 + tid];
        task_offsets(ma2, mb2);
      }
      fetch(ma2, mb2, kc2, nxt.i0, nxt.j0);
    }
    // a wave whose 32 x 32 block lies wholly outside the output block only
    // helps to stage operands (its MFMAs and flushes are skipped)
    const bool active = (cur.i0 + wm < cur.seg.m) && (cur.j0 + wn < cur.seg.n);
    if (active) {
      // real part: ar*br + (-ai)*bi; imaginary part: ar*bi + ai*br.  Slice 0
      // of A is read first, then B slice by slice: the first MFMAs wait for
      // three reads, and the later reads land while the earlier MFMAs run.
      v8bf ar = *reinterpret_cast<const v8bf *>(&As[buf][0][0][wm + lrow][lk]);
      v8bf ai = *reinterpret_cast<const v8bf *>(&As[buf][1][0][wm + lrow][lk]);
      v8bf b0[S], b1[S];
#pragma unroll
      for (int s2 = 0; s2 < S; ++s2) {
        b0[s2] = *reinterpret_cast<const v8bf *>(&Bs[buf][comp][s2][wn + lrow][lk]);
        b1[s2] = *reinterpret_cast<const v8bf *>(&Bs[buf][1 - comp][s2][wn + lrow][lk]);
      }
      // a raised wave priority while the MFMAs issue: this wave wins the
      // issue arbitration against the fragment reads of the other wave on
      // its SIMD (the kernel runs about 1 % faster)
      __builtin_amdgcn_s_setprio(1);
#pragma unroll
      for (int sa = 0; sa < S; ++sa) {
        if (sa > 0) {
          ar = *reinterpret_cast<const v8bf *>(&As[buf][0][sa][wm + lrow][lk]);
          ai = *reinterpret_cast<const v8bf *>(&As[buf][1][sa][wm + lrow][lk]);
        }
        if (comp == 0) ai = neg8(ai);
#pragma unroll
        for (int sb = 0; sb < S - sa; ++sb) {
          const int q = group_of(sa, sb);
          g[q] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ar, b0[sb], g[q], 0, 0, 0);
          g[q] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ai, b1[sb], g[q], 0, 0, 0);
        }
      }
      __builtin_amdgcn_s_setprio(0);
      const int kk = (kc + 1) * KC;                // k since the task start
      const bool last = (kc == nkc - 1);
      if (last || kk % 32 == 0) {
        if (last || kk % 128 == 0) {
          flush(std::integral_constant<int, 4>{});
        } else if (kk % 64 == 0) {
          flush(std::integral_constant<int, 3>{});
        } else {
          flush(std::integral_constant<int, 2>{});
        }
      }
    }
    // ---- a finished tile: each wave stores its own output component ----
    if (tile_end) {
      if (active) {
        const int col = cur.j0 + wn + lrow;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const int row = cur.i0 + wm + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
          if (row < cur.seg.m && col < cur.seg.n) {
            double *cp = C + 2 * (cur.seg.c_off + static_cast<int64_t>(row) * cur.seg.n + col) + comp;
            double y = alpha * acc64[r];
            if (beta != 0.0) y += beta * *cp;
            *cp = y;
          }
        }
      }
#pragma unroll
      for (int r = 0; r < 16; ++r) acc64[r] = 0.0;
    }
    // ---- stage the prefetched chunk into the other buffer ----
    if (has_next) {
      stage(buf ^ 1);
      if (tid < TM) s_ea[buf ^ 1][tid] = new_task ? ea_next : s_ea[buf][tid];
    }
    __syncthreads();
    if (!has_next) break;
    if (tile_end) {
      cur = nxt;
      bt = bt2;
    }
    if (new_task) {
      ma = ma2;
      mb = mb2;
      nkc = ma.k_pad / KC;
      eb = eb2;
    }
    t = t2;
    kc = kc2;
    buf ^= 1;
  }
}

// ---- int8 digit form ----

typedef int32_t v4i __attribute__((ext_vector_type(4)));
typedef int32_t v16i __attribute__((ext_vector_type(16)));

// Bytewise negation of 16 int8 digits in [-127, 127] (no digit overflows):
// per byte, (0x80 - (x & 0x7f)) cannot borrow, and xor-ing the complement of
// x's sign bit gives -x.
__device__ __forceinline__ v4i neg_digits(v4i x) {
  const v4i h = v4i{static_cast<int32_t>(0x80808080u), static_cast<int32_t>(0x80808080u),
                    static_cast<int32_t>(0x80808080u), static_cast<int32_t>(0x80808080u)};
  return (h - (x & ~h)) ^ (~x & h);
}

// LDS position of the 16-byte half `half` of row `row` of a 32-byte digit row:
// the halves of rows 4..7 (mod 8) are swapped, so the 16-byte reads of eight
// consecutive rows cover every bank once.
__device__ __forceinline__ int i8_half(int row, int half) { return 16 * (half ^ ((row >> 2) & 1)); }

// Recombines the six int32 digit levels of accumulator element r in FP64.
// The caller adds ldexp(sum * 255^5, e - 92) to its FP64 accumulator: the
// power of two comes last, so a result in the subnormal range is rounded once
// instead of passing through a subnormal intermediate.
__device__ __forceinline__ double i8_level_sum(const v16i (&g)[S8], int r) {
  double sum = static_cast<double>(g[0][r]);
#pragma unroll
  for (int l = 1; l < S8; ++l) sum = fma(sum, 255.0, static_cast<double>(g[l][r]));
  return sum;
}
constexpr double k255p5 = 1078203909375.0;   // 255^5, exact in FP64

// Real int8 split GEMM: one 256-thread workgroup per 64 x 64 output tile,
// 2 x 2 waves of 32 x 32, the 21 digit pairs of a 32-k chunk as one
// v_mfma_i32_32x32x32_i8 each, with the software pipeline of
// k_split_gemm_f64 (the next chunk's digits in flight while the MFMAs run).
// Bounded to two workgroups per CU: the registers then fit two waves per
// SIMD (1.8x faster than one wave per SIMD with the level sums in AGPRs).
__global__ __launch_bounds__(256, 2) void k_split_i8_f64(
    const Tile *__restrict__ tiles, const Seg *__restrict__ segs, const ExecTask *__restrict__ exec,
    const PackMeta *__restrict__ meta_a, const PackMeta *__restrict__ meta_b,
    const int8_t *__restrict__ planes, const int32_t *__restrict__ exps, double *__restrict__ C,
    double alpha, double beta) {
  constexpr int PIECES = S8 * 128;                 // 16-byte pieces per operand per k-chunk
  constexpr int PER_THREAD = PIECES / 256;
  static_assert(PIECES % 256 == 0, "pieces per thread");
  __shared__ __attribute__((aligned(16))) int8_t As[2][S8][TM][KC8];
  __shared__ __attribute__((aligned(16))) int8_t Bs[2][S8][TN][KC8];
  __shared__ int32_t s_ea[2][TM];

  const Tile tile = tiles[split_xcd_tile(static_cast<int>(blockIdx.x), static_cast<int>(gridDim.x))];
  const Seg seg = segs[tile.seg];
  const int i0 = tile.tm * TM, j0 = tile.tn * TN;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
  const int lrow = lane & 31, lh = lane >> 5;
  const bool active = (i0 + wm < seg.m) && (j0 + wn < seg.n);

  double acc64[16];
#pragma unroll
  for (int r = 0; r < 16; ++r) acc64[r] = 0.0;
  v16i g[S8];
#pragma unroll
  for (int l = 0; l < S8; ++l) g[l] = v16i{};

  v4u pa[PER_THREAD], pb[PER_THREAD];
  int32_t ea_next = 0;
  auto fetch = [&](const PackMeta &ma, const PackMeta &mb, int kc) {
    const int nkc = ma.k_pad / KC8;
#pragma unroll
    for (int u = 0; u < PER_THREAD; ++u) {
      const int c = tid + 256 * u;
      const int ps = c >> 7, rr = (c & 127) >> 1, half = c & 1;
      const int64_t ia = ma.plane_off + ((static_cast<int64_t>(ps) * nkc + kc) * ma.r_pad + i0 + rr) * KC8 + half * 16;
      const int64_t ib = mb.plane_off + ((static_cast<int64_t>(ps) * nkc + kc) * mb.r_pad + j0 + rr) * KC8 + half * 16;
      pa[u] = *reinterpret_cast<const v4u *>(planes + ia);
      pb[u] = *reinterpret_cast<const v4u *>(planes + ib);
    }
  };
  auto stage = [&](int buf) {
#pragma unroll
    for (int u = 0; u < PER_THREAD; ++u) {
      const int c = tid + 256 * u;
      const int ps = c >> 7, rr = (c & 127) >> 1, half = c & 1;
      *reinterpret_cast<v4u *>(&As[buf][ps][rr][i8_half(rr, half)]) = pa[u];
      *reinterpret_cast<v4u *>(&Bs[buf][ps][rr][i8_half(rr, half)]) = pb[u];
    }
  };

  int t = seg.t0, kc = 0;
  PackMeta ma = meta_a[exec[t].a_blk], mb = meta_b[exec[t].b_blk];
  int nkc = ma.k_pad / KC8;
  int32_t eb = exps[mb.exp_off + j0 + wn + lrow];
  fetch(ma, mb, 0);
  if (tid < TM) s_ea[0][tid] = exps[ma.exp_off + i0 + tid];
  stage(0);
  __syncthreads();
  int buf = 0;

  while (t < seg.t1) {
    int t2 = t, kc2 = kc + 1;
    if (kc2 == nkc) {
      t2 = t + 1;
      kc2 = 0;
    }
    const bool has_next = t2 < seg.t1;
    PackMeta ma2 = ma, mb2 = mb;
    int32_t eb2 = eb;
    if (has_next) {
      if (t2 != t) {
        const ExecTask e2 = exec[t2];
        ma2 = meta_a[e2.a_blk];
        mb2 = meta_b[e2.b_blk];
        eb2 = exps[mb2.exp_off + j0 + wn + lrow];
        if (tid < TM) ea_next = exps[ma2.exp_off + i0 + tid];
      }
      fetch(ma2, mb2, kc2);
    }
    if (active) {
      const int brow = wn + lrow, arow = wm + lrow;
      v4i bfr[S8];
#pragma unroll
      for (int s2 = 0; s2 < S8; ++s2) bfr[s2] = *reinterpret_cast<const v4i *>(&Bs[buf][s2][brow][i8_half(brow, lh)]);
#pragma unroll
      for (int sa = 0; sa < S8; ++sa) {
        const v4i ar = *reinterpret_cast<const v4i *>(&As[buf][sa][arow][i8_half(arow, lh)]);
#pragma unroll
        for (int sb = 0; sb < S8 - sa; ++sb)
          g[sa + sb] = __builtin_amdgcn_mfma_i32_32x32x32_i8(ar, bfr[sb], g[sa + sb], 0, 0, 0);
      }
      const int kk = (kc + 1) * KC8;                // k since the task start
      if (kc == nkc - 1 || kk % kI8FlushK == 0) {
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const int row = wm + 8 * (r >> 2) + 4 * lh + (r & 3);
          acc64[r] += ldexp(i8_level_sum(g, r) * k255p5, s_ea[buf][row] + eb - 92);
        }
#pragma unroll
        for (int l = 0; l < S8; ++l) g[l] = v16i{};
      }
    }
    if (has_next) {
      stage(buf ^ 1);
      if (tid < TM) s_ea[buf ^ 1][tid] = (t2 != t) ? ea_next : s_ea[buf][tid];
    }
    __syncthreads();
    if (t2 != t) {
      ma = ma2;
      mb = mb2;
      nkc = ma.k_pad / KC8;
      eb = eb2;
    }
    t = t2;
    kc = kc2;
    buf ^= 1;
  }

  const int col = j0 + wn + lrow;
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int row = i0 + wm + 8 * (r >> 2) + 4 * lh + (r & 3);
    if (row < seg.m && col < seg.n) {
      const int64_t idx = seg.c_off + static_cast<int64_t>(row) * seg.n + col;
      double y = alpha * acc64[r];
      if (beta != 0.0) y += beta * C[idx];
      C[idx] = y;
    }
  }
}

// Complex int8 split GEMM: the persistent structure of k_split_gemm_c2 (one
// 512-thread workgroup per CU walking tiles blockIdx.x, blockIdx.x + gridDim.x,
// ...; waves 0-3 the real part ar*br + (-ai)*bi, waves 4-7 the imaginary part
// ar*bi + ai*br of the same 2 x 2 arrangement of 32 x 32 blocks; the pipeline
// crosses task and tile boundaries; each wave stores its own component of a
// finished tile; real scalars) on int8 digits.
__global__ __launch_bounds__(512) void k_split_i8_c2(
    const SplitTile *__restrict__ stiles, int32_t n_tiles, const SplitTask *__restrict__ stasks,
    const int8_t *__restrict__ planes, const int32_t *__restrict__ exps, double *__restrict__ C, double alpha,
    double beta) {
  constexpr int PARTS = 2;
  constexpr int PIECES = PARTS * S8 * 128;          // 16-byte pieces per operand per k-chunk
  constexpr int NT = 512;
  constexpr int PER_THREAD = PIECES / NT;
  static_assert(PIECES % NT == 0, "pieces per thread");
  __shared__ __attribute__((aligned(16))) int8_t As[2][PARTS][S8][TM][KC8];
  __shared__ __attribute__((aligned(16))) int8_t Bs[2][PARTS][S8][TN][KC8];
  __shared__ int32_t s_ea[2][TM];

  const int tid = threadIdx.x, lane = tid & 63, w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int ws = w & 3, comp = w >> 2;             // spatial block, output component
  const int wm = (ws >> 1) * 32, wn = (ws & 1) * 32;
  const int lrow = lane & 31, lh = lane >> 5;
  const int G = static_cast<int>(gridDim.x);

  double acc64[16];
#pragma unroll
  for (int r = 0; r < 16; ++r) acc64[r] = 0.0;
  v16i g[S8];
#pragma unroll
  for (int l = 0; l < S8; ++l) g[l] = v16i{};

  v4u pa[PER_THREAD], pb[PER_THREAD];
  int32_t ea_next = 0;
  // byte offsets of a thread's pieces within a task's k-chunk 0 (a block's
  // digit planes stay below 2^31 bytes), recomputed when the task changes
  uint32_t oa[PER_THREAD], ob[PER_THREAD];
  auto task_offsets = [&](const SplitTask &tk) __attribute__((always_inline)) {
#pragma unroll
    for (int u = 0; u < PER_THREAD; ++u) {
      const int c = tid + NT * u;
      const int ps = c >> 7, rr = (c & 127) >> 1, half = c & 1;
      oa[u] = static_cast<uint32_t>((ps * tk.nkc * tk.a_rpad + rr) * KC8 + half * 16);
      ob[u] = static_cast<uint32_t>((ps * tk.nkc * tk.b_rpad + rr) * KC8 + half * 16);
    }
  };
  auto fetch = [&](const SplitTask &tk, int kc, int i0, int j0) __attribute__((always_inline)) {
    const int8_t *ba = planes + tk.a_plane + (static_cast<int64_t>(kc) * tk.a_rpad + i0) * KC8;
    const int8_t *bb = planes + tk.b_plane + (static_cast<int64_t>(kc) * tk.b_rpad + j0) * KC8;
#pragma unroll
    for (int u = 0; u < PER_THREAD; ++u) {
      pa[u] = *reinterpret_cast<const v4u *>(ba + oa[u]);
      pb[u] = *reinterpret_cast<const v4u *>(bb + ob[u]);
    }
  };
  auto stage = [&](int buf) __attribute__((always_inline)) {
#pragma unroll
    for (int u = 0; u < PER_THREAD; ++u) {
      const int c = tid + NT * u;
      const int ps = c >> 7, rr = (c & 127) >> 1, half = c & 1;
      *reinterpret_cast<v4u *>(&As[buf][ps / S8][ps % S8][rr][i8_half(rr, half)]) = pa[u];
      *reinterpret_cast<v4u *>(&Bs[buf][ps / S8][ps % S8][rr][i8_half(rr, half)]) = pb[u];
    }
  };

  // The tile and task records are read one step ahead of their use: `nxt`
  // is the workgroup's next tile, `tk_n` the task after the current one (in
  // this tile or the next), so a boundary waits only on plane loads.
  int bt = static_cast<int>(blockIdx.x);
  SplitTile cur = load_tile(stiles, split_xcd_tile(bt, n_tiles));
  const bool more0 = bt + G < n_tiles;
  SplitTile nxt = more0 ? load_tile(stiles, split_xcd_tile(bt + G, n_tiles)) : cur;
  int t = cur.t0, kc = 0;
  SplitTask tk = load_task(stasks, t);
  SplitTask tk_n = (t + 1 < cur.t1) ? load_task(stasks, t + 1) : (more0 ? load_task(stasks, nxt.t0) : tk);
  int32_t eb = exps[tk.b_exp + cur.j0 + wn + lrow];
  task_offsets(tk);
  fetch(tk, 0, cur.i0, cur.j0);
  if (tid < TM) s_ea[0][tid] = exps[tk.a_exp + cur.i0 + tid];
  stage(0);
  __syncthreads();
  int buf = 0;

  while (true) {
    const bool new_task = kc + 1 == tk.nkc;
    const bool tile_end = new_task && t + 1 >= cur.t1;
    const int bt2 = tile_end ? bt + G : bt;
    const bool has_next = bt2 < n_tiles;
    const int i0n = tile_end ? nxt.i0 : cur.i0, j0n = tile_end ? nxt.j0 : cur.j0;
    const int kc2 = new_task ? 0 : kc + 1;
    int32_t eb2 = eb;
    if (has_next) {
      if (new_task) {
        eb2 = exps[tk_n.b_exp + j0n + wn + lrow];
        if (tid < TM) ea_next = exps[tk_n.a_exp + i0n + tid];
        task_offsets(tk_n);
        fetch(tk_n, 0, i0n, j0n);
      } else {
        fetch(tk, kc2, i0n, j0n);
      }
    }
    const bool active = (cur.i0 + wm < cur.m) && (cur.j0 + wn < cur.n);
    if (active) {
      const int brow = wn + lrow, arow = wm + lrow;
      // real part ar*br + (-ai)*bi, imaginary part ar*bi + ai*br.  Digit 0
      // of A first, then B digit by digit: the first MFMAs wait for three
      // reads, and the later reads land while the earlier MFMAs run.
      const int p0 = comp, p1 = 1 - comp;
      v4i ar = *reinterpret_cast<const v4i *>(&As[buf][0][0][arow][i8_half(arow, lh)]);
      v4i ai = *reinterpret_cast<const v4i *>(&As[buf][1][0][arow][i8_half(arow, lh)]);
      v4i b0[S8], b1[S8];
#pragma unroll
      for (int s2 = 0; s2 < S8; ++s2) {
        b0[s2] = *reinterpret_cast<const v4i *>(&Bs[buf][p0][s2][brow][i8_half(brow, lh)]);
        b1[s2] = *reinterpret_cast<const v4i *>(&Bs[buf][p1][s2][brow][i8_half(brow, lh)]);
      }
#pragma unroll
      for (int sa = 0; sa < S8; ++sa) {
        if (sa > 0) {
          ar = *reinterpret_cast<const v4i *>(&As[buf][0][sa][arow][i8_half(arow, lh)]);
          ai = *reinterpret_cast<const v4i *>(&As[buf][1][sa][arow][i8_half(arow, lh)]);
        }
        if (comp == 0) ai = neg_digits(ai);
#pragma unroll
        for (int sb = 0; sb < S8 - sa; ++sb) {
          g[sa + sb] = __builtin_amdgcn_mfma_i32_32x32x32_i8(ar, b0[sb], g[sa + sb], 0, 0, 0);
          g[sa + sb] = __builtin_amdgcn_mfma_i32_32x32x32_i8(ai, b1[sb], g[sa + sb], 0, 0, 0);
        }
      }
      const int kk = (kc + 1) * KC8;                // k since the task start
      if (new_task || kk % kI8FlushK == 0) {
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const int row = wm + 8 * (r >> 2) + 4 * lh + (r & 3);
          acc64[r] += ldexp(i8_level_sum(g, r) * k255p5, s_ea[buf][row] + eb - 92);
        }
#pragma unroll
        for (int l = 0; l < S8; ++l) g[l] = v16i{};
      }
    }
    // ---- a finished tile: each wave stores its own output component ----
    if (tile_end) {
      if (active) {
        const int col = cur.j0 + wn + lrow;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const int row = cur.i0 + wm + 8 * (r >> 2) + 4 * lh + (r & 3);
          if (row < cur.m && col < cur.n) {
            double *cp = C + 2 * (cur.c_off + static_cast<int64_t>(row) * cur.n + col) + comp;
            double y = alpha * acc64[r];
            if (beta != 0.0) y += beta * *cp;
            *cp = y;
          }
        }
      }
#pragma unroll
      for (int r = 0; r < 16; ++r) acc64[r] = 0.0;
    }
    if (has_next) {
      stage(buf ^ 1);
      if (tid < TM) s_ea[buf ^ 1][tid] = new_task ? ea_next : s_ea[buf][tid];
    }
    __syncthreads();
    if (!has_next) break;
    // ---- advance, and read the records one step ahead ----
    if (new_task) {
      t = tile_end ? nxt.t0 : t + 1;
      tk = tk_n;
      eb = eb2;
      if (tile_end) {
        bt = bt2;
        cur = nxt;
        if (bt + G < n_tiles) nxt = load_tile(stiles, split_xcd_tile(bt + G, n_tiles));
      }
      const bool more = bt + G < n_tiles;
      if (t + 1 < cur.t1)
        tk_n = load_task(stasks, t + 1);
      else if (more)
        tk_n = load_task(stasks, nxt.t0);
    }
    kc = kc2;
    buf ^= 1;
  }
}

// Scales the tiles the complex split kernels write by a complex factor: a
// complex alpha or beta is applied around those kernels, which take real
// scalars, as C = alpha (X + (beta / alpha) C).
__global__ __launch_bounds__(256) void k_scale_tiles(const SplitTile *__restrict__ tiles, double sr, double si,
                                                     double *__restrict__ C) {
  const SplitTile t = tiles[blockIdx.x];
  for (int e = static_cast<int>(threadIdx.x); e < TM * TN; e += 256) {
    const int row = t.i0 + e / TN, col = t.j0 + e % TN;
    if (row < t.m && col < t.n) {
      double2 *p = reinterpret_cast<double2 *>(C) + t.c_off + static_cast<int64_t>(row) * t.n + col;
      const double2 v = *p;
      *p = double2{sr * v.x - si * v.y, sr * v.y + si * v.x};
    }
  }
}

inline int64_t round_up(int64_t x, int64_t m) { return (x + m - 1) / m * m; }

// Digit format of a split mode: slices per element, k per packed chunk,
// bytes per plane element.
struct SplitFmt {
  int slices, kc, elem_bytes;
};
inline SplitFmt split_fmt(int32_t mode) { return mode == QLT_MODE_SPLIT_I8 ? SplitFmt{S8, KC8, 1} : SplitFmt{S, KC, 2}; }

// Host-side layout of the split workspace of a plan.
struct SplitLayout {
  std::vector<PackMeta> meta_a, meta_b;
  std::vector<PackItem> items;    // phase 1: (side, block, 64-row group)
  std::vector<PackItem> citems;   // phase 2: (side, block, 64-row group, k-chunk)
  std::vector<PackItem> sitems;   // single-read pack of k_pad <= kShortK: (side, block, first row, 0)
  std::vector<SplitTile> rtiles;  // self-contained records of `tiles`
  std::vector<SplitTask> rtasks;  // self-contained records of the plan's tasks
  std::vector<Tile> tiles;        // 64 x 64 output tiles of the MFMA-class blocks, longest first
  int64_t plane_elems = 0;
  int64_t exp_elems = 0;
  size_t off_meta_a = 0, off_meta_b = 0, off_items = 0, off_aoffs = 0, off_boffs = 0, off_exps = 0,
         off_planes = 0, off_tiles = 0, off_citems = 0, off_sitems = 0, off_rtiles = 0, off_rtasks = 0, total = 0;
};

SplitLayout compute_layout(const qlt_plan *plan, int32_t dtype, int32_t mode) {
  SplitLayout L;
  const int parts = (dtype == QLT_C128) ? 2 : 1;
  const SplitFmt fmt = split_fmt(mode);
  const int32_t kp = (dtype == QLT_C128) ? 32 : 64;   // k per phase-2 pack chunk
  L.meta_a.assign(plan->n_a_blks, PackMeta{0, 0, 0, 0, 0, 0});
  L.meta_b.assign(plan->n_b_blks, PackMeta{0, 0, 0, 0, 0, 0});
  std::vector<char> used_a(plan->n_a_blks, 0), used_b(plan->n_b_blks, 0);
  for (size_t o = 0; o < plan->h_segs.size(); ++o) {
    const Seg &s = plan->h_segs[o];
    if (!is_mfma_class(s.cls)) continue;   // row-class segments always run natively
    for (int32_t t = s.t0; t < s.t1; ++t) {
      const ExecTask &e = plan->h_exec[t];
      if (!used_a[e.a_blk]) {
        used_a[e.a_blk] = 1;
        PackMeta &m = L.meta_a[e.a_blk];
        m.rows = s.m;
        m.cols = e.k;
        m.r_pad = static_cast<int32_t>(round_up(s.m, RPAD));
        m.k_pad = static_cast<int32_t>(round_up(e.k, fmt.kc));
      }
      if (!used_b[e.b_blk]) {
        used_b[e.b_blk] = 1;
        PackMeta &m = L.meta_b[e.b_blk];
        m.rows = s.n;
        m.cols = e.k;
        m.r_pad = static_cast<int32_t>(round_up(s.n, RPAD));
        m.k_pad = static_cast<int32_t>(round_up(e.k, fmt.kc));
      }
    }
  }
  auto assign = [&](std::vector<PackMeta> &meta, std::vector<char> &used, int side) {
    for (size_t i = 0; i < meta.size(); ++i) {
      if (!used[i]) continue;
      PackMeta &m = meta[i];
      m.plane_off = L.plane_elems;
      m.exp_off = L.exp_elems;
      L.plane_elems += static_cast<int64_t>(parts) * fmt.slices * m.r_pad * m.k_pad;
      L.exp_elems += m.r_pad;
      if (m.k_pad <= kShortK) {
        // groups of RW = floor(256 / (k_pad / 8)) real rows; the padding rows stay zero
        const int32_t rw = 256 / (m.k_pad / 8);
        for (int32_t r0 = 0; r0 < m.rows; r0 += rw) L.sitems.push_back(PackItem{side, static_cast<int32_t>(i), r0, 0});
        continue;
      }
      for (int32_t r0 = 0; r0 < m.r_pad; r0 += 64) {
        L.items.push_back(PackItem{side, static_cast<int32_t>(i), r0, 0});
        for (int32_t k0 = 0; k0 < m.k_pad; k0 += kp)
          L.citems.push_back(PackItem{side, static_cast<int32_t>(i), r0, k0});
      }
    }
  };
  assign(L.meta_a, used_a, 0);
  assign(L.meta_b, used_b, 1);
  {
    std::vector<std::pair<int64_t, int32_t>> order;   // (-work, segment)
    for (size_t o = 0; o < plan->h_segs.size(); ++o)
      if (is_mfma_class(plan->h_segs[o].cls)) order.push_back({-plan->h_segs[o].k_total, static_cast<int32_t>(o)});
    std::stable_sort(order.begin(), order.end(),
                     [](const std::pair<int64_t, int32_t> &x, const std::pair<int64_t, int32_t> &y) {
                       return x.first < y.first;
                     });
    for (const auto &e : order) {
      const Seg &sg = plan->h_segs[e.second];
      const int ntm = (sg.m + TM - 1) / TM, ntn = (sg.n + TN - 1) / TN;
      for (int a = 0; a < ntm; ++a)
        for (int b = 0; b < ntn; ++b) L.tiles.push_back(Tile{e.second, a, b, 0, 0, 0, 0, -1});
    }
  }
  for (const Tile &tl : L.tiles) {
    const Seg &sg = plan->h_segs[tl.seg];
    L.rtiles.push_back(SplitTile{sg.c_off, sg.m, sg.n, tl.tm * TM, tl.tn * TN, sg.t0, sg.t1});
  }
  L.rtasks.assign(plan->h_exec.size(), SplitTask{0, 0, 0, 0, 0, 0, 0, 0});
  for (size_t t = 0; t < plan->h_exec.size(); ++t) {
    const ExecTask &e = plan->h_exec[t];
    const PackMeta &ma = L.meta_a[e.a_blk], &mb = L.meta_b[e.b_blk];
    if (ma.k_pad == 0 || mb.k_pad == 0) continue;   // a row-class task: never split
    L.rtasks[t] = SplitTask{ma.plane_off, mb.plane_off, static_cast<int32_t>(ma.exp_off),
                            static_cast<int32_t>(mb.exp_off), ma.r_pad, mb.r_pad, ma.k_pad / fmt.kc, 0};
  }
  auto align = [](size_t x) { return (x + 255) / 256 * 256; };
  size_t off = 0;
  L.off_meta_a = off;
  off = align(off + sizeof(PackMeta) * L.meta_a.size());
  L.off_meta_b = off;
  off = align(off + sizeof(PackMeta) * L.meta_b.size());
  L.off_items = off;
  off = align(off + sizeof(PackItem) * L.items.size());
  L.off_aoffs = off;
  off = align(off + sizeof(int64_t) * L.meta_a.size());
  L.off_boffs = off;
  off = align(off + sizeof(int64_t) * L.meta_b.size());
  L.off_exps = off;
  off = align(off + sizeof(int32_t) * L.exp_elems);
  L.off_planes = off;
  off = align(off + static_cast<size_t>(fmt.elem_bytes) * L.plane_elems);
  L.off_tiles = off;
  off = align(off + sizeof(Tile) * L.tiles.size());
  L.off_citems = off;
  off = align(off + sizeof(PackItem) * L.citems.size());
  L.off_sitems = off;
  off = align(off + sizeof(PackItem) * L.sitems.size());
  L.off_rtiles = off;
  off = align(off + sizeof(SplitTile) * L.rtiles.size());
  L.off_rtasks = off;
  off = align(off + sizeof(SplitTask) * L.rtasks.size());
  L.total = off;
  return L;
}

}  // namespace

namespace qlt {

qlt_status launch_native(const qlt_plan *plan, int32_t dtype, const void *a, const void *b, void *c,
                         const double *alpha, const double *beta, hipStream_t st);
qlt_status launch_native_classes(const qlt_plan *plan, int32_t dtype, const void *a, const void *b, void *c,
                                 const double *alpha, const double *beta, hipStream_t st, bool mfma_class,
                                 bool probe, ClassStreams *shared);

// Flop-weighted mean k below which a plan runs natively even in split mode:
// such contractions are bound by memory traffic, not by the MFMA rate, and
// slicing them would only add traffic.
static double g_split_min_mean_k = 48.0;

// Algorithmic flops per byte of the distinct operand blocks below which the
// split path loses to the native one: packing reads every operand block
// twice and writes 5/4 of it back as slice planes, so a contraction whose
// operands are large relative to its flops is cheaper natively.
constexpr double kSplitMinFlopsPerByte = 96.0;

bool split_profitable(const qlt_plan *plan, int32_t dtype, int32_t mode) {
  int64_t mfma_tiles = 0;
  for (int c = 0; c < kNumMfmaClasses; ++c) mfma_tiles += plan->cls_tiles[c];
  if (mfma_tiles == 0) return false;
  double w = 0.0, wk = 0.0, elems = 0.0;
  std::vector<char> seen_a(plan->n_a_blks, 0), seen_b(plan->n_b_blks, 0);
  const int64_t parts = (dtype == QLT_C128) ? 2 : 1;
  const SplitFmt fmt = split_fmt(mode);
  for (const Seg &s : plan->h_segs) {
    if (!is_mfma_class(s.cls)) continue;
    for (int32_t t = s.t0; t < s.t1; ++t) {
      const ExecTask &e = plan->h_exec[t];
      // the split kernels address a block's slice planes with 32-bit offsets
      const int64_t kp = round_up(e.k, fmt.kc);
      if (parts * fmt.slices * round_up(s.m, RPAD) * kp >= (int64_t{1} << 31) ||
          parts * fmt.slices * round_up(s.n, RPAD) * kp >= (int64_t{1} << 31))
        return false;
      const double k = e.k;
      const double f = static_cast<double>(s.m) * s.n * k;
      w += f;
      wk += f * k;
      if (!seen_a[e.a_blk]) {
        seen_a[e.a_blk] = 1;
        elems += static_cast<double>(s.m) * k;
      }
      if (!seen_b[e.b_blk]) {
        seen_b[e.b_blk] = 1;
        elems += k * static_cast<double>(s.n);
      }
    }
  }
  if (g_split_min_mean_k <= 0.0) return true;   // forced
  if (w <= 0.0 || wk / w < g_split_min_mean_k) return false;
  const double flops = (dtype == QLT_C128 ? 8.0 : 2.0) * w;
  const double bytes = (dtype == QLT_C128 ? 16.0 : 8.0) * elems;
  return flops / bytes >= kSplitMinFlopsPerByte;
}

qlt_status split_workspace_bytes(const qlt_plan *plan, int32_t dtype, int32_t mode, size_t *bytes) {
  if (!split_profitable(plan, dtype, mode)) {
    *bytes = 0;
    return QLT_OK;
  }
  *bytes = compute_layout(plan, dtype, mode).total;
  return QLT_OK;
}

qlt_status split_prepare_metadata(qlt_plan *plan, int32_t dtype, int32_t mode) {
  // Writes the pack metadata into the (already allocated) workspace.
  if (!split_profitable(plan, dtype, mode)) return QLT_OK;
  SplitLayout L = compute_layout(plan, dtype, mode);
  if (plan->work_bytes < L.total) return fail(QLT_ERR_INTERNAL, "split workspace too small");
  if (L.exp_elems >= (int64_t{1} << 31))   // task records hold 32-bit exponent offsets
    return fail(QLT_ERR_UNSUPPORTED, "split workspace: more than 2^31 packed rows");
  char *base = static_cast<char *>(plan->d_work);
  std::vector<int64_t> aoffs(plan->n_a_blks, 0), boffs(plan->n_b_blks, 0);
  for (const ExecTask &e : plan->h_exec) {
    aoffs[e.a_blk] = e.a_off;
    boffs[e.b_blk] = e.b_off;
  }
  // The single-read pack writes the real rows only: the padding rows of the
  // planes and their exponents are zeroed here, once per layout.
  QLT_HIP(hipMemset(base + L.off_exps, 0, sizeof(int32_t) * L.exp_elems));
  QLT_HIP(hipMemset(base + L.off_planes, 0, static_cast<size_t>(split_fmt(mode).elem_bytes) * L.plane_elems));
  QLT_HIP(hipDeviceSynchronize());
  QLT_HIP(hipMemcpy(base + L.off_meta_a, L.meta_a.data(), sizeof(PackMeta) * L.meta_a.size(),
                    hipMemcpyHostToDevice));
  QLT_HIP(hipMemcpy(base + L.off_meta_b, L.meta_b.data(), sizeof(PackMeta) * L.meta_b.size(),
                    hipMemcpyHostToDevice));
  QLT_HIP(hipMemcpy(base + L.off_items, L.items.data(), sizeof(PackItem) * L.items.size(),
                    hipMemcpyHostToDevice));
  QLT_HIP(hipMemcpy(base + L.off_aoffs, aoffs.data(), sizeof(int64_t) * aoffs.size(), hipMemcpyHostToDevice));
  QLT_HIP(hipMemcpy(base + L.off_boffs, boffs.data(), sizeof(int64_t) * boffs.size(), hipMemcpyHostToDevice));
  QLT_HIP(hipMemcpy(base + L.off_tiles, L.tiles.data(), sizeof(Tile) * L.tiles.size(), hipMemcpyHostToDevice));
  QLT_HIP(hipMemcpy(base + L.off_citems, L.citems.data(), sizeof(PackItem) * L.citems.size(),
                    hipMemcpyHostToDevice));
  QLT_HIP(hipMemcpy(base + L.off_sitems, L.sitems.data(), sizeof(PackItem) * L.sitems.size(),
                    hipMemcpyHostToDevice));
  QLT_HIP(hipMemcpy(base + L.off_rtiles, L.rtiles.data(), sizeof(SplitTile) * L.rtiles.size(),
                    hipMemcpyHostToDevice));
  QLT_HIP(hipMemcpy(base + L.off_rtasks, L.rtasks.data(), sizeof(SplitTask) * L.rtasks.size(),
                    hipMemcpyHostToDevice));
  plan->split_dtype = dtype;
  plan->split_mode = mode;
  plan->split_items = static_cast<int64_t>(L.items.size());
  plan->split_tiles = static_cast<int64_t>(L.tiles.size());
  plan->split_citems = static_cast<int64_t>(L.citems.size());
  plan->split_sitems = static_cast<int64_t>(L.sitems.size());
  const size_t offs[12] = {L.off_meta_a, L.off_meta_b, L.off_items,  L.off_aoffs,  L.off_boffs,  L.off_exps,
                           L.off_planes, L.off_tiles,  L.off_citems, L.off_sitems, L.off_rtiles, L.off_rtasks};
  std::memcpy(plan->split_off, offs, sizeof(offs));
  return QLT_OK;
}

qlt_status launch_split(const qlt_plan *plan, int32_t dtype, int32_t mode, const void *a, const void *b, void *c,
                        const double *alpha, const double *beta, hipStream_t st) {
  if (!split_profitable(plan, dtype, mode)) return launch_native(plan, dtype, a, b, c, alpha, beta, st);
  if (plan->split_dtype != dtype || plan->split_mode != mode)
    return fail(QLT_ERR_INVALID_ARG,
                "split workspace was prepared for another dtype or mode; call qlt_plan_prepare");
  if (plan->d_work == nullptr) return fail(QLT_ERR_INTERNAL, "split workspace missing");
  char *base = static_cast<char *>(plan->d_work);
  const size_t *so = plan->split_off;
  const auto *meta_a = reinterpret_cast<const PackMeta *>(base + so[0]);
  const auto *meta_b = reinterpret_cast<const PackMeta *>(base + so[1]);
  const auto *items = reinterpret_cast<const PackItem *>(base + so[2]);
  const auto *aoffs = reinterpret_cast<const int64_t *>(base + so[3]);
  const auto *boffs = reinterpret_cast<const int64_t *>(base + so[4]);
  auto *exps = reinterpret_cast<int32_t *>(base + so[5]);
  void *planes = base + so[6];
  const auto *stiles = reinterpret_cast<const Tile *>(base + so[7]);
  const auto *citems = reinterpret_cast<const PackItem *>(base + so[8]);
  const unsigned n_citems = static_cast<unsigned>(plan->split_citems);
  const auto *sitems = reinterpret_cast<const PackItem *>(base + so[9]);
  const auto *rtiles = reinterpret_cast<const SplitTile *>(base + so[10]);
  const auto *rtasks = reinterpret_cast<const SplitTask *>(base + so[11]);
  const unsigned n_sitems = static_cast<unsigned>(plan->split_sitems);
  const unsigned n_items = static_cast<unsigned>(plan->split_items);
  const unsigned n_mfma = static_cast<unsigned>(plan->split_tiles);
  const double ar = alpha[0], ai = (dtype == QLT_C128) ? alpha[1] : 0.0;
  const double br = beta[0], bi = (dtype == QLT_C128) ? beta[1] : 0.0;
  const dim3 blk(256);
  // the row classes (if any) go on the class streams, forked before the pack
  std::unique_ptr<ClassStreams> rows_cs;
  if (plan->cls_tiles[kTileRows4] > 0 || plan->cls_tiles[kTileRows16] > 0 || plan->n_ptiles > 0)
    rows_cs = std::make_unique<ClassStreams>(st);
  // a kernel probe times the split GEMM unless the row classes carry more work
  double mfma_mkn = 0.0, rows_mkn = 0.0;
  for (int32_t cls = 0; cls < kNumTileClasses; ++cls) (is_mfma_class(cls) ? mfma_mkn : rows_mkn) += plan->cls_mkn[cls];
  const bool probe_split = probe_armed(st) && mfma_mkn >= rows_mkn;
  const bool cplx = dtype == QLT_C128;
  const bool i8 = mode == QLT_MODE_SPLIT_I8;
  const auto *planes16 = static_cast<const uint16_t *>(planes);
  const auto *planes8 = static_cast<const int8_t *>(planes);
  auto *cd = static_cast<double *>(c);
  const auto *ad = static_cast<const double *>(a);
  const auto *bd = static_cast<const double *>(b);
  // blocks with k_pad <= kShortK: single-read pack; longer ones: row exponents, then digits per k-chunk
  auto launch_pack = [&](auto k_short, auto k_exp, auto k_dig) {
    if (n_sitems > 0)
      hipLaunchKernelGGL(k_short, dim3(n_sitems), blk, 0, st, sitems, meta_a, meta_b, plan->d_ga, plan->d_gb, aoffs,
                         boffs, plan->layout, ad, bd, planes, exps);
    if (n_items > 0) {
      hipLaunchKernelGGL(k_exp, dim3(n_items), blk, 0, st, items, meta_a, meta_b, plan->d_ga, plan->d_gb, aoffs,
                         boffs, plan->layout, ad, bd, planes, exps);
      hipLaunchKernelGGL(k_dig, dim3(n_citems), blk, 0, st, citems, meta_a, meta_b, plan->d_ga, plan->d_gb, aoffs,
                         boffs, plan->layout, ad, bd, planes, exps);
    }
  };
  // the complex kernels are persistent: one workgroup per CU (their LDS allows one)
  const unsigned grid = std::min<unsigned>(n_mfma, static_cast<unsigned>(device_cu_count()));
  if (!cplx && !i8) {
    launch_pack(k_pack_short<false, false>, k_pack<false, false, 1>, k_pack<false, false, 2>);
    if (probe_split) probe_begin(st);
    hipLaunchKernelGGL(k_split_gemm_f64, dim3(n_mfma), blk, 0, st, stiles, plan->d_segs, plan->d_exec, meta_a,
                       meta_b, planes16, exps, cd, ar, br);
  } else if (!cplx) {
    launch_pack(k_pack_short<false, true>, k_pack<false, true, 1>, k_pack<false, true, 2>);
    if (probe_split) probe_begin(st);
    hipLaunchKernelGGL(k_split_i8_f64, dim3(n_mfma), blk, 0, st, stiles, plan->d_segs, plan->d_exec, meta_a,
                       meta_b, planes8, exps, cd, ar, br);
  } else {
    // The complex kernels take real scalars; a complex alpha or beta is
    // applied on their tiles around them: C = alpha (X + (beta / alpha) C),
    // C = alpha X when beta = 0 (C is not read), C = beta C when alpha = 0.
    double gar = ar, gbr = br;
    bool pre = false, post = false, gemm = true;
    double pre_r = 0.0, pre_i = 0.0;
    if (ai != 0.0 || bi != 0.0) {
      if (ar == 0.0 && ai == 0.0) {
        gemm = false;
        pre = true;
        pre_r = br;
        pre_i = bi;
      } else if (br == 0.0 && bi == 0.0) {
        gar = 1.0;
        post = true;
      } else {
        const double d = ar * ar + ai * ai;
        pre = true;
        pre_r = (br * ar + bi * ai) / d;
        pre_i = (bi * ar - br * ai) / d;
        gar = 1.0;
        gbr = 1.0;
        post = true;
      }
    }
    if (pre) hipLaunchKernelGGL(k_scale_tiles, dim3(n_mfma), blk, 0, st, rtiles, pre_r, pre_i, cd);
    if (gemm) {
      if (i8)
        launch_pack(k_pack_short<true, true>, k_pack<true, true, 1>, k_pack<true, true, 2>);
      else
        launch_pack(k_pack_short<true, false>, k_pack<true, false, 1>, k_pack<true, false, 2>);
      if (probe_split) probe_begin(st);
      if (i8)
        hipLaunchKernelGGL(k_split_i8_c2, dim3(grid), dim3(512), 0, st, rtiles, static_cast<int32_t>(n_mfma),
                           rtasks, planes8, exps, cd, gar, gbr);
      else
        hipLaunchKernelGGL(k_split_gemm_c2, dim3(grid), dim3(512), 0, st, stiles, static_cast<int32_t>(n_mfma),
                           plan->d_segs, plan->d_exec, meta_a, meta_b, planes16, exps, cd, gar, gbr);
    } else if (probe_split) {
      probe_begin(st);
    }
    if (post) hipLaunchKernelGGL(k_scale_tiles, dim3(n_mfma), blk, 0, st, rtiles, ar, ai, cd);
  }
  if (probe_split) {
    const double elems = plan->mfma_elems;   // distinct A and B blocks of all classes, plus C
    const char *name = i8 ? (cplx ? "k_split_i8_c2 c128" : "k_split_i8_f64 f64")
                          : (cplx ? "k_split_gemm_c2 c128" : "k_split_gemm_f64 f64");
    probe_end(st, name, (cplx ? 8.0 : 2.0) * mfma_mkn,
              (cplx ? 16.0 : 8.0) * elems);
  }
  QLT_HIP(hipGetLastError());
  // narrow outputs: the native row-streaming kernel, on the class streams so
  // that it overlaps the split GEMM (the output blocks are disjoint)
  return launch_native_classes(plan, dtype, a, b, c, alpha, beta, st, false, !probe_split, rows_cs.get());
}

}  // namespace qlt

extern "C" {

qlt_status qlt_set_split_min_mean_k(double min_mean_k) {
  QLT_CHECK_ARG(min_mean_k >= 0.0, "qlt_set_split_min_mean_k: negative threshold");
  qlt::g_split_min_mean_k = min_mean_k;
  return QLT_OK;
}

double qlt_get_split_min_mean_k(void) { return qlt::g_split_min_mean_k; }

}  // extern "C"
