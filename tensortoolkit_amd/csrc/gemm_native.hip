// Grouped block GEMM over a contraction plan, native FP64 / complex128 MFMA.
//
// Replaces the execution loop of CtrctTwoBSDTAndAssignIn
// (include/qlten/qltensor/blk_spar_data_ten/global_operations.h:895-992):
// there, every task lazily transposes its A and B blocks with HPTT/cuTENSOR
// (hp_numeric::TensorTranspose, framework/hp_numeric/ten_trans.h:56-114,245-349)
// and then issues one BLAS call C_blk = A'(m x k) B'(k x n) + beta C_blk
// (hp_numeric::MatMultiply, framework/hp_numeric/blas_level3.h:35-56,798-856).
// Here one launch covers the whole plan.  Each workgroup owns one 64 x 64
// tile of one output block and walks every task of that block, so the
// per-block beta chain (0 for the first task, 1 after) becomes a single
// write of the finished tile.  The transposes are never materialised: the
// prologue gathers A'(i, p) = A[a_off + row_off(i) + col_off(p)] straight
// from the original block layout through per-tile offset tables, staging the
// k-chunk in LDS, with the thread mapping chosen so that the block's
// innermost memory axis is read contiguously.

#include "qlt_internal.h"

using namespace qlt;

namespace {

typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int TM = 64;
constexpr int TN = 64;
constexpr int KC = 16;
constexpr int LDP = TM + 1;   // padded LDS row (doubles)

// Offset of linear index `x` inside a block, decomposed row-major over
// dims[0..n-1] (last fastest) with the given memory strides.
__device__ __forceinline__ int32_t split_offset(int32_t x, const int32_t *dims, const int32_t *strides,
                                                int n) {
  int32_t off = 0;
  for (int j = n - 1; j >= 0; --j) {
    const int32_t d = dims[j];
    const int32_t q = x / d;
    off += (x - q * d) * strides[j];
    x = q;
  }
  return off;
}

template <bool CPLX>
__global__ __launch_bounds__(256) void k_grouped_gemm_f64(
    const Tile *__restrict__ tiles, const Seg *__restrict__ segs, const ExecTask *__restrict__ exec,
    const BlkGather *__restrict__ ga, const BlkGather *__restrict__ gb, const GatherLayout L,
    const double *__restrict__ A, const double *__restrict__ B, double *__restrict__ C, double alpha_r,
    double alpha_i, double beta_r, double beta_i) {
  __shared__ double As[CPLX ? 2 : 1][KC][LDP];
  __shared__ double Bs[CPLX ? 2 : 1][KC][LDP];
  __shared__ int32_t s_arow[TM], s_bcol[TN], s_acol[KC], s_brow[KC];

  const Tile tile = tiles[blockIdx.x];
  const Seg seg = segs[tile.seg];
  const int i0 = tile.tm * TM, j0 = tile.tn * TN;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
  const int lr = lane & 15, lk = lane >> 4;

  v4d acc[2][2], acc_i[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      acc[a][b] = v4d{0.0, 0.0, 0.0, 0.0};
      acc_i[a][b] = v4d{0.0, 0.0, 0.0, 0.0};
    }

  for (int t = seg.t0; t < seg.t1; ++t) {
    const ExecTask et = exec[t];
    const BlkGather &gA = ga[et.a_blk];
    const BlkGather &gB = gb[et.b_blk];
    __syncthreads();
    if (tid < TM) {
      const int r = i0 + tid;
      s_arow[tid] = (r < seg.m) ? split_offset(r, gA.rdim, gA.rstr, L.a_nr) : -1;
    } else if (tid < TM + TN) {
      const int c = j0 + tid - TM;
      s_bcol[tid - TM] = (c < seg.n) ? split_offset(c, gB.cdim, gB.cstr, L.b_nc) : -1;
    }
    for (int kk = 0; kk < et.k; kk += KC) {
      if (tid >= 128 && tid < 128 + KC) {
        const int p = kk + tid - 128;
        s_acol[tid - 128] = (p < et.k) ? split_offset(p, gA.cdim, gA.cstr, L.a_nc) : -1;
      } else if (tid >= 192 && tid < 192 + KC) {
        const int p = kk + tid - 192;
        s_brow[tid - 192] = (p < et.k) ? split_offset(p, gB.rdim, gB.rstr, L.b_nr) : -1;
      }
      __syncthreads();
      // ---- gather A'(i0.., kk..) -> As[p][r] ----
#pragma unroll
      for (int q = 0; q < (TM * KC) / 256; ++q) {
        int r, p;
        if (L.a_kfast) {
          p = tid & (KC - 1);
          r = (tid >> 4) + 16 * q;
        } else {
          r = tid & (TM - 1);
          p = (tid >> 6) + 4 * q;
        }
        const int32_t ro = s_arow[r], co = s_acol[p];
        const bool ok = (ro >= 0) && (co >= 0);
        const int64_t e = et.a_off + ro + co;
        if (CPLX) {
          double2 v = ok ? reinterpret_cast<const double2 *>(A)[e] : double2{0.0, 0.0};
          As[0][p][r] = v.x;
          As[1][p][r] = v.y;
        } else {
          As[0][p][r] = ok ? A[e] : 0.0;
        }
      }
      // ---- gather B'(kk.., j0..) -> Bs[p][c] ----
#pragma unroll
      for (int q = 0; q < (TN * KC) / 256; ++q) {
        int c, p;
        if (L.b_nfast) {
          c = tid & (TN - 1);
          p = (tid >> 6) + 4 * q;
        } else {
          p = tid & (KC - 1);
          c = (tid >> 4) + 16 * q;
        }
        const int32_t ro = s_brow[p], co = s_bcol[c];
        const bool ok = (ro >= 0) && (co >= 0);
        const int64_t e = et.b_off + ro + co;
        if (CPLX) {
          double2 v = ok ? reinterpret_cast<const double2 *>(B)[e] : double2{0.0, 0.0};
          Bs[0][p][c] = v.x;
          Bs[1][p][c] = v.y;
        } else {
          Bs[0][p][c] = ok ? B[e] : 0.0;
        }
      }
      __syncthreads();
      // ---- MFMA: each wave 32 x 32 as 2 x 2 tiles of 16 x 16 x 4 ----
#pragma unroll
      for (int ks = 0; ks < KC; ks += 4) {
        const double a0 = As[0][ks + lk][wm + lr];
        const double a1 = As[0][ks + lk][wm + 16 + lr];
        const double b0 = Bs[0][ks + lk][wn + lr];
        const double b1 = Bs[0][ks + lk][wn + 16 + lr];
        acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
        acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
        acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
        acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
        if (CPLX) {
          const double ai0 = As[1][ks + lk][wm + lr];
          const double ai1 = As[1][ks + lk][wm + 16 + lr];
          const double bi0 = Bs[1][ks + lk][wn + lr];
          const double bi1 = Bs[1][ks + lk][wn + 16 + lr];
          const double nai0 = -ai0, nai1 = -ai1;
          // re += ar*br - ai*bi ; im += ar*bi + ai*br
          acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(nai0, bi0, acc[0][0], 0, 0, 0);
          acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(nai0, bi1, acc[0][1], 0, 0, 0);
          acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(nai1, bi0, acc[1][0], 0, 0, 0);
          acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(nai1, bi1, acc[1][1], 0, 0, 0);
          acc_i[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, bi0, acc_i[0][0], 0, 0, 0);
          acc_i[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, bi1, acc_i[0][1], 0, 0, 0);
          acc_i[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, bi0, acc_i[1][0], 0, 0, 0);
          acc_i[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, bi1, acc_i[1][1], 0, 0, 0);
          acc_i[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai0, b0, acc_i[0][0], 0, 0, 0);
          acc_i[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai0, b1, acc_i[0][1], 0, 0, 0);
          acc_i[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai1, b0, acc_i[1][0], 0, 0, 0);
          acc_i[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai1, b1, acc_i[1][1], 0, 0, 0);
        }
      }
      __syncthreads();
    }
  }

  // ---- epilogue: C = alpha * acc + beta * C (C untouched when beta == 0) ----
  const bool use_beta = (beta_r != 0.0) || (beta_i != 0.0);
#pragma unroll
  for (int mt = 0; mt < 2; ++mt)
#pragma unroll
    for (int nt = 0; nt < 2; ++nt)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = i0 + wm + mt * 16 + 4 * v + lk;   // f64 16x16x4 output: row 4*reg + lane/16
        const int col = j0 + wn + nt * 16 + lr;
        if (row < seg.m && col < seg.n) {
          const int64_t idx = seg.c_off + static_cast<int64_t>(row) * seg.n + col;
          if (CPLX) {
            const double xr = acc[mt][nt][v], xi = acc_i[mt][nt][v];
            double yr = alpha_r * xr - alpha_i * xi, yi = alpha_r * xi + alpha_i * xr;
            double2 *cp = reinterpret_cast<double2 *>(C) + idx;
            if (use_beta) {
              const double2 o = *cp;
              yr += beta_r * o.x - beta_i * o.y;
              yi += beta_r * o.y + beta_i * o.x;
            }
            *cp = double2{yr, yi};
          } else {
            double y = alpha_r * acc[mt][nt][v];
            if (use_beta) y += beta_r * C[idx];
            C[idx] = y;
          }
        }
      }
}

// Row-streaming kernel for narrow outputs (n <= kRowsMaxN): one thread per
// output row, 256 rows per workgroup.  Each k-chunk of A' (256 rows x 16) is
// staged in LDS with loads contiguous along A's innermost memory axis (rows
// or k), the matching 16 x n chunk of B' is staged once per workgroup, and
// every thread accumulates its n outputs with FP64 FMAs.  The work is
// bound by reading A and writing C once.
constexpr int RK = 16;            // k per stage
constexpr int RLD = kRowsPerTile + 1;

template <bool CPLX>
__global__ __launch_bounds__(256) void k_rows_f64(
    const Tile *__restrict__ tiles, const Seg *__restrict__ segs, const ExecTask *__restrict__ exec,
    const BlkGather *__restrict__ ga, const BlkGather *__restrict__ gb, const GatherLayout L,
    const double *__restrict__ A, const double *__restrict__ B, double *__restrict__ C, double alpha_r,
    double alpha_i, double beta_r, double beta_i) {
  constexpr int P = CPLX ? 2 : 1;
  __shared__ double As[P][RK][RLD];
  __shared__ double Bs[P][RK][kRowsMaxN];
  __shared__ int32_t s_arow[kRowsPerTile];
  __shared__ int32_t s_acol[RK], s_brow[RK], s_bcol[kRowsMaxN];

  const Tile tile = tiles[blockIdx.x];
  const Seg seg = segs[tile.seg];
  const int tid = threadIdx.x;
  const int r0 = tile.tm * kRowsPerTile;
  const int row = r0 + tid;
  const int n = seg.n;

  double acc[kRowsMaxN], acc_i[kRowsMaxN];
#pragma unroll
  for (int j = 0; j < kRowsMaxN; ++j) {
    acc[j] = 0.0;
    acc_i[j] = 0.0;
  }

  for (int t = seg.t0; t < seg.t1; ++t) {
    const ExecTask et = exec[t];
    const BlkGather &gA = ga[et.a_blk];
    const BlkGather &gB = gb[et.b_blk];
    __syncthreads();
    s_arow[tid] = (row < seg.m) ? split_offset(row, gA.rdim, gA.rstr, L.a_nr) : -1;
    if (tid < kRowsMaxN) s_bcol[tid] = (tid < n) ? split_offset(tid, gB.cdim, gB.cstr, L.b_nc) : -1;
    for (int kk = 0; kk < et.k; kk += RK) {
      if (tid >= 128 && tid < 128 + RK) {
        const int p = kk + tid - 128;
        s_acol[tid - 128] = (p < et.k) ? split_offset(p, gA.cdim, gA.cstr, L.a_nc) : -1;
      } else if (tid >= 192 && tid < 192 + RK) {
        const int p = kk + tid - 192;
        s_brow[tid - 192] = (p < et.k) ? split_offset(p, gB.rdim, gB.rstr, L.b_nr) : -1;
      }
      __syncthreads();
      // A' chunk: 256 rows x RK
#pragma unroll 4
      for (int q = 0; q < RK; ++q) {
        int r, p;
        if (L.a_kfast) {
          p = tid & (RK - 1);
          r = (tid >> 4) + 16 * q;
        } else {
          r = tid;
          p = q;
        }
        const int32_t ro = s_arow[r], co = s_acol[p];
        const bool ok = (ro >= 0) && (co >= 0);
        const int64_t e = et.a_off + ro + co;
        if (CPLX) {
          const double2 v = ok ? reinterpret_cast<const double2 *>(A)[e] : double2{0.0, 0.0};
          As[0][p][r] = v.x;
          As[P - 1][p][r] = v.y;
        } else {
          As[0][p][r] = ok ? A[e] : 0.0;
        }
      }
      // B' chunk: RK x n
      {
        const int p = tid / kRowsMaxN, j = tid % kRowsMaxN;
        const int32_t ro = s_brow[p], co = s_bcol[j];
        const bool ok = (ro >= 0) && (co >= 0);
        const int64_t e = et.b_off + ro + co;
        if (CPLX) {
          const double2 v = ok ? reinterpret_cast<const double2 *>(B)[e] : double2{0.0, 0.0};
          Bs[0][p][j] = v.x;
          Bs[P - 1][p][j] = v.y;
        } else {
          Bs[0][p][j] = ok ? B[e] : 0.0;
        }
      }
      __syncthreads();
      const int kc = min(RK, et.k - kk);
      for (int p = 0; p < kc; ++p) {
        const double a = As[0][p][tid];
        const double a_im = As[P - 1][p][tid];
#pragma unroll
        for (int j = 0; j < kRowsMaxN; ++j) {
          if (j < n) {
            const double b = Bs[0][p][j];
            if (CPLX) {
              const double b_im = Bs[P - 1][p][j];
              acc[j] = fma(a, b, acc[j]);
              acc[j] = fma(-a_im, b_im, acc[j]);
              acc_i[j] = fma(a, b_im, acc_i[j]);
              acc_i[j] = fma(a_im, b, acc_i[j]);
            } else {
              acc[j] = fma(a, b, acc[j]);
            }
          }
        }
      }
      __syncthreads();
    }
  }

  if (row >= seg.m) return;
  const bool use_beta = (beta_r != 0.0) || (beta_i != 0.0);
  const int64_t base = seg.c_off + static_cast<int64_t>(row) * n;
#pragma unroll
  for (int j = 0; j < kRowsMaxN; ++j) {
    if (j < n) {
      if (CPLX) {
        double yr = alpha_r * acc[j] - alpha_i * acc_i[j], yi = alpha_r * acc_i[j] + alpha_i * acc[j];
        double2 *cp = reinterpret_cast<double2 *>(C) + base + j;
        if (use_beta) {
          const double2 o = *cp;
          yr += beta_r * o.x - beta_i * o.y;
          yi += beta_r * o.y + beta_i * o.x;
        }
        *cp = double2{yr, yi};
      } else {
        double y = alpha_r * acc[j];
        if (use_beta) y += beta_r * C[base + j];
        C[base + j] = y;
      }
    }
  }
}

}  // namespace

namespace qlt {

// Native grouped execution of a plan: MFMA tiles for the first tile class,
// the row-streaming kernel for the second.  `mfma_class` = false skips the
// MFMA class (the split path runs it on its own kernel).
qlt_status launch_native_classes(const qlt_plan *plan, int32_t dtype, const void *a, const void *b, void *c,
                                 const double *alpha, const double *beta, hipStream_t st, bool mfma_class) {
  const double ar = alpha[0], ai = (dtype == QLT_C128) ? alpha[1] : 0.0;
  const double br = beta[0], bi = (dtype == QLT_C128) ? beta[1] : 0.0;
  const dim3 block(256);
  const int64_t n0 = plan->cls_tiles[kTileMfma], n1 = plan->cls_tiles[kTileRows];
  const auto *A = static_cast<const double *>(a);
  const auto *B = static_cast<const double *>(b);
  auto *Cp = static_cast<double *>(c);
  if (mfma_class && n0 > 0) {
    if (dtype == QLT_F64)
      hipLaunchKernelGGL(k_grouped_gemm_f64<false>, dim3(static_cast<unsigned>(n0)), block, 0, st, plan->d_tiles,
                         plan->d_segs, plan->d_exec, plan->d_ga, plan->d_gb, plan->layout, A, B, Cp, ar, ai, br, bi);
    else
      hipLaunchKernelGGL(k_grouped_gemm_f64<true>, dim3(static_cast<unsigned>(n0)), block, 0, st, plan->d_tiles,
                         plan->d_segs, plan->d_exec, plan->d_ga, plan->d_gb, plan->layout, A, B, Cp, ar, ai, br, bi);
  }
  if (n1 > 0) {
    const Tile *t1 = plan->d_tiles + n0;
    if (dtype == QLT_F64)
      hipLaunchKernelGGL(k_rows_f64<false>, dim3(static_cast<unsigned>(n1)), block, 0, st, t1, plan->d_segs,
                         plan->d_exec, plan->d_ga, plan->d_gb, plan->layout, A, B, Cp, ar, ai, br, bi);
    else
      hipLaunchKernelGGL(k_rows_f64<true>, dim3(static_cast<unsigned>(n1)), block, 0, st, t1, plan->d_segs,
                         plan->d_exec, plan->d_ga, plan->d_gb, plan->layout, A, B, Cp, ar, ai, br, bi);
  }
  QLT_HIP(hipGetLastError());
  return QLT_OK;
}

qlt_status launch_native(const qlt_plan *plan, int32_t dtype, const void *a, const void *b, void *c,
                         const double *alpha, const double *beta, hipStream_t st) {
  return launch_native_classes(plan, dtype, a, b, c, alpha, beta, st, true);
}

}  // namespace qlt
