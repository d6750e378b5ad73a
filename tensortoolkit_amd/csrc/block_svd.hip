// Batched dense decompositions of the block matrices of a block-sparse
// tensor, in FP64 and complex128: SVD by one-sided block Jacobi, QR by
// Householder reflections.  They replace MatSVD (framework/hp_numeric/
// lapack.h:41-82 gesvd, GPU :811-1011) and MatQR (lapack.h:332-487 geqrf +
// orgqr/ungqr) called once per block matrix by DataBlkDecompSVD /
// DataBlkDecompQR (qltensor/blk_spar_data_ten/data_blk_operations.h:630-663,
// 1028-1046); here every block matrix of the tensor is decomposed in one
// batched launch sequence.
//
// SVD.  The working matrix X is A (rows >= cols) or A^H (rows < cols), stored
// column-major so that columns are contiguous; V starts as the identity.  The
// columns are split into blocks of kB; every round of a sweep pairs the
// blocks by the circle method, and for every block pair
//   1. one workgroup per kRG-row group forms that group's part of the Gram
//      matrix G = P^H P of the panel P = [X_I X_J],
//   2. one workgroup sums the parts in group order, skips the pair when
//      max |G_pq| / sqrt(G_pp G_qq) < tol, and otherwise diagonalises G in
//      LDS by cyclic two-sided Jacobi (at most kInnerSweeps sweeps),
//      accumulating the unitary Q (G <- J^H G J, Q <- Q J per rotation),
//   3. one workgroup per kRG-row group updates P <- P Q and the matching
//      columns of V <- V Q.
// Spreading the rows over workgroups keeps the chip busy on the few large
// block matrices, whose pairs set the length of a round.
// Before every sweep the columns are ordered by decreasing norm (the block
// order follows it), which keeps graded spectra to a few sweeps.  A matrix
// whose sweep rotates nothing has converged; then sigma_j = |X_j|,
// U_j = X_j / sigma_j (zero when sigma_j = 0) and A = U diag(sigma) V^H,
// sorted by decreasing sigma (ties by column).
//
// QR.  One workgroup per matrix runs the unblocked Householder QR with
// LAPACK's reflector convention (xLARFG: beta = -sign(Re alpha) * |(alpha, x)|,
// tau = (beta - alpha) / beta, v = x / (alpha - beta); no reflection when x = 0
// and alpha is real), so R has the same diagonal signs as geqrf, then forms
// Q = H_0 ... H_{k-1} I[:, :k] as orgqr / ungqr do.

#include <algorithm>
#include <cmath>
#include <vector>

#include "qlt_internal.h"

namespace {

struct cplx {
  double x, y;
};

__device__ __forceinline__ double re_(double a) { return a; }
__device__ __forceinline__ double re_(cplx a) { return a.x; }
__device__ __forceinline__ double conj_(double a) { return a; }
__device__ __forceinline__ cplx conj_(cplx a) { return {a.x, -a.y}; }
__device__ __forceinline__ double mul_(double a, double b) { return a * b; }
__device__ __forceinline__ cplx mul_(cplx a, cplx b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
// conj(a) * b
__device__ __forceinline__ double cmul_(double a, double b) { return a * b; }
__device__ __forceinline__ cplx cmul_(cplx a, cplx b) { return {a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x}; }
__device__ __forceinline__ double add_(double a, double b) { return a + b; }
__device__ __forceinline__ cplx add_(cplx a, cplx b) { return {a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ double sub_(double a, double b) { return a - b; }
__device__ __forceinline__ cplx sub_(cplx a, cplx b) { return {a.x - b.x, a.y - b.y}; }
__device__ __forceinline__ double scl_(double s, double a) { return s * a; }
__device__ __forceinline__ cplx scl_(double s, cplx a) { return {s * a.x, s * a.y}; }
__device__ __forceinline__ double abs2_(double a) { return a * a; }
__device__ __forceinline__ double abs2_(cplx a) { return a.x * a.x + a.y * a.y; }
template <typename T> __device__ __forceinline__ T zero_();
template <> __device__ __forceinline__ double zero_<double>() { return 0.0; }
template <> __device__ __forceinline__ cplx zero_<cplx>() { return {0.0, 0.0}; }
template <typename T> __device__ __forceinline__ T one_();
template <> __device__ __forceinline__ double one_<double>() { return 1.0; }
template <> __device__ __forceinline__ cplx one_<cplx>() { return {1.0, 0.0}; }
// unit phase of g (|g| > 0), conjugated: e^{-i arg g}, from 1 / |g|
__device__ __forceinline__ double uphase_conj_inv(double g, double inv) { return g * inv; }
__device__ __forceinline__ cplx uphase_conj_inv(cplx g, double inv) { return {g.x * inv, -g.y * inv}; }
// complex reciprocal
__device__ __forceinline__ double recip_(double a) { return 1.0 / a; }
__device__ __forceinline__ cplx recip_(cplx a) {
  const double d = a.x * a.x + a.y * a.y;
  return {a.x / d, -a.y / d};
}

template <bool C> struct ElemOf { using T = double; };
template <> struct ElemOf<true> { using T = cplx; };

__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}
__device__ __forceinline__ double shfl_(double v, int src) { return __shfl(v, src); }
__device__ __forceinline__ cplx shfl_(cplx v, int src) { return {__shfl(v.x, src), __shfl(v.y, src)}; }
__device__ __forceinline__ double wave_max(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  return v;
}

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;
constexpr int kB = 16;          // columns per Jacobi block
constexpr int kW = 2 * kB;      // panel width
constexpr int kRG = 64;         // rows per Gram / update workgroup
constexpr int kMaxSweeps = 40;
// Sweeps of the pair's Jacobi eigen-solve per visit: the outer sweeps repeat
// the pair until its columns are orthogonal, so the solve need not finish.
constexpr int kInnerSweeps = 1;
constexpr int kMaxSortCols = 4096;   // LDS bound of the per-matrix column sort

struct SvdMat {
  int64_t M, N;       // working shape (M >= N)
  int64_t x_off;      // X: column-major M x N, in elements of the workspace
  int64_t v_off;      // V: column-major N x N
  int64_t p_off;      // column order (int32), in int32 elements
  int64_t m, n;       // caller shape
  int64_t a_off, u_off, s_off, vt_off;
  int32_t nb, P;      // column blocks; players of the circle method (even)
  int32_t tr;         // X = A^H
  int32_t ng;         // row groups of X (kRG rows each)
  int32_t pair0;      // first pair slot of the matrix (P / 2 slots)
  int32_t pad;
  int64_t gram0;      // first partial-Gram slot (P / 2 * ng slots)
};

struct SvdItem {
  int32_t mat;
  int32_t k;          // pair index within a round
};

// One row group of one pair: rows g * kRG .. of X (kind 0) or of V (kind 1).
struct SvdRowItem {
  int32_t mat;
  int32_t k;
  int32_t g;
  int32_t kind;
};

// Circle-method pairing: `players` even, round r in [0, players - 1), pair
// k in [0, players / 2).  Both sums lie in [0, 2 (players - 1)), so one
// conditional subtraction reduces them modulo players - 1.
__device__ __forceinline__ void circle_pair(int players, int r, int k, int &a, int &b) {
  const int m = players - 1;
  if (k == 0) {
    a = m;
    b = r;
  } else {
    a = r + k;
    if (a >= m) a -= m;
    b = r - k + m;
    if (b >= m) b -= m;
  }
}

template <bool C>
__global__ void __launch_bounds__(kThreads) k_svd_load(const SvdMat *mats, const void *a_, void *ws_, int32_t *perm) {
  using T = typename ElemOf<C>::T;
  const SvdMat mt = mats[blockIdx.y];
  const T *a = static_cast<const T *>(a_) + mt.a_off;
  T *X = static_cast<T *>(ws_) + mt.x_off;
  T *V = static_cast<T *>(ws_) + mt.v_off;
  const int64_t MN = mt.M * mt.N, NN = mt.N * mt.N;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * kThreads;
  for (int64_t e = blockIdx.x * kThreads + threadIdx.x; e < MN; e += stride) {
    const int64_t j = e / mt.M, i = e - j * mt.M;   // X(i, j) column-major
    X[e] = mt.tr ? conj_(a[e]) : a[i * mt.n + j];
  }
  for (int64_t e = blockIdx.x * kThreads + threadIdx.x; e < NN; e += stride) {
    const int64_t j = e / mt.N, i = e - j * mt.N;
    V[e] = (i == j) ? one_<T>() : zero_<T>();
  }
  for (int64_t e = blockIdx.x * kThreads + threadIdx.x; e < mt.N; e += stride)
    perm[mt.p_off + e] = static_cast<int32_t>(e);
}

// Squared column norms of the working matrices, one wave per column:
// grid (column groups of kWaves, matrices); `all` = 0 skips the matrices
// that have converged.  norms[p_off + c] = |X_c|^2.
template <bool C>
__global__ void __launch_bounds__(kThreads) k_svd_norms(const SvdMat *mats, const void *ws_, const int32_t *active,
                                                         int32_t all, double *norms) {
  using T = typename ElemOf<C>::T;
  const int mi = blockIdx.y;
  if (!all && !active[mi]) return;
  const SvdMat mt = mats[mi];
  const int64_t c = int64_t(blockIdx.x) * kWaves + (threadIdx.x >> 6);
  if (c >= mt.N) return;
  const int lane = threadIdx.x & 63;
  const T *col = static_cast<const T *>(ws_) + mt.x_off + c * mt.M;
  double acc = 0.0;
  for (int64_t i = lane; i < mt.M; i += 64) acc += abs2_(col[i]);
  acc = wave_sum(acc);
  if (lane == 0) norms[mt.p_off + c] = acc;
}

// The matrix's squared column norms into sn[0..N), then ord[rank] = column
// by decreasing norm (ties by column index).
__device__ void order_by_norms(const double *norms, int64_t N, double *sn, int32_t *ord) {
  for (int64_t c = threadIdx.x; c < N; c += kThreads) sn[c] = norms[c];
  __syncthreads();
  for (int64_t c = threadIdx.x; c < N; c += kThreads) {
    const double v = sn[c];
    int32_t rank = 0;
    for (int64_t j = 0; j < N; ++j) {
      const double w = sn[j];
      rank += (w > v) || (w == v && j < c);
    }
    ord[rank] = static_cast<int32_t>(c);
  }
  __syncthreads();
}

// Column order of the next sweep: the physical columns by decreasing norm
// (the new order is a permutation of physical columns, so the previous order
// is not read).
__global__ void __launch_bounds__(kThreads) k_svd_sort(const SvdMat *mats, const double *norms, int32_t *perm,
                                                        const int32_t *active, int32_t all) {
  __shared__ double sn[kMaxSortCols];
  __shared__ int32_t so[kMaxSortCols];
  const int mi = blockIdx.x;
  if (!all && !active[mi]) return;
  const SvdMat mt = mats[mi];
  if (mt.N > kMaxSortCols) return;   // keep the previous order
  order_by_norms(norms + mt.p_off, mt.N, sn, so);
  for (int64_t c = threadIdx.x; c < mt.N; c += kThreads) perm[mt.p_off + c] = so[c];
}

// The pair (bi, bj) of round `round`, its column count w and its columns.
// Returns false when the matrix sits this round out or the pair has < 2 columns.
__device__ __forceinline__ bool pair_cols(const SvdMat &mt, int32_t round, int32_t k, const int32_t *perm,
                                          int32_t *sCol, int &w) {
  if (round >= mt.P - 1) return false;
  int bi, bj;
  circle_pair(mt.P, round, k, bi, bj);
  const int wi = bi < mt.nb ? static_cast<int>(min<int64_t>(kB, mt.N - int64_t(bi) * kB)) : 0;
  const int wj = bj < mt.nb ? static_cast<int>(min<int64_t>(kB, mt.N - int64_t(bj) * kB)) : 0;
  w = wi + wj;
  if (w < 2) return false;
  const int32_t *pm = perm + mt.p_off;
  const int tid = threadIdx.x;
  if (tid < w) sCol[tid] = tid < wi ? pm[int64_t(bi) * kB + tid] : pm[int64_t(bj) * kB + (tid - wi)];
  return true;
}

// Round phase 1: the partial Gram matrix P^H P of one kRG-row group of the
// pair's panel P = [X_I X_J], stored w x w with row stride kW.
// Stage rows rb .. re of the panel's w columns (column c = X column sCol[c])
// into LDS as real and imaginary planes sP[part][c][row], zero beyond the
// rows and columns.
template <bool C, typename T>
__device__ __forceinline__ void stage_panel(const T *base, int64_t rows, int64_t rb, int64_t re, int w,
                                            const int32_t *sCol, double (*sP)[kW][kRG + 1]) {
  for (int idx = threadIdx.x; idx < kW * kRG; idx += kThreads) {
    const int c = idx / kRG, rr = idx - c * kRG;
    const int64_t row = rb + rr;
    T v = zero_<T>();
    if (c < w && row < re) v = base[int64_t(sCol[c]) * rows + row];
    if constexpr (C) {
      sP[0][c][rr] = v.x;
      sP[1][c][rr] = v.y;
    } else {
      sP[0][c][rr] = v;
    }
  }
}

// One row group's part of the pair's Gram matrix, G = P^H P over rows
// g kRG .. (g + 1) kRG - 1, on v_mfma_f64_16x16x4f64: wave wv computes the
// 16 x 16 tile (wv / 2, wv % 2) of the 32 x 32 result over 16 k-steps of 4
// rows.  Complex: Re G = Pr^T Pr + Pi^T Pi, Im G = Pr^T Pi - Pi^T Pr.
template <bool C>
__global__ void __launch_bounds__(kThreads) k_svd_gram(const SvdMat *mats, const SvdRowItem *items, int32_t round,
                                                        const void *ws_, const int32_t *perm, const int32_t *active,
                                                        void *part_) {
  using T = typename ElemOf<C>::T;
  constexpr int PARTS = C ? 2 : 1;
  __shared__ double sP[PARTS][kW][kRG + 1];
  __shared__ int32_t sCol[kW];
  const SvdRowItem it = items[blockIdx.x];
  if (!active[it.mat]) return;
  const SvdMat mt = mats[it.mat];
  int w = 0;
  if (!pair_cols(mt, round, it.k, perm, sCol, w)) return;
  const int tid = threadIdx.x;
  const T *X = static_cast<const T *>(ws_) + mt.x_off;
  __syncthreads();
  const int64_t rb = int64_t(it.g) * kRG, re = min<int64_t>(mt.M, rb + kRG);
  stage_panel<C>(X, mt.M, rb, re, w, sCol, sP);
  __syncthreads();
  const int lane = tid & 63, wv = tid >> 6, lr = lane & 15, lk = lane >> 4;
  const int p0 = 16 * (wv >> 1), q0 = 16 * (wv & 1);
  typedef double v4d __attribute__((ext_vector_type(4)));
  v4d acc = {0.0, 0.0, 0.0, 0.0}, acc_i = {0.0, 0.0, 0.0, 0.0};
  if (p0 < w && q0 < w) {
#pragma unroll 4
    for (int ks = 0; ks < kRG; ks += 4) {
      const double ar = sP[0][p0 + lr][ks + lk], br = sP[0][q0 + lr][ks + lk];
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, br, acc, 0, 0, 0);
      if constexpr (C) {
        const double ai = sP[1][p0 + lr][ks + lk], bi = sP[1][q0 + lr][ks + lk];
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(ai, bi, acc, 0, 0, 0);
        acc_i = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, bi, acc_i, 0, 0, 0);
        acc_i = __builtin_amdgcn_mfma_f64_16x16x4f64(-ai, br, acc_i, 0, 0, 0);
      }
    }
  }
  T *part = static_cast<T *>(part_) + (mt.gram0 + int64_t(it.k) * mt.ng + it.g) * (kW * kW);
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int p = p0 + 4 * v + lk, q = q0 + lr;
    if (p < w && q < w) {
      if constexpr (C) part[p * kW + q] = cplx{acc[v], acc_i[v]};
      else part[p * kW + q] = acc[v];
    }
  }
}

// Round phase 2, one workgroup per pair: the Gram matrix as the sum of the
// row groups' partials (in group order), the pair's convergence test, and the
// cyclic two-sided Jacobi eigen-solve of the Gram in LDS, whose unitary Q
// (w x w, row stride kW) phase 3 applies.  qflag = 1 when Q is to be applied.
template <bool C>
__global__ void __launch_bounds__(kThreads) k_svd_solve(const SvdMat *mats, const SvdItem *items, int32_t round,
                                                         const void *part_, const int32_t *perm,
                                                         const int32_t *active, void *q_, int32_t *qflag,
                                                         int32_t *rotated) {
  using T = typename ElemOf<C>::T;
  __shared__ T sG[2][kW][kW + 1];   // G of the current round, and the next
  __shared__ T sQ[kW][kW + 1];
  __shared__ int32_t sCol[kW];
  __shared__ double sRed[kWaves];

  const SvdItem it = items[blockIdx.x];
  if (!active[it.mat]) return;
  const SvdMat mt = mats[it.mat];
  const int tid = threadIdx.x;
  int w = 0;
  if (!pair_cols(mt, round, it.k, perm, sCol, w)) return;
  int32_t *flag = qflag + mt.pair0 + it.k;
  const T *part = static_cast<const T *>(part_) + (mt.gram0 + int64_t(it.k) * mt.ng) * (kW * kW);
  for (int e = tid; e < w * w; e += kThreads) {
    const int p = e / w, q = e - p * w;
    T s = zero_<T>();
    for (int g = 0; g < mt.ng; ++g) s = add_(s, part[int64_t(g) * (kW * kW) + p * kW + q]);
    sG[0][p][q] = s;
  }
  __syncthreads();

  // ---- convergence test of the pair ----
  double off = 0.0;
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int e = tid + u * kThreads;
    if (e < w * w) {
      const int p = e / w, q = e - p * w;
      if (p < q) {
        const double dp = re_(sG[0][p][p]), dq = re_(sG[0][q][q]);
        if (dp > 0.0 && dq > 0.0) off = fmax(off, abs2_(sG[0][p][q]) / (dp * dq));
      }
    }
  }
  off = wave_max(off);
  if ((tid & 63) == 0) sRed[tid >> 6] = off;
  __syncthreads();
  const double tol = fmax(16.0, static_cast<double>(mt.M)) * 2.220446049250313e-16;
  const double offmax = fmax(fmax(sRed[0], sRed[1]), fmax(sRed[2], sRed[3]));
  if (offmax < tol * tol) {
    if (tid == 0) *flag = 0;
    return;
  }
  if (tid == 0) {
    rotated[it.mat] = 1;
    *flag = 1;
  }

  // ---- Jacobi diagonalisation of G in LDS ----
  // A round applies the rotations of P2 / 2 disjoint index pairs,
  // G <- J^H G J and Q <- Q J, as 2 x 2 blocks: thread (bi, bj) takes the
  // block of G at rows (p, q) of pair bi and columns of pair bj, applies the
  // rotation of pair bj to its columns and then that of pair bi to its rows
  // (the same operations, in the same order, as a column pass followed by a
  // row pass), and writes the result into the other G buffer; it also
  // rotates rows 2 bi, 2 bi + 1 of Q in the columns of pair bj, in place.
  // Every wave computes the round's rotations from the round's G, pair k in
  // lane k (and its copies k + 16, ...), and each lane takes those of pairs
  // bi and bj by shuffle, so a round takes one barrier.
  for (int e = tid; e < w * w; e += kThreads) {
    const int p = e / w, q = e - p * w;
    sQ[p][q] = (p == q) ? one_<T>() : zero_<T>();
  }
  const int P2 = w + (w & 1);
  const double eps = 2.220446049250313e-16;
  static_assert(kW == 32 && kThreads == 256, "one thread per 2 x 2 block of pairs");
  const int bi = tid / (kW / 2), bj = tid % (kW / 2);
  const bool blk_on = bi < P2 / 2 && bj < P2 / 2;
  // rotation of pair k of round r from the round's G: indexes, c, s and the
  // unit phase; false (identity) when the pair is orthogonal or padded
  auto rotation = [&](const T (*G)[kW + 1], int r, int k, int &p, int &q, double &c, double &sn, T &em) {
    circle_pair(P2, r, k, p, q);
    c = 1.0;
    sn = 0.0;
    em = one_<T>();
    if (p >= w || q >= w) return false;
    const double al = re_(G[p][p]), be = re_(G[q][q]);
    const T g = G[p][q];
    // |g| > eps sqrt(al be), squared: three square roots and three divisions
    // per rotation, the round's latency chain
    const double ag2 = abs2_(g);
    if (!(ag2 > 0.0 && al > 0.0 && be > 0.0 && ag2 > (eps * eps) * (al * be))) return false;
    const double ag = sqrt(ag2);
    const double inv = 1.0 / ag;
    const double zeta = (be - al) * (0.5 * inv);
    double t;
    if (fabs(zeta) > 1e150) t = 0.5 / zeta;
    else t = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
    c = 1.0 / sqrt(1.0 + t * t);
    sn = c * t;
    em = uphase_conj_inv(g, inv);
    return true;
  };
  __syncthreads();
  int cur = 0;
  for (int sweep = 0; sweep < kInnerSweeps; ++sweep) {
    int any = 0;
    for (int r = 0; r < P2 - 1; ++r) {
      const T(*G)[kW + 1] = sG[cur];
      const int kr = (tid & 63) % (kW / 2);
      int rp = 0, rq = 0;
      double rc = 1.0, rs = 0.0;
      T rem = one_<T>();
      const bool ron = kr < P2 / 2 && rotation(G, r, kr, rp, rq, rc, rs, rem);
      any |= ron ? 1 : 0;
      const int pi = __shfl(rp, bi), qi = __shfl(rq, bi), pj = __shfl(rp, bj), qj = __shfl(rq, bj);
      const bool oni = __shfl(ron ? 1 : 0, bi) != 0, onj = __shfl(ron ? 1 : 0, bj) != 0;
      const double ci = __shfl(rc, bi), si = __shfl(rs, bi), cj = __shfl(rc, bj), sj = __shfl(rs, bj);
      const T emi = shfl_(rem, bi), emj = shfl_(rem, bj);
      if (blk_on) {
        T(*Gn)[kW + 1] = sG[cur ^ 1];
        const int rows[2] = {pi, qi}, cols[2] = {pj, qj};
        T x[2][2];
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
          for (int b = 0; b < 2; ++b) x[a][b] = (rows[a] < w && cols[b] < w) ? G[rows[a]][cols[b]] : zero_<T>();
        // columns:  [x_p, x_q] <- [x_p, x_q] J,  J = [[c, s], [-s e^{-i phi}, c e^{-i phi}]]
        if (onj) {
#pragma unroll
          for (int a = 0; a < 2; ++a) {
            const T xp = x[a][0], emxq = mul_(emj, x[a][1]);
            x[a][0] = sub_(scl_(cj, xp), scl_(sj, emxq));
            x[a][1] = add_(scl_(sj, xp), scl_(cj, emxq));
          }
        }
        // rows:  [g_p; g_q] <- J^H [g_p; g_q]; the pair's own off-diagonal entries become 0
        if (oni) {
          const T ep = conj_(emi);
#pragma unroll
          for (int b = 0; b < 2; ++b) {
            const T gp = x[0][b], epgq = mul_(ep, x[1][b]);
            T np = sub_(scl_(ci, gp), scl_(si, epgq));
            T nq = add_(scl_(si, gp), scl_(ci, epgq));
            if (cols[b] == qi) np = zero_<T>();
            if (cols[b] == pi) nq = zero_<T>();
            x[0][b] = np;
            x[1][b] = nq;
          }
        }
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
          for (int b = 0; b < 2; ++b)
            if (rows[a] < w && cols[b] < w) Gn[rows[a]][cols[b]] = x[a][b];
        // Q: rows 2 bi, 2 bi + 1, columns of pair bj
        if (onj) {
#pragma unroll
          for (int a = 0; a < 2; ++a) {
            const int i = 2 * bi + a;
            if (i >= w) continue;
            const T xp = sQ[i][pj], emxq = mul_(emj, sQ[i][qj]);
            sQ[i][pj] = sub_(scl_(cj, xp), scl_(sj, emxq));
            sQ[i][qj] = add_(scl_(sj, xp), scl_(cj, emxq));
          }
        }
      }
      __syncthreads();
      cur ^= 1;
    }
    if (!__syncthreads_or(any)) break;
  }
  T *Q = static_cast<T *>(q_) + int64_t(mt.pair0 + it.k) * (kW * kW);
  for (int e = tid; e < w * w; e += kThreads) {
    const int p = e / w, q = e - p * w;
    Q[p * kW + q] = sQ[p][q];
  }
}

// Round phase 3: P <- P Q on one kRG-row group of the pair's columns of X
// (kind 0) or of V (kind 1).
// P <- P Q (and V <- V Q) for one row group of the pair, on
// v_mfma_f64_16x16x4f64: wave wv computes rows 16 wv .. 16 wv + 15 of the
// kRG-row group, in two 16-column tiles, over k-steps of 4 of the pair's w
// columns.  Complex: Re = Pr Qr - Pi Qi, Im = Pr Qi + Pi Qr.
template <bool C>
__global__ void __launch_bounds__(kThreads) k_svd_update(const SvdMat *mats, const SvdRowItem *items, int32_t round,
                                                          void *ws_, const int32_t *perm, const int32_t *active,
                                                          const void *q_, const int32_t *qflag) {
  using T = typename ElemOf<C>::T;
  constexpr int PARTS = C ? 2 : 1;
  __shared__ double sP[PARTS][kW][kRG + 1];
  __shared__ double sQ[PARTS][kW][kW];
  __shared__ int32_t sCol[kW];
  const SvdRowItem it = items[blockIdx.x];
  if (!active[it.mat]) return;
  const SvdMat mt = mats[it.mat];
  int w = 0;
  if (!pair_cols(mt, round, it.k, perm, sCol, w)) return;
  if (!qflag[mt.pair0 + it.k]) return;
  const int tid = threadIdx.x;
  const T *Q = static_cast<const T *>(q_) + int64_t(mt.pair0 + it.k) * (kW * kW);
  for (int e = tid; e < kW * kW; e += kThreads) {
    const int p = e / kW, q = e - p * kW;
    const T v = (p < w && q < w) ? Q[p * kW + q] : zero_<T>();
    if constexpr (C) {
      sQ[0][p][q] = v.x;
      sQ[1][p][q] = v.y;
    } else {
      sQ[0][p][q] = v;
    }
  }
  T *base = static_cast<T *>(ws_) + (it.kind ? mt.v_off : mt.x_off);
  const int64_t rows = it.kind ? mt.N : mt.M;
  const int64_t rb = int64_t(it.g) * kRG, re = min<int64_t>(rows, rb + kRG);
  __syncthreads();
  stage_panel<C>(base, rows, rb, re, w, sCol, sP);
  __syncthreads();
  const int lane = tid & 63, wv = tid >> 6, lr = lane & 15, lk = lane >> 4;
  typedef double v4d __attribute__((ext_vector_type(4)));
  v4d acc[2] = {{0.0, 0.0, 0.0, 0.0}, {0.0, 0.0, 0.0, 0.0}};
  v4d acc_i[2] = {{0.0, 0.0, 0.0, 0.0}, {0.0, 0.0, 0.0, 0.0}};
  if (rb + 16 * wv < re) {
    for (int ks = 0; ks < w; ks += 4) {
      const double ar = sP[0][ks + lk][16 * wv + lr];
      const double ai = C ? sP[PARTS - 1][ks + lk][16 * wv + lr] : 0.0;
#pragma unroll
      for (int nt = 0; nt < 2; ++nt) {
        const double br = sQ[0][ks + lk][16 * nt + lr];
        acc[nt] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, br, acc[nt], 0, 0, 0);
        if constexpr (C) {
          const double bi = sQ[1][ks + lk][16 * nt + lr];
          acc[nt] = __builtin_amdgcn_mfma_f64_16x16x4f64(-ai, bi, acc[nt], 0, 0, 0);
          acc_i[nt] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, bi, acc_i[nt], 0, 0, 0);
          acc_i[nt] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai, br, acc_i[nt], 0, 0, 0);
        }
      }
    }
  }
  // every read of the rows is done (they sit in LDS): write the new columns
#pragma unroll
  for (int nt = 0; nt < 2; ++nt)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int64_t row = rb + 16 * wv + 4 * v + lk;
      const int j = 16 * nt + lr;
      if (row < re && j < w) {
        T *dst = base + int64_t(sCol[j]) * rows + row;
        if constexpr (C) *dst = cplx{acc[nt][v], acc_i[nt][v]};
        else *dst = acc[nt][v];
      }
    }
}

// Singular values and vectors, grid (element groups, matrices): the columns
// in the order of `ord` (decreasing norm, written by k_svd_sort), sigma_k =
// sqrt(norm of column ord[k]).
template <bool C>
__global__ void __launch_bounds__(kThreads) k_svd_finish(const SvdMat *mats, const void *ws_, const double *norms,
                                                          const int32_t *ord, void *u_, double *s_, void *vt_) {
  using T = typename ElemOf<C>::T;
  const SvdMat mt = mats[blockIdx.y];
  const T *X = static_cast<const T *>(ws_) + mt.x_off;
  const T *V = static_cast<const T *>(ws_) + mt.v_off;
  const double *nm = norms + mt.p_off;
  const int32_t *so = ord + mt.p_off;
  T *U = static_cast<T *>(u_) + mt.u_off;
  T *Vt = static_cast<T *>(vt_) + mt.vt_off;
  double *S = s_ + mt.s_off;
  const int64_t r = mt.N;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * kThreads;
  const int64_t e0 = blockIdx.x * kThreads + threadIdx.x;
  for (int64_t k = e0; k < r; k += stride) S[k] = sqrt(nm[so[k]]);
  if (!mt.tr) {
    // A = X_final V^H:  U (m x r) = X_col / sigma, Vt (r x n) = V^H
    for (int64_t e = e0; e < mt.m * r; e += stride) {
      const int64_t i = e / r, k = e - i * r;
      const int32_t c = so[k];
      const double sg = sqrt(nm[c]);
      U[e] = sg > 0.0 ? scl_(1.0 / sg, X[int64_t(c) * mt.M + i]) : zero_<T>();
    }
    for (int64_t e = e0; e < r * mt.n; e += stride) {
      const int64_t k = e / mt.n, j = e - k * mt.n;
      Vt[e] = conj_(V[int64_t(so[k]) * mt.N + j]);
    }
  } else {
    // A^H = X_final V^H  =>  A = V X_final^H:  U (m x r) = V, Vt (r x n) = (X_col / sigma)^H
    for (int64_t e = e0; e < mt.m * r; e += stride) {
      const int64_t i = e / r, k = e - i * r;
      U[e] = V[int64_t(so[k]) * mt.N + i];
    }
    for (int64_t e = e0; e < r * mt.n; e += stride) {
      const int64_t k = e / mt.n, j = e - k * mt.n;
      const int32_t c = so[k];
      const double sg = sqrt(nm[c]);
      Vt[e] = sg > 0.0 ? scl_(1.0 / sg, conj_(X[int64_t(c) * mt.M + j])) : zero_<T>();
    }
  }
}

// ---- QR ----

struct QrMat {
  int64_t m, n, k;
  int64_t w_off;      // W: column-major m x n (becomes R and the reflectors)
  int64_t q_off_ws;   // Qw: column-major m x k
  int64_t t_off;      // tau (k elements)
  int64_t a_off, q_off, r_off;
};

template <typename T>
__device__ __forceinline__ double block_sum(double v, double *red) {
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  const double s = red[0] + red[1] + red[2] + red[3];
  __syncthreads();
  return s;
}

template <bool C>
__global__ void __launch_bounds__(kThreads) k_qr(const QrMat *mats, const void *a_, void *ws_, void *q_, void *r_) {
  using T = typename ElemOf<C>::T;
  __shared__ double sRed[kWaves];
  __shared__ T sTau, sScale;
  const QrMat mt = mats[blockIdx.x];
  const T *A = static_cast<const T *>(a_) + mt.a_off;
  T *Wm = static_cast<T *>(ws_) + mt.w_off;
  T *Qw = static_cast<T *>(ws_) + mt.q_off_ws;
  T *tau = static_cast<T *>(ws_) + mt.t_off;
  const int64_t m = mt.m, n = mt.n, k = mt.k;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int64_t e = threadIdx.x; e < m * n; e += kThreads) {
    const int64_t j = e / m, i = e - j * m;
    Wm[e] = A[i * n + j];
  }
  __syncthreads();
  for (int64_t j = 0; j < k; ++j) {
    T *col = Wm + j * m;
    double part = 0.0;
    for (int64_t i = j + 1 + threadIdx.x; i < m; i += kThreads) part += abs2_(col[i]);
    const double xn2 = block_sum<T>(part, sRed);
    if (threadIdx.x == 0) {
      const T alpha = col[j];
      T t = zero_<T>(), sc = zero_<T>();
      const double ai = C ? (abs2_(alpha) - re_(alpha) * re_(alpha)) : 0.0;   // (Im alpha)^2
      if (!(xn2 == 0.0 && ai == 0.0)) {
        const double nrm = sqrt(abs2_(alpha) + xn2);
        const double beta = re_(alpha) >= 0.0 ? -nrm : nrm;
        T bt = zero_<T>();
        bt = add_(bt, scl_(beta, one_<T>()));
        t = scl_(1.0 / beta, sub_(bt, alpha));
        sc = recip_(sub_(alpha, bt));
        col[j] = bt;
      }
      sTau = t;
      sScale = sc;
      tau[j] = t;
    }
    __syncthreads();
    const T t = sTau;
    if (abs2_(t) == 0.0) continue;
    const T sc = sScale;
    for (int64_t i = j + 1 + threadIdx.x; i < m; i += kThreads) col[i] = mul_(sc, col[i]);
    __syncthreads();
    // trailing columns: x <- H^H x = x - conj(tau) v (v^H x)
    const T ct = conj_(t);
    for (int64_t c = j + 1 + wave; c < n; c += kWaves) {
      T *x = Wm + c * m;
      T d = zero_<T>();
      for (int64_t i = j + lane; i < m; i += 64) {
        const T vi = (i == j) ? one_<T>() : col[i];
        d = add_(d, cmul_(vi, x[i]));
      }
      if constexpr (C) {
        d.x = wave_sum(d.x);
        d.y = wave_sum(d.y);
      } else {
        d = wave_sum(d);
      }
      const T f = mul_(ct, d);
      for (int64_t i = j + lane; i < m; i += 64) {
        const T vi = (i == j) ? one_<T>() : col[i];
        x[i] = sub_(x[i], mul_(vi, f));
      }
    }
    __syncthreads();
  }
  // R (k x n), upper triangle of W
  T *R = static_cast<T *>(r_) + mt.r_off;
  for (int64_t e = threadIdx.x; e < k * n; e += kThreads) {
    const int64_t i = e / n, c = e - i * n;
    R[e] = c >= i ? Wm[c * m + i] : zero_<T>();
  }
  // Q = H_0 ... H_{k-1} I[:, :k]
  for (int64_t e = threadIdx.x; e < m * k; e += kThreads) {
    const int64_t c = e / m, i = e - c * m;
    Qw[e] = (i == c) ? one_<T>() : zero_<T>();
  }
  __syncthreads();
  for (int64_t j = k - 1; j >= 0; --j) {
    const T t = tau[j];
    if (abs2_(t) == 0.0) continue;
    const T *v = Wm + j * m;
    for (int64_t c = j + wave; c < k; c += kWaves) {
      T *x = Qw + c * m;
      T d = zero_<T>();
      for (int64_t i = j + lane; i < m; i += 64) {
        const T vi = (i == j) ? one_<T>() : v[i];
        d = add_(d, cmul_(vi, x[i]));
      }
      if constexpr (C) {
        d.x = wave_sum(d.x);
        d.y = wave_sum(d.y);
      } else {
        d = wave_sum(d);
      }
      const T f = mul_(t, d);
      for (int64_t i = j + lane; i < m; i += 64) {
        const T vi = (i == j) ? one_<T>() : v[i];
        x[i] = sub_(x[i], mul_(vi, f));
      }
    }
    __syncthreads();
  }
  T *Q = static_cast<T *>(q_) + mt.q_off;
  for (int64_t e = threadIdx.x; e < m * k; e += kThreads) {
    const int64_t i = e / k, c = e - i * k;
    Q[e] = Qw[c * m + i];
  }
}

struct DevBuf {
  void *p = nullptr;
  ~DevBuf() {
    if (p != nullptr) (void)hipFree(p);
  }
};

qlt_status check_shapes(const char *who, int64_t n_mats, const int64_t *rows, const int64_t *cols) {
  QLT_CHECK_ARG(n_mats >= 0, std::string(who) + ": negative matrix count");
  QLT_CHECK_ARG(n_mats == 0 || (rows != nullptr && cols != nullptr), std::string(who) + ": null shape array");
  QLT_CHECK_ARG(n_mats < (int64_t(1) << 31), std::string(who) + ": too many matrices");
  for (int64_t i = 0; i < n_mats; ++i)
    QLT_CHECK_ARG(rows[i] >= 1 && cols[i] >= 1, std::string(who) + ": matrix " + std::to_string(i) + " is empty");
  return QLT_OK;
}

template <bool C>
qlt_status svd_impl(int64_t n_mats, const int64_t *rows, const int64_t *cols, const void *a, const int64_t *a_offsets,
                    void *u, const int64_t *u_offsets, double *s, const int64_t *s_offsets, void *vt,
                    const int64_t *vt_offsets, int32_t *sweeps_out, hipStream_t st) {
  const size_t esz = C ? 16 : 8;
  QLT_CHECK_ARG(n_mats <= 65535, "qlt_block_svd: more than 65535 matrices in one call");
  std::vector<SvdMat> mats(static_cast<size_t>(n_mats));
  std::vector<SvdItem> items;
  std::vector<SvdRowItem> gitems, uitems;
  int64_t ws_elems = 0, perm_elems = 0, max_elems = 0, gram_slots = 0;
  int32_t max_players = 0, pair_slots = 0;
  for (int64_t i = 0; i < n_mats; ++i) {
    SvdMat &mt = mats[static_cast<size_t>(i)];
    mt.m = rows[i];
    mt.n = cols[i];
    mt.tr = rows[i] < cols[i];
    mt.M = std::max(rows[i], cols[i]);
    mt.N = std::min(rows[i], cols[i]);
    QLT_CHECK_ARG(mt.N <= kMaxSortCols, "qlt_block_svd: min(rows, cols) of matrix " + std::to_string(i) +
                                            " exceeds " + std::to_string(kMaxSortCols));
    mt.x_off = ws_elems;
    ws_elems += mt.M * mt.N;
    mt.v_off = ws_elems;
    ws_elems += mt.N * mt.N;
    mt.p_off = perm_elems;
    perm_elems += mt.N;
    mt.a_off = a_offsets[i];
    mt.u_off = u_offsets[i];
    mt.s_off = s_offsets[i];
    mt.vt_off = vt_offsets[i];
    mt.nb = static_cast<int32_t>((mt.N + kB - 1) / kB);
    mt.P = mt.nb == 1 ? 2 : mt.nb + (mt.nb & 1);
    mt.ng = static_cast<int32_t>((mt.M + kRG - 1) / kRG);
    mt.pair0 = pair_slots;
    mt.pad = 0;
    mt.gram0 = gram_slots;
    max_elems = std::max(max_elems, mt.M * mt.N);
    if (mt.N >= 2) {
      max_players = std::max(max_players, mt.P);
      const int32_t ngv = static_cast<int32_t>((mt.N + kRG - 1) / kRG);
      for (int32_t k = 0; k < mt.P / 2; ++k) {
        items.push_back(SvdItem{static_cast<int32_t>(i), k});
        for (int32_t g = 0; g < mt.ng; ++g) {
          gitems.push_back(SvdRowItem{static_cast<int32_t>(i), k, g, 0});
          uitems.push_back(SvdRowItem{static_cast<int32_t>(i), k, g, 0});
        }
        for (int32_t g = 0; g < ngv; ++g) uitems.push_back(SvdRowItem{static_cast<int32_t>(i), k, g, 1});
      }
      pair_slots += mt.P / 2;
      gram_slots += int64_t(mt.P / 2) * mt.ng;
    }
  }
  QLT_CHECK_ARG(gitems.size() < (size_t(1) << 31) && uitems.size() < (size_t(1) << 31),
                "qlt_block_svd: too many row groups in one call");
  DevBuf d_ws, d_mats, d_items, d_perm, d_flags, d_gitems, d_uitems, d_part, d_q, d_qflag, d_norms;
  QLT_HIP(hipMalloc(&d_ws.p, std::max<size_t>(1, static_cast<size_t>(ws_elems) * esz)));
  QLT_HIP(hipMalloc(&d_mats.p, sizeof(SvdMat) * mats.size()));
  QLT_HIP(hipMalloc(&d_items.p, sizeof(SvdItem) * std::max<size_t>(1, items.size())));
  QLT_HIP(hipMalloc(&d_gitems.p, sizeof(SvdRowItem) * std::max<size_t>(1, gitems.size())));
  QLT_HIP(hipMalloc(&d_uitems.p, sizeof(SvdRowItem) * std::max<size_t>(1, uitems.size())));
  QLT_HIP(hipMalloc(&d_part.p, esz * kW * kW * static_cast<size_t>(std::max<int64_t>(1, gram_slots))));
  QLT_HIP(hipMalloc(&d_q.p, esz * kW * kW * static_cast<size_t>(std::max<int32_t>(1, pair_slots))));
  QLT_HIP(hipMalloc(&d_qflag.p, sizeof(int32_t) * static_cast<size_t>(std::max<int32_t>(1, pair_slots))));
  QLT_HIP(hipMalloc(&d_perm.p, sizeof(int32_t) * std::max<int64_t>(1, perm_elems)));
  QLT_HIP(hipMalloc(&d_norms.p, sizeof(double) * std::max<int64_t>(1, perm_elems)));
  QLT_HIP(hipMalloc(&d_flags.p, sizeof(int32_t) * 2 * mats.size()));
  auto *active = static_cast<int32_t *>(d_flags.p);
  auto *rotated = active + mats.size();
  QLT_HIP(hipMemcpyAsync(d_mats.p, mats.data(), sizeof(SvdMat) * mats.size(), hipMemcpyHostToDevice, st));
  if (!items.empty()) {
    QLT_HIP(hipMemcpyAsync(d_items.p, items.data(), sizeof(SvdItem) * items.size(), hipMemcpyHostToDevice, st));
    QLT_HIP(hipMemcpyAsync(d_gitems.p, gitems.data(), sizeof(SvdRowItem) * gitems.size(), hipMemcpyHostToDevice,
                           st));
    QLT_HIP(hipMemcpyAsync(d_uitems.p, uitems.data(), sizeof(SvdRowItem) * uitems.size(), hipMemcpyHostToDevice,
                           st));
  }
  const auto *gi = static_cast<const SvdRowItem *>(d_gitems.p);
  const auto *ui = static_cast<const SvdRowItem *>(d_uitems.p);
  auto *qflag = static_cast<int32_t *>(d_qflag.p);
  const auto *dm = static_cast<const SvdMat *>(d_mats.p);
  auto *perm = static_cast<int32_t *>(d_perm.p);
  auto *norms = static_cast<double *>(d_norms.p);
  int64_t max_n = 1;
  for (const SvdMat &mt : mats) max_n = std::max(max_n, mt.N);
  const dim3 norms_grid(static_cast<unsigned>((max_n + kWaves - 1) / kWaves), static_cast<unsigned>(n_mats));
  const unsigned gx = static_cast<unsigned>(std::min<int64_t>(1024, std::max<int64_t>(1, (max_elems + kThreads - 1) / kThreads)));
  hipLaunchKernelGGL(k_svd_load<C>, dim3(gx, static_cast<unsigned>(n_mats)), dim3(kThreads), 0, st, dm, a,
                     d_ws.p, perm);
  QLT_HIP(hipGetLastError());
  std::vector<int32_t> h_active(mats.size()), h_rot(mats.size(), 0);
  for (size_t i = 0; i < mats.size(); ++i) h_active[i] = mats[i].N >= 2;
  int32_t sweeps = 0;
  bool any = !items.empty();
  while (any && sweeps < kMaxSweeps) {
    QLT_HIP(hipMemcpyAsync(active, h_active.data(), sizeof(int32_t) * mats.size(), hipMemcpyHostToDevice, st));
    std::fill(h_rot.begin(), h_rot.end(), 0);
    QLT_HIP(hipMemcpyAsync(rotated, h_rot.data(), sizeof(int32_t) * mats.size(), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_svd_norms<C>, norms_grid, dim3(kThreads), 0, st, dm, d_ws.p, active, 0, norms);
    hipLaunchKernelGGL(k_svd_sort, dim3(static_cast<unsigned>(n_mats)), dim3(kThreads), 0, st, dm, norms, perm,
                       active, 0);
    for (int32_t r = 0; r < max_players - 1; ++r) {
      hipLaunchKernelGGL(k_svd_gram<C>, dim3(static_cast<unsigned>(gitems.size())), dim3(kThreads), 0, st, dm, gi,
                         r, d_ws.p, perm, active, d_part.p);
      hipLaunchKernelGGL(k_svd_solve<C>, dim3(static_cast<unsigned>(items.size())), dim3(kThreads), 0, st, dm,
                         static_cast<const SvdItem *>(d_items.p), r, d_part.p, perm, active, d_q.p, qflag, rotated);
      hipLaunchKernelGGL(k_svd_update<C>, dim3(static_cast<unsigned>(uitems.size())), dim3(kThreads), 0, st, dm, ui,
                         r, d_ws.p, perm, active, d_q.p, qflag);
    }
    QLT_HIP(hipGetLastError());
    QLT_HIP(hipMemcpyAsync(h_rot.data(), rotated, sizeof(int32_t) * mats.size(), hipMemcpyDeviceToHost, st));
    QLT_HIP(hipStreamSynchronize(st));
    ++sweeps;
    any = false;
    for (size_t i = 0; i < mats.size(); ++i) {
      h_active[i] = h_active[i] && h_rot[i];
      any = any || h_active[i];
    }
  }
  // final order of every matrix (into the column-order buffer), then U, S, V^H
  hipLaunchKernelGGL(k_svd_norms<C>, norms_grid, dim3(kThreads), 0, st, dm, d_ws.p, active, 1, norms);
  hipLaunchKernelGGL(k_svd_sort, dim3(static_cast<unsigned>(n_mats)), dim3(kThreads), 0, st, dm, norms, perm, active,
                     1);
  hipLaunchKernelGGL(k_svd_finish<C>, dim3(gx, static_cast<unsigned>(n_mats)), dim3(kThreads), 0, st, dm, d_ws.p,
                     norms, perm, u, s, vt);
  QLT_HIP(hipGetLastError());
  QLT_HIP(hipStreamSynchronize(st));
  if (sweeps_out != nullptr) *sweeps_out = sweeps;
  if (any) return qlt::fail(QLT_ERR_NOT_CONVERGED, "qlt_block_svd: Jacobi sweeps did not converge in " +
                                                        std::to_string(kMaxSweeps) + " sweeps");
  return QLT_OK;
}

template <bool C>
qlt_status qr_impl(int64_t n_mats, const int64_t *rows, const int64_t *cols, const void *a, const int64_t *a_offsets,
                   void *q, const int64_t *q_offsets, void *r, const int64_t *r_offsets, hipStream_t st) {
  const size_t esz = C ? 16 : 8;
  std::vector<QrMat> mats(static_cast<size_t>(n_mats));
  int64_t ws = 0;
  for (int64_t i = 0; i < n_mats; ++i) {
    QrMat &mt = mats[static_cast<size_t>(i)];
    mt.m = rows[i];
    mt.n = cols[i];
    mt.k = std::min(rows[i], cols[i]);
    mt.w_off = ws;
    ws += mt.m * mt.n;
    mt.q_off_ws = ws;
    ws += mt.m * mt.k;
    mt.t_off = ws;
    ws += mt.k;
    mt.a_off = a_offsets[i];
    mt.q_off = q_offsets[i];
    mt.r_off = r_offsets[i];
  }
  DevBuf d_ws, d_mats;
  QLT_HIP(hipMalloc(&d_ws.p, std::max<size_t>(1, static_cast<size_t>(ws) * esz)));
  QLT_HIP(hipMalloc(&d_mats.p, sizeof(QrMat) * mats.size()));
  QLT_HIP(hipMemcpyAsync(d_mats.p, mats.data(), sizeof(QrMat) * mats.size(), hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_qr<C>, dim3(static_cast<unsigned>(n_mats)), dim3(kThreads), 0, st,
                     static_cast<const QrMat *>(d_mats.p), a, d_ws.p, q, r);
  QLT_HIP(hipGetLastError());
  QLT_HIP(hipStreamSynchronize(st));
  return QLT_OK;
}

// Single-precision element types: the block matrices are widened to FP64
// (complex128) device buffers, decomposed there, and the factors rounded
// back; `ext` is the extent (elements) of a buffer addressed by offsets.
int64_t extent(int64_t n, const int64_t *offs, const int64_t *r, const int64_t *c) {
  int64_t e = 0;
  for (int64_t i = 0; i < n; ++i) e = std::max(e, offs[i] + r[i] * c[i]);
  return e;
}

struct DevTmp {
  void *p = nullptr;
  ~DevTmp() {
    if (p != nullptr) (void)hipFree(p);
  }
};

}  // namespace

extern "C" {

qlt_status qlt_block_svd(int32_t dtype, int64_t n_mats, const int64_t *rows, const int64_t *cols, const void *a,
                         const int64_t *a_offsets, void *u, const int64_t *u_offsets, double *s,
                         const int64_t *s_offsets, void *vt, const int64_t *vt_offsets, int32_t *sweeps,
                         void *stream) {
  QLT_REQUIRE_INIT();
  QLT_CHECK_ARG(qlt::valid_dtype(dtype), "qlt_block_svd: bad dtype");
  QLT_TRY(check_shapes("qlt_block_svd", n_mats, rows, cols));
  if (sweeps != nullptr) *sweeps = 0;
  if (n_mats == 0) return QLT_OK;
  QLT_CHECK_ARG(a != nullptr && u != nullptr && s != nullptr && vt != nullptr, "qlt_block_svd: null data pointer");
  QLT_CHECK_ARG(a_offsets != nullptr && u_offsets != nullptr && s_offsets != nullptr && vt_offsets != nullptr,
                "qlt_block_svd: null offset array");
  QLT_CHECK_ARG(!qlt::stream_capturing(qlt::as_stream(stream)), "qlt_block_svd: not capturable (host-driven sweeps)");
  if (qlt::dtype_is_single(dtype)) {
    const int32_t wide = dtype == QLT_F32 ? QLT_F64 : QLT_C128;
    const int64_t esz = qlt::dtype_bytes(wide);
    std::vector<int64_t> kk(static_cast<size_t>(n_mats));
    for (int64_t i = 0; i < n_mats; ++i) kk[i] = std::min(rows[i], cols[i]);
    const int64_t na = extent(n_mats, a_offsets, rows, cols);
    const int64_t nu = extent(n_mats, u_offsets, rows, kk.data());
    const int64_t nv = extent(n_mats, vt_offsets, kk.data(), cols);
    DevTmp wa, wu, wv;
    QLT_HIP(hipMalloc(&wa.p, std::max<int64_t>(1, na) * esz));
    QLT_HIP(hipMalloc(&wu.p, std::max<int64_t>(1, nu) * esz));
    QLT_HIP(hipMalloc(&wv.p, std::max<int64_t>(1, nv) * esz));
    QLT_TRY(qlt_convert(dtype, wide, na, a, wa.p, stream));
    QLT_TRY(qlt_block_svd(wide, n_mats, rows, cols, wa.p, a_offsets, wu.p, u_offsets, s, s_offsets, wv.p,
                          vt_offsets, sweeps, stream));
    QLT_TRY(qlt_convert(wide, dtype, nu, wu.p, u, stream));
    QLT_TRY(qlt_convert(wide, dtype, nv, wv.p, vt, stream));
    QLT_HIP(hipStreamSynchronize(qlt::as_stream(stream)));
    return QLT_OK;
  }
  if (dtype == QLT_F64)
    return svd_impl<false>(n_mats, rows, cols, a, a_offsets, u, u_offsets, s, s_offsets, vt, vt_offsets, sweeps,
                           qlt::as_stream(stream));
  return svd_impl<true>(n_mats, rows, cols, a, a_offsets, u, u_offsets, s, s_offsets, vt, vt_offsets, sweeps,
                        qlt::as_stream(stream));
}

qlt_status qlt_block_qr(int32_t dtype, int64_t n_mats, const int64_t *rows, const int64_t *cols, const void *a,
                        const int64_t *a_offsets, void *q, const int64_t *q_offsets, void *r,
                        const int64_t *r_offsets, void *stream) {
  QLT_REQUIRE_INIT();
  QLT_CHECK_ARG(qlt::valid_dtype(dtype), "qlt_block_qr: bad dtype");
  QLT_TRY(check_shapes("qlt_block_qr", n_mats, rows, cols));
  if (n_mats == 0) return QLT_OK;
  QLT_CHECK_ARG(a != nullptr && q != nullptr && r != nullptr, "qlt_block_qr: null data pointer");
  QLT_CHECK_ARG(a_offsets != nullptr && q_offsets != nullptr && r_offsets != nullptr,
                "qlt_block_qr: null offset array");
  QLT_CHECK_ARG(!qlt::stream_capturing(qlt::as_stream(stream)), "qlt_block_qr: not capturable (allocates)");
  if (qlt::dtype_is_single(dtype)) {
    const int32_t wide = dtype == QLT_F32 ? QLT_F64 : QLT_C128;
    const int64_t esz = qlt::dtype_bytes(wide);
    std::vector<int64_t> kk(static_cast<size_t>(n_mats));
    for (int64_t i = 0; i < n_mats; ++i) kk[i] = std::min(rows[i], cols[i]);
    const int64_t na = extent(n_mats, a_offsets, rows, cols);
    const int64_t nq = extent(n_mats, q_offsets, rows, kk.data());
    const int64_t nr = extent(n_mats, r_offsets, kk.data(), cols);
    DevTmp wa, wq, wr;
    QLT_HIP(hipMalloc(&wa.p, std::max<int64_t>(1, na) * esz));
    QLT_HIP(hipMalloc(&wq.p, std::max<int64_t>(1, nq) * esz));
    QLT_HIP(hipMalloc(&wr.p, std::max<int64_t>(1, nr) * esz));
    QLT_TRY(qlt_convert(dtype, wide, na, a, wa.p, stream));
    QLT_TRY(qlt_block_qr(wide, n_mats, rows, cols, wa.p, a_offsets, wq.p, q_offsets, wr.p, r_offsets, stream));
    QLT_TRY(qlt_convert(wide, dtype, nq, wq.p, q, stream));
    QLT_TRY(qlt_convert(wide, dtype, nr, wr.p, r, stream));
    QLT_HIP(hipStreamSynchronize(qlt::as_stream(stream)));
    return QLT_OK;
  }
  if (dtype == QLT_F64)
    return qr_impl<false>(n_mats, rows, cols, a, a_offsets, q, q_offsets, r, r_offsets, qlt::as_stream(stream));
  return qr_impl<true>(n_mats, rows, cols, a, a_offsets, q, q_offsets, r, r_offsets, qlt::as_stream(stream));
}

}  // extern "C"
