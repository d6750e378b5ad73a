// Device-resident Krylov primitives over a basis arena: block inner products,
// block updates and the Lanczos extension step, with the number of active
// basis vectors kept in device memory.  They let one Lanczos iteration --
// the matvec chain, two passes of classical Gram-Schmidt against the whole
// basis, the normalisation of the new vector -- be launched with no host
// round trip, so it can be captured once as a graph and replayed.  They
// restate the vector operations the reference's Lanczos solver uses on
// tensors: InnerProduct (tensor_manipulation/basic_operations.h:146-203),
// LinearCombine (ten_linear_combine.h:186) and Normalize
// (qltensor/qltensor_impl.h:568-573), on raw data.
//
// Every reduction is two-pass with a fixed grid and a fixed summation order,
// so results are bitwise reproducible (a replayed graph equals the eager
// launches bit for bit).

#include <algorithm>
#include <cmath>
#include <string>

#include "qlt_internal.h"

namespace {

constexpr int kParts = 512;          // partial sums per reduction (fixed grid)
constexpr int kGroup = 8;            // basis vectors per pass over w
constexpr int kFuseMax = 16;         // basis vectors whose values k_basis_update_dot keeps per element
constexpr int32_t kMaxBasis = 1024;  // arena slots a call may address

__device__ __forceinline__ int active_count(const int32_t *count, int32_t k_max) {
  const int m = *count;
  return m < 0 ? 0 : (m > k_max ? k_max : m);
}

// Sum over the 256 threads of a workgroup, in a fixed order: wave shuffles,
// then the four wave sums in order.  Every thread gets the result.
__device__ __forceinline__ double block_sum(double v, double *s_red) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int tid = threadIdx.x;
  __syncthreads();
  if ((tid & 63) == 0) s_red[tid >> 6] = v;
  __syncthreads();
  return ((s_red[0] + s_red[1]) + s_red[2]) + s_red[3];
}

// work[(i * kParts + b) * 2 + {0, 1}] = partial <V_i, w> of workgroup b
// (conj(V_i) * w), for i < m.
// With m <= skip_upto the kernel does nothing (k_basis_update_dot has
// written the partials already).
template <bool CPLX>
__global__ __launch_bounds__(256) void k_basis_dot_partial(int64_t n, int32_t k_max, const int32_t *count,
                                                           const double *__restrict__ V, int64_t ldv,
                                                           const double *__restrict__ w, double *work,
                                                           int32_t skip_upto) {
  __shared__ double s_red[4];
  const int m = active_count(count, k_max);
  if (m <= skip_upto) return;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  const int64_t first = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  for (int i0 = 0; i0 < m; i0 += kGroup) {
    const int nq = min(kGroup, m - i0);
    double ar[kGroup], ai[kGroup];
#pragma unroll
    for (int q = 0; q < kGroup; ++q) ar[q] = ai[q] = 0.0;
    for (int64_t e = first; e < n; e += stride) {
      // the group's loads go out together (a vector past the count reads
      // the group's first one and is not used)
      if (CPLX) {
        const double2 y = reinterpret_cast<const double2 *>(w)[e];
        double2 x[kGroup];
#pragma unroll
        for (int q = 0; q < kGroup; ++q) x[q] = reinterpret_cast<const double2 *>(V)[(i0 + (q < nq ? q : 0)) * ldv + e];
#pragma unroll
        for (int q = 0; q < kGroup; ++q) {
          if (q < nq) {
            ar[q] += x[q].x * y.x + x[q].y * y.y;
            ai[q] += x[q].x * y.y - x[q].y * y.x;
          }
        }
      } else {
        const double y = w[e];
        double x[kGroup];
#pragma unroll
        for (int q = 0; q < kGroup; ++q) x[q] = V[(i0 + (q < nq ? q : 0)) * ldv + e];
#pragma unroll
        for (int q = 0; q < kGroup; ++q)
          if (q < nq) ar[q] += x[q] * y;
      }
    }
#pragma unroll
    for (int q = 0; q < kGroup; ++q) {
      if (i0 + q < m) {
        const double sr = block_sum(ar[q], s_red);
        const double si = CPLX ? block_sum(ai[q], s_red) : 0.0;
        if (threadIdx.x == 0) {
          work[(static_cast<int64_t>(i0 + q) * kParts + blockIdx.x) * 2] = sr;
          work[(static_cast<int64_t>(i0 + q) * kParts + blockIdx.x) * 2 + 1] = si;
        }
      }
    }
  }
}

// c[2i], c[2i+1] = sum over the partials of basis vector i (i < m), else 0.
// One workgroup per basis vector.
__global__ __launch_bounds__(256) void k_basis_dot_final(int32_t k_max, const int32_t *count,
                                                         const double *work, double *c) {
  __shared__ double s_red[4];
  const int i = blockIdx.x;
  const int m = active_count(count, k_max);
  double sr = 0.0, si = 0.0;
  if (i < m) {
    for (int b = threadIdx.x; b < kParts; b += blockDim.x) {
      sr += work[(static_cast<int64_t>(i) * kParts + b) * 2];
      si += work[(static_cast<int64_t>(i) * kParts + b) * 2 + 1];
    }
  }
  sr = block_sum(sr, s_red);
  si = block_sum(si, s_red);
  if (threadIdx.x == 0) {
    c[2 * i] = sr;
    c[2 * i + 1] = si;
  }
}

// w[e] += scale * sum_{i < m} c_i V_i[e], the terms added in order of i.
template <bool CPLX>
__global__ __launch_bounds__(256) void k_basis_update(int64_t n, int32_t k_max, const int32_t *count, double scale,
                                                      const double *__restrict__ c, const double *__restrict__ V,
                                                      int64_t ldv, double *__restrict__ w) {
  __shared__ double s_c[2 * kMaxBasis];
  const int m = active_count(count, k_max);
  for (int i = threadIdx.x; i < m; i += blockDim.x) {
    s_c[2 * i] = scale * c[2 * i];
    s_c[2 * i + 1] = scale * c[2 * i + 1];
  }
  __syncthreads();
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t e = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; e < n; e += stride) {
    // kGroup basis loads in flight at a time, the terms still added in order of i
    if (CPLX) {
      double2 y = reinterpret_cast<const double2 *>(w)[e];
      for (int i0 = 0; i0 < m; i0 += kGroup) {
        const int nq = min(kGroup, m - i0);
        double2 x[kGroup];
#pragma unroll
        for (int q = 0; q < kGroup; ++q) x[q] = reinterpret_cast<const double2 *>(V)[(i0 + (q < nq ? q : 0)) * ldv + e];
#pragma unroll
        for (int q = 0; q < kGroup; ++q) {
          if (q < nq) {
            const double cr = s_c[2 * (i0 + q)], ci = s_c[2 * (i0 + q) + 1];
            y.x += cr * x[q].x - ci * x[q].y;
            y.y += cr * x[q].y + ci * x[q].x;
          }
        }
      }
      reinterpret_cast<double2 *>(w)[e] = y;
    } else {
      double y = w[e];
      for (int i0 = 0; i0 < m; i0 += kGroup) {
        const int nq = min(kGroup, m - i0);
        double x[kGroup];
#pragma unroll
        for (int q = 0; q < kGroup; ++q) x[q] = V[(i0 + (q < nq ? q : 0)) * ldv + e];
#pragma unroll
        for (int q = 0; q < kGroup; ++q)
          if (q < nq) y += s_c[2 * (i0 + q)] * x[q];
      }
      w[e] = y;
    }
  }
}

// The update of k_basis_update followed by the partial products of
// k_basis_dot_partial on the updated w, in one pass over the basis when
// m <= kFuseMax: each thread keeps its elements' basis values between the
// two.  Same workgroup partition as k_basis_dot_partial (kParts workgroups),
// same terms in the same order per element and per partial sum, so w and
// the partials are bitwise those of the two kernels run one after the other.
// With m > kFuseMax only the update is done here (k_basis_dot_partial then
// runs on the updated w).
template <bool CPLX>
__global__ __launch_bounds__(256) void k_basis_update_dot(int64_t n, int32_t k_max, const int32_t *count,
                                                          double scale, const double *__restrict__ c,
                                                          const double *__restrict__ V, int64_t ldv,
                                                          double *__restrict__ w, double *work) {
  __shared__ double s_c[2 * kMaxBasis];
  __shared__ double s_red[4];
  const int m = active_count(count, k_max);
  for (int i = threadIdx.x; i < m; i += blockDim.x) {
    s_c[2 * i] = scale * c[2 * i];
    s_c[2 * i + 1] = scale * c[2 * i + 1];
  }
  __syncthreads();
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  const int64_t first = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (m > kFuseMax) {
    for (int64_t e = first; e < n; e += stride) {
      if (CPLX) {
        double2 y = reinterpret_cast<const double2 *>(w)[e];
        for (int i0 = 0; i0 < m; i0 += kGroup) {
          const int nq = min(kGroup, m - i0);
          double2 x[kGroup];
#pragma unroll
          for (int q = 0; q < kGroup; ++q) x[q] = reinterpret_cast<const double2 *>(V)[(i0 + (q < nq ? q : 0)) * ldv + e];
#pragma unroll
          for (int q = 0; q < kGroup; ++q) {
            if (q < nq) {
              const double cr = s_c[2 * (i0 + q)], ci = s_c[2 * (i0 + q) + 1];
              y.x += cr * x[q].x - ci * x[q].y;
              y.y += cr * x[q].y + ci * x[q].x;
            }
          }
        }
        reinterpret_cast<double2 *>(w)[e] = y;
      } else {
        double y = w[e];
        for (int i0 = 0; i0 < m; i0 += kGroup) {
          const int nq = min(kGroup, m - i0);
          double x[kGroup];
#pragma unroll
          for (int q = 0; q < kGroup; ++q) x[q] = V[(i0 + (q < nq ? q : 0)) * ldv + e];
#pragma unroll
          for (int q = 0; q < kGroup; ++q)
            if (q < nq) y += s_c[2 * (i0 + q)] * x[q];
        }
        w[e] = y;
      }
    }
    return;
  }
  double ar[kFuseMax], ai[kFuseMax];
#pragma unroll
  for (int q = 0; q < kFuseMax; ++q) ar[q] = ai[q] = 0.0;
  for (int64_t e = first; e < n; e += stride) {
    // all m basis values of the element go out together (a vector past the
    // count reads vector 0 and is not used)
    if (CPLX) {
      double2 y = reinterpret_cast<const double2 *>(w)[e];
      double2 x[kFuseMax];
#pragma unroll
      for (int q = 0; q < kFuseMax; ++q) x[q] = reinterpret_cast<const double2 *>(V)[(q < m ? q : 0) * ldv + e];
#pragma unroll
      for (int q = 0; q < kFuseMax; ++q) {
        if (q < m) {
          const double cr = s_c[2 * q], ci = s_c[2 * q + 1];
          y.x += cr * x[q].x - ci * x[q].y;
          y.y += cr * x[q].y + ci * x[q].x;
        }
      }
      reinterpret_cast<double2 *>(w)[e] = y;
#pragma unroll
      for (int q = 0; q < kFuseMax; ++q) {
        if (q < m) {
          ar[q] += x[q].x * y.x + x[q].y * y.y;
          ai[q] += x[q].x * y.y - x[q].y * y.x;
        }
      }
    } else {
      double y = w[e];
      double x[kFuseMax];
#pragma unroll
      for (int q = 0; q < kFuseMax; ++q) x[q] = V[(q < m ? q : 0) * ldv + e];
#pragma unroll
      for (int q = 0; q < kFuseMax; ++q)
        if (q < m) y += s_c[2 * q] * x[q];
      w[e] = y;
#pragma unroll
      for (int q = 0; q < kFuseMax; ++q)
        if (q < m) ar[q] += x[q] * y;
    }
  }
#pragma unroll
  for (int q = 0; q < kFuseMax; ++q) {
    if (q < m) {
      const double sr = block_sum(ar[q], s_red);
      const double si = CPLX ? block_sum(ai[q], s_red) : 0.0;
      if (threadIdx.x == 0) {
        work[(static_cast<int64_t>(q) * kParts + blockIdx.x) * 2] = sr;
        work[(static_cast<int64_t>(q) * kParts + blockIdx.x) * 2 + 1] = si;
      }
    }
  }
}

// work[b] = partial sum of |w_e|^2 of workgroup b.
template <bool CPLX>
__global__ __launch_bounds__(256) void k_norm2_partial(int64_t n, const double *__restrict__ w, double *work) {
  __shared__ double s_red[4];
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  double acc = 0.0;
  for (int64_t e = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; e < n; e += stride) {
    if (CPLX) {
      const double2 y = reinterpret_cast<const double2 *>(w)[e];
      acc += y.x * y.x + y.y * y.y;
    } else {
      acc += w[e] * w[e];
    }
  }
  acc = block_sum(acc, s_red);
  if (threadIdx.x == 0) work[blockIdx.x] = acc;
}

// The Lanczos coefficients of vector j = m - 1: alpha_j = Re c_j, beta_j =
// |w|; work[kParts] = 1 / beta_j (0 when beta_j = 0).
__global__ __launch_bounds__(256) void k_lanczos_coeffs(int32_t k_max, const int32_t *count, const double *c,
                                                        double *work, double *alphas, double *betas) {
  __shared__ double s_red[4];
  double acc = 0.0;
  for (int b = threadIdx.x; b < kParts; b += blockDim.x) acc += work[b];
  acc = block_sum(acc, s_red);
  if (threadIdx.x == 0) {
    const int m = active_count(count, k_max);
    const double beta = sqrt(acc);
    if (m > 0) {
      alphas[m - 1] = c[2 * (m - 1)];
      betas[m - 1] = beta;
    }
    work[kParts] = beta > 0.0 ? 1.0 / beta : 0.0;
  }
}

// V_m = w / beta when slot m exists (m < k_max).
template <bool CPLX>
__global__ __launch_bounds__(256) void k_lanczos_store(int64_t n, int32_t k_max, const int32_t *count,
                                                       const double *work, const double *__restrict__ w,
                                                       double *__restrict__ V, int64_t ldv) {
  const int m = active_count(count, k_max);
  if (m >= k_max) return;
  const double inv = work[kParts];
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t e = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; e < n; e += stride) {
    if (CPLX) {
      const double2 y = reinterpret_cast<const double2 *>(w)[e];
      reinterpret_cast<double2 *>(V)[m * ldv + e] = double2{y.x * inv, y.y * inv};
    } else {
      V[m * ldv + e] = w[e] * inv;
    }
  }
}

__global__ void k_count_advance(int32_t k_max, int32_t *count) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    const int m = active_count(count, k_max);
    if (m < k_max) *count = m + 1;
  }
}

// x = V_{*index + offset} (nothing when that slot is outside 0 .. k_max - 1).
template <bool CPLX>
__global__ __launch_bounds__(256) void k_basis_select(int64_t n, int32_t k_max, const int32_t *index, int32_t offset,
                                                      const double *__restrict__ V, int64_t ldv,
                                                      double *__restrict__ x) {
  const int s = *index + offset;
  if (s < 0 || s >= k_max) return;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t e = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; e < n; e += stride) {
    if (CPLX) reinterpret_cast<double2 *>(x)[e] = reinterpret_cast<const double2 *>(V)[s * ldv + e];
    else x[e] = V[s * ldv + e];
  }
}

inline unsigned stream_grid(int64_t n) {
  return static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 8192)));
}

qlt_status check_basis_args(const char *fn, int32_t dtype, int64_t n, int32_t k_max, int64_t ldv) {
  using qlt::fail;
  QLT_CHECK_ARG(dtype == QLT_F64 || dtype == QLT_C128, std::string(fn) + ": bad dtype");
  QLT_CHECK_ARG(n >= 0, std::string(fn) + ": negative length");
  QLT_CHECK_ARG(k_max >= 1 && k_max <= kMaxBasis, std::string(fn) + ": k_max must be in 1 .. 1024");
  QLT_CHECK_ARG(ldv >= n, std::string(fn) + ": slot stride below the vector length");
  return QLT_OK;
}

}  // namespace

using namespace qlt;

extern "C" {

qlt_status qlt_basis_dot(int32_t dtype, int64_t n, int32_t k_max, const int32_t *count, const void *V, int64_t ldv,
                         const void *w, double *c, double *work, void *stream) {
  QLT_REQUIRE_INIT();
  QLT_TRY(check_basis_args("qlt_basis_dot", dtype, n, k_max, ldv));
  QLT_CHECK_ARG(count && V && w && c && work, "qlt_basis_dot: null pointer");
  hipStream_t st = as_stream(stream);
  const auto *vd = static_cast<const double *>(V);
  const auto *wd = static_cast<const double *>(w);
  if (dtype == QLT_C128)
    hipLaunchKernelGGL(k_basis_dot_partial<true>, dim3(kParts), dim3(256), 0, st, n, k_max, count, vd, ldv, wd, work,
                       -1);
  else
    hipLaunchKernelGGL(k_basis_dot_partial<false>, dim3(kParts), dim3(256), 0, st, n, k_max, count, vd, ldv, wd, work,
                       -1);
  hipLaunchKernelGGL(k_basis_dot_final, dim3(k_max), dim3(256), 0, st, k_max, count, work, c);
  QLT_HIP(hipGetLastError());
  return QLT_OK;
}

qlt_status qlt_basis_update_dot(int32_t dtype, int64_t n, int32_t k_max, const int32_t *count, double scale,
                                const double *c, const void *V, int64_t ldv, void *w, double *c_out, double *work,
                                void *stream) {
  QLT_REQUIRE_INIT();
  QLT_TRY(check_basis_args("qlt_basis_update_dot", dtype, n, k_max, ldv));
  QLT_CHECK_ARG(count && V && w && c && c_out && work, "qlt_basis_update_dot: null pointer");
  QLT_CHECK_ARG(c != c_out, "qlt_basis_update_dot: c and c_out must not alias");
  hipStream_t st = as_stream(stream);
  const auto *vd = static_cast<const double *>(V);
  auto *wd = static_cast<double *>(w);
  if (dtype == QLT_C128) {
    hipLaunchKernelGGL(k_basis_update_dot<true>, dim3(kParts), dim3(256), 0, st, n, k_max, count, scale, c, vd, ldv,
                       wd, work);
    hipLaunchKernelGGL(k_basis_dot_partial<true>, dim3(kParts), dim3(256), 0, st, n, k_max, count, vd, ldv, wd, work,
                       kFuseMax);
  } else {
    hipLaunchKernelGGL(k_basis_update_dot<false>, dim3(kParts), dim3(256), 0, st, n, k_max, count, scale, c, vd, ldv,
                       wd, work);
    hipLaunchKernelGGL(k_basis_dot_partial<false>, dim3(kParts), dim3(256), 0, st, n, k_max, count, vd, ldv, wd, work,
                       kFuseMax);
  }
  hipLaunchKernelGGL(k_basis_dot_final, dim3(k_max), dim3(256), 0, st, k_max, count, work, c_out);
  QLT_HIP(hipGetLastError());
  return QLT_OK;
}

qlt_status qlt_basis_update(int32_t dtype, int64_t n, int32_t k_max, const int32_t *count, double scale,
                            const double *c, const void *V, int64_t ldv, void *w, void *stream) {
  QLT_REQUIRE_INIT();
  QLT_TRY(check_basis_args("qlt_basis_update", dtype, n, k_max, ldv));
  QLT_CHECK_ARG(count && V && w && c, "qlt_basis_update: null pointer");
  if (n == 0) return QLT_OK;
  hipStream_t st = as_stream(stream);
  const auto *vd = static_cast<const double *>(V);
  auto *wd = static_cast<double *>(w);
  if (dtype == QLT_C128)
    hipLaunchKernelGGL(k_basis_update<true>, dim3(stream_grid(n)), dim3(256), 0, st, n, k_max, count, scale, c, vd,
                       ldv, wd);
  else
    hipLaunchKernelGGL(k_basis_update<false>, dim3(stream_grid(n)), dim3(256), 0, st, n, k_max, count, scale, c, vd,
                       ldv, wd);
  QLT_HIP(hipGetLastError());
  return QLT_OK;
}

qlt_status qlt_lanczos_extend(int32_t dtype, int64_t n, int32_t k_max, int32_t *count, const double *c,
                              const void *w, void *V, int64_t ldv, double *alphas, double *betas, double *work,
                              void *stream) {
  QLT_REQUIRE_INIT();
  QLT_TRY(check_basis_args("qlt_lanczos_extend", dtype, n, k_max, ldv));
  QLT_CHECK_ARG(count && c && w && V && alphas && betas && work, "qlt_lanczos_extend: null pointer");
  hipStream_t st = as_stream(stream);
  const auto *wd = static_cast<const double *>(w);
  auto *vd = static_cast<double *>(V);
  if (dtype == QLT_C128) {
    hipLaunchKernelGGL(k_norm2_partial<true>, dim3(kParts), dim3(256), 0, st, n, wd, work);
  } else {
    hipLaunchKernelGGL(k_norm2_partial<false>, dim3(kParts), dim3(256), 0, st, n, wd, work);
  }
  hipLaunchKernelGGL(k_lanczos_coeffs, dim3(1), dim3(256), 0, st, k_max, count, c, work, alphas, betas);
  if (n > 0) {
    if (dtype == QLT_C128)
      hipLaunchKernelGGL(k_lanczos_store<true>, dim3(stream_grid(n)), dim3(256), 0, st, n, k_max, count, work, wd, vd,
                         ldv);
    else
      hipLaunchKernelGGL(k_lanczos_store<false>, dim3(stream_grid(n)), dim3(256), 0, st, n, k_max, count, work, wd,
                         vd, ldv);
  }
  hipLaunchKernelGGL(k_count_advance, dim3(1), dim3(64), 0, st, k_max, count);
  QLT_HIP(hipGetLastError());
  return QLT_OK;
}

qlt_status qlt_basis_select(int32_t dtype, int64_t n, int32_t k_max, const int32_t *index, int32_t offset,
                            const void *V, int64_t ldv, void *x, void *stream) {
  QLT_REQUIRE_INIT();
  QLT_TRY(check_basis_args("qlt_basis_select", dtype, n, k_max, ldv));
  QLT_CHECK_ARG(index && V && x, "qlt_basis_select: null pointer");
  if (n == 0) return QLT_OK;
  hipStream_t st = as_stream(stream);
  const auto *vd = static_cast<const double *>(V);
  auto *xd = static_cast<double *>(x);
  if (dtype == QLT_C128)
    hipLaunchKernelGGL(k_basis_select<true>, dim3(stream_grid(n)), dim3(256), 0, st, n, k_max, index, offset, vd, ldv,
                       xd);
  else
    hipLaunchKernelGGL(k_basis_select<false>, dim3(stream_grid(n)), dim3(256), 0, st, n, k_max, index, offset, vd,
                       ldv, xd);
  QLT_HIP(hipGetLastError());
  return QLT_OK;
}

}  // extern "C"
