// Contraction plan built on the device.
//
// Reproduces BlockSparseDataTensor::DataBlkGenForTenCtrct
// (include/qlten/qltensor/blk_spar_data_ten/data_blk_operations.h:411-578)
// bit for bit, as integer kernels:
//   1. part hash of each block's contracted coordinates, taken in the given
//      axis order (VecPartHasher, include/qlten/framework/vec_hash.h:39-47,
//      via GenBlkIdxQNBlkCoorPartHashMap, blk_spar_data_ten.h:959-973);
//   2. A-major / B-minor matching on equal hashes, both in ascending block
//      index (data_blk_operations.h:494-567); for a scalar output the B loop
//      stops at the first match (data_blk_operations.h:451-491);
//   3. output block creation on first touch -> beta 0, later touches beta 1;
//   4. output offsets = exclusive prefix sum of block sizes in ascending
//      output block index (DataBlksOffsetRefresh, data_blk_operations.h:137-145,568);
//   5. c_data_offset fill (data_blk_operations.h:571-574).
// The same kernels also emit the grouped-GEMM work list: tasks sorted by
// output block (an order SortTasksByCBlkIdx, raw_data_operation_tasks.h:250-269,
// can produce, with generation order kept inside a block), one segment per
// output block, and the output tiles ordered longest first.

#include <algorithm>
#include <atomic>
#include <memory>
#include <cstring>
#include <numeric>
#include <string>

#include "qlt_internal.h"

using namespace qlt;

namespace {

constexpr uint64_t kP1 = 11400714785074694791ull;
constexpr uint64_t kP2 = 14029467366897019727ull;
constexpr uint64_t kP5 = 2870177450012600261ull;

struct PlanArgs {
  int32_t a_rank, b_rank, n_ctrct, n_a_saved, n_b_saved, c_rank;
  int32_t a_ctrct[QLT_MAX_RANK], b_ctrct[QLT_MAX_RANK];
  int32_t a_saved[QLT_MAX_RANK], b_saved[QLT_MAX_RANK];
  int64_t c_stride[QLT_MAX_RANK];
  int32_t c_is_scalar;
  int32_t tile_m, tile_n;
  int8_t cls_table[QLT_TUNE_BUCKETS][QLT_TUNE_BUCKETS];   // tile class per size bucket (qlt_tuning)
};

struct DevTask {
  int32_t a_pos;
  int32_t b_pos;
};

__device__ __forceinline__ uint64_t rotl31(uint64_t x) { return (x << 31) | (x >> 33); }

__global__ void k_part_hash(const int32_t *coords, int64_t n, int32_t rank, const PlanArgs args,
                            int32_t is_a, uint64_t *h) {
  const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  const int32_t *c = coords + i * rank;
  const int32_t *axes = is_a ? args.a_ctrct : args.b_ctrct;
  uint64_t hv = kP5;
  for (int j = 0; j < args.n_ctrct; ++j) {
    hv += static_cast<uint64_t>(c[axes[j]]) * kP2;
    hv = rotl31(hv);
    hv *= kP1;
  }
  hv += static_cast<uint64_t>(args.n_ctrct) ^ kP5;
  h[i] = hv;
}

// counts[i] = number of B blocks whose hash equals A block i's (at most 1 for
// a scalar output).  B hashes are staged through LDS.
__global__ __launch_bounds__(256) void k_count(const uint64_t *ha, int64_t na, const uint64_t *hb,
                                               int64_t nb, int32_t scalar, int64_t *counts) {
  __shared__ uint64_t sh[1024];
  const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  const uint64_t mine = (i < na) ? ha[i] : 0;
  int64_t cnt = 0;
  for (int64_t base = 0; base < nb; base += 1024) {
    const int64_t len = (nb - base < 1024) ? (nb - base) : 1024;
    __syncthreads();
    for (int64_t j = threadIdx.x; j < len; j += blockDim.x) sh[j] = hb[base + j];
    __syncthreads();
    if (i < na)
      for (int64_t j = 0; j < len; ++j) cnt += (sh[j] == mine) ? 1 : 0;
  }
  if (i < na) counts[i] = (scalar && cnt > 1) ? 1 : cnt;
}

// Single-workgroup exclusive scan of n int64 values; total written to *total.
__global__ __launch_bounds__(1024) void k_scan_excl(const int64_t *in, int64_t *out, int64_t n,
                                                    int64_t *total) {
  __shared__ int64_t wsum[16];
  __shared__ int64_t carry;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  if (tid == 0) carry = 0;
  __syncthreads();
  for (int64_t base = 0; base < n; base += 1024) {
    const int64_t i = base + tid;
    const int64_t v = (i < n) ? in[i] : 0;
    int64_t x = v;
    for (int off = 1; off < 64; off <<= 1) {
      const int64_t y = __shfl_up(x, off, 64);
      if (lane >= off) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    int64_t wprefix = 0;
    for (int j = 0; j < w; ++j) wprefix += wsum[j];
    const int64_t c = carry;
    if (i < n) out[i] = c + wprefix + x - v;
    __syncthreads();
    if (tid == 1023) carry = c + wprefix + x;
    __syncthreads();
  }
  if (tid == 0) *total = carry;
}

// One wavefront per A block: walk B in ascending order, ballot the matches
// and write them at consecutive task slots, so the task list comes out in
// the reference's A-major / B-minor order.
__global__ __launch_bounds__(256) void k_emit(const uint64_t *ha, int64_t na, const uint64_t *hb,
                                              int64_t nb, const int64_t *starts, int32_t scalar,
                                              DevTask *tasks) {
  const int lane = threadIdx.x & 63;
  const int64_t i = (blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x) >> 6;
  if (i >= na) return;
  const uint64_t mine = ha[i];
  int64_t pos = starts[i];
  for (int64_t base = 0; base < nb; base += 64) {
    const int64_t j = base + lane;
    const bool match = (j < nb) && (hb[j] == mine);
    const uint64_t bal = __ballot(match);
    if (bal == 0) continue;
    if (scalar) {
      const int first = __ffsll(static_cast<unsigned long long>(bal)) - 1;
      if (lane == first) tasks[pos] = DevTask{static_cast<int32_t>(i), static_cast<int32_t>(j)};
      return;
    }
    const int before = __popcll(bal & ((1ull << lane) - 1ull));
    if (match) tasks[pos + before] = DevTask{static_cast<int32_t>(i), static_cast<int32_t>(j)};
    pos += __popcll(bal);
  }
}

// Output block index (sort key) and generation position (payload) of every task.
__global__ void k_task_keys(const DevTask *tasks, int64_t nt, const int32_t *a_coords,
                            const int32_t *b_coords, const PlanArgs args, uint64_t *key,
                            int32_t *val) {
  const int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (t >= nt) return;
  uint64_t c = 0;
  if (!args.c_is_scalar) {
    const int32_t *ac = a_coords + static_cast<int64_t>(tasks[t].a_pos) * args.a_rank;
    const int32_t *bc = b_coords + static_cast<int64_t>(tasks[t].b_pos) * args.b_rank;
    int ci = 0;
    for (int j = 0; j < args.n_a_saved; ++j, ++ci)
      c += static_cast<uint64_t>(ac[args.a_saved[j]]) * static_cast<uint64_t>(args.c_stride[ci]);
    for (int j = 0; j < args.n_b_saved; ++j, ++ci)
      c += static_cast<uint64_t>(bc[args.b_saved[j]]) * static_cast<uint64_t>(args.c_stride[ci]);
  }
  key[t] = c;
  val[t] = static_cast<int32_t>(t);
}

// Single-workgroup stable LSD radix sort (8-bit digits) of (key, val) pairs.
// Each 1024-element chunk is ranked with a wavefront multisplit (eight
// ballots per element) and chunks are placed in input order, so equal keys
// keep their input order.  The result is in (k0, v0) after an even number of
// passes and in (k1, v1) after an odd one.
__global__ __launch_bounds__(1024) void k_radix_sort(uint64_t *k0, int32_t *v0, uint64_t *k1,
                                                     int32_t *v1, int64_t n, int32_t passes) {
  __shared__ int64_t digit_base[256];
  __shared__ int32_t chunk_cnt[256];
  __shared__ int32_t wave_cnt[16][256];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  for (int pass = 0; pass < passes; ++pass) {
    const uint64_t *sk = (pass & 1) ? k1 : k0;
    const int32_t *sv = (pass & 1) ? v1 : v0;
    uint64_t *dk = (pass & 1) ? k0 : k1;
    int32_t *dv = (pass & 1) ? v0 : v1;
    const int shift = 8 * pass;
    if (tid < 256) chunk_cnt[tid] = 0;
    __syncthreads();
    for (int64_t i = tid; i < n; i += 1024) atomicAdd(&chunk_cnt[(sk[i] >> shift) & 255u], 1);
    __syncthreads();
    if (tid == 0) {
      int64_t run = 0;
      for (int d = 0; d < 256; ++d) {
        digit_base[d] = run;
        run += chunk_cnt[d];
      }
    }
    __syncthreads();
    for (int64_t chunk = 0; chunk < n; chunk += 1024) {
      const int64_t i = chunk + tid;
      const bool valid = i < n;
      const uint64_t key = valid ? sk[i] : 0;
      const int32_t val = valid ? sv[i] : 0;
      const int d = static_cast<int>((key >> shift) & 255u);
      for (int j = tid; j < 16 * 256; j += 1024) wave_cnt[j >> 8][j & 255] = 0;
      __syncthreads();
      uint64_t mask = __ballot(valid);
      for (int b = 0; b < 8; ++b) {
        const bool bit = (d >> b) & 1;
        const uint64_t bal = __ballot(bit);
        mask &= bit ? bal : ~bal;
      }
      const int rank = __popcll(mask & ((1ull << lane) - 1ull));
      if (valid && rank == 0) wave_cnt[w][d] = __popcll(mask);
      __syncthreads();
      if (tid < 256) {
        int32_t run = 0;
        for (int ww = 0; ww < 16; ++ww) {
          const int32_t c = wave_cnt[ww][tid];
          wave_cnt[ww][tid] = run;
          run += c;
        }
        chunk_cnt[tid] = run;
      }
      __syncthreads();
      if (valid) {
        const int64_t pos = digit_base[d] + wave_cnt[w][d] + rank;
        dk[pos] = key;
        dv[pos] = val;
      }
      __syncthreads();
      if (tid < 256) digit_base[tid] += chunk_cnt[tid];
      __syncthreads();
    }
  }
}

// head[p] = 1 where sorted position p starts a new output block.
__global__ void k_heads(const uint64_t *sk, int64_t n, int64_t *head) {
  const int64_t p = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (p >= n) return;
  head[p] = (p == 0 || sk[p] != sk[p - 1]) ? 1 : 0;
}

// For each output block (one per head): index, first sorted position, size,
// coordinates and degeneracy shape.
__global__ void k_cblocks(const uint64_t *sk, const int32_t *sv, const int64_t *head,
                          const int64_t *ord, int64_t n, const DevTask *tasks,
                          const int32_t *a_coords, const int32_t *a_dims, const int32_t *b_coords,
                          const int32_t *b_dims, const PlanArgs args, int64_t *c_idx,
                          int64_t *c_first, int64_t *c_size, int32_t *c_coords, int32_t *c_dims) {
  const int64_t p = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (p >= n || !head[p]) return;
  const int64_t o = ord[p];
  const DevTask t = tasks[sv[p]];
  c_idx[o] = static_cast<int64_t>(sk[p]);
  c_first[o] = p;
  const int32_t *ac = a_coords + static_cast<int64_t>(t.a_pos) * args.a_rank;
  const int32_t *ad = a_dims + static_cast<int64_t>(t.a_pos) * args.a_rank;
  const int32_t *bc = b_coords + static_cast<int64_t>(t.b_pos) * args.b_rank;
  const int32_t *bd = b_dims + static_cast<int64_t>(t.b_pos) * args.b_rank;
  int64_t size = 1;
  int ci = 0;
  for (int j = 0; j < args.n_a_saved; ++j, ++ci) {
    c_coords[o * QLT_MAX_RANK + ci] = ac[args.a_saved[j]];
    c_dims[o * QLT_MAX_RANK + ci] = ad[args.a_saved[j]];
    size *= ad[args.a_saved[j]];
  }
  for (int j = 0; j < args.n_b_saved; ++j, ++ci) {
    c_coords[o * QLT_MAX_RANK + ci] = bc[args.b_saved[j]];
    c_dims[o * QLT_MAX_RANK + ci] = bd[args.b_saved[j]];
    size *= bd[args.b_saved[j]];
  }
  c_size[o] = size;
}

// Finalise every task: output offset, beta, and its execution-order record.
__global__ void k_finalize(const int32_t *sv, const int64_t *head, const int64_t *ord, int64_t n,
                           const DevTask *tasks, const int64_t *c_off, const int64_t *c_idx,
                           const int32_t *a_dims, const int64_t *a_offs, const int64_t *a_idx,
                           const int32_t *b_dims, const int64_t *b_offs, const int64_t *b_idx,
                           const PlanArgs args, qlt_task *out, ExecTask *exec) {
  const int64_t p = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (p >= n) return;
  const int32_t g = sv[p];
  const DevTask t = tasks[g];
  const int32_t *ad = a_dims + static_cast<int64_t>(t.a_pos) * args.a_rank;
  const int32_t *bd = b_dims + static_cast<int64_t>(t.b_pos) * args.b_rank;
  int64_t m = 1, k = 1, nn = 1;
  for (int j = 0; j < args.n_a_saved; ++j) m *= ad[args.a_saved[j]];
  for (int j = 0; j < args.n_ctrct; ++j) k *= ad[args.a_ctrct[j]];
  for (int j = 0; j < args.n_b_saved; ++j) nn *= bd[args.b_saved[j]];
  const int64_t o = ord[p];
  qlt_task r;
  r.a_blk_idx = a_idx[t.a_pos];
  r.a_data_offset = a_offs[t.a_pos];
  r.b_blk_idx = b_idx[t.b_pos];
  r.b_data_offset = b_offs[t.b_pos];
  r.c_blk_idx = c_idx[o];
  r.c_data_offset = c_off[o];
  r.m = m;
  r.k = k;
  r.n = nn;
  r.f_ex_sign = 1.0;
  r.beta = head[p] ? 0.0 : 1.0;
  out[g] = r;
  ExecTask e;
  e.a_off = a_offs[t.a_pos];
  e.b_off = b_offs[t.b_pos];
  e.a_blk = t.a_pos;
  e.b_blk = t.b_pos;
  e.k = static_cast<int32_t>(k);
  e.pad = 0;
  exec[p] = e;
}

__global__ void k_segs(const int64_t *c_first, const int64_t *c_off, const int32_t *c_dims,
                       int64_t n_c, int64_t n_tasks, const ExecTask *exec, const PlanArgs args,
                       Seg *segs, int64_t *tile_cnt, int64_t *cost) {
  const int64_t o = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (o >= n_c) return;
  int64_t m = 1, n = 1;
  for (int j = 0; j < args.n_a_saved; ++j) m *= c_dims[o * QLT_MAX_RANK + j];
  for (int j = 0; j < args.n_b_saved; ++j) n *= c_dims[o * QLT_MAX_RANK + args.n_a_saved + j];
  Seg s;
  s.c_off = c_off[o];
  s.m = static_cast<int32_t>(m);
  s.n = static_cast<int32_t>(n);
  s.t0 = static_cast<int32_t>(c_first[o]);
  s.t1 = static_cast<int32_t>((o + 1 < n_c) ? c_first[o + 1] : n_tasks);
  int64_t kt = 0;
  for (int32_t t = s.t0; t < s.t1; ++t) kt += exec[t].k;
  s.k_total = kt;
  s.cls = tuned_tile_class(m, n, args.cls_table);
  s.flags = 0;      // plain flags are set on the host once the gather descriptors exist
  segs[o] = s;
  tile_cnt[o] = tiles_of(s.cls, m, n);
  cost[o] = kt;
}

__global__ void k_tiles(const Seg *segs, const int64_t *tile_start, const int64_t *tile_cnt,
                        int64_t n_c, int64_t kmax, const PlanArgs args, uint64_t *tkey,
                        int32_t *tval, Tile *tiles) {
  const int64_t o = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (o >= n_c) return;
  const Seg s = segs[o];
  const int ntn = is_mfma_class(s.cls) ? (s.n + mfma_tile_n(s.cls) - 1) / mfma_tile_n(s.cls) : 1;
  const int64_t base = tile_start[o];
  for (int64_t q = 0; q < tile_cnt[o]; ++q) {
    Tile t;
    t.seg = static_cast<int32_t>(o);
    t.tm = static_cast<int32_t>(q / ntn);
    t.tn = static_cast<int32_t>(q % ntn);
    t.cls = s.cls;
    t.t_begin = s.t0;
    t.k_begin = 0;
    t.n_chunks = 0;       // set on the host for the MFMA classes
    t.part = -1;
    tiles[base + q] = t;
    // class-major, then longest first inside a class
    tkey[base + q] = (static_cast<uint64_t>(s.cls) << 40) | static_cast<uint64_t>(kmax - s.k_total);
    tval[base + q] = static_cast<int32_t>(base + q);
  }
}

// Gather descriptor of every input block: A blocks are read as
// (saved ++ contracted), B blocks as (contracted ++ saved).
__global__ void k_gather_desc(const int32_t *dims, int64_t n, const PlanArgs args, int32_t is_a,
                              BlkGather *out) {
  const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  const int32_t rank = is_a ? args.a_rank : args.b_rank;
  const int32_t nr = is_a ? args.n_a_saved : args.n_ctrct;
  const int32_t nc = is_a ? args.n_ctrct : args.n_b_saved;
  const int32_t *raxes = is_a ? args.a_saved : args.b_ctrct;
  const int32_t *caxes = is_a ? args.a_ctrct : args.b_saved;
  const int32_t *d = dims + i * rank;
  int32_t stride[QLT_MAX_RANK];
  int32_t s = 1;
  for (int j = rank - 1; j >= 0; --j) {
    stride[j] = s;
    s *= d[j];
  }
  BlkGather g;
  for (int j = 0; j < QLT_MAX_RANK; ++j) {
    g.rdim[j] = 1;
    g.rstr[j] = 0;
    g.cdim[j] = 1;
    g.cstr[j] = 0;
  }
  int32_t nrows = 1, ncols = 1, last_axis = -1, plain = 1;
  for (int j = 0; j < nr; ++j) {
    g.rdim[j] = d[raxes[j]];
    g.rstr[j] = stride[raxes[j]];
    nrows *= d[raxes[j]];
    if (d[raxes[j]] > 1) {
      if (raxes[j] < last_axis) plain = 0;
      last_axis = raxes[j];
    }
  }
  for (int j = 0; j < nc; ++j) {
    g.cdim[j] = d[caxes[j]];
    g.cstr[j] = stride[caxes[j]];
    ncols *= d[caxes[j]];
    if (d[caxes[j]] > 1) {
      if (caxes[j] < last_axis) plain = 0;
      last_axis = caxes[j];
    }
  }
  g.plain = plain;
  g.ncols = ncols;
  g.nrows = nrows;
  // fastcol: the block's innermost axis of extent > 1 (stride 1) is one of
  // the col axes, so walking the col index walks memory contiguously
  int32_t inner = -1;
  for (int j = rank - 1; j >= 0; --j)
    if (d[j] > 1) {
      inner = j;
      break;
    }
  int32_t fastcol = 0;
  for (int j = 0; j < nc; ++j)
    if (caxes[j] == inner) fastcol = 1;
  g.fastcol = fastcol;
  // merge the axes of extent > 1 of each part: affine when every adjacent
  // pair satisfies stride[j] == stride[j + 1] * dim[j + 1]
  auto affine = [](const int32_t *dim, const int32_t *str, int n, int32_t &step) {
    int32_t last = -1;
    int ok = 1;
    step = 1;
    for (int j = 0; j < n; ++j) {
      if (dim[j] <= 1) continue;
      if (last >= 0 && str[last] != str[j] * dim[j]) ok = 0;
      last = j;
    }
    if (last >= 0) step = str[last];
    return ok;
  };
  g.caff = affine(g.cdim, g.cstr, nc, g.cstep);
  g.raff = affine(g.rdim, g.rstr, nr, g.rstep);
  for (int q = 0; q < QLT_MAX_RANK; ++q) {
    magic_of(static_cast<uint32_t>(g.rdim[q]), g.rmag[q], g.rsh[q]);
    magic_of(static_cast<uint32_t>(g.cdim[q]), g.cmag[q], g.csh[q]);
  }
  out[i] = g;
}

// ord[p] = ordinal of the output block that sorted position p belongs to.
__global__ void k_ord(const int64_t *excl_heads, const int64_t *head, int64_t n, int64_t *ord) {
  const int64_t p = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (p < n) ord[p] = excl_heads[p] + head[p] - 1;
}

__global__ void k_permute_tiles(const int32_t *order, const Tile *in, int64_t n, Tile *out) {
  const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (i < n) out[i] = in[order[i]];
}

inline unsigned blocks_for(int64_t n, int threads = 256) {
  int64_t g = (n + threads - 1) / threads;
  return static_cast<unsigned>(g < 1 ? 1 : g);
}

int bits_for(uint64_t maxval) {
  int b = 0;
  while (b < 64 && (maxval >> b) != 0) ++b;
  return b;
}

// RAII list of device allocations released at scope exit.
struct DevPool {
  std::vector<void *> ptrs;
  ~DevPool() {
    for (void *p : ptrs) (void)hipFree(p);
  }
  template <typename T>
  qlt_status alloc(T **p, int64_t count) {
    *p = nullptr;
    if (count <= 0) count = 1;
    hipError_t e = hipMalloc(reinterpret_cast<void **>(p), sizeof(T) * static_cast<size_t>(count));
    if (e != hipSuccess) return hip_fail(e, "hipMalloc (plan scratch)", __FILE__, __LINE__);
    ptrs.push_back(*p);
    return QLT_OK;
  }
};

// Marks the output blocks whose every task reads a plain A block and/or a
// plain B block; the MFMA kernel runs those on its division-free path.
qlt_status mark_plain_segments(qlt_plan *plan) {
  if (plan->n_segs == 0) return QLT_OK;
  std::vector<BlkGather> ga(static_cast<size_t>(plan->n_a_blks)), gb(static_cast<size_t>(plan->n_b_blks));
  QLT_HIP(hipMemcpy(ga.data(), plan->d_ga, sizeof(BlkGather) * ga.size(), hipMemcpyDeviceToHost));
  QLT_HIP(hipMemcpy(gb.data(), plan->d_gb, sizeof(BlkGather) * gb.size(), hipMemcpyDeviceToHost));
  for (Seg &sg : plan->h_segs) {
    int32_t f = kSegAPlain | kSegBPlain | kSegAColAffine | kSegBRowAffine;
    for (int32_t t = sg.t0; t < sg.t1; ++t) {
      const ExecTask &e = plan->h_exec[static_cast<size_t>(t)];
      const BlkGather &a = ga[static_cast<size_t>(e.a_blk)], &b = gb[static_cast<size_t>(e.b_blk)];
      if (!a.plain) f &= ~kSegAPlain;
      if (!b.plain) f &= ~kSegBPlain;
      if (!a.plain && !a.caff) f &= ~kSegAColAffine;
      if (!b.plain && !b.raff) f &= ~kSegBRowAffine;
    }
    sg.flags = f;
  }
  QLT_HIP(hipMemcpy(plan->d_segs, plan->h_segs.data(), sizeof(Seg) * plan->h_segs.size(), hipMemcpyHostToDevice));
  return QLT_OK;
}

// Row panels on (1) or off (0) for plans built from now on (qlt_set_row_panels).
std::atomic<int> g_row_panels{1};

// Host form of split_offset (plain division).
int64_t host_split_offset(int64_t x, const int32_t *dims, const int32_t *strides, int n) {
  int64_t off = 0;
  for (int j = n - 1; j >= 0; --j) {
    const int64_t q = x / dims[j];
    off += (x - q * dims[j]) * strides[j];
    x = q;
  }
  return off;
}

// Row panels of the n <= 4 row class (PanelTile).  The output blocks are in
// ascending block order with A's saved axes leading, so the blocks sharing
// their coordinates on those axes are consecutive.  Each run of such blocks
// whose n <= 4 members together touch at most kPanelCols A' columns, with at
// most kPanelMaxOuts output columns and kPanelMaxTerms terms, becomes one
// panel; in_panel[o] marks its output blocks, which leave the row class.
qlt_status build_panels(qlt_plan *plan, std::vector<char> &in_panel) {
  const int64_t n_c = plan->n_segs;
  in_panel.assign(static_cast<size_t>(n_c), 0);
  const int nr = plan->layout.a_nr;
  if (n_c == 0 || nr < 1 || nr > kPanelRowAxes || plan->c_rank < nr || plan->c_coords.empty()) return QLT_OK;
  const auto &segs = plan->h_segs;
  const auto &ex = plan->h_exec;
  std::vector<BlkGather> ga(static_cast<size_t>(plan->n_a_blks)), gb(static_cast<size_t>(plan->n_b_blks));
  QLT_HIP(hipMemcpy(ga.data(), plan->d_ga, sizeof(BlkGather) * ga.size(), hipMemcpyDeviceToHost));
  QLT_HIP(hipMemcpy(gb.data(), plan->d_gb, sizeof(BlkGather) * gb.size(), hipMemcpyDeviceToHost));
  const int32_t cr = plan->c_rank;
  const GatherLayout &L = plan->layout;
  std::vector<PanelTile> tiles;
  std::vector<PanelOut> outs;
  std::vector<PanelTerm> terms;
  int64_t o0 = 0;
  while (o0 < n_c) {
    // the run of output blocks with the coordinates of block o0 on A's saved axes
    int64_t o1 = o0 + 1;
    auto same_rows = [&](int64_t a, int64_t b) {
      for (int j = 0; j < nr; ++j)
        if (plan->c_coords[a * cr + j] != plan->c_coords[b * cr + j]) return false;
      return true;
    };
    while (o1 < n_c && same_rows(o0, o1)) ++o1;
    // its n <= 4 blocks, their A' columns (block, contracted index) in
    // ascending order, and per output column the terms in (task, k) order,
    // which must be ascending in the column
    std::vector<std::pair<int32_t, int32_t>> cols;
    std::vector<int64_t> members;
    std::vector<PanelOut> p_outs;
    std::vector<PanelTerm> p_terms;
    bool ok = true;
    int32_t m = -1;
    for (int64_t o = o0; o < o1 && ok; ++o) {
      const Seg &sg = segs[static_cast<size_t>(o)];
      if (sg.cls != kTileRows4) continue;
      if (m >= 0 && sg.m != m) ok = false;
      m = sg.m;
      members.push_back(o);
      for (int32_t t = sg.t0; t < sg.t1; ++t) {
        const ExecTask &e = ex[static_cast<size_t>(t)];
        for (int32_t p = 0; p < e.k; ++p) cols.emplace_back(e.a_blk, p);
      }
    }
    std::sort(cols.begin(), cols.end());
    cols.erase(std::unique(cols.begin(), cols.end()), cols.end());
    ok = ok && !members.empty() && m > 0 && static_cast<int64_t>(cols.size()) <= kPanelCols;
    for (size_t mi = 0; mi < members.size() && ok; ++mi) {
      const Seg &sg = segs[static_cast<size_t>(members[mi])];
      for (int32_t j = 0; j < sg.n && ok; ++j) {
        PanelOut po{static_cast<int32_t>(members[mi]), j, static_cast<int32_t>(p_terms.size()), 0};
        int32_t last = -1;
        for (int32_t t = sg.t0; t < sg.t1 && ok; ++t) {
          const ExecTask &e = ex[static_cast<size_t>(t)];
          const BlkGather &gB = gb[static_cast<size_t>(e.b_blk)];
          const int64_t bcol = host_split_offset(j, gB.cdim, gB.cstr, L.b_nc);
          for (int32_t p = 0; p < e.k && ok; ++p) {
            const int32_t c = static_cast<int32_t>(
                std::lower_bound(cols.begin(), cols.end(), std::make_pair(e.a_blk, p)) - cols.begin());
            if (c <= last) {
              ok = false;
              break;
            }
            last = c;
            po.mask |= 1 << c;
            const int64_t brow = host_split_offset(p, gB.rdim, gB.rstr, L.b_nr);
            p_terms.push_back(PanelTerm{e.b_off + brow + bcol, c, 0});
          }
        }
        p_outs.push_back(po);
      }
    }
    ok = ok && static_cast<int64_t>(p_outs.size()) <= kPanelMaxOuts &&
         static_cast<int64_t>(p_terms.size()) <= kPanelMaxTerms;
    if (ok) {
      PanelTile pt{};
      const BlkGather &g0 = ga[static_cast<size_t>(cols[0].first)];
      for (int l = 0; l < kPanelRowAxes; ++l) {
        pt.rdim[l] = l < nr ? g0.rdim[l] : 1;
        pt.rmag[l] = g0.rmag[l];
        pt.rsh[l] = g0.rsh[l];
      }
      for (int c = 0; c < kPanelCols; ++c) {
        pt.cbase[c] = 0;
        for (int l = 0; l < kPanelRowAxes; ++l) pt.cstr[c][l] = 0;
      }
      for (size_t c = 0; c < cols.size(); ++c) {
        const BlkGather &g = ga[static_cast<size_t>(cols[c].first)];
        // the A block's element offset comes with any task that reads it
        int64_t a_off = -1;
        for (int64_t o : members) {
          const Seg &sg = segs[static_cast<size_t>(o)];
          for (int32_t t = sg.t0; t < sg.t1 && a_off < 0; ++t)
            if (ex[static_cast<size_t>(t)].a_blk == cols[c].first) a_off = ex[static_cast<size_t>(t)].a_off;
        }
        pt.cbase[c] = a_off + host_split_offset(cols[c].second, g.cdim, g.cstr, L.a_nc);
        for (int l = 0; l < nr; ++l) pt.cstr[c][l] = g.rstr[l];
      }
      pt.pair1 = 0;
      for (int c = 0; c + 1 < static_cast<int>(cols.size()) && pt.pair1 == 0; ++c) {
        bool adj = pt.cbase[c + 1] == pt.cbase[c] + 1;
        for (int l = 0; l < kPanelRowAxes; ++l) adj = adj && pt.cstr[c][l] == pt.cstr[c + 1][l];
        if (adj) pt.pair1 = c + 1;
      }
      pt.m = m;
      pt.n_cols = static_cast<int32_t>(cols.size());
      pt.n_outs = static_cast<int32_t>(p_outs.size());
      pt.n_terms = static_cast<int32_t>(p_terms.size());
      pt.out0 = static_cast<int32_t>(outs.size());
      pt.term0 = static_cast<int32_t>(terms.size());
      for (int32_t r0 = 0; r0 < m; r0 += kPanelRows) {
        pt.r0 = r0;
        tiles.push_back(pt);
      }
      outs.insert(outs.end(), p_outs.begin(), p_outs.end());
      terms.insert(terms.end(), p_terms.begin(), p_terms.end());
      for (int64_t o : members) {
        in_panel[static_cast<size_t>(o)] = 1;
        const Seg &sg = segs[static_cast<size_t>(o)];
        for (int32_t t = sg.t0; t < sg.t1; ++t)
          plan->panel_mkn += static_cast<double>(sg.m) * static_cast<double>(ex[static_cast<size_t>(t)].k) * sg.n;
        plan->panel_elems += static_cast<double>(sg.m) * sg.n;
      }
      plan->panel_elems += static_cast<double>(m) * static_cast<double>(cols.size());
    }
    o0 = o1;
  }
  QLT_CHECK_ARG(static_cast<int64_t>(terms.size()) < (int64_t(1) << 31) &&
                    static_cast<int64_t>(tiles.size()) < (int64_t(1) << 31),
                "qlt_plan_build: too many row panels");
  plan->n_ptiles = static_cast<int64_t>(tiles.size());
  if (tiles.empty()) return QLT_OK;
  QLT_HIP(hipMalloc(&plan->d_ptiles, sizeof(PanelTile) * tiles.size()));
  QLT_HIP(hipMalloc(&plan->d_pouts, sizeof(PanelOut) * outs.size()));
  QLT_HIP(hipMalloc(&plan->d_pterms, sizeof(PanelTerm) * terms.size()));
  QLT_HIP(hipMemcpy(plan->d_ptiles, tiles.data(), sizeof(PanelTile) * tiles.size(), hipMemcpyHostToDevice));
  QLT_HIP(hipMemcpy(plan->d_pouts, outs.data(), sizeof(PanelOut) * outs.size(), hipMemcpyHostToDevice));
  QLT_HIP(hipMemcpy(plan->d_pterms, terms.data(), sizeof(PanelTerm) * terms.size(), hipMemcpyHostToDevice));
  // what the row class keeps
  plan->cls_mkn[kTileRows4] = 0.0;
  plan->cls_elems[kTileRows4] = 0.0;
  std::vector<char> seen_a(static_cast<size_t>(plan->n_a_blks), 0), seen_b(static_cast<size_t>(plan->n_b_blks), 0);
  for (int64_t o = 0; o < n_c; ++o) {
    const Seg &sg = segs[static_cast<size_t>(o)];
    if (sg.cls != kTileRows4 || in_panel[static_cast<size_t>(o)]) continue;
    plan->cls_mkn[kTileRows4] += static_cast<double>(sg.m) * sg.n * static_cast<double>(sg.k_total);
    plan->cls_elems[kTileRows4] += static_cast<double>(sg.m) * sg.n;
    for (int32_t t = sg.t0; t < sg.t1; ++t) {
      const ExecTask &e = ex[static_cast<size_t>(t)];
      if (!seen_a[e.a_blk]) {
        seen_a[e.a_blk] = 1;
        plan->cls_elems[kTileRows4] += static_cast<double>(sg.m) * e.k;
      }
      if (!seen_b[e.b_blk]) {
        seen_b[e.b_blk] = 1;
        plan->cls_elems[kTileRows4] += static_cast<double>(e.k) * sg.n;
      }
    }
  }
  return QLT_OK;
}

// k-slicing of long tiles (MFMA and row classes alike).  Per class, a step budget Q = max(kMinSliceChunks,
// ceil(total steps / T)) is set, T = kSliceTargetWGsMfma (MFMA classes) or kSliceTargetWGs (row classes)
// unless tuned (qlt_tuning); a tile of w > Q steps is cut into
// ceil(w / Q) (at most kMaxSlices) slices of nearly equal length along its
// (task, k-chunk) sequence, each writing an unscaled partial sum that a
// reduction kernel adds in slice order.  Classes whose tiles are many and even
// stay unsliced; few or long tiles (a contraction over a long k onto a small
// output) spread over the chip.  Slices of >= kMinSliceChunks steps keep the
// partial-sum traffic small against the slice's MFMA work.
constexpr int64_t kSliceTargetWGs = 2048;
// The MFMA classes aim lower: fewer, longer slices (and less partial-sum
// traffic) measured faster on every matvec chain here than 2048.
constexpr int64_t kSliceTargetWGsMfma = 512;
constexpr int64_t kMinSliceChunks = 8;
constexpr int64_t kMaxSlices = 64;

qlt_status slice_long_tiles(qlt_plan *plan) {
  const int64_t n = plan->n_tiles;
  if (n == 0) return QLT_OK;
  std::vector<Tile> h(static_cast<size_t>(n));
  QLT_HIP(hipMemcpy(h.data(), plan->d_tiles, sizeof(Tile) * n, hipMemcpyDeviceToHost));
  const auto &segs = plan->h_segs;
  const auto &ex = plan->h_exec;
  auto steps_of = [](int64_t k) { return (k + kKChunk - 1) / kKChunk; };
  std::vector<int64_t> seg_steps(segs.size(), 0);
  for (size_t o = 0; o < segs.size(); ++o)
    for (int32_t t = segs[o].t0; t < segs[o].t1; ++t) seg_steps[o] += steps_of(ex[static_cast<size_t>(t)].k);
  std::vector<char> in_panel;
  if (plan->panels_enabled) QLT_TRY(build_panels(plan, in_panel));
  std::vector<Tile> out;
  std::vector<ReduceItem> red;
  out.reserve(static_cast<size_t>(n));
  int64_t base = 0, part_doubles = 0;
  for (int32_t cls = 0; cls < kNumTileClasses; ++cls) {
    const int64_t cnt = plan->cls_tiles[cls];
    // partial-sum doubles of one slice (complex), in the kernels' thread order
    const int64_t slice_doubles = is_mfma_class(cls) ? int64_t(mfma_tile_m(cls)) * mfma_tile_n(cls) * 2
                                                     : int64_t(rows_per_tile(cls)) * (cls == kTileRows4 ? 4 : 16) * 2;
    int64_t total = 0;
    for (int64_t i = base; i < base + cnt; ++i)
      if (in_panel.empty() || !in_panel[static_cast<size_t>(h[i].seg)]) total += seg_steps[static_cast<size_t>(h[i].seg)];
    // tuned workgroup target of the class (qlt_tuning; -1: unsliced)
    const bool unsliced = plan->slice_target[cls] < 0;
    const int64_t target = plan->slice_target[cls] > 0 ? plan->slice_target[cls]
                                                       : (is_mfma_class(cls) ? kSliceTargetWGsMfma : kSliceTargetWGs);
    const int64_t q = std::max(kMinSliceChunks, (total + target - 1) / target);
    std::vector<Tile> cl;
    int64_t parts = 0, n_red = 0;
    for (int64_t i = base; i < base + cnt; ++i) {
      Tile t = h[static_cast<size_t>(i)];
      if (!in_panel.empty() && in_panel[static_cast<size_t>(t.seg)]) continue;   // runs as a row panel
      const Seg &s = segs[static_cast<size_t>(t.seg)];
      const int64_t w = seg_steps[static_cast<size_t>(t.seg)];
      const int64_t ns = unsliced ? 1 : std::min(kMaxSlices, (w + q - 1) / q);
      t.t_begin = s.t0;
      t.k_begin = 0;
      t.n_chunks = static_cast<int32_t>(w);
      t.part = -1;
      if (ns <= 1) {
        cl.push_back(t);
        continue;
      }
      red.push_back(ReduceItem{t.seg, t.tm, t.tn, cls, static_cast<int32_t>(parts), static_cast<int32_t>(ns)});
      ++n_red;
      int32_t task = s.t0;
      int64_t in_task = 0;     // steps already taken inside `task`
      for (int64_t sl = 0; sl < ns; ++sl) {
        const int64_t len = w / ns + (sl < w % ns ? 1 : 0);
        Tile u = t;
        u.t_begin = task;
        u.k_begin = static_cast<int32_t>(in_task * kKChunk);
        u.n_chunks = static_cast<int32_t>(len);
        u.part = static_cast<int32_t>(parts + sl);
        cl.push_back(u);
        for (int64_t rem = len; rem > 0;) {
          const int64_t left = steps_of(ex[static_cast<size_t>(task)].k) - in_task;
          if (rem < left) {
            in_task += rem;
            rem = 0;
          } else {
            rem -= left;
            ++task;
            in_task = 0;
          }
        }
      }
      parts += ns;
    }
    if (is_mfma_class(cls)) {
      std::stable_sort(cl.begin(), cl.end(), [](const Tile &a, const Tile &b) { return a.n_chunks > b.n_chunks; });
    } else {
      // row tiles are short and bound by reading A: tiles that start on the
      // same A block rows run next to each other (and on one XCD), so the
      // output blocks that re-read an A block find it in L2
      auto key = [&](const Tile &t) {
        const ExecTask &e = ex[static_cast<size_t>(t.t_begin)];
        return std::make_pair(static_cast<int64_t>(e.a_blk), static_cast<int64_t>(t.tm));
      };
      std::stable_sort(cl.begin(), cl.end(), [&](const Tile &a, const Tile &b) { return key(a) < key(b); });
    }
    out.insert(out.end(), cl.begin(), cl.end());
    plan->cls_tiles[cls] = static_cast<int64_t>(cl.size());
    plan->cls_red[cls] = n_red;
    // every class gets its own region: the class kernels run concurrently
    plan->cls_part_off[cls] = part_doubles;
    part_doubles += parts * slice_doubles;
    base += cnt;
  }
  QLT_CHECK_ARG(static_cast<int64_t>(out.size()) < (int64_t(1) << 31), "qlt_plan_build: too many tiles after k-slicing");
  if (static_cast<int64_t>(out.size()) != n) {
    Tile *nt = nullptr;
    QLT_HIP(hipMalloc(&nt, sizeof(Tile) * out.size()));
    QLT_HIP(hipFree(plan->d_tiles));
    plan->d_tiles = nt;
  }
  QLT_HIP(hipMemcpy(plan->d_tiles, out.data(), sizeof(Tile) * out.size(), hipMemcpyHostToDevice));
  plan->n_tiles = static_cast<int64_t>(out.size());
  if (!red.empty()) {
    QLT_HIP(hipMalloc(&plan->d_red, sizeof(ReduceItem) * red.size()));
    QLT_HIP(hipMemcpy(plan->d_red, red.data(), sizeof(ReduceItem) * red.size(), hipMemcpyHostToDevice));
    QLT_HIP(hipMalloc(&plan->d_part, sizeof(double) * static_cast<size_t>(part_doubles)));
    plan->part_doubles = static_cast<size_t>(part_doubles);
  }
  return QLT_OK;
}

void free_plan_device(qlt_plan *p) {
  if (p->d_exec) (void)hipFree(p->d_exec);
  if (p->d_segs) (void)hipFree(p->d_segs);
  if (p->d_tiles) (void)hipFree(p->d_tiles);
  if (p->d_ga) (void)hipFree(p->d_ga);
  if (p->d_gb) (void)hipFree(p->d_gb);
  if (p->d_work) (void)hipFree(p->d_work);
  if (p->d_red) (void)hipFree(p->d_red);
  if (p->d_part) (void)hipFree(p->d_part);
  if (p->d_ptiles) (void)hipFree(p->d_ptiles);
  if (p->d_pouts) (void)hipFree(p->d_pouts);
  if (p->d_pterms) (void)hipFree(p->d_pterms);
  p->d_ptiles = nullptr;
  p->d_pouts = nullptr;
  p->d_pterms = nullptr;
  p->d_red = nullptr;
  p->d_part = nullptr;
  p->d_exec = nullptr;
  p->d_segs = nullptr;
  p->d_tiles = nullptr;
  p->d_ga = nullptr;
  p->d_gb = nullptr;
  p->d_work = nullptr;
}

qlt_status validate_desc(const qlt_bsdt_desc *d, const char *name) {
  const std::string n(name);
  QLT_CHECK_ARG(d != nullptr, "qlt_plan_build: null descriptor " + n);
  QLT_CHECK_ARG(d->rank >= 1 && d->rank <= QLT_MAX_RANK,
                "qlt_plan_build: rank of " + n + " must be in 1.." + std::to_string(QLT_MAX_RANK));
  QLT_CHECK_ARG(d->n_blks >= 0 && d->n_blks < (int64_t(1) << 31),
                "qlt_plan_build: bad block count of " + n);
  if (d->n_blks == 0) return QLT_OK;
  QLT_CHECK_ARG(d->blk_shape && d->blk_coords && d->blk_dims && d->blk_offsets,
                "qlt_plan_build: null array in descriptor " + n);
  for (int j = 0; j < d->rank; ++j)
    QLT_CHECK_ARG(d->blk_shape[j] >= 1 && d->blk_shape[j] < (int64_t(1) << 31),
                  "qlt_plan_build: bad sector count in " + n);
  uint64_t prev_idx = 0;
  for (int64_t i = 0; i < d->n_blks; ++i) {
    uint64_t idx = 0;
    int64_t size = 1;
    for (int j = 0; j < d->rank; ++j) {
      const int64_t c = d->blk_coords[i * d->rank + j];
      const int64_t dim = d->blk_dims[i * d->rank + j];
      QLT_CHECK_ARG(c >= 0 && c < d->blk_shape[j], "qlt_plan_build: block coordinate out of range in " + n);
      QLT_CHECK_ARG(dim >= 1 && dim < (int64_t(1) << 31), "qlt_plan_build: bad block dimension in " + n);
      idx = idx * static_cast<uint64_t>(d->blk_shape[j]) + static_cast<uint64_t>(c);
      size *= dim;
      QLT_CHECK_ARG(size < (int64_t(1) << 31), "qlt_plan_build: block of " + n + " exceeds 2^31 elements");
    }
    QLT_CHECK_ARG(i == 0 || idx > prev_idx,
                  "qlt_plan_build: blocks of " + n + " must be in strictly ascending block index");
    QLT_CHECK_ARG(d->blk_offsets[i] >= 0, "qlt_plan_build: negative block offset in " + n);
    prev_idx = idx;
  }
  return QLT_OK;
}

}  // namespace


namespace {

// Block index of every block of a descriptor (row-major over blk_shape).
std::vector<int64_t> block_indices(const qlt_bsdt_desc *d) {
  std::vector<int64_t> idx(d->n_blks);
  for (int64_t i = 0; i < d->n_blks; ++i) {
    uint64_t v = 0;
    for (int j = 0; j < d->rank; ++j)
      v = v * static_cast<uint64_t>(d->blk_shape[j]) + static_cast<uint64_t>(d->blk_coords[i * d->rank + j]);
    idx[i] = static_cast<int64_t>(v);
  }
  return idx;
}

qlt_status check_axes(const qlt_bsdt_desc *a, const qlt_bsdt_desc *b, int32_t n_ctrct,
                      const int32_t *a_ctrct_axes, const int32_t *b_ctrct_axes, int32_t n_a_saved,
                      const int32_t *a_saved_axes, int32_t n_b_saved, const int32_t *b_saved_axes) {
  QLT_CHECK_ARG(n_ctrct >= 0 && n_a_saved >= 0 && n_b_saved >= 0, "qlt_plan_build: negative axis count");
  QLT_CHECK_ARG(n_ctrct + n_a_saved == a->rank, "qlt_plan_build: A axes do not cover A's rank");
  QLT_CHECK_ARG(n_ctrct + n_b_saved == b->rank, "qlt_plan_build: B axes do not cover B's rank");
  QLT_CHECK_ARG(n_ctrct == 0 || (a_ctrct_axes && b_ctrct_axes), "qlt_plan_build: null contracted axes");
  QLT_CHECK_ARG(n_a_saved == 0 || a_saved_axes, "qlt_plan_build: null A saved axes");
  QLT_CHECK_ARG(n_b_saved == 0 || b_saved_axes, "qlt_plan_build: null B saved axes");
  QLT_CHECK_ARG(n_a_saved + n_b_saved <= QLT_MAX_RANK, "qlt_plan_build: output rank exceeds QLT_MAX_RANK");
  std::vector<int> seen_a(a->rank, 0), seen_b(b->rank, 0);
  for (int j = 0; j < n_ctrct; ++j) {
    QLT_CHECK_ARG(a_ctrct_axes[j] >= 0 && a_ctrct_axes[j] < a->rank, "qlt_plan_build: bad A axis");
    QLT_CHECK_ARG(b_ctrct_axes[j] >= 0 && b_ctrct_axes[j] < b->rank, "qlt_plan_build: bad B axis");
    seen_a[a_ctrct_axes[j]]++;
    seen_b[b_ctrct_axes[j]]++;
  }
  for (int j = 0; j < n_a_saved; ++j) {
    QLT_CHECK_ARG(a_saved_axes[j] >= 0 && a_saved_axes[j] < a->rank, "qlt_plan_build: bad A saved axis");
    seen_a[a_saved_axes[j]]++;
  }
  for (int j = 0; j < n_b_saved; ++j) {
    QLT_CHECK_ARG(b_saved_axes[j] >= 0 && b_saved_axes[j] < b->rank, "qlt_plan_build: bad B saved axis");
    seen_b[b_saved_axes[j]]++;
  }
  for (int v : seen_a) QLT_CHECK_ARG(v == 1, "qlt_plan_build: A axes are not a permutation of A's axes");
  for (int v : seen_b) QLT_CHECK_ARG(v == 1, "qlt_plan_build: B axes are not a permutation of B's axes");
  if (a->n_blks > 0 && b->n_blks > 0)
    for (int j = 0; j < n_ctrct; ++j)
      QLT_CHECK_ARG(a->blk_shape[a_ctrct_axes[j]] == b->blk_shape[b_ctrct_axes[j]],
                    "qlt_plan_build: contracted indexes have different sector counts");
  return QLT_OK;
}

}  // namespace

extern "C" {

qlt_status qlt_plan_build(const qlt_bsdt_desc *a, const qlt_bsdt_desc *b, int32_t n_ctrct,
                          const int32_t *a_ctrct_axes, const int32_t *b_ctrct_axes,
                          int32_t n_a_saved, const int32_t *a_saved_axes, int32_t n_b_saved,
                          const int32_t *b_saved_axes, void *stream, qlt_plan **plan_out) {
  QLT_REQUIRE_INIT();
  QLT_CHECK_ARG(plan_out != nullptr, "qlt_plan_build: null output");
  *plan_out = nullptr;
  QLT_TRY(validate_desc(a, "A"));
  QLT_TRY(validate_desc(b, "B"));
  QLT_TRY(check_axes(a, b, n_ctrct, a_ctrct_axes, b_ctrct_axes, n_a_saved, a_saved_axes, n_b_saved,
                     b_saved_axes));
  hipStream_t st = as_stream(stream);
  QLT_CHECK_ARG(!stream_capturing(st), "qlt_plan_build: cannot build a plan under graph capture");
  const int32_t c_rank = n_a_saved + n_b_saved;

  PlanArgs args{};
  args.a_rank = a->rank;
  args.b_rank = b->rank;
  args.n_ctrct = n_ctrct;
  args.n_a_saved = n_a_saved;
  args.n_b_saved = n_b_saved;
  args.c_rank = c_rank;
  for (int j = 0; j < n_ctrct; ++j) {
    args.a_ctrct[j] = a_ctrct_axes[j];
    args.b_ctrct[j] = b_ctrct_axes[j];
  }
  for (int j = 0; j < n_a_saved; ++j) args.a_saved[j] = a_saved_axes[j];
  for (int j = 0; j < n_b_saved; ++j) args.b_saved[j] = b_saved_axes[j];
  args.c_is_scalar = (c_rank == 0) ? 1 : 0;
  args.tile_m = 64;
  const qlt_tuning tun = tuning_snapshot();
  std::memcpy(args.cls_table, tun.tile_class, sizeof(args.cls_table));
  args.tile_n = 64;

  auto plan = std::make_unique<qlt_plan>();
  plan->device = current_device();
  plan->c_rank = c_rank;
  plan->c_is_scalar = args.c_is_scalar;
  plan->tile_m = args.tile_m;
  for (int c = 0; c < kNumTileClasses; ++c) plan->slice_target[c] = tun.slice_target[c];
  plan->tile_n = args.tile_n;
  plan->n_a_blks = a->n_blks;
  plan->n_b_blks = b->n_blks;
  {
    GatherLayout &L = plan->layout;
    L.a_nr = n_a_saved;
    L.a_nc = n_ctrct;
    L.b_nr = n_ctrct;
    L.b_nc = n_b_saved;
    L.a_kfast = (n_ctrct > 0 && a_ctrct_axes[n_ctrct - 1] == a->rank - 1) ? 1 : 0;
    L.b_nfast = (n_b_saved > 0 && b_saved_axes[n_b_saved - 1] == b->rank - 1) ? 1 : 0;
  }
  const int64_t na = a->n_blks, nb = b->n_blks;
  for (int64_t i = 0; i < na; ++i) {
    int64_t sz = 1;
    for (int j = 0; j < a->rank; ++j) sz *= a->blk_dims[i * a->rank + j];
    plan->max_blk_elems_a = std::max(plan->max_blk_elems_a, sz);
    plan->a_raw_extent = std::max(plan->a_raw_extent, a->blk_offsets[i] + sz);
  }
  for (int64_t i = 0; i < nb; ++i) {
    int64_t sz = 1;
    for (int j = 0; j < b->rank; ++j) sz *= b->blk_dims[i * b->rank + j];
    plan->max_blk_elems_b = std::max(plan->max_blk_elems_b, sz);
    plan->b_raw_extent = std::max(plan->b_raw_extent, b->blk_offsets[i] + sz);
  }
  if (na == 0 || nb == 0) {
    *plan_out = plan.release();
    return QLT_OK;
  }

  // Output block strides, row-major over the output sector counts.
  uint64_t c_space = 1;
  {
    std::vector<int64_t> cshape;
    for (int j = 0; j < n_a_saved; ++j) cshape.push_back(a->blk_shape[a_saved_axes[j]]);
    for (int j = 0; j < n_b_saved; ++j) cshape.push_back(b->blk_shape[b_saved_axes[j]]);
    unsigned __int128 s = 1;
    for (int j = c_rank - 1; j >= 0; --j) {
      args.c_stride[j] = static_cast<int64_t>(s);
      s *= static_cast<uint64_t>(cshape[j]);
      QLT_CHECK_ARG(s < (static_cast<unsigned __int128>(1) << 63),
                    "qlt_plan_build: output block index space exceeds 2^63");
    }
    c_space = static_cast<uint64_t>(s);
  }

  // ---- upload block metadata ----
  DevPool pool;
  const std::vector<int64_t> h_aidx = block_indices(a), h_bidx = block_indices(b);
  std::vector<int32_t> h_ac(na * a->rank), h_ad(na * a->rank), h_bc(nb * b->rank), h_bd(nb * b->rank);
  for (int64_t i = 0; i < na * a->rank; ++i) {
    h_ac[i] = static_cast<int32_t>(a->blk_coords[i]);
    h_ad[i] = static_cast<int32_t>(a->blk_dims[i]);
  }
  for (int64_t i = 0; i < nb * b->rank; ++i) {
    h_bc[i] = static_cast<int32_t>(b->blk_coords[i]);
    h_bd[i] = static_cast<int32_t>(b->blk_dims[i]);
  }
  int32_t *d_ac, *d_ad, *d_bc, *d_bd;
  int64_t *d_aoff, *d_boff, *d_aidx, *d_bidx;
  QLT_TRY(pool.alloc(&d_ac, na * a->rank));
  QLT_TRY(pool.alloc(&d_ad, na * a->rank));
  QLT_TRY(pool.alloc(&d_bc, nb * b->rank));
  QLT_TRY(pool.alloc(&d_bd, nb * b->rank));
  QLT_TRY(pool.alloc(&d_aoff, na));
  QLT_TRY(pool.alloc(&d_boff, nb));
  QLT_TRY(pool.alloc(&d_aidx, na));
  QLT_TRY(pool.alloc(&d_bidx, nb));
  QLT_HIP(hipMemcpyAsync(d_ac, h_ac.data(), h_ac.size() * 4, hipMemcpyHostToDevice, st));
  QLT_HIP(hipMemcpyAsync(d_ad, h_ad.data(), h_ad.size() * 4, hipMemcpyHostToDevice, st));
  QLT_HIP(hipMemcpyAsync(d_bc, h_bc.data(), h_bc.size() * 4, hipMemcpyHostToDevice, st));
  QLT_HIP(hipMemcpyAsync(d_bd, h_bd.data(), h_bd.size() * 4, hipMemcpyHostToDevice, st));
  QLT_HIP(hipMemcpyAsync(d_aoff, a->blk_offsets, na * 8, hipMemcpyHostToDevice, st));
  QLT_HIP(hipMemcpyAsync(d_boff, b->blk_offsets, nb * 8, hipMemcpyHostToDevice, st));
  QLT_HIP(hipMemcpyAsync(d_aidx, h_aidx.data(), na * 8, hipMemcpyHostToDevice, st));
  QLT_HIP(hipMemcpyAsync(d_bidx, h_bidx.data(), nb * 8, hipMemcpyHostToDevice, st));

  // ---- 1. part hashes ----
  uint64_t *d_ha, *d_hb;
  QLT_TRY(pool.alloc(&d_ha, na));
  QLT_TRY(pool.alloc(&d_hb, nb));
  hipLaunchKernelGGL(k_part_hash, dim3(blocks_for(na)), dim3(256), 0, st, d_ac, na, a->rank, args, 1, d_ha);
  hipLaunchKernelGGL(k_part_hash, dim3(blocks_for(nb)), dim3(256), 0, st, d_bc, nb, b->rank, args, 0, d_hb);

  // ---- 2. matches per A block and task slots ----
  int64_t *d_cnt, *d_start, *d_scal;
  QLT_TRY(pool.alloc(&d_cnt, na));
  QLT_TRY(pool.alloc(&d_start, na));
  QLT_TRY(pool.alloc(&d_scal, 8));
  hipLaunchKernelGGL(k_count, dim3(blocks_for(na)), dim3(256), 0, st, d_ha, na, d_hb, nb,
                     args.c_is_scalar, d_cnt);
  hipLaunchKernelGGL(k_scan_excl, dim3(1), dim3(1024), 0, st, d_cnt, d_start, na, d_scal);
  QLT_HIP(hipGetLastError());
  int64_t nt = 0;
  QLT_HIP(hipMemcpyAsync(&nt, d_scal, 8, hipMemcpyDeviceToHost, st));
  QLT_HIP(hipStreamSynchronize(st));
  if (nt == 0) {
    *plan_out = plan.release();
    return QLT_OK;
  }
  QLT_CHECK_ARG(nt < (int64_t(1) << 31), "qlt_plan_build: more than 2^31 tasks");
  DevTask *d_tasks;
  QLT_TRY(pool.alloc(&d_tasks, nt));
  hipLaunchKernelGGL(k_emit, dim3(blocks_for(na * 64)), dim3(256), 0, st, d_ha, na, d_hb, nb, d_start,
                     args.c_is_scalar, d_tasks);

  // ---- 3. stable sort of tasks by output block index ----
  uint64_t *d_k0, *d_k1;
  int32_t *d_v0, *d_v1;
  QLT_TRY(pool.alloc(&d_k0, nt));
  QLT_TRY(pool.alloc(&d_k1, nt));
  QLT_TRY(pool.alloc(&d_v0, nt));
  QLT_TRY(pool.alloc(&d_v1, nt));
  hipLaunchKernelGGL(k_task_keys, dim3(blocks_for(nt)), dim3(256), 0, st, d_tasks, nt, d_ac, d_bc, args,
                     d_k0, d_v0);
  const int passes = (bits_for(c_space - 1) + 7) / 8;
  if (passes > 0)
    hipLaunchKernelGGL(k_radix_sort, dim3(1), dim3(1024), 0, st, d_k0, d_v0, d_k1, d_v1, nt, passes);
  const uint64_t *d_sk = (passes & 1) ? d_k1 : d_k0;
  const int32_t *d_sv = (passes & 1) ? d_v1 : d_v0;

  // ---- 4. output blocks (first touch -> beta 0) ----
  int64_t *d_head, *d_exh, *d_ord;
  QLT_TRY(pool.alloc(&d_head, nt));
  QLT_TRY(pool.alloc(&d_exh, nt));
  QLT_TRY(pool.alloc(&d_ord, nt));
  hipLaunchKernelGGL(k_heads, dim3(blocks_for(nt)), dim3(256), 0, st, d_sk, nt, d_head);
  hipLaunchKernelGGL(k_scan_excl, dim3(1), dim3(1024), 0, st, d_head, d_exh, nt, d_scal + 1);
  hipLaunchKernelGGL(k_ord, dim3(blocks_for(nt)), dim3(256), 0, st, d_exh, d_head, nt, d_ord);
  QLT_HIP(hipGetLastError());
  int64_t n_c = 0;
  QLT_HIP(hipMemcpyAsync(&n_c, d_scal + 1, 8, hipMemcpyDeviceToHost, st));
  QLT_HIP(hipStreamSynchronize(st));
  int64_t *d_cidx, *d_cfirst, *d_csize, *d_coff;
  int32_t *d_ccoords, *d_cdims;
  QLT_TRY(pool.alloc(&d_cidx, n_c));
  QLT_TRY(pool.alloc(&d_cfirst, n_c));
  QLT_TRY(pool.alloc(&d_csize, n_c));
  QLT_TRY(pool.alloc(&d_coff, n_c));
  QLT_TRY(pool.alloc(&d_ccoords, n_c * QLT_MAX_RANK));
  QLT_TRY(pool.alloc(&d_cdims, n_c * QLT_MAX_RANK));
  hipLaunchKernelGGL(k_cblocks, dim3(blocks_for(nt)), dim3(256), 0, st, d_sk, d_sv, d_head, d_ord, nt,
                     d_tasks, d_ac, d_ad, d_bc, d_bd, args, d_cidx, d_cfirst, d_csize, d_ccoords, d_cdims);
  // ---- 5. output offsets: prefix sum in ascending output block index ----
  hipLaunchKernelGGL(k_scan_excl, dim3(1), dim3(1024), 0, st, d_csize, d_coff, n_c, d_scal + 2);

  // ---- 6. task records and execution order ----
  qlt_task *d_out;
  QLT_TRY(pool.alloc(&d_out, nt));
  hipError_t e = hipMalloc(&plan->d_exec, sizeof(ExecTask) * nt);
  if (e != hipSuccess) return hip_fail(e, "hipMalloc (exec list)", __FILE__, __LINE__);
  hipLaunchKernelGGL(k_finalize, dim3(blocks_for(nt)), dim3(256), 0, st, d_sv, d_head, d_ord, nt, d_tasks,
                     d_coff, d_cidx, d_ad, d_aoff, d_aidx, d_bd, d_boff, d_bidx, args, d_out, plan->d_exec);

  // ---- 7. segments and tiles ----
  int64_t *d_tcnt, *d_tstart, *d_cost;
  QLT_TRY(pool.alloc(&d_tcnt, n_c));
  QLT_TRY(pool.alloc(&d_tstart, n_c));
  QLT_TRY(pool.alloc(&d_cost, n_c));
  e = hipMalloc(&plan->d_segs, sizeof(Seg) * n_c);
  if (e != hipSuccess) {
    free_plan_device(plan.get());
    return hip_fail(e, "hipMalloc (segments)", __FILE__, __LINE__);
  }
  hipLaunchKernelGGL(k_segs, dim3(blocks_for(n_c)), dim3(256), 0, st, d_cfirst, d_coff, d_cdims, n_c, nt,
                     plan->d_exec, args, plan->d_segs, d_tcnt, d_cost);
  hipLaunchKernelGGL(k_scan_excl, dim3(1), dim3(1024), 0, st, d_tcnt, d_tstart, n_c, d_scal + 3);
  QLT_HIP(hipGetLastError());

  // ---- 8. copy the plan back to the host ----
  int64_t scal[4] = {0, 0, 0, 0};
  QLT_HIP(hipMemcpyAsync(scal, d_scal, sizeof(scal), hipMemcpyDeviceToHost, st));
  plan->tasks.resize(nt);
  QLT_HIP(hipMemcpyAsync(plan->tasks.data(), d_out, sizeof(qlt_task) * nt, hipMemcpyDeviceToHost, st));
  plan->h_segs.resize(n_c);
  QLT_HIP(hipMemcpyAsync(plan->h_segs.data(), plan->d_segs, sizeof(Seg) * n_c, hipMemcpyDeviceToHost, st));
  plan->h_exec.resize(nt);
  QLT_HIP(hipMemcpyAsync(plan->h_exec.data(), plan->d_exec, sizeof(ExecTask) * nt, hipMemcpyDeviceToHost, st));
  std::vector<int64_t> h_cidx(n_c), h_coff(n_c);
  std::vector<int32_t> h_ccoords(n_c * QLT_MAX_RANK), h_cdims(n_c * QLT_MAX_RANK);
  QLT_HIP(hipMemcpyAsync(h_cidx.data(), d_cidx, 8 * n_c, hipMemcpyDeviceToHost, st));
  QLT_HIP(hipMemcpyAsync(h_coff.data(), d_coff, 8 * n_c, hipMemcpyDeviceToHost, st));
  QLT_HIP(hipMemcpyAsync(h_ccoords.data(), d_ccoords, 4 * n_c * QLT_MAX_RANK, hipMemcpyDeviceToHost, st));
  QLT_HIP(hipMemcpyAsync(h_cdims.data(), d_cdims, 4 * n_c * QLT_MAX_RANK, hipMemcpyDeviceToHost, st));
  QLT_HIP(hipStreamSynchronize(st));
  const int64_t c_raw = scal[2];
  const int64_t n_tiles = scal[3];
  QLT_CHECK_ARG(n_tiles < (int64_t(1) << 31), "qlt_plan_build: more than 2^31 output tiles");

  // Tiles, grouped by class and ordered by decreasing work inside a class
  // so the longest start first.
  int64_t kmax = 0;
  for (const Seg &s : plan->h_segs) {
    kmax = std::max(kmax, s.k_total);
    plan->cls_tiles[s.cls] += tiles_of(s.cls, s.m, s.n);
  }
  // Per-class algorithmic work (m*k*n summed) and compulsory elements
  // (distinct A and B blocks a class reads, plus the C blocks it writes).
  {
    std::vector<uint8_t> seen_a(static_cast<size_t>(plan->n_a_blks), 0), seen_b(static_cast<size_t>(plan->n_b_blks), 0);
    for (const Seg &s : plan->h_segs) {
      plan->cls_mkn[s.cls] += static_cast<double>(s.m) * static_cast<double>(s.n) * static_cast<double>(s.k_total);
      plan->cls_elems[s.cls] += static_cast<double>(s.m) * static_cast<double>(s.n);
      const uint8_t bit = static_cast<uint8_t>(1u << s.cls);
      // bit 7: read by any MFMA class (the split GEMM runs all of them in one launch)
      constexpr uint8_t kAnyMfma = 0x80;
      const bool mfma = is_mfma_class(s.cls);
      if (mfma) plan->mfma_elems += static_cast<double>(s.m) * static_cast<double>(s.n);
      for (int32_t t = s.t0; t < s.t1; ++t) {
        const ExecTask &e = plan->h_exec[static_cast<size_t>(t)];
        if (!(seen_a[e.a_blk] & bit)) {
          seen_a[e.a_blk] |= bit;
          plan->cls_elems[s.cls] += static_cast<double>(s.m) * static_cast<double>(e.k);
        }
        if (!(seen_b[e.b_blk] & bit)) {
          seen_b[e.b_blk] |= bit;
          plan->cls_elems[s.cls] += static_cast<double>(e.k) * static_cast<double>(s.n);
        }
        if (mfma && !(seen_a[e.a_blk] & kAnyMfma)) {
          seen_a[e.a_blk] |= kAnyMfma;
          plan->mfma_elems += static_cast<double>(s.m) * static_cast<double>(e.k);
        }
        if (mfma && !(seen_b[e.b_blk] & kAnyMfma)) {
          seen_b[e.b_blk] |= kAnyMfma;
          plan->mfma_elems += static_cast<double>(e.k) * static_cast<double>(s.n);
        }
      }
    }
  }
  {
    int64_t total = 0;
    for (int c = 0; c < kNumTileClasses; ++c) total += plan->cls_tiles[c];
    QLT_CHECK_ARG(total == n_tiles, "qlt_plan_build: tile count mismatch");
  }
  Tile *d_tiles_raw;
  uint64_t *d_tk0, *d_tk1;
  int32_t *d_tv0, *d_tv1;
  QLT_TRY(pool.alloc(&d_tiles_raw, n_tiles));
  QLT_TRY(pool.alloc(&d_tk0, n_tiles));
  QLT_TRY(pool.alloc(&d_tk1, n_tiles));
  QLT_TRY(pool.alloc(&d_tv0, n_tiles));
  QLT_TRY(pool.alloc(&d_tv1, n_tiles));
  hipLaunchKernelGGL(k_tiles, dim3(blocks_for(n_c)), dim3(256), 0, st, plan->d_segs, d_tstart, d_tcnt, n_c,
                     kmax, args, d_tk0, d_tv0, d_tiles_raw);
  QLT_CHECK_ARG(kmax < (int64_t(1) << 40), "qlt_plan_build: contracted extent of an output block exceeds 2^40");
  const int tpasses = (bits_for((static_cast<uint64_t>(kNumTileClasses - 1) << 40) | static_cast<uint64_t>(kmax)) + 7) / 8;
  if (tpasses > 0)
    hipLaunchKernelGGL(k_radix_sort, dim3(1), dim3(1024), 0, st, d_tk0, d_tv0, d_tk1, d_tv1, n_tiles,
                       tpasses);
  e = hipMalloc(&plan->d_tiles, sizeof(Tile) * n_tiles);
  if (e != hipSuccess) {
    free_plan_device(plan.get());
    return hip_fail(e, "hipMalloc (tiles)", __FILE__, __LINE__);
  }
  hipLaunchKernelGGL(k_permute_tiles, dim3(blocks_for(n_tiles)), dim3(256), 0, st,
                     (tpasses & 1) ? d_tv1 : d_tv0, d_tiles_raw, n_tiles, plan->d_tiles);
  // Gather descriptors of the input blocks.
  e = hipMalloc(&plan->d_ga, sizeof(BlkGather) * na);
  if (e == hipSuccess) e = hipMalloc(&plan->d_gb, sizeof(BlkGather) * nb);
  if (e != hipSuccess) {
    free_plan_device(plan.get());
    return hip_fail(e, "hipMalloc (gather descriptors)", __FILE__, __LINE__);
  }
  hipLaunchKernelGGL(k_gather_desc, dim3(blocks_for(na)), dim3(256), 0, st, d_ad, na, args, 1, plan->d_ga);
  hipLaunchKernelGGL(k_gather_desc, dim3(blocks_for(nb)), dim3(256), 0, st, d_bd, nb, args, 0, plan->d_gb);
  QLT_HIP(hipGetLastError());
  QLT_HIP(hipStreamSynchronize(st));

  // ---- host-side summary ----
  plan->n_exec = nt;
  plan->n_segs = n_c;
  plan->n_tiles = n_tiles;
  plan->c_raw_size = c_raw;
  if (args.c_is_scalar) {
    // The reference keeps no block for a rank-0 output; its raw size is 1.
    plan->c_blk_idx.clear();
    plan->c_offsets.clear();
  } else {
    plan->c_blk_idx = h_cidx;
    plan->c_offsets = h_coff;
    plan->c_coords.resize(n_c * c_rank);
    plan->c_dims.resize(n_c * c_rank);
    for (int64_t o = 0; o < n_c; ++o)
      for (int j = 0; j < c_rank; ++j) {
        plan->c_coords[o * c_rank + j] = h_ccoords[o * QLT_MAX_RANK + j];
        plan->c_dims[o * c_rank + j] = h_cdims[o * QLT_MAX_RANK + j];
      }
  }
  std::vector<char> a_used(na, 0), b_used(nb, 0);
  std::vector<int64_t> a_pos_of(0);
  {
    // map block index -> position to mark used blocks
    std::vector<std::pair<int64_t, int64_t>> amap(na), bmap(nb);
    for (int64_t i = 0; i < na; ++i) amap[i] = {h_aidx[i], i};
    for (int64_t i = 0; i < nb; ++i) bmap[i] = {h_bidx[i], i};
    auto find = [](const std::vector<std::pair<int64_t, int64_t>> &m, int64_t key) {
      auto it = std::lower_bound(m.begin(), m.end(), std::make_pair(key, int64_t(-1)));
      return it->second;
    };
    int64_t sum = 0;
    for (const qlt_task &t : plan->tasks) {
      sum += t.m * t.k * t.n;
      const int64_t ia = find(amap, t.a_blk_idx), ib = find(bmap, t.b_blk_idx);
      if (!a_used[ia]) {
        a_used[ia] = 1;
        plan->a_unique_elems += t.m * t.k;
      }
      if (!b_used[ib]) {
        b_used[ib] = 1;
        plan->b_unique_elems += t.k * t.n;
      }
    }
    plan->sum_mkn = sum;
  }
  QLT_TRY(mark_plain_segments(plan.get()));
  plan->panels_enabled = g_row_panels.load() != 0;
  QLT_TRY(slice_long_tiles(plan.get()));
  *plan_out = plan.release();
  return QLT_OK;
}

qlt_status qlt_set_row_panels(int32_t enable) {
  QLT_CHECK_ARG(enable == 0 || enable == 1, "qlt_set_row_panels: enable must be 0 or 1");
  g_row_panels.store(enable);
  return QLT_OK;
}

int32_t qlt_get_row_panels(void) { return g_row_panels.load(); }

qlt_status qlt_plan_destroy(qlt_plan *plan) {
  delete plan;   // the destructor releases the device work list
  return QLT_OK;
}

qlt_status qlt_plan_get_info(const qlt_plan *plan, qlt_plan_info *info) {
  QLT_CHECK_ARG(plan != nullptr && info != nullptr, "qlt_plan_get_info: null argument");
  info->n_tasks = static_cast<int64_t>(plan->tasks.size());
  info->n_c_blks = static_cast<int64_t>(plan->c_blk_idx.size());
  info->c_raw_size = plan->c_raw_size;
  info->c_rank = plan->c_rank;
  info->c_is_scalar = plan->c_is_scalar;
  info->sum_mkn = plan->sum_mkn;
  info->n_tiles = plan->n_tiles;
  info->a_unique_elems = plan->a_unique_elems;
  info->b_unique_elems = plan->b_unique_elems;
  info->n_panel_tiles = plan->n_ptiles;
  return QLT_OK;
}

qlt_status qlt_plan_get_tasks(const qlt_plan *plan, qlt_task *tasks) {
  QLT_CHECK_ARG(plan != nullptr, "qlt_plan_get_tasks: null plan");
  if (plan->tasks.empty()) return QLT_OK;
  QLT_CHECK_ARG(tasks != nullptr, "qlt_plan_get_tasks: null output");
  std::memcpy(tasks, plan->tasks.data(), sizeof(qlt_task) * plan->tasks.size());
  return QLT_OK;
}

qlt_status qlt_plan_get_c_blocks(const qlt_plan *plan, int64_t *blk_idx, int64_t *coords, int64_t *dims,
                                 int64_t *offsets) {
  QLT_CHECK_ARG(plan != nullptr, "qlt_plan_get_c_blocks: null plan");
  const size_t n = plan->c_blk_idx.size();
  if (n == 0) return QLT_OK;
  if (blk_idx) std::memcpy(blk_idx, plan->c_blk_idx.data(), 8 * n);
  if (offsets) std::memcpy(offsets, plan->c_offsets.data(), 8 * n);
  if (coords) std::memcpy(coords, plan->c_coords.data(), 8 * plan->c_coords.size());
  if (dims) std::memcpy(dims, plan->c_dims.data(), 8 * plan->c_dims.size());
  return QLT_OK;
}

qlt_status qlt_plan_set_c_offsets(qlt_plan *plan, const int64_t *c_offsets) {
  QLT_REQUIRE_INIT();
  QLT_CHECK_ARG(plan != nullptr, "qlt_plan_set_c_offsets: null plan");
  const int64_t n = static_cast<int64_t>(plan->h_segs.size());
  if (n == 0) return QLT_OK;
  QLT_CHECK_ARG(c_offsets != nullptr, "qlt_plan_set_c_offsets: null offsets");
  bool same = true;
  for (int64_t o = 0; o < n && same; ++o) same = plan->h_segs[o].c_off == c_offsets[o];
  if (same) return QLT_OK;
  // The kernels of an earlier execute of this plan may still be running on a
  // non-blocking stream and read the segment records (and the split tile
  // records, which carry copies of the output offsets): wait for the whole
  // device before the records change.
  QLT_HIP(hipDeviceSynchronize());
  if (plan->c_is_scalar) {
    plan->h_segs[0].c_off = c_offsets[0];
    for (qlt_task &t : plan->tasks) t.c_data_offset = c_offsets[0];
  } else {
    std::vector<int64_t> by_idx_order(plan->c_blk_idx.size());
    for (int64_t o = 0; o < n; ++o) {
      QLT_CHECK_ARG(c_offsets[o] >= 0, "qlt_plan_set_c_offsets: negative offset");
      plan->h_segs[o].c_off = c_offsets[o];
      plan->c_offsets[o] = c_offsets[o];
    }
    for (qlt_task &t : plan->tasks) {
      auto it = std::lower_bound(plan->c_blk_idx.begin(), plan->c_blk_idx.end(), t.c_blk_idx);
      t.c_data_offset = c_offsets[it - plan->c_blk_idx.begin()];
    }
  }
  QLT_HIP(hipMemcpy(plan->d_segs, plan->h_segs.data(), sizeof(Seg) * n, hipMemcpyHostToDevice));
  // the split metadata holds the old offsets: the next split execute rewrites it
  plan->split_dtype = -1;
  plan->split_mode = -1;
  return QLT_OK;
}

}  // extern "C"
