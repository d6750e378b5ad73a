// Internal declarations shared by the HIP translation units of libqlten_amd.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "qlten_amd.h"

namespace qlt {

// ---- error reporting ------------------------------------------------------
void set_error(const std::string &msg);
qlt_status fail(qlt_status code, const std::string &msg);
qlt_status hip_fail(hipError_t e, const char *what, const char *file, int line);
bool initialized();
int current_device();
// Compute units of the library's device (read once).
int device_cu_count();

#define QLT_HIP(call)                                                   \
  do {                                                                  \
    hipError_t qlt_e_ = (call);                                         \
    if (qlt_e_ != hipSuccess)                                           \
      return ::qlt::hip_fail(qlt_e_, #call, __FILE__, __LINE__);        \
  } while (0)

#define QLT_CHECK_ARG(cond, msg)                                        \
  do {                                                                  \
    if (!(cond)) return ::qlt::fail(QLT_ERR_INVALID_ARG, (msg));        \
  } while (0)

#define QLT_TRY(expr)                                                   \
  do {                                                                  \
    qlt_status qlt_s_ = (expr);                                         \
    if (qlt_s_ != QLT_OK) return qlt_s_;                                \
  } while (0)

#define QLT_REQUIRE_INIT()                                              \
  do {                                                                  \
    if (!::qlt::initialized())                                          \
      return ::qlt::fail(QLT_ERR_NOT_INITIALIZED,                       \
                         "qlt_init() has not been called");             \
  } while (0)

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

// True when `s` is currently being captured into a graph.
bool stream_capturing(hipStream_t s);

// ---- device work-list records -------------------------------------------

// Gather descriptor of one input block: the block permuted to
// (row axes ++ col axes), each part decomposed row-major, last axis fastest.
// For A the rows are the saved axes and the cols the contracted axes; for B
// the rows are the contracted axes and the cols the saved axes.
// Each dim also carries a multiply-shift reciprocal: for 0 <= x < 2^31,
// x / dim == (x * mag) >> sh with sh = 31 + ceil(log2 dim) and
// mag = ceil(2^sh / dim) (< 2^33).
struct BlkGather {
  uint64_t rmag[QLT_MAX_RANK];
  uint64_t cmag[QLT_MAX_RANK];
  int32_t rsh[QLT_MAX_RANK];
  int32_t csh[QLT_MAX_RANK];
  int32_t rdim[QLT_MAX_RANK];
  int32_t rstr[QLT_MAX_RANK];
  int32_t cdim[QLT_MAX_RANK];
  int32_t cstr[QLT_MAX_RANK];
  // 1 when the permuted block is the block itself in memory order, i.e. the
  // element (row i, col p) sits at i * ncols + p: the axes of extent > 1,
  // taken in (row axes ++ col axes) order, appear in increasing memory order.
  int32_t plain;
  int32_t ncols;   // product of cdim (the leading dimension when plain)
  int32_t nrows;   // product of rdim
  int32_t fastcol; // 1 when the block's innermost axis of extent > 1 is a col axis
  // 1 when the col (row) axes of extent > 1 merge into one axis, so that the
  // offset of col p (row p) is p * cstep (p * rstep)
  int32_t caff, cstep;
  int32_t raff, rstep;
};

// One task in execution order (sorted by output block, generation order
// inside a block).
struct ExecTask {
  int64_t a_off;   // element offset of the A block
  int64_t b_off;   // element offset of the B block
  int32_t a_blk;   // position of the A block in A's block list
  int32_t b_blk;   // position of the B block in B's block list
  int32_t k;
  int32_t pad;
};

// One output block (segment of consecutive ExecTasks).
struct Seg {
  int64_t c_off;
  int32_t m;
  int32_t n;
  int32_t t0;      // first ExecTask
  int32_t t1;      // one past the last ExecTask
  int64_t k_total; // sum of k over the segment's tasks
  int32_t cls;     // tile class of the segment (kTileMfma, kTileRows4, kTileRows16)
  int32_t flags;   // kSegAPlain | kSegBPlain when every task's A (B) block is plain
};
constexpr int32_t kSegAPlain = 1;
constexpr int32_t kSegBPlain = 2;
constexpr int32_t kSegAColAffine = 4;   // every task's A block has affine cols (contracted axes)
constexpr int32_t kSegBRowAffine = 8;   // every task's B block has affine rows (contracted axes)

// Offset of linear index `x` (0 <= x < 2^31) inside a block, decomposed
// row-major over dims[0..n-1] (last fastest) with the given memory strides;
// the divisions use the dims' multiply-shift reciprocals.
__device__ __forceinline__ int32_t split_offset(int32_t x, const int32_t *dims, const int32_t *strides,
                                                const uint64_t *mags, const int32_t *shs, int n) {
  int32_t off = 0;
  for (int j = n - 1; j >= 0; --j) {
    const int32_t q = static_cast<int32_t>((static_cast<uint64_t>(x) * mags[j]) >> shs[j]);
    off += (x - q * dims[j]) * strides[j];
    x = q;
  }
  return off;
}

// Multiply-shift reciprocal of d (1 <= d < 2^31), exact for 0 <= x < 2^31.
__host__ __device__ inline void magic_of(uint32_t d, uint64_t &mag, int32_t &sh) {
  int32_t l = 0;
  while ((uint64_t(1) << l) < d) ++l;
  sh = 31 + l;
  mag = ((uint64_t(1) << sh) + d - 1) / d;
}

// Tile classes.  Classes 0..3: MFMA output tiles of 64x64, 32x64, 64x32 and
// 32x32 (rows x cols), chosen per output block to limit padding.
// kTileRows4 / kTileRows16: narrow outputs (n <= 4 / n <= 16) on a
// row-streaming kernel (bandwidth-bound work), 1024 / 256 output rows per
// workgroup (4 / 1 rows per thread).
constexpr int32_t kTileMfma = 0;       // 64 x 64 (first MFMA class)
constexpr int32_t kNumMfmaClasses = 4;
constexpr int32_t kTileRows4 = 4;
constexpr int32_t kTileRows16 = 5;
constexpr int32_t kNumTileClasses = 6;

__host__ __device__ constexpr int32_t mfma_tile_m(int32_t cls) { return (cls == 1 || cls == 3) ? 32 : 64; }
__host__ __device__ constexpr int32_t mfma_tile_n(int32_t cls) { return (cls == 2 || cls == 3) ? 32 : 64; }
__host__ __device__ constexpr bool is_mfma_class(int32_t cls) { return cls < kNumMfmaClasses; }
// Element types: FP64 / complex128, and FP32 / complex64 (stored in single
// precision, contracted in FP64 arithmetic).
constexpr bool valid_dtype(int32_t d) { return d == QLT_F64 || d == QLT_C128 || d == QLT_F32 || d == QLT_C64; }
constexpr bool dtype_is_complex(int32_t d) { return d == QLT_C128 || d == QLT_C64; }
constexpr bool dtype_is_single(int32_t d) { return d == QLT_F32 || d == QLT_C64; }
constexpr int32_t dtype_bytes(int32_t d) { return d == QLT_C128 ? 16 : (d == QLT_F32 ? 4 : 8); }
constexpr bool valid_mode(int32_t mode) {
  return mode == QLT_MODE_NATIVE || mode == QLT_MODE_SPLIT_BF16 || mode == QLT_MODE_SPLIT_I8;
}

// 32 along a dimension when it cuts the padded extent by at least 10%.
__host__ __device__ constexpr bool prefer_32(int64_t x) {
  return ((x + 31) / 32) * 32 * 10 <= ((x + 63) / 64) * 64 * 9;
}
// Built-in rule: 32-row tiles for every MFMA-class block (more workgroups,
// three or four resident per CU, less row padding; measured faster than
// 64-row tiles on every workload here and on dense blocks up to 2048), 32
// columns when that cuts the padded width by >= 10%, else 64.  m is kept in
// the signature for the tuning table's buckets.
__host__ __device__ constexpr int32_t tile_class_of(int64_t m, int64_t n) {
  return (void)m, n <= 4 ? kTileRows4 : (n <= 16 ? kTileRows16 : 1 + (prefer_32(n) ? 2 : 0));
}
// Tile class of an output block under a tuning table (qlt_tuning): the
// narrow-output row classes are fixed; an MFMA-class block takes the table
// entry of its size bucket (ceil(m/32), ceil(n/32), each capped at
// QLT_TUNE_BUCKETS) when that entry is >= 0, else the built-in rule.
__host__ __device__ inline int32_t tuned_tile_class(int64_t m, int64_t n,
                                                     const int8_t (&table)[QLT_TUNE_BUCKETS][QLT_TUNE_BUCKETS]) {
  const int32_t c = tile_class_of(m, n);
  if (!is_mfma_class(c)) return c;
  const int64_t bm = (m + 31) / 32 < QLT_TUNE_BUCKETS ? (m + 31) / 32 : QLT_TUNE_BUCKETS;
  const int64_t bn = (n + 31) / 32 < QLT_TUNE_BUCKETS ? (n + 31) / 32 : QLT_TUNE_BUCKETS;
  const int8_t t = table[bm - 1][bn - 1];
  return (t >= 0 && t < kNumMfmaClasses) ? t : c;
}
__host__ __device__ constexpr int32_t rows_per_tile(int32_t cls) { return cls == kTileRows4 ? 1024 : 256; }
__host__ __device__ constexpr int64_t tiles_of(int32_t cls, int64_t m, int64_t n) {
  return is_mfma_class(cls)
             ? ((m + mfma_tile_m(cls) - 1) / mfma_tile_m(cls)) * ((n + mfma_tile_n(cls) - 1) / mfma_tile_n(cls))
             : (m + rows_per_tile(cls) - 1) / rows_per_tile(cls);
}

// k extent of one pipeline step of the MFMA kernels; a task of contracted
// extent k takes ceil(k / kKChunk) steps.
constexpr int32_t kKChunk = 16;

// One output tile of the grouped kernel, or one k-slice of it.  A k-slice
// covers `n_chunks` consecutive steps of the output block's (task, k-chunk)
// sequence starting at task `t_begin`, k offset `k_begin`, and writes its
// unscaled partial sum to partial buffer `part`; unsliced tiles have part = -1
// and cover the whole sequence.
struct Tile {
  int32_t seg;
  int32_t tm;      // tile row index within the block
  int32_t tn;      // tile column index within the block
  int32_t cls;     // tile-size class
  int32_t t_begin;
  int32_t k_begin;
  int32_t n_chunks;
  int32_t part;
};

// Row panels.  The output blocks of the n <= 4 row class that share their
// coordinates on A's saved axes read the same rows of A' (each from its own
// A blocks, with the same row digits and per-block strides).  When those
// blocks together touch at most kPanelCols distinct A' columns (block,
// contracted index), they run as one panel: a workgroup loads each A element
// of its rows once into registers and forms every output column of every
// block of the panel from them, where the row kernel would re-read an A
// block once per output block that uses it.  An output column is the sum of
// its terms (A' column, B' element) in (task, k) order, the order of the row
// kernel, so both give bitwise the same values.
constexpr int kPanelCols = 4;        // distinct A' columns per panel
constexpr int kPanelRowAxes = 4;     // A saved axes
constexpr int kPanelMaxOuts = 16;    // output columns per panel
constexpr int kPanelMaxTerms = 64;   // terms per panel
constexpr int kPanelRpt = 2;         // rows per thread (256 threads)
constexpr int kPanelRows = 256 * kPanelRpt;

// One 512-row tile of a panel.  Column c's element of row i sits at
// A + cbase[c] + sum_l digit_l(i) * cstr[c][l], the digits of i taken over
// rdim (last fastest) with the multiply-shift reciprocals rmag / rsh.
struct PanelTile {
  uint64_t rmag[kPanelRowAxes];
  int32_t rdim[kPanelRowAxes];
  int32_t rsh[kPanelRowAxes];
  int64_t cbase[kPanelCols];
  int32_t cstr[kPanelCols][kPanelRowAxes];
  int32_t m;        // rows of the panel's blocks
  int32_t r0;       // first row of this tile
  int32_t n_cols;   // A' columns
  int32_t n_outs;   // output columns (PanelOut records from out0)
  int32_t n_terms;  // terms (PanelTerm records from term0)
  int32_t out0;
  int32_t term0;
  // c + 1 when columns c and c + 1 are adjacent elements of one A block (the
  // same row strides, bases one element apart; the first such c), else 0
  int32_t pair1;
};
// Output column j of output block `seg`: one term per set bit c of `mask`
// (A' column c of the panel), in ascending c, stored from term0 on (indices
// relative to the panel's first term).  The panel's columns are numbered in
// ascending (A block, contracted index) order, and an output block's tasks
// run in ascending A block order, so ascending c is the (task, k) order.
struct PanelOut {
  int32_t seg;
  int32_t j;
  int32_t term0;
  int32_t mask;
};
// One term: an A' column of the panel times the B element at `b`.
struct PanelTerm {
  int64_t b;
  int32_t col;
  int32_t pad;
};

// Output tile whose k-slices are summed by the reduction kernel: partial
// buffers part0 .. part0 + n_parts - 1, in slice order.
struct ReduceItem {
  int32_t seg;
  int32_t tm;
  int32_t tn;
  int32_t cls;
  int32_t part0;
  int32_t n_parts;
};

// Axis-split layout of a plan (same permutation for every block pair).
struct GatherLayout {
  int32_t a_nr, a_nc;  // A: saved count, contracted count
  int32_t b_nr, b_nc;  // B: contracted count, saved count
  int32_t a_kfast;     // A's innermost memory axis is its last contracted axis
  int32_t b_nfast;     // B's innermost memory axis is its last saved axis
};

// Class streams: the kernels of the tile classes of one execute call write
// disjoint output blocks, so they run on the library's class streams (default
// priority), forked from and joined back into the caller's stream.  A
// ClassStreams object records the fork point on construction, hands out the
// stream of a class (making it wait for the fork on first use) and makes the
// caller's stream wait for every used class stream on destruction.  Executes
// that fork are serialised by a process-wide mutex held for the object's life.
// Graph capture records the fork and join as dependencies.
class ClassStreams {
 public:
  explicit ClassStreams(hipStream_t st);
  ~ClassStreams();
  ClassStreams(const ClassStreams &) = delete;
  ClassStreams &operator=(const ClassStreams &) = delete;
  hipStream_t get(int32_t cls);
  bool ok() const { return ok_; }

 private:
  hipStream_t st_;
  uint32_t used_ = 0;
  bool ok_ = false;
};

// Kernel probe (qlt_kernel_probe_set): events recorded around the dominant
// kernel of an execute call.
// The library tuning in effect (qlt_tuning_set), read when a plan is built.
qlt_tuning tuning_snapshot();
bool probe_armed(hipStream_t st);
void probe_begin(hipStream_t st);
void probe_end(hipStream_t st, const char *name, double flops, double bytes);

}  // namespace qlt

// Opaque plan (definition shared between plan.hip and the GEMM units).
struct qlt_plan {
  int device = 0;
  // ---- host copy of the reference plan ----
  std::vector<qlt_task> tasks;          // generation order
  std::vector<int64_t> c_blk_idx;       // ascending
  std::vector<int64_t> c_coords;        // n_c * c_rank
  std::vector<int64_t> c_dims;          // n_c * c_rank
  std::vector<int64_t> c_offsets;       // n_c
  int64_t c_raw_size = 0;
  int32_t c_rank = 0;
  int32_t c_is_scalar = 0;
  int64_t sum_mkn = 0;
  int64_t a_unique_elems = 0;
  int64_t b_unique_elems = 0;
  int64_t max_blk_elems_a = 0;
  int64_t max_blk_elems_b = 0;
  int64_t a_raw_extent = 0;             // max over A blocks of offset + size
  int64_t b_raw_extent = 0;
  // ---- device work list ----
  qlt::GatherLayout layout{};
  int64_t n_exec = 0;                   // == tasks.size()
  int64_t n_segs = 0;
  int64_t n_tiles = 0;
  int64_t n_a_blks = 0;
  int64_t n_b_blks = 0;
  qlt::ExecTask *d_exec = nullptr;
  qlt::Seg *d_segs = nullptr;
  qlt::Tile *d_tiles = nullptr;
  qlt::BlkGather *d_ga = nullptr;
  qlt::BlkGather *d_gb = nullptr;
  std::vector<qlt::Seg> h_segs;         // host mirror (used to rebuild offsets)
  std::vector<qlt::ExecTask> h_exec;    // host mirror of the execution list
  // tile-size class table chosen at build time
  int32_t tile_m = 64;
  int32_t tile_n = 64;
  // k-slicing workgroup target per tile class in effect at build time
  // (0: the built-in target, -1: unsliced)
  int32_t slice_target[qlt::kNumTileClasses] = {0, 0, 0, 0, 0, 0};
  int64_t cls_tiles[qlt::kNumTileClasses] = {0, 0, 0, 0, 0, 0};  // tiles per class, class-major in d_tiles
  double cls_mkn[qlt::kNumTileClasses] = {0, 0, 0, 0, 0, 0};     // sum of m*k*n over the class's tasks
  double cls_elems[qlt::kNumTileClasses] = {0, 0, 0, 0, 0, 0};   // distinct A + B block elements + C elements
  double mfma_elems = 0.0;              // the same over all MFMA classes together (blocks counted once)
  // ---- k-slicing of long MFMA tiles (native path) ----
  qlt::ReduceItem *d_red = nullptr;     // class-major reduction items
  int64_t cls_red[qlt::kNumTileClasses] = {0, 0, 0, 0, 0, 0};
  // k-slice partial sums, sized for complex128; each tile class owns the
  // region starting at cls_part_off[cls] doubles (the classes run concurrently)
  double *d_part = nullptr;
  size_t part_doubles = 0;
  int64_t cls_part_off[qlt::kNumTileClasses] = {0, 0, 0, 0, 0, 0};
  // ---- row panels (output blocks of the n <= 4 row class sharing A rows) ----
  bool panels_enabled = true;           // qlt_set_row_panels in effect at build time
  int64_t n_ptiles = 0;
  qlt::PanelTile *d_ptiles = nullptr;
  qlt::PanelOut *d_pouts = nullptr;
  qlt::PanelTerm *d_pterms = nullptr;
  double panel_mkn = 0.0;               // sum of m*k*n over the panels' tasks
  double panel_elems = 0.0;             // distinct A elements + C elements of the panels
  // ---- split-precision workspace ----
  void *d_work = nullptr;
  size_t work_bytes = 0;
  int32_t split_dtype = -1;             // dtype the split metadata was written for
  int32_t split_mode = -1;              // split mode (digit format) it was written for
  int64_t split_items = 0;              // pack work items
  size_t split_off[12] = {};            // workspace sub-buffer offsets
  int64_t split_tiles = 0;              // 64 x 64 tiles of the split kernel
  int64_t split_citems = 0;             // phase-2 pack work items (row group x k-chunk)
  int64_t split_sitems = 0;             // single-read pack work items (row groups of short blocks)

  qlt_plan() = default;
  qlt_plan(const qlt_plan &) = delete;
  qlt_plan &operator=(const qlt_plan &) = delete;
  ~qlt_plan() {
    for (void *p : {static_cast<void *>(d_exec), static_cast<void *>(d_segs), static_cast<void *>(d_tiles),
                    static_cast<void *>(d_ga), static_cast<void *>(d_gb), d_work, static_cast<void *>(d_red),
                    static_cast<void *>(d_part), static_cast<void *>(d_ptiles), static_cast<void *>(d_pouts),
                    static_cast<void *>(d_pterms)})
      if (p != nullptr) (void)hipFree(p);
  }
};
