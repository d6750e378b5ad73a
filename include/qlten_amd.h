/*
 * qlten_amd.h -- C ABI of the MI355X (gfx950) block-sparse contraction library.
 *
 * The reference (QuantumLiquids/TensorToolkit v0.1.3) has no FFI layer: its
 * contraction is a header-only C++ template path whose seams are
 *   - the plan builder      BlockSparseDataTensor::DataBlkGenForTenCtrct
 *                           (include/qlten/qltensor/blk_spar_data_ten/data_blk_operations.h:411-578)
 *   - the plan executor     BlockSparseDataTensor::CtrctTwoBSDTAndAssignIn
 *                           (include/qlten/qltensor/blk_spar_data_ten/global_operations.h:895-992)
 *                           and CtrctAccordingTask (global_operations.h:994-1080)
 *   - the memory primitives QLMalloc / QLFree / QLMemcpy (include/qlten/framework/mem_ops.h:38-104)
 * This ABI sits on those whole-plan seams (never on the per-GEMM
 * hp_numeric::MatMultiply seam, which would rule out grouping and graph
 * capture).  The host C++ layer in include/qlten/ keeps the reference's
 * qlten::Contract / QLTensor<ElemT,QNT> API and calls only these functions.
 *
 * Conventions
 *   - every function returns a qlt_status; QLT_OK == 0.  On failure the
 *     message is available from qlt_last_error() (thread-local).  Nothing here
 *     ever calls exit() (the reference exits on CUDA errors, mem_ops.h:28-35).
 *   - device buffers hold raw block-sparse data exactly as the reference's
 *     pactual_raw_data_ (blk_spar_data_ten.h:481-487): blocks contiguous in
 *     ascending block index, each block row-major over its degeneracy shape;
 *     complex128 is interleaved (re, im) (framework/value_t.h:40,58).
 *   - `stream` is a hipStream_t passed as void*; NULL means the null stream.
 */
#ifndef QLTEN_AMD_H
#define QLTEN_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t qlt_status;

enum {
  QLT_OK = 0,
  QLT_ERR_INVALID_ARG = 1,
  QLT_ERR_HIP = 2,
  QLT_ERR_NOT_INITIALIZED = 3,
  QLT_ERR_OUT_OF_MEMORY = 4,
  QLT_ERR_UNSUPPORTED = 5,
  QLT_ERR_CAPTURE = 6,
  QLT_ERR_INTERNAL = 7,
  QLT_ERR_NOT_CONVERGED = 8
};

/* Element types (reference QLTEN_Double / QLTEN_Complex / QLTEN_Float /
 * QLTEN_ComplexFloat, framework/value_t.h:37-81).  The single-precision types
 * are stored as float (complex: interleaved float pairs) and contracted with
 * FP64 arithmetic, rounded to float once per output element; the split modes
 * (which emulate FP64) run them on the native kernels. */
enum { QLT_F64 = 0, QLT_C128 = 1, QLT_F32 = 2, QLT_C64 = 3 };

/* Arithmetic of the block GEMMs.
 *   QLT_MODE_NATIVE      FP64 MFMA (complex128 as four real FP64 MFMA products).
 *   QLT_MODE_SPLIT_BF16  FP64 emulated by exactly scaled multi-term BF16 slices on
 *                        the BF16 MFMA units, FP32 accumulation over k-chunks short
 *                        enough to stay exact, FP64 recombination.
 *   QLT_MODE_SPLIT_I8    FP64 emulated by exactly scaled int8 digit slices (six
 *                        base-255 digits per element) on the int8 MFMA units, which
 *                        run at twice the BF16 rate; exact int32 accumulation per
 *                        digit level, FP64 recombination.
 * Both split modes keep every output block within 1e-12 relative error of FP64;
 * neither has a counterpart in the reference, whose Contract computes in FP64. */
enum { QLT_MODE_NATIVE = 0, QLT_MODE_SPLIT_BF16 = 1, QLT_MODE_SPLIT_I8 = 2 };

#define QLT_MAX_RANK 8

typedef struct qlt_plan qlt_plan;
typedef struct qlt_graph qlt_graph;

/* Block metadata of one BlockSparseDataTensor (blk_spar_data_ten.h:53-595).
 * Blocks must be listed in ascending block index (the std::map order of
 * blk_idx_data_blk_map_); block index = sum(coords[i] * stride[i]) with
 * row-major strides over blk_shape (utils_inl.h:125-148). */
typedef struct {
  int32_t rank;
  int64_t n_blks;
  const int64_t *blk_shape;   /* [rank]           number of QN sectors of each index     */
  const int64_t *blk_coords;  /* [n_blks * rank]  sector coordinates of each block       */
  const int64_t *blk_dims;    /* [n_blks * rank]  degeneracy shape of each block         */
  const int64_t *blk_offsets; /* [n_blks]         element offset of each block (raw data)*/
} qlt_bsdt_desc;

/* One contraction task; mirrors RawDataCtrctTask
 * (include/qlten/qltensor/blk_spar_data_ten/raw_data_operation_tasks.h:196-248). */
typedef struct {
  int64_t a_blk_idx;
  int64_t a_data_offset;
  int64_t b_blk_idx;
  int64_t b_data_offset;
  int64_t c_blk_idx;
  int64_t c_data_offset;
  int64_t m;
  int64_t k;
  int64_t n;
  double f_ex_sign;
  double beta;
} qlt_task;

typedef struct {
  int64_t n_tasks;       /* number of RawDataCtrctTask records                     */
  int64_t n_c_blks;      /* number of output blocks (0 for a scalar output)        */
  int64_t c_raw_size;    /* output raw data size in elements                      */
  int32_t c_rank;        /* output rank                                            */
  int32_t c_is_scalar;   /* 1 if every axis was contracted                         */
  int64_t sum_mkn;       /* sum over tasks of m*k*n (flops = 2x real, 8x complex;
                            tensor_op_cost.h:384-406)                             */
  int64_t n_tiles;       /* output tiles of the grouped kernels (row panels aside)  */
  int64_t a_unique_elems;/* elements of A blocks touched by at least one task      */
  int64_t b_unique_elems;/* elements of B blocks touched by at least one task      */
  int64_t n_panel_tiles; /* 512-row tiles of the row panels (qlt_set_row_panels)   */
} qlt_plan_info;

/* ---- runtime ---------------------------------------------------------- */

const char *qlt_last_error(void);
qlt_status qlt_init(int32_t device);
qlt_status qlt_finalize(void);
qlt_status qlt_device_sync(void);
qlt_status qlt_stream_create(void **stream);
qlt_status qlt_stream_destroy(void *stream);
qlt_status qlt_stream_sync(void *stream);

/* Device memory behind QLMalloc / QLFree / QLMemcpy (mem_ops.h:38-104).
 * The copies are complete when they return (so work queued afterwards on any
 * stream sees them); they do not wait for work queued earlier on non-null
 * streams, which the caller synchronises first. */
qlt_status qlt_malloc(void **ptr, size_t bytes);
qlt_status qlt_free(void *ptr);
qlt_status qlt_memcpy_h2d(void *dst, const void *src, size_t bytes);
qlt_status qlt_memcpy_d2h(void *dst, const void *src, size_t bytes);
qlt_status qlt_memcpy_d2d(void *dst, const void *src, size_t bytes);
qlt_status qlt_memset_zero(void *dst, size_t bytes, void *stream);

/* ---- contraction plan (replaces DataBlkGenForTenCtrct, data_blk_operations.h:411-578) ---
 * Builds the reference plan on the device with integer kernels: part-hash of
 * the contracted coordinates (vec_hash.h:39-47, blk_spar_data_ten.h:959-973),
 * A-major/B-minor matching in ascending block index, output block creation
 * with beta = 0 on first touch, prefix-sum output offsets in ascending output
 * block index, c_data_offset fill.  The output index order is
 * A[a_saved...] ++ B[b_saved...] (ten_ctrct.h:372-394); the general Contract
 * passes ascending saved axes (data_blk_operations.h:268-296), the contiguous
 * path passes cyclic ones (contract_contiguous_axes.h:342-353).
 * The plan also holds the device work list of the grouped GEMM. */
qlt_status qlt_plan_build(const qlt_bsdt_desc *a, const qlt_bsdt_desc *b,
                          int32_t n_ctrct, const int32_t *a_ctrct_axes,
                          const int32_t *b_ctrct_axes,
                          int32_t n_a_saved, const int32_t *a_saved_axes,
                          int32_t n_b_saved, const int32_t *b_saved_axes,
                          void *stream, qlt_plan **plan);
qlt_status qlt_plan_destroy(qlt_plan *plan);
qlt_status qlt_plan_get_info(const qlt_plan *plan, qlt_plan_info *info);
/* Tasks in the reference's generation order (A-major, B-minor). */
qlt_status qlt_plan_get_tasks(const qlt_plan *plan, qlt_task *tasks);
/* Output block map in ascending block index: blk_idx[n], coords[n*c_rank],
 * dims[n*c_rank], offsets[n].  Any pointer may be NULL. */
qlt_status qlt_plan_get_c_blocks(const qlt_plan *plan, int64_t *blk_idx,
                                 int64_t *coords, int64_t *dims, int64_t *offsets);
/* Point the plan's output offsets at an existing output layout (the
 * accumulate path, contract_contiguous_axes.h:389-470): c_offsets[i] is the
 * element offset, in the caller's output buffer, of the plan's i-th output
 * block (ascending block index). */
qlt_status qlt_plan_set_c_offsets(qlt_plan *plan, const int64_t *c_offsets);
/* Allocate the work buffers an execute in `mode` needs, so that a later
 * execute allocates nothing (required before stream capture). */
qlt_status qlt_plan_prepare(qlt_plan *plan, int32_t dtype, int32_t mode);

/* ---- execution (replaces CtrctTwoBSDTAndAssignIn, global_operations.h:895-992) ---
 * For every output block Cb of the plan:
 *     Cb = alpha * sum_{tasks t -> Cb} A'_t (m x k) * B'_t (k x n) + beta * Cb
 * where A'_t is A's block permuted to (saved ++ contracted) and B'_t is B's
 * block permuted to (contracted ++ saved), both row-major -- the transposes of
 * global_operations.h:919-967 are folded into the kernel's gather prologue.
 * beta == 0 never reads C (C may hold garbage, as after the reference's
 * non-zeroing Allocate).  alpha/beta point to 1 (f64) or 2 (c128: re, im)
 * doubles.  a, b, c are device pointers to the raw data buffers. */
qlt_status qlt_contract_execute(const qlt_plan *plan, int32_t dtype, int32_t mode,
                                const void *a, const void *b, void *c,
                                const double *alpha, const double *beta,
                                void *stream);

/* Split-mode dispatch threshold: a plan whose flop-weighted mean contracted
 * extent k is below `min_mean_k`, or whose flops per byte of distinct operand
 * blocks are below 96, runs on the native kernel even in a split mode
 * (such contractions are bound by memory traffic, where slicing only adds
 * traffic).  Default 48; 0 forces the split kernel.  A
 * runtime knob in the spirit of the reference's SetCpuNumericsThreads
 * (framework/hp_numeric/omp_set.h:67-83); it has no reference counterpart. */
qlt_status qlt_set_split_min_mean_k(double min_mean_k);
double qlt_get_split_min_mean_k(void);

/* Row panels (default on, 1), read when a plan is built: the narrow output
 * blocks (n <= 4) that share their coordinates on A's saved axes, and
 * together read at most 4 contracted columns of A', run as one unit that
 * reads each A element once for all of them.  0 keeps every such block on
 * the per-block row kernel.  The values are bitwise the same either way
 * (same sums in the same order).  An execution-strategy knob with no
 * reference counterpart (the reference runs one BLAS call per task,
 * global_operations.h:919-978). */
qlt_status qlt_set_row_panels(int32_t enable);
int32_t qlt_get_row_panels(void);

/* ---- tile-shape and split-k tuning; no counterpart in the reference (its
 * per-task BLAS calls have no tiles) ----
 * Output blocks of the MFMA tile classes (n > 16) fall into size buckets
 * (ceil(m/32), ceil(n/32)), each capped at QLT_TUNE_BUCKETS.  A table entry
 * tile_class[bm-1][bn-1] selects the tile shape of its bucket (0: 64x64,
 * 1: 32x64, 2: 64x32, 3: 32x32; -1: the built-in rule, which takes 32 in a
 * dimension when that pads at most 9/10 as much as 64).  slice_target[c] is the
 * workgroup count that k-slicing of tile class c aims at (0: the built-in
 * 512 for the MFMA classes 0-3, 2048 for classes 4 and 5, the narrow-output
 * row classes).  A plan keeps
 * the tuning in effect when it was built. */
#define QLT_TUNE_BUCKETS 16
#define QLT_TILE_CLASSES 6
typedef struct {
  int8_t tile_class[QLT_TUNE_BUCKETS][QLT_TUNE_BUCKETS];
  int32_t slice_target[QLT_TILE_CLASSES];
} qlt_tuning;
qlt_status qlt_tuning_get(qlt_tuning *tuning);
/* NULL restores the built-in rules. */
qlt_status qlt_tuning_set(const qlt_tuning *tuning);

/* One contraction of an autotuning workload: the qlt_plan_build arguments
 * and device operands (c is overwritten). */
typedef struct {
  const qlt_bsdt_desc *a, *b;
  int32_t n_ctrct;
  const int32_t *a_ctrct_axes, *b_ctrct_axes;
  int32_t n_a_saved;
  const int32_t *a_saved_axes;
  int32_t n_b_saved;
  const int32_t *b_saved_axes;
  const void *a_data, *b_data;
  void *c_data;
} qlt_tune_item;

/* Autotuner: times the whole workload (every item executed in order on
 * `stream`, best of `reps` timed runs after a warm-up) under candidate
 * settings and installs the fastest as the library's tuning.  Coordinate
 * descent: the size buckets present, heaviest first (by algorithmic flops),
 * each try the four tile shapes; then each tile class tries the slice targets
 * 512, 1024, 2048, 4096 and unsliced.  At most `max_buckets` buckets are
 * searched (<= 0: all).  Returns the installed tuning, the workload time
 * under the tuning in effect on entry and under the result (milliseconds),
 * and the number of candidate settings timed. */
qlt_status qlt_autotune(int32_t n_items, const qlt_tune_item *items, int32_t dtype, int32_t mode,
                        int32_t reps, int32_t max_buckets, void *stream, qlt_tuning *best,
                        double *ms_before, double *ms_after, int32_t *n_trials);

/* Kernel probe, for measurement tools; no counterpart in the reference.
 * With both events non-null (hipEvent_t created on the current device),
 * every later qlt_contract_execute outside graph capture records
 * `start_event` on its stream just before the launch of its dominant kernel
 * (the tile-class kernel, or the split GEMM, with the most algorithmic
 * flops) and `stop_event` just after it.  qlt_kernel_probe_last returns that
 * kernel's name, its algorithmic flops (sum of c*m*k*n, c = 2 real or 8
 * complex) and its compulsory bytes (distinct A and B blocks it reads plus
 * the C elements it writes).  Null events disarm the probe.  Process-wide;
 * meant for one measuring thread. */
qlt_status qlt_kernel_probe_set(void *start_event, void *stop_event);
qlt_status qlt_kernel_probe_last(char *name, int32_t name_len, double *flops, double *bytes);

/* ---- graph capture of a repeated contraction chain (Lanczos / DMRG sweep) ---- */
qlt_status qlt_graph_begin(void *stream);
qlt_status qlt_graph_end(void *stream, qlt_graph **graph);
qlt_status qlt_graph_replay(qlt_graph *graph, void *stream);
qlt_status qlt_graph_destroy(qlt_graph *graph);

/* ---- raw-data elementwise primitives --------------------------------------
 * Seeded fill: double i of x (complex data: re = double 2j, im = double 2j+1)
 * is lo + (hi - lo) * u with u = (SplitMix64(seed + (first_index+i+1) * 0x9E3779B97F4A7C15) >> 11) * 2^-53.
 * (lo, hi) = (0, 1) gives the reference's Random range [0, 1); (-1, 1) gives
 * 2u - 1.  It replaces the curand / mt19937 fill of Random
 * (raw_data_operations.h:141-166), whose stream is implementation-defined. */
qlt_status qlt_fill_uniform(void *x, int64_t n_doubles, uint64_t seed, int64_t first_index,
                            double lo, double hi, void *stream);
/* y = a*x + b*y over n elements of any element type, FP64 arithmetic
 * (LinearCombine building block, ten_linear_combine.h:186). */
qlt_status qlt_axpby(int32_t dtype, int64_t n, const double *a, const void *x,
                     const double *b, void *y, void *stream);
/* out[0..1] = sum conj(x_i) * y_i (InnerProduct, basic_operations.h:203); real: out[1]=0.
 * `out` is a host pointer; the call synchronises the stream. */
qlt_status qlt_dot(int32_t dtype, int64_t n, const void *x, const void *y, double *out,
                   void *stream);
/* Real -> complex promotion (VectorRealToCplxKernel, framework/hp_numeric/blas_level1.h:825-843). */
qlt_status qlt_real_to_complex(int64_t n, const double *x, double *z, void *stream);
/* Complex conjugation in place (ConjugateKernel, utility/utils_inl.h:556-571);
 * n complex elements, complex128 or (qlt_conjugate_c64) complex64. */
qlt_status qlt_conjugate(int64_t n, double *z, void *stream);
qlt_status qlt_conjugate_c64(int64_t n, float *z, void *stream);
/* The seeded fill of qlt_fill_uniform with each value rounded to float
 * (Random of QLTEN_Float / QLTEN_ComplexFloat tensors). */
qlt_status qlt_fill_uniform_f32(void *x, int64_t n_floats, uint64_t seed, int64_t first_index,
                                double lo, double hi, void *stream);
/* Element conversion of n elements between element types: precision changes
 * (rounded to nearest) and real -> complex (imaginary part 0; ToComplex,
 * tensor_manipulation/basic_operations.h:98-130); complex -> real is
 * rejected. */
qlt_status qlt_convert(int32_t src_dtype, int32_t dst_dtype, int64_t n, const void *src, void *dst,
                       void *stream);
/* Scale selected blocks: x[off_i .. off_i+len_i) *= s  (untouched-block scaling of the
 * accumulate path, contract_contiguous_axes.h ScaleUntouchedOutputBlocks_). */
qlt_status qlt_scale_segments(int32_t dtype, int64_t n_segs, const int64_t *offsets,
                              const int64_t *lengths, const double *s, void *x, void *stream);
/* qlt_scale_segments without any device allocation; it has no counterpart in
 * the reference.  For callers that cannot allocate device memory (tight
 * sweep loops, harnesses that forbid hipMalloc): the segment list is staged in
 * a fixed 1 MiB device array of the library, one per GPU, held until the
 * process ends, instead of a per-call allocation.  At most 65536 segments
 * (QLT_ERR_INVALID_ARG beyond).  Same arguments and result as
 * qlt_scale_segments; returns once the scaling is complete.  Calls that
 * overlap on one GPU (from several host threads) share the array and race. */
qlt_status qlt_scale_segments_noalloc(int32_t dtype, int64_t n_segs, const int64_t *offsets,
                                      const int64_t *lengths, const double *s, void *x, void *stream);

/* ---- device-resident Krylov primitives over a basis arena ----
 * The vector operations of the reference's Lanczos solver on tensors --
 * InnerProduct (tensor_manipulation/basic_operations.h:146-203), LinearCombine
 * (ten_linear_combine.h:186), Normalize (qltensor/qltensor_impl.h:568-573) --
 * on raw data, with every scalar and the number of active basis vectors in
 * device memory, so that a whole Lanczos iteration (matvec chain, two passes
 * of classical Gram-Schmidt against the basis, normalisation) launches with
 * no host round trip and can be captured once as a graph and replayed.
 * The arena holds k_max vectors of n elements, vector i at V + i * ldv
 * (elements of dtype; ldv >= n); `count` points to an int32 in device memory,
 * the number m of active vectors (clamped to 0 .. k_max), 1 <= k_max <= 1024.
 * Reductions are two-pass with a fixed grid and summation order: results are
 * bitwise reproducible.  None of these calls synchronises.
 *
 * qlt_basis_dot:    c[2i], c[2i+1] = <V_i, w> = sum conj(V_i) w for i < m,
 *                   0 for m <= i < k_max (c: 2 k_max doubles, device);
 *                   work: QLT_BASIS_WORK_DOUBLES(k_max) doubles, device.
 * qlt_basis_update: w += scale * sum_{i<m} c_i V_i (c complex pairs, device).
 * qlt_basis_update_dot: qlt_basis_update, then qlt_basis_dot of the updated w
 *                   into c_out (c_out must not alias c; work as for
 *                   qlt_basis_dot), with bitwise the same results as the two
 *                   calls; for m <= 16 in one pass over the basis (the second
 *                   Gram-Schmidt pass's products formed while the first pass
 *                   updates w).
 * qlt_lanczos_extend: with j = m - 1: alphas[j] = Re c[2j], betas[j] = |w|,
 *                   V_m = w / |w| (zero when |w| = 0) if m < k_max, then
 *                   m = m + 1 (when m < k_max).  work as for qlt_basis_dot.
 * qlt_basis_select: x = V_{*index + offset} (no-op outside 0 .. k_max-1). */
#define QLT_BASIS_WORK_DOUBLES(k_max) (2 * 512 * (k_max) + 1)
qlt_status qlt_basis_dot(int32_t dtype, int64_t n, int32_t k_max, const int32_t *count, const void *V, int64_t ldv,
                         const void *w, double *c, double *work, void *stream);
qlt_status qlt_basis_update(int32_t dtype, int64_t n, int32_t k_max, const int32_t *count, double scale,
                            const double *c, const void *V, int64_t ldv, void *w, void *stream);
qlt_status qlt_basis_update_dot(int32_t dtype, int64_t n, int32_t k_max, const int32_t *count, double scale,
                                const double *c, const void *V, int64_t ldv, void *w, double *c_out, double *work,
                                void *stream);
qlt_status qlt_lanczos_extend(int32_t dtype, int64_t n, int32_t k_max, int32_t *count, const double *c,
                              const void *w, void *V, int64_t ldv, double *alphas, double *betas, double *work,
                              void *stream);
qlt_status qlt_basis_select(int32_t dtype, int64_t n, int32_t k_max, const int32_t *index, int32_t offset,
                            const void *V, int64_t ldv, void *x, void *stream);

/* Batched 2-D copy: rectangle i (rows[i] x cols[i] elements of dtype) from
 * src + src_offsets[i] with row stride src_ld[i] to dst + dst_offsets[i] with
 * row stride dst_ld[i]; host arrays, counted in elements.  Assembles block
 * matrices from blocks and scatters decomposition results back, as
 * RawDataGenDenseDataBlkMat_ / SubMatMemCpy (raw_data_operations.h:550-575)
 * and DataBlkCopySVDUdata / DataBlkCopySVDVtData / DataBlkCopyQRQdata
 * (data_blk_operations.h:953-1078) do per block.  Synchronous. */
qlt_status qlt_copy_blocks_2d(int32_t dtype, int64_t n, const int64_t *rows, const int64_t *cols,
                              const void *src, const int64_t *src_offsets, const int64_t *src_ld,
                              void *dst, const int64_t *dst_offsets, const int64_t *dst_ld, void *stream);

/* ---- batched FP64 / complex128 block decompositions ----
 * Replace MatSVD (framework/hp_numeric/lapack.h:41-82 gesvd; GPU :811-1011)
 * and MatQR (lapack.h:332-487 geqrf + orgqr/ungqr), which the reference calls
 * once per block matrix from DataBlkDecompSVD / DataBlkDecompQR
 * (qltensor/blk_spar_data_ten/data_blk_operations.h:630-663,1028-1046).
 * n_mats dense row-major matrices of `dtype`; matrix i is rows[i] x cols[i]
 * at a + a_offsets[i].  Offsets are host arrays counted in elements of dtype
 * (complex elements for QLT_C128); s and its offsets are in doubles.  Both
 * calls are synchronous: they return after the outputs are written.  They
 * allocate device workspace and cannot be graph-captured.
 *
 * SVD: A = U diag(S) Vt with r = min(rows, cols), U rows x r, S r (descending),
 * Vt r x cols, all row-major, by one-sided block Jacobi.  For sigma_k == 0 the
 * k-th column of U (rows >= cols) or row of Vt (rows < cols) is zero rather
 * than an orthonormal completion.  `sweeps` (nullable) receives the Jacobi
 * sweeps run; QLT_ERR_NOT_CONVERGED if 40 sweeps did not converge (outputs
 * are still written).  min(rows, cols) <= 4096 per matrix.
 *
 * QR: A = Q R with k = min(rows, cols), Q rows x k with orthonormal columns,
 * R k x cols upper triangular, by Householder reflections with LAPACK's
 * reflector convention (the diagonal of R has geqrf's signs). */
qlt_status qlt_block_svd(int32_t dtype, int64_t n_mats, const int64_t *rows, const int64_t *cols,
                         const void *a, const int64_t *a_offsets,
                         void *u, const int64_t *u_offsets,
                         double *s, const int64_t *s_offsets,
                         void *vt, const int64_t *vt_offsets,
                         int32_t *sweeps, void *stream);
qlt_status qlt_block_qr(int32_t dtype, int64_t n_mats, const int64_t *rows, const int64_t *cols,
                        const void *a, const int64_t *a_offsets,
                        void *q, const int64_t *q_offsets,
                        void *r, const int64_t *r_offsets, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* QLTEN_AMD_H */
