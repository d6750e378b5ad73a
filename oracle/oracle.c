/*
 * CPU oracle -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference's block-sparse contraction
 * (QuantumLiquids/TensorToolkit v0.1.3).  Only the tests, the smoke check and
 * the CPU-baseline leg of the benchmark load this library, always as the
 * checker or the timed baseline, never as part of the product path.
 *
 * What each function follows (paths relative to the reference's include/qlten/):
 *   ora_part_hash        framework/vec_hash.h:39-47 (VecPartHasher) as called by
 *                        qltensor/blk_spar_data_ten/blk_spar_data_ten.h:959-973
 *   ora_plan_build       qltensor/blk_spar_data_ten/data_blk_operations.h:411-578
 *                        (DataBlkGenForTenCtrct), offsets as DataBlksOffsetRefresh
 *                        (data_blk_operations.h:137-145), block index as
 *                        BlkCoorsToBlkIdx (blk_spar_data_ten.h:380-382,
 *                        utility/utils_inl.h:125-148)
 *   ora_execute          qltensor/blk_spar_data_ten/global_operations.h:895-992
 *                        (CtrctTwoBSDTAndAssignIn): sort by (c_blk_idx, beta)
 *                        (raw_data_operation_tasks.h:250-269), per-block
 *                        transposes to (saved ++ contracted) for A and
 *                        (contracted ++ saved) for B (ten_trans.h:130-186, the
 *                        plain transpose), one GEMM per task C = f*A'B' + beta*C,
 *                        row-major NN with lda=k, ldb=n, ldc=n
 *                        (framework/hp_numeric/blas_level3.h:35-108).
 *                        With sorted == 0 the same function runs the tasks in
 *                        generation order without the sort, as
 *                        CtrctAccordingTask (global_operations.h:994-1080).
 * Pinning (each checked in tests/test_oracle.py):
 *   - plan: brute-force enumeration of all block pairs by coordinates on tiny
 *     inputs; the reference's cost known-answer test
 *     (tests/test_tensor_manipulation/test_tensor_op_cost.cc:19-42: one task,
 *     output block {2,4}, 8 elements, 48 flops); closed-form task/block counts
 *     of the rank-3 U(1) case;
 *   - values: dense equivalence (densify, numpy tensordot), the reference's own
 *     method (test_ten_ctrct.cc:226-494), and the closed forms t.t^dagger = sum|t|^2
 *     (test_ten_ctrct.cc:187-206) and full contraction = dot product (:272-306).
 * The GEMM is either a plain triple loop or a caller-supplied cblas_dgemm /
 * cblas_zgemm (OpenBLAS), selected with ora_set_blas.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORA_MAX_RANK 8

static const uint64_t P1 = 11400714785074694791ull;
static const uint64_t P2 = 14029467366897019727ull;
static const uint64_t P5 = 2870177450012600261ull;

typedef struct {
  int32_t rank;
  int64_t n_blks;
  const int64_t *blk_shape;
  const int64_t *blk_coords;
  const int64_t *blk_dims;
  const int64_t *blk_offsets;
} ora_bsdt;

typedef struct {
  int64_t a_blk_idx, a_data_offset, b_blk_idx, b_data_offset, c_blk_idx, c_data_offset, m, k, n;
  double f_ex_sign, beta;
  /* positions of the A and B blocks in their descriptors (oracle-internal) */
  int64_t a_pos, b_pos;
} ora_task;

typedef struct {
  int64_t n_tasks;
  ora_task *tasks;          /* generation order */
  int64_t n_c;
  int64_t *c_blk_idx;       /* ascending */
  int64_t *c_coords;        /* n_c * c_rank */
  int64_t *c_dims;          /* n_c * c_rank */
  int64_t *c_offsets;
  int64_t c_raw_size;
  int32_t c_rank;
  int32_t c_is_scalar;
  /* copies needed by the executor */
  int32_t a_rank, b_rank, n_ctrct, n_a_saved, n_b_saved;
  int32_t a_ctrct[ORA_MAX_RANK], b_ctrct[ORA_MAX_RANK], a_saved[ORA_MAX_RANK], b_saved[ORA_MAX_RANK];
  int64_t *a_dims;          /* copy of A block dims (n_a * a_rank) */
  int64_t *b_dims;
} ora_plan;

typedef void (*dgemm_fn)(int, int, int, int, int, int, double, const double *, int, const double *, int,
                         double, double *, int);
typedef void (*zgemm_fn)(int, int, int, int, int, int, const void *, const void *, int, const void *, int,
                         const void *, void *, int);

static dgemm_fn g_dgemm = NULL;
static zgemm_fn g_zgemm = NULL;
static int g_threads = 1;

void ora_set_blas(void *dgemm, void *zgemm) {
  g_dgemm = (dgemm_fn)dgemm;
  g_zgemm = (zgemm_fn)zgemm;
}

void ora_set_threads(int n) { g_threads = n > 0 ? n : 1; }

/* VecPartHasher over the selected coordinates (vec_hash.h:39-47). */
uint64_t ora_part_hash(const int64_t *coords, const int32_t *axes, int32_t n_axes) {
  uint64_t h = P5;
  for (int32_t i = 0; i < n_axes; ++i) {
    h += (uint64_t)coords[axes[i]] * P2;
    h = (h << 31) | (h >> 33);
    h *= P1;
  }
  h += (uint64_t)n_axes ^ P5;
  return h;
}

/* ---- small ordered set of output block indexes (sorted array + open hash) ---- */
typedef struct {
  int64_t *keys;
  int64_t *first_task; /* generation position of the task that created the block */
  int64_t cap, n;
  int64_t *slots;      /* open addressing: index into keys, -1 empty */
  int64_t nslots;
} cset;

static void cset_init(cset *s, int64_t cap) {
  s->cap = cap > 16 ? cap : 16;
  s->n = 0;
  s->keys = (int64_t *)malloc(sizeof(int64_t) * s->cap);
  s->first_task = (int64_t *)malloc(sizeof(int64_t) * s->cap);
  s->nslots = 1;
  while (s->nslots < 2 * s->cap) s->nslots <<= 1;
  s->slots = (int64_t *)malloc(sizeof(int64_t) * s->nslots);
  for (int64_t i = 0; i < s->nslots; ++i) s->slots[i] = -1;
}

static void cset_free(cset *s) {
  free(s->keys);
  free(s->first_task);
  free(s->slots);
}

static uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

static void cset_rehash(cset *s) {
  for (int64_t i = 0; i < s->nslots; ++i) s->slots[i] = -1;
  for (int64_t j = 0; j < s->n; ++j) {
    uint64_t h = mix64((uint64_t)s->keys[j]) & (uint64_t)(s->nslots - 1);
    while (s->slots[h] >= 0) h = (h + 1) & (uint64_t)(s->nslots - 1);
    s->slots[h] = j;
  }
}

/* returns 1 if inserted (new), 0 if already present */
static int cset_insert(cset *s, int64_t key, int64_t task) {
  uint64_t h = mix64((uint64_t)key) & (uint64_t)(s->nslots - 1);
  while (s->slots[h] >= 0) {
    if (s->keys[s->slots[h]] == key) return 0;
    h = (h + 1) & (uint64_t)(s->nslots - 1);
  }
  if (s->n == s->cap) {
    s->cap *= 2;
    s->keys = (int64_t *)realloc(s->keys, sizeof(int64_t) * s->cap);
    s->first_task = (int64_t *)realloc(s->first_task, sizeof(int64_t) * s->cap);
    free(s->slots);
    s->nslots = 1;
    while (s->nslots < 2 * s->cap) s->nslots <<= 1;
    s->slots = (int64_t *)malloc(sizeof(int64_t) * s->nslots);
    s->keys[s->n] = key;
    s->first_task[s->n] = task;
    s->n++;
    cset_rehash(s);
    return 1;
  }
  s->slots[h] = s->n;
  s->keys[s->n] = key;
  s->first_task[s->n] = task;
  s->n++;
  return 1;
}

static int cmp_i64(const void *x, const void *y) {
  const int64_t a = *(const int64_t *)x, b = *(const int64_t *)y;
  return (a > b) - (a < b);
}

static int64_t bsearch_i64(const int64_t *arr, int64_t n, int64_t key) {
  int64_t lo = 0, hi = n - 1;
  while (lo <= hi) {
    int64_t mid = (lo + hi) / 2;
    if (arr[mid] == key) return mid;
    if (arr[mid] < key) lo = mid + 1; else hi = mid - 1;
  }
  return -1;
}

static int64_t prod_sel(const int64_t *v, const int32_t *sel, int32_t n) {
  int64_t p = 1;
  for (int32_t i = 0; i < n; ++i) p *= v[sel[i]];
  return p;
}

void ora_plan_free(ora_plan *p) {
  if (!p) return;
  free(p->tasks);
  free(p->c_blk_idx);
  free(p->c_coords);
  free(p->c_dims);
  free(p->c_offsets);
  free(p->a_dims);
  free(p->b_dims);
  free(p);
}

/* DataBlkGenForTenCtrct (data_blk_operations.h:411-578). Returns NULL on bad input. */
ora_plan *ora_plan_build(const ora_bsdt *a, const ora_bsdt *b, int32_t n_ctrct, const int32_t *a_ctrct,
                         const int32_t *b_ctrct, int32_t n_a_saved, const int32_t *a_saved,
                         int32_t n_b_saved, const int32_t *b_saved) {
  if (a->rank > ORA_MAX_RANK || b->rank > ORA_MAX_RANK) return NULL;
  if (n_ctrct + n_a_saved != a->rank || n_ctrct + n_b_saved != b->rank) return NULL;
  ora_plan *p = (ora_plan *)calloc(1, sizeof(ora_plan));
  p->a_rank = a->rank;
  p->b_rank = b->rank;
  p->n_ctrct = n_ctrct;
  p->n_a_saved = n_a_saved;
  p->n_b_saved = n_b_saved;
  for (int i = 0; i < n_ctrct; ++i) {
    p->a_ctrct[i] = a_ctrct[i];
    p->b_ctrct[i] = b_ctrct[i];
  }
  for (int i = 0; i < n_a_saved; ++i) p->a_saved[i] = a_saved[i];
  for (int i = 0; i < n_b_saved; ++i) p->b_saved[i] = b_saved[i];
  p->c_rank = n_a_saved + n_b_saved;
  p->c_is_scalar = (p->c_rank == 0);
  p->a_dims = (int64_t *)malloc(sizeof(int64_t) * (a->n_blks * a->rank + 1));
  p->b_dims = (int64_t *)malloc(sizeof(int64_t) * (b->n_blks * b->rank + 1));
  if (a->n_blks) memcpy(p->a_dims, a->blk_dims, sizeof(int64_t) * a->n_blks * a->rank);
  if (b->n_blks) memcpy(p->b_dims, b->blk_dims, sizeof(int64_t) * b->n_blks * b->rank);

  /* block indexes of A and B (ascending by precondition) */
  int64_t *a_idx = (int64_t *)malloc(sizeof(int64_t) * (a->n_blks + 1));
  int64_t *b_idx = (int64_t *)malloc(sizeof(int64_t) * (b->n_blks + 1));
  for (int64_t i = 0; i < a->n_blks; ++i) {
    int64_t v = 0;
    for (int j = 0; j < a->rank; ++j) v = v * a->blk_shape[j] + a->blk_coords[i * a->rank + j];
    a_idx[i] = v;
  }
  for (int64_t i = 0; i < b->n_blks; ++i) {
    int64_t v = 0;
    for (int j = 0; j < b->rank; ++j) v = v * b->blk_shape[j] + b->blk_coords[i * b->rank + j];
    b_idx[i] = v;
  }
  /* GenBlkIdxQNBlkCoorPartHashMap: (blk_idx, part hash) in ascending blk_idx */
  uint64_t *ha = (uint64_t *)malloc(sizeof(uint64_t) * (a->n_blks + 1));
  uint64_t *hb = (uint64_t *)malloc(sizeof(uint64_t) * (b->n_blks + 1));
  for (int64_t i = 0; i < a->n_blks; ++i) ha[i] = ora_part_hash(a->blk_coords + i * a->rank, a_ctrct, n_ctrct);
  for (int64_t i = 0; i < b->n_blks; ++i) hb[i] = ora_part_hash(b->blk_coords + i * b->rank, b_ctrct, n_ctrct);

  /* C strides (row-major over C's sector counts) */
  int64_t c_stride[ORA_MAX_RANK];
  {
    int64_t cshape[ORA_MAX_RANK];
    int ci = 0;
    for (int j = 0; j < n_a_saved; ++j) cshape[ci++] = a->n_blks ? a->blk_shape[a_saved[j]] : 1;
    for (int j = 0; j < n_b_saved; ++j) cshape[ci++] = b->n_blks ? b->blk_shape[b_saved[j]] : 1;
    int64_t s = 1;
    for (int j = p->c_rank - 1; j >= 0; --j) {
      c_stride[j] = s;
      s *= cshape[j];
    }
  }

  int64_t cap = 64, nt = 0;
  ora_task *tasks = (ora_task *)malloc(sizeof(ora_task) * cap);
  cset cs;
  cset_init(&cs, 64);
  if (p->c_is_scalar) {
    /* data_blk_operations.h:451-491: m = n = 1, break after the first match */
    for (int64_t i = 0; i < a->n_blks; ++i) {
      for (int64_t j = 0; j < b->n_blks; ++j) {
        if (ha[i] == hb[j]) {
          if (nt == cap) {
            cap *= 2;
            tasks = (ora_task *)realloc(tasks, sizeof(ora_task) * cap);
          }
          ora_task t;
          t.a_blk_idx = a_idx[i];
          t.a_data_offset = a->blk_offsets[i];
          t.b_blk_idx = b_idx[j];
          t.b_data_offset = b->blk_offsets[j];
          t.c_blk_idx = 0;
          t.c_data_offset = 0;
          t.m = 1;
          t.n = 1;
          t.k = prod_sel(a->blk_dims + i * a->rank, a_ctrct, n_ctrct);
          t.f_ex_sign = 1.0;
          t.beta = 1.0;
          t.a_pos = i;
          t.b_pos = j;
          tasks[nt++] = t;
          break;
        }
      }
    }
    if (nt > 0) {
      tasks[0].beta = 0.0;
      p->c_raw_size = 1;
    }
    p->n_c = 0;
  } else {
    for (int64_t i = 0; i < a->n_blks; ++i) {
      const int64_t *ac = a->blk_coords + i * a->rank;
      const int64_t *ad = a->blk_dims + i * a->rank;
      int64_t m = 0, k = 0;
      for (int64_t j = 0; j < b->n_blks; ++j) {
        if (ha[i] != hb[j]) continue;
        const int64_t *bc = b->blk_coords + j * b->rank;
        const int64_t *bd = b->blk_dims + j * b->rank;
        if (m == 0) {
          m = prod_sel(ad, a_saved, n_a_saved);
          k = prod_sel(ad, a_ctrct, n_ctrct);
        }
        const int64_t n = prod_sel(bd, b_saved, n_b_saved);
        int64_t c_idx = 0;
        int ci = 0;
        for (int q = 0; q < n_a_saved; ++q, ++ci) c_idx += ac[a_saved[q]] * c_stride[ci];
        for (int q = 0; q < n_b_saved; ++q, ++ci) c_idx += bc[b_saved[q]] * c_stride[ci];
        const int inserted = cset_insert(&cs, c_idx, nt);
        if (nt == cap) {
          cap *= 2;
          tasks = (ora_task *)realloc(tasks, sizeof(ora_task) * cap);
        }
        ora_task t;
        t.a_blk_idx = a_idx[i];
        t.a_data_offset = a->blk_offsets[i];
        t.b_blk_idx = b_idx[j];
        t.b_data_offset = b->blk_offsets[j];
        t.c_blk_idx = c_idx;
        t.c_data_offset = -1;
        t.m = m;
        t.k = k;
        t.n = n;
        t.f_ex_sign = 1.0;
        t.beta = inserted ? 0.0 : 1.0;
        t.a_pos = i;
        t.b_pos = j;
        tasks[nt++] = t;
      }
    }
    /* C blocks in ascending block index, offsets by prefix sum (DataBlksOffsetRefresh) */
    p->n_c = cs.n;
    p->c_blk_idx = (int64_t *)malloc(sizeof(int64_t) * (cs.n + 1));
    memcpy(p->c_blk_idx, cs.keys, sizeof(int64_t) * cs.n);
    qsort(p->c_blk_idx, (size_t)cs.n, sizeof(int64_t), cmp_i64);
    p->c_coords = (int64_t *)malloc(sizeof(int64_t) * (cs.n * p->c_rank + 1));
    p->c_dims = (int64_t *)malloc(sizeof(int64_t) * (cs.n * p->c_rank + 1));
    p->c_offsets = (int64_t *)malloc(sizeof(int64_t) * (cs.n + 1));
    /* block coordinates and shape from the creating task */
    for (int64_t q = 0; q < cs.n; ++q) {
      const int64_t pos = bsearch_i64(p->c_blk_idx, cs.n, cs.keys[q]);
      const ora_task *t = &tasks[cs.first_task[q]];
      const int64_t *ac = a->blk_coords + t->a_pos * a->rank, *ad = a->blk_dims + t->a_pos * a->rank;
      const int64_t *bc = b->blk_coords + t->b_pos * b->rank, *bd = b->blk_dims + t->b_pos * b->rank;
      int ci = 0;
      for (int u = 0; u < n_a_saved; ++u, ++ci) {
        p->c_coords[pos * p->c_rank + ci] = ac[a_saved[u]];
        p->c_dims[pos * p->c_rank + ci] = ad[a_saved[u]];
      }
      for (int u = 0; u < n_b_saved; ++u, ++ci) {
        p->c_coords[pos * p->c_rank + ci] = bc[b_saved[u]];
        p->c_dims[pos * p->c_rank + ci] = bd[b_saved[u]];
      }
    }
    int64_t run = 0;
    for (int64_t q = 0; q < cs.n; ++q) {
      p->c_offsets[q] = run;
      int64_t sz = 1;
      for (int u = 0; u < p->c_rank; ++u) sz *= p->c_dims[q * p->c_rank + u];
      run += sz;
    }
    p->c_raw_size = run;
    for (int64_t t = 0; t < nt; ++t)
      tasks[t].c_data_offset = p->c_offsets[bsearch_i64(p->c_blk_idx, cs.n, tasks[t].c_blk_idx)];
  }
  cset_free(&cs);
  p->n_tasks = nt;
  p->tasks = tasks;
  free(a_idx);
  free(b_idx);
  free(ha);
  free(hb);
  return p;
}

int64_t ora_plan_n_tasks(const ora_plan *p) { return p->n_tasks; }
int64_t ora_plan_n_c(const ora_plan *p) { return p->n_c; }
int64_t ora_plan_c_raw_size(const ora_plan *p) { return p->c_raw_size; }

/* tasks as 11 doubles-worth of int64 records: a_idx,a_off,b_idx,b_off,c_idx,c_off,m,k,n then f,beta */
void ora_plan_get_tasks(const ora_plan *p, int64_t *ints /* n*9 */, double *fb /* n*2 */) {
  for (int64_t t = 0; t < p->n_tasks; ++t) {
    const ora_task *x = &p->tasks[t];
    int64_t *o = ints + 9 * t;
    o[0] = x->a_blk_idx; o[1] = x->a_data_offset; o[2] = x->b_blk_idx; o[3] = x->b_data_offset;
    o[4] = x->c_blk_idx; o[5] = x->c_data_offset; o[6] = x->m; o[7] = x->k; o[8] = x->n;
    fb[2 * t] = x->f_ex_sign;
    fb[2 * t + 1] = x->beta;
  }
}

void ora_plan_get_c_blocks(const ora_plan *p, int64_t *idx, int64_t *coords, int64_t *dims, int64_t *offs) {
  if (p->n_c == 0) return;
  memcpy(idx, p->c_blk_idx, sizeof(int64_t) * p->n_c);
  memcpy(coords, p->c_coords, sizeof(int64_t) * p->n_c * p->c_rank);
  memcpy(dims, p->c_dims, sizeof(int64_t) * p->n_c * p->c_rank);
  memcpy(offs, p->c_offsets, sizeof(int64_t) * p->n_c);
}

/* ---- plain transpose (ten_trans.h:130-186): out[perm-ordered] = in, alpha = 1 ----
 * in has shape dims[0..r-1] row-major; out has shape dims[perm[0]], ..., dims[perm[r-1]]. */
static void transpose_block(const double *in, double *out, const int64_t *dims, const int32_t *perm, int r,
                            int cplx) {
  int64_t in_stride[ORA_MAX_RANK], out_dims[ORA_MAX_RANK], total = 1;
  int64_t s = 1;
  for (int j = r - 1; j >= 0; --j) {
    in_stride[j] = s;
    s *= dims[j];
  }
  total = s;
  for (int j = 0; j < r; ++j) out_dims[j] = dims[perm[j]];
  /* strides of the input seen in output order */
  int64_t gstr[ORA_MAX_RANK];
  for (int j = 0; j < r; ++j) gstr[j] = in_stride[perm[j]];
  const int64_t inner = out_dims[r - 1];
  const int64_t outer = total / inner;
#pragma omp parallel for num_threads(g_threads) schedule(static) if (total > 65536)
  for (int64_t o = 0; o < outer; ++o) {
    int64_t rem = o, src = 0;
    for (int j = r - 2; j >= 0; --j) {
      const int64_t c = rem % out_dims[j];
      rem /= out_dims[j];
      src += c * gstr[j];
    }
    const int64_t g = gstr[r - 1];
    if (cplx) {
      for (int64_t i = 0; i < inner; ++i) {
        out[2 * (o * inner + i)] = in[2 * (src + i * g)];
        out[2 * (o * inner + i) + 1] = in[2 * (src + i * g) + 1];
      }
    } else {
      for (int64_t i = 0; i < inner; ++i) out[o * inner + i] = in[src + i * g];
    }
  }
}

static void gemm_naive(int cplx, int64_t m, int64_t k, int64_t n, double f, const double *A, const double *B,
                       double beta, double *C) {
  if (!cplx) {
    for (int64_t i = 0; i < m; ++i)
      for (int64_t j = 0; j < n; ++j) {
        double s = 0.0;
        for (int64_t p = 0; p < k; ++p) s += A[i * k + p] * B[p * n + j];
        C[i * n + j] = (beta == 0.0) ? f * s : f * s + beta * C[i * n + j];
      }
  } else {
    for (int64_t i = 0; i < m; ++i)
      for (int64_t j = 0; j < n; ++j) {
        double sr = 0.0, si = 0.0;
        for (int64_t p = 0; p < k; ++p) {
          const double ar = A[2 * (i * k + p)], ai = A[2 * (i * k + p) + 1];
          const double br = B[2 * (p * n + j)], bi = B[2 * (p * n + j) + 1];
          sr += ar * br - ai * bi;
          si += ar * bi + ai * br;
        }
        double *c = C + 2 * (i * n + j);
        if (beta == 0.0) {
          c[0] = f * sr;
          c[1] = f * si;
        } else {
          c[0] = f * sr + beta * c[0];
          c[1] = f * si + beta * c[1];
        }
      }
  }
}

static void gemm_task(int cplx, int64_t m, int64_t k, int64_t n, double f, const double *A, const double *B,
                      double beta, double *C) {
  /* hp_numeric::MatMultiply, blas_level3.h:35-108: row-major NN, lda=k, ldb=n, ldc=n */
  if (!cplx && g_dgemm) {
    g_dgemm(101, 111, 111, (int)m, (int)n, (int)k, f, A, (int)k, B, (int)n, beta, C, (int)n);
  } else if (cplx && g_zgemm) {
    const double alpha[2] = {f, 0.0}, bet[2] = {beta, 0.0};
    g_zgemm(101, 111, 111, (int)m, (int)n, (int)k, alpha, A, (int)k, B, (int)n, bet, C, (int)n);
  } else {
    gemm_naive(cplx, m, k, n, f, A, B, beta, C);
  }
}

typedef struct {
  int64_t c_blk_idx;
  double beta;
  int64_t gen;
} sort_rec;

static int cmp_sort_rec(const void *x, const void *y) {
  const sort_rec *a = (const sort_rec *)x, *b = (const sort_rec *)y;
  if (a->c_blk_idx != b->c_blk_idx) return (a->c_blk_idx > b->c_blk_idx) - (a->c_blk_idx < b->c_blk_idx);
  if (a->beta != b->beta) return (a->beta > b->beta) - (a->beta < b->beta);
  return (a->gen > b->gen) - (a->gen < b->gen);
}

/* CtrctTwoBSDTAndAssignIn (global_operations.h:895-992).  `sorted` = 1 sorts
 * by (c_blk_idx, beta) first (ties in generation order, one outcome of the
 * reference's unstable std::sort); 0 keeps generation order as
 * CtrctAccordingTask does (global_operations.h:994-1080).  C must hold
 * c_raw_size elements; it is written, never read, where beta == 0. */
int ora_execute(const ora_plan *p, int cplx, const double *A, const double *B, double *C, int sorted) {
  const int64_t nt = p->n_tasks;
  if (nt == 0) return 0;
  const int esz = cplx ? 2 : 1;
  int64_t *order = (int64_t *)malloc(sizeof(int64_t) * nt);
  if (sorted) {
    sort_rec *r = (sort_rec *)malloc(sizeof(sort_rec) * nt);
    for (int64_t t = 0; t < nt; ++t) {
      r[t].c_blk_idx = p->tasks[t].c_blk_idx;
      r[t].beta = p->tasks[t].beta;
      r[t].gen = t;
    }
    qsort(r, (size_t)nt, sizeof(sort_rec), cmp_sort_rec);
    for (int64_t t = 0; t < nt; ++t) order[t] = r[t].gen;
    free(r);
  } else {
    for (int64_t t = 0; t < nt; ++t) order[t] = t;
  }
  /* TenCtrctNeedTransCheck (ten_ctrct.h:396-443): A -> saved ++ ctrct, B -> ctrct ++ saved */
  int32_t a_perm[ORA_MAX_RANK], b_perm[ORA_MAX_RANK];
  int a_need = 0, b_need = 0;
  for (int j = 0; j < p->n_a_saved; ++j) a_perm[j] = p->a_saved[j];
  for (int j = 0; j < p->n_ctrct; ++j) a_perm[p->n_a_saved + j] = p->a_ctrct[j];
  for (int j = 0; j < p->n_ctrct; ++j) b_perm[j] = p->b_ctrct[j];
  for (int j = 0; j < p->n_b_saved; ++j) b_perm[p->n_ctrct + j] = p->b_saved[j];
  for (int j = 0; j < p->a_rank; ++j) a_need |= (a_perm[j] != j);
  for (int j = 0; j < p->b_rank; ++j) b_need |= (b_perm[j] != j);
  /* transposed copies, cached per block (global_operations.h:922-964) */
  double **a_cache = NULL, **b_cache = NULL;
  int64_t max_a = 0, max_b = 0;
  for (int64_t t = 0; t < nt; ++t) {
    if (p->tasks[t].a_pos + 1 > max_a) max_a = p->tasks[t].a_pos + 1;
    if (p->tasks[t].b_pos + 1 > max_b) max_b = p->tasks[t].b_pos + 1;
  }
  if (a_need) a_cache = (double **)calloc((size_t)max_a, sizeof(double *));
  if (b_need) b_cache = (double **)calloc((size_t)max_b, sizeof(double *));
  for (int64_t q = 0; q < nt; ++q) {
    const ora_task *t = &p->tasks[order[q]];
    const double *a_data = A + esz * t->a_data_offset;
    const double *b_data = B + esz * t->b_data_offset;
    if (a_need) {
      if (!a_cache[t->a_pos]) {
        a_cache[t->a_pos] = (double *)malloc(sizeof(double) * esz * t->m * t->k);
        transpose_block(a_data, a_cache[t->a_pos], p->a_dims + t->a_pos * p->a_rank, a_perm, p->a_rank, cplx);
      }
      a_data = a_cache[t->a_pos];
    }
    if (b_need) {
      if (!b_cache[t->b_pos]) {
        b_cache[t->b_pos] = (double *)malloc(sizeof(double) * esz * t->k * t->n);
        transpose_block(b_data, b_cache[t->b_pos], p->b_dims + t->b_pos * p->b_rank, b_perm, p->b_rank, cplx);
      }
      b_data = b_cache[t->b_pos];
    }
    gemm_task(cplx, t->m, t->k, t->n, t->f_ex_sign, a_data, b_data, t->beta, C + esz * t->c_data_offset);
  }
  if (a_cache) {
    for (int64_t i = 0; i < max_a; ++i) free(a_cache[i]);
    free(a_cache);
  }
  if (b_cache) {
    for (int64_t i = 0; i < max_b; ++i) free(b_cache[i]);
    free(b_cache);
  }
  free(order);
  return 0;
}

/* SplitMix64-based uniform(-1,1) fill, the input generator shared with the
 * device (qlt_fill_uniform with lo = -1, hi = 1): value i = -1 + 2 * ((z >> 11) * 2^-53). */
void ora_fill_uniform(double *x, int64_t n, uint64_t seed, int64_t first) {
#pragma omp parallel for num_threads(g_threads) schedule(static) if (n > 1000000)
  for (int64_t i = 0; i < n; ++i) {
    uint64_t z = seed + ((uint64_t)(first + i) + 1ull) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    x[i] = -1.0 + 2.0 * ((double)(z >> 11) * 0x1.0p-53);
  }
}
