// Tile-shape and split-k tuning of the grouped kernel, and the autotuner
// that chooses them by timing a workload on the device.
//
// The reference has no counterpart: its contraction issues one BLAS call per
// task (CtrctTwoBSDTAndAssignIn, global_operations.h:895-992), so there are no
// tiles to shape.  Here an output block of the MFMA tile classes falls into a
// size bucket (ceil(m/32), ceil(n/32)); the tuning table picks the tile shape
// of each bucket, and a workgroup target per tile class sets how far k-slicing
// spreads long tiles.  The plan builder reads the tuning when it builds a plan.
#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include "qlt_internal.h"

namespace qlt {
namespace {

qlt_tuning builtin_tuning() {
  qlt_tuning t;
  std::memset(t.tile_class, 0xff, sizeof(t.tile_class));   // -1: built-in rule
  for (int c = 0; c < QLT_TILE_CLASSES; ++c) t.slice_target[c] = 0;
  return t;
}

std::mutex g_tune_mutex;
qlt_tuning g_tuning = builtin_tuning();

void install(const qlt_tuning &t) {
  std::lock_guard<std::mutex> lock(g_tune_mutex);
  g_tuning = t;
}

bool valid_tuning(const qlt_tuning &t) {
  for (int i = 0; i < QLT_TUNE_BUCKETS; ++i)
    for (int j = 0; j < QLT_TUNE_BUCKETS; ++j)
      if (t.tile_class[i][j] < -1 || t.tile_class[i][j] >= kNumMfmaClasses) return false;
  for (int c = 0; c < QLT_TILE_CLASSES; ++c)
    if (t.slice_target[c] < -1) return false;
  return true;
}

using Bucket = std::pair<int, int>;

// Builds the workload's plans under the installed tuning, runs it once, then
// times `reps` runs and returns the fastest (ms).  Optionally returns the
// algorithmic weight (sum of m*n*k) of every size bucket of the MFMA classes
// and the tile count of every class.
qlt_status time_workload(int32_t n_items, const qlt_tune_item *items, int32_t dtype, int32_t mode, int32_t reps,
                         hipStream_t st, double *ms, std::map<Bucket, double> *weights,
                         int64_t *cls_tiles) {
  std::vector<qlt_plan *> plans(static_cast<size_t>(n_items), nullptr);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  auto cleanup = [&]() {
    for (qlt_plan *p : plans)
      if (p != nullptr) qlt_plan_destroy(p);
    if (e0 != nullptr) (void)hipEventDestroy(e0);
    if (e1 != nullptr) (void)hipEventDestroy(e1);
  };
  qlt_status s = QLT_OK;
  for (int32_t i = 0; i < n_items && s == QLT_OK; ++i) {
    const qlt_tune_item &it = items[i];
    s = qlt_plan_build(it.a, it.b, it.n_ctrct, it.a_ctrct_axes, it.b_ctrct_axes, it.n_a_saved, it.a_saved_axes,
                       it.n_b_saved, it.b_saved_axes, st, &plans[static_cast<size_t>(i)]);
    if (s == QLT_OK) s = qlt_plan_prepare(plans[static_cast<size_t>(i)], dtype, mode);
  }
  if (s != QLT_OK) {
    cleanup();
    return s;
  }
  for (qlt_plan *p : plans) {
    if (weights != nullptr)
      for (const Seg &sg : p->h_segs) {
        if (!is_mfma_class(tile_class_of(sg.m, sg.n))) continue;
        const int bm = static_cast<int>(std::min<int64_t>((sg.m + 31) / 32, QLT_TUNE_BUCKETS));
        const int bn = static_cast<int>(std::min<int64_t>((sg.n + 31) / 32, QLT_TUNE_BUCKETS));
        (*weights)[{bm, bn}] += static_cast<double>(sg.m) * sg.n * static_cast<double>(sg.k_total);
      }
    if (cls_tiles != nullptr)
      for (int c = 0; c < kNumTileClasses; ++c) cls_tiles[c] += p->cls_tiles[c];
  }
  const double alpha[2] = {1.0, 0.0}, beta[2] = {0.0, 0.0};
  auto run_all = [&]() -> qlt_status {
    for (int32_t i = 0; i < n_items; ++i) {
      const qlt_tune_item &it = items[i];
      QLT_TRY(qlt_contract_execute(plans[static_cast<size_t>(i)], dtype, mode, it.a_data, it.b_data, it.c_data,
                                   alpha, beta, st));
    }
    return QLT_OK;
  };
  s = run_all();
  hipError_t e = hipSuccess;
  if (s == QLT_OK) e = hipEventCreate(&e0);
  if (s == QLT_OK && e == hipSuccess) e = hipEventCreate(&e1);
  double best = 1e300;
  for (int32_t r = 0; r < reps && s == QLT_OK && e == hipSuccess; ++r) {
    e = hipEventRecord(e0, st);
    if (e != hipSuccess) break;
    s = run_all();
    if (s != QLT_OK) break;
    e = hipEventRecord(e1, st);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float t = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
    if (e == hipSuccess) best = std::min(best, static_cast<double>(t));
  }
  cleanup();
  if (s != QLT_OK) return s;
  if (e != hipSuccess) return hip_fail(e, "qlt_autotune timing", __FILE__, __LINE__);
  *ms = best;
  return QLT_OK;
}

}  // namespace

qlt_tuning tuning_snapshot() {
  std::lock_guard<std::mutex> lock(g_tune_mutex);
  return g_tuning;
}

}  // namespace qlt

using namespace qlt;

extern "C" {

qlt_status qlt_tuning_get(qlt_tuning *tuning) {
  QLT_CHECK_ARG(tuning != nullptr, "qlt_tuning_get: null argument");
  *tuning = tuning_snapshot();
  return QLT_OK;
}

qlt_status qlt_tuning_set(const qlt_tuning *tuning) {
  if (tuning == nullptr) {
    install(builtin_tuning());
    return QLT_OK;
  }
  QLT_CHECK_ARG(valid_tuning(*tuning),
                "qlt_tuning_set: tile classes must be -1..3 and slice targets >= -1");
  install(*tuning);
  return QLT_OK;
}

qlt_status qlt_autotune(int32_t n_items, const qlt_tune_item *items, int32_t dtype, int32_t mode, int32_t reps,
                        int32_t max_buckets, void *stream, qlt_tuning *best, double *ms_before, double *ms_after,
                        int32_t *n_trials) {
  QLT_CHECK_ARG(n_items > 0 && items != nullptr, "qlt_autotune: empty workload");
  QLT_CHECK_ARG(reps >= 1, "qlt_autotune: reps must be >= 1");
  QLT_CHECK_ARG(dtype == QLT_F64 || dtype == QLT_C128, "qlt_autotune: dtype must be QLT_F64 or QLT_C128");
  QLT_CHECK_ARG(valid_mode(mode), "qlt_autotune: unknown mode");
  hipStream_t st = static_cast<hipStream_t>(stream);
  QLT_CHECK_ARG(!stream_capturing(st), "qlt_autotune: cannot tune under graph capture");
  const qlt_tuning entry = tuning_snapshot();

  std::map<Bucket, double> weights;
  int64_t cls_tiles[kNumTileClasses] = {0, 0, 0, 0, 0, 0};
  double base_ms = 0.0;
  QLT_TRY(time_workload(n_items, items, dtype, mode, reps, st, &base_ms, &weights, cls_tiles));
  int32_t trials = 1;
  qlt_tuning cur = entry;
  double cur_ms = base_ms;
  // A candidate replaces the current setting only when it is faster by more
  // than the run-to-run spread of best-of-reps timings (about half a percent).
  auto trial = [&](const qlt_tuning &t) -> qlt_status {
    install(t);
    double ms = 0.0;
    const qlt_status s = time_workload(n_items, items, dtype, mode, reps, st, &ms, nullptr, nullptr);
    ++trials;
    if (s != QLT_OK) {
      install(entry);
      return s;
    }
    if (ms < cur_ms * 0.995) {
      cur = t;
      cur_ms = ms;
    }
    return QLT_OK;
  };

  std::vector<std::pair<double, Bucket>> order;
  for (const auto &kv : weights) order.push_back({-kv.second, kv.first});
  std::sort(order.begin(), order.end());
  if (max_buckets > 0 && static_cast<int32_t>(order.size()) > max_buckets) order.resize(static_cast<size_t>(max_buckets));
  for (const auto &ob : order) {
    const int i = ob.second.first - 1, j = ob.second.second - 1;
    const int8_t orig = cur.tile_class[i][j];
    for (int8_t c = -1; c < kNumMfmaClasses; ++c) {
      if (c == orig) continue;
      qlt_tuning t = cur;
      t.tile_class[i][j] = c;
      QLT_TRY(trial(t));
    }
  }
  const int32_t targets[] = {512, 1024, 2048, 4096, -1};
  for (int c = 0; c < kNumTileClasses; ++c) {
    if (cls_tiles[c] == 0) continue;
    for (int32_t tg : targets) {
      const int32_t now = cur.slice_target[c] == 0 ? 2048 : cur.slice_target[c];
      if (tg == now) continue;
      qlt_tuning t = cur;
      t.slice_target[c] = tg;
      QLT_TRY(trial(t));
    }
  }
  install(cur);
  double final_ms = cur_ms;
  QLT_TRY(time_workload(n_items, items, dtype, mode, reps, st, &final_ms, nullptr, nullptr));
  ++trials;
  if (best != nullptr) *best = cur;
  if (ms_before != nullptr) *ms_before = base_ms;
  if (ms_after != nullptr) *ms_after = final_ms;
  if (n_trials != nullptr) *n_trials = trials;
  return QLT_OK;
}

}  // extern "C"
