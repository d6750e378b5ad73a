// Runtime part of libqlten_amd: error state, device selection, memory,
// streams, graph capture and the raw-data elementwise primitives.
//
// Error convention: every entry point returns a status and records a
// thread-local message; nothing exits the process (the reference prints and
// calls std::exit on CUDA errors, include/qlten/framework/mem_ops.h:28-35).

#include <cstdio>
#include <cstring>
#include <mutex>
#include <type_traits>
#include <string>
#include <vector>

#include "qlt_internal.h"

// Segment list of qlt_scale_segments_noalloc (offsets, then lengths), for
// callers that cannot allocate device memory.  One copy per GPU, 1 MiB, held
// until the process ends; its device address is looked up once per device in
// qlt_init (the name itself is a host-side stand-in, never passed to a kernel).
constexpr int64_t kNoAllocMaxSegs = 65536;
static __device__ __align__(256) int64_t scratch[2 * kNoAllocMaxSegs];

namespace qlt {

namespace {
constexpr int kMaxDevices = 64;
int64_t *g_scratch_ptr[kMaxDevices] = {};   // device address of `scratch`, per device
thread_local std::string g_last_error;
bool g_initialized = false;
int g_device = 0;
double *g_reduce_buf = nullptr;          // partial sums of qlt_dot
constexpr int kReduceBlocks = 1024;
std::mutex g_init_mutex;
}  // namespace

void set_error(const std::string &msg) { g_last_error = msg; }

qlt_status fail(qlt_status code, const std::string &msg) {
  g_last_error = msg;
  return code;
}

qlt_status hip_fail(hipError_t e, const char *what, const char *file, int line) {
  char buf[512];
  std::snprintf(buf, sizeof(buf), "HIP error %d (%s) in %s at %s:%d", static_cast<int>(e),
                hipGetErrorString(e), what, file, line);
  g_last_error = buf;
  return e == hipErrorOutOfMemory ? QLT_ERR_OUT_OF_MEMORY : QLT_ERR_HIP;
}

bool initialized() { return g_initialized; }
int current_device() { return g_device; }

int device_cu_count() {
  // read once per process for the library's device (qlt_init picks one device)
  static std::once_flag once;
  static int cus = 0;
  std::call_once(once, [] {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, g_device) == hipSuccess && v > 0) cus = v;
  });
  return cus > 0 ? cus : 256;
}

bool stream_capturing(hipStream_t s) {
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &st) != hipSuccess) return false;
  return st == hipStreamCaptureStatusActive;
}

// ---- class streams ----
namespace {
std::mutex g_cls_mutex;
hipStream_t g_cls_stream[kNumTileClasses] = {};
hipEvent_t g_cls_fork = nullptr;
hipEvent_t g_cls_join[kNumTileClasses] = {};

hipError_t class_streams_create() {
  for (int c = 0; c < kNumTileClasses; ++c) {
    hipError_t e = hipStreamCreateWithFlags(&g_cls_stream[c], hipStreamNonBlocking);
    if (e != hipSuccess) return e;
    e = hipEventCreateWithFlags(&g_cls_join[c], hipEventDisableTiming);
    if (e != hipSuccess) return e;
  }
  return hipEventCreateWithFlags(&g_cls_fork, hipEventDisableTiming);
}

void class_streams_destroy() {
  for (int c = 0; c < kNumTileClasses; ++c) {
    if (g_cls_stream[c] != nullptr) (void)hipStreamDestroy(g_cls_stream[c]);
    if (g_cls_join[c] != nullptr) (void)hipEventDestroy(g_cls_join[c]);
    g_cls_stream[c] = nullptr;
    g_cls_join[c] = nullptr;
  }
  if (g_cls_fork != nullptr) (void)hipEventDestroy(g_cls_fork);
  g_cls_fork = nullptr;
}
}  // namespace

ClassStreams::ClassStreams(hipStream_t st) : st_(st) {
  g_cls_mutex.lock();
  ok_ = g_cls_fork != nullptr && hipEventRecord(g_cls_fork, st) == hipSuccess;
}

hipStream_t ClassStreams::get(int32_t cls) {
  if (!ok_) return st_;
  if (!(used_ & (1u << cls))) {
    if (hipStreamWaitEvent(g_cls_stream[cls], g_cls_fork, 0) != hipSuccess) return st_;
    used_ |= 1u << cls;
  }
  return g_cls_stream[cls];
}

ClassStreams::~ClassStreams() {
  for (int c = 0; c < kNumTileClasses; ++c)
    if (used_ & (1u << c)) {
      (void)hipEventRecord(g_cls_join[c], g_cls_stream[c]);
      (void)hipStreamWaitEvent(st_, g_cls_join[c], 0);
    }
  g_cls_mutex.unlock();
}

// ---- kernel probe ----
namespace {
struct KernelProbe {
  hipEvent_t start = nullptr;
  hipEvent_t stop = nullptr;
  std::string name;
  double flops = 0.0;
  double bytes = 0.0;
  bool fired = false;
};
KernelProbe g_probe;
std::mutex g_probe_mutex;
}  // namespace

bool probe_armed(hipStream_t st) { return g_probe.start != nullptr && !stream_capturing(st); }

void probe_begin(hipStream_t st) { (void)hipEventRecord(g_probe.start, st); }

void probe_end(hipStream_t st, const char *name, double flops, double bytes) {
  (void)hipEventRecord(g_probe.stop, st);
  std::lock_guard<std::mutex> lock(g_probe_mutex);
  g_probe.name = name;
  g_probe.flops = flops;
  g_probe.bytes = bytes;
  g_probe.fired = true;
}

}  // namespace qlt

using namespace qlt;

// ---------------------------------------------------------------------------
// elementwise kernels
// ---------------------------------------------------------------------------
namespace {

__device__ __forceinline__ double splitmix_unit(uint64_t seed, uint64_t i) {
  uint64_t z = seed + (i + 1ull) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return static_cast<double>(z >> 11) * 0x1.0p-53;
}

__global__ void k_fill_uniform(double *x, int64_t n, uint64_t seed, int64_t first, double lo, double width) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n; i += stride)
    x[i] = lo + width * splitmix_unit(seed, static_cast<uint64_t>(first + i));
}

__global__ void k_zero(uint8_t *p, int64_t bytes) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  const int64_t nwords = bytes / 8;
  uint64_t *w = reinterpret_cast<uint64_t *>(p);
  for (int64_t j = i; j < nwords; j += stride) w[j] = 0;
  for (int64_t j = nwords * 8 + i; j < bytes; j += stride) p[j] = 0;
}

// y = a x + b y over n elements stored as E (double or float), FP64 arithmetic.
template <typename E>
__global__ void k_axpby_real(int64_t n, double a, const E *x, double b, E *y) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n; i += stride) {
    const double xv = static_cast<double>(x[i]);
    y[i] = static_cast<E>((b == 0.0) ? a * xv : a * xv + b * static_cast<double>(y[i]));
  }
}

template <typename E>
__global__ void k_axpby_cplx(int64_t n, double ar, double ai, const E *x, double br, double bi, E *y) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  const bool bzero = (br == 0.0 && bi == 0.0);
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n; i += stride) {
    const double xr = x[2 * i], xi = x[2 * i + 1];
    double yr = ar * xr - ai * xi, yi = ar * xi + ai * xr;
    if (!bzero) {
      const double vr = y[2 * i], vi = y[2 * i + 1];
      yr += br * vr - bi * vi;
      yi += br * vi + bi * vr;
    }
    y[2 * i] = static_cast<E>(yr);
    y[2 * i + 1] = static_cast<E>(yi);
  }
}

// Two-level reduction of sum conj(x) * y: block partials, then one block.
template <bool CPLX, typename E = double>
__global__ __launch_bounds__(256) void k_dot_partial(int64_t n, const E *x, const E *y, double *partial) {
  __shared__ double sr[256], si[256];
  double accr = 0.0, acci = 0.0;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n; i += stride) {
    if (CPLX) {
      const double xr = x[2 * i], xi = x[2 * i + 1], yr = y[2 * i], yi = y[2 * i + 1];
      accr += xr * yr + xi * yi;
      acci += xr * yi - xi * yr;
    } else {
      accr += static_cast<double>(x[i]) * static_cast<double>(y[i]);
    }
  }
  sr[threadIdx.x] = accr;
  si[threadIdx.x] = acci;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      sr[threadIdx.x] += sr[threadIdx.x + s];
      si[threadIdx.x] += si[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = sr[0];
    partial[2 * blockIdx.x + 1] = si[0];
  }
}

__global__ __launch_bounds__(256) void k_dot_final(int nparts, double *partial) {
  __shared__ double sr[256], si[256];
  double accr = 0.0, acci = 0.0;
  for (int i = threadIdx.x; i < nparts; i += 256) {
    accr += partial[2 * i];
    acci += partial[2 * i + 1];
  }
  sr[threadIdx.x] = accr;
  si[threadIdx.x] = acci;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      sr[threadIdx.x] += sr[threadIdx.x + s];
      si[threadIdx.x] += si[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partial[2 * kReduceBlocks] = sr[0];
    partial[2 * kReduceBlocks + 1] = si[0];
  }
}

__global__ void k_real_to_complex(int64_t n, const double *x, double *z) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n; i += stride) {
    z[2 * i] = x[i];
    z[2 * i + 1] = 0.0;
  }
}

template <typename E>
__global__ void k_conjugate(int64_t n, E *z) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n; i += stride)
    z[2 * i + 1] = -z[2 * i + 1];
}

// Element conversion between storage types: FS/FD the source/destination
// real types, CS/CD whether the element is complex (real -> complex sets the
// imaginary part to 0; complex -> real is not offered).
template <typename FS, bool CS, typename FD, bool CD>
__global__ void k_convert(int64_t n, const FS *src, FD *dst) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n; i += stride) {
    if constexpr (CD) {
      dst[2 * i] = static_cast<FD>(CS ? src[2 * i] : src[i]);
      dst[2 * i + 1] = CS ? static_cast<FD>(src[2 * i + 1]) : FD(0);
    } else {
      dst[i] = static_cast<FD>(src[i]);
    }
  }
}

// The seeded uniform fill of qlt_fill_uniform, rounded to float.
__global__ void k_fill_uniform_f32(float *x, int64_t n, uint64_t seed, int64_t first, double lo, double width) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n; i += stride)
    x[i] = static_cast<float>(lo + width * splitmix_unit(seed, static_cast<uint64_t>(first + i)));
}

// One block per segment; scale by a real or complex factor.
template <bool CPLX, typename E = double>
__global__ void k_scale_segments(const int64_t *offs, const int64_t *lens, double sr, double si, E *x) {
  const int64_t off = offs[blockIdx.x], len = lens[blockIdx.x];
  for (int64_t i = threadIdx.x; i < len; i += blockDim.x) {
    if (CPLX) {
      E *p = x + 2 * (off + i);
      const double r = p[0], im = p[1];
      p[0] = static_cast<E>(sr * r - si * im);
      p[1] = static_cast<E>(sr * im + si * r);
    } else {
      x[off + i] = static_cast<E>(sr * static_cast<double>(x[off + i]));
    }
  }
}

void launch_scale_segments(int32_t dtype, unsigned n_segs, const int64_t *offs, const int64_t *lens, const double *s,
                           void *x, hipStream_t st) {
  if (dtype == QLT_F64)
    hipLaunchKernelGGL((k_scale_segments<false, double>), dim3(n_segs), dim3(256), 0, st, offs, lens, s[0], 0.0,
                       static_cast<double *>(x));
  else if (dtype == QLT_C128)
    hipLaunchKernelGGL((k_scale_segments<true, double>), dim3(n_segs), dim3(256), 0, st, offs, lens, s[0], s[1],
                       static_cast<double *>(x));
  else if (dtype == QLT_F32)
    hipLaunchKernelGGL((k_scale_segments<false, float>), dim3(n_segs), dim3(256), 0, st, offs, lens, s[0], 0.0,
                       static_cast<float *>(x));
  else
    hipLaunchKernelGGL((k_scale_segments<true, float>), dim3(n_segs), dim3(256), 0, st, offs, lens, s[0], s[1],
                       static_cast<float *>(x));
}

inline unsigned grid_for(int64_t n, int threads = 256, int64_t cap = 8192) {
  int64_t g = (n + threads - 1) / threads;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return static_cast<unsigned>(g);
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

const char *qlt_last_error(void) { return g_last_error.c_str(); }

qlt_status qlt_kernel_probe_set(void *start_event, void *stop_event) {
  QLT_CHECK_ARG((start_event == nullptr) == (stop_event == nullptr),
                "qlt_kernel_probe_set: give both events or neither");
  std::lock_guard<std::mutex> lock(g_probe_mutex);
  g_probe.start = static_cast<hipEvent_t>(start_event);
  g_probe.stop = static_cast<hipEvent_t>(stop_event);
  g_probe.fired = false;
  return QLT_OK;
}

qlt_status qlt_kernel_probe_last(char *name, int32_t name_len, double *flops, double *bytes) {
  QLT_CHECK_ARG(name != nullptr && name_len > 0 && flops != nullptr && bytes != nullptr,
                "qlt_kernel_probe_last: null output");
  std::lock_guard<std::mutex> lock(g_probe_mutex);
  if (!g_probe.fired) return fail(QLT_ERR_INVALID_ARG, "qlt_kernel_probe_last: no probed launch since arming");
  std::snprintf(name, static_cast<size_t>(name_len), "%s", g_probe.name.c_str());
  *flops = g_probe.flops;
  *bytes = g_probe.bytes;
  return QLT_OK;
}

qlt_status qlt_init(int32_t device) {
  std::lock_guard<std::mutex> lock(g_init_mutex);
  int count = 0;
  QLT_HIP(hipGetDeviceCount(&count));
  if (device < 0 || device >= count)
    return fail(QLT_ERR_INVALID_ARG, "qlt_init: device " + std::to_string(device) +
                                         " out of range (" + std::to_string(count) + " devices)");
  QLT_HIP(hipSetDevice(device));
  if (device >= kMaxDevices)
    return fail(QLT_ERR_UNSUPPORTED, "qlt_init: device " + std::to_string(device) + " beyond the supported 64");
  if (g_scratch_ptr[device] == nullptr) {
    void *p = nullptr;
    QLT_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[device] = static_cast<int64_t *>(p);
  }
  if (g_initialized && g_device == device) return QLT_OK;
  if (g_reduce_buf != nullptr) {
    (void)hipFree(g_reduce_buf);
    g_reduce_buf = nullptr;
  }
  QLT_HIP(hipMalloc(&g_reduce_buf, sizeof(double) * (2 * kReduceBlocks + 2)));
  {
    std::lock_guard<std::mutex> cl(g_cls_mutex);
    class_streams_destroy();
    QLT_HIP(class_streams_create());
  }
  g_device = device;
  g_initialized = true;
  return QLT_OK;
}

qlt_status qlt_finalize(void) {
  std::lock_guard<std::mutex> lock(g_init_mutex);
  if (!g_initialized) return QLT_OK;
  QLT_HIP(hipDeviceSynchronize());
  if (g_reduce_buf != nullptr) QLT_HIP(hipFree(g_reduce_buf));
  g_reduce_buf = nullptr;
  {
    std::lock_guard<std::mutex> cl(g_cls_mutex);
    class_streams_destroy();
  }
  g_initialized = false;
  return QLT_OK;
}

qlt_status qlt_device_sync(void) {
  QLT_REQUIRE_INIT();
  QLT_HIP(hipDeviceSynchronize());
  return QLT_OK;
}

qlt_status qlt_stream_create(void **stream) {
  QLT_REQUIRE_INIT();
  QLT_CHECK_ARG(stream != nullptr, "qlt_stream_create: null output");
  hipStream_t s;
  QLT_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  *stream = s;
  return QLT_OK;
}

qlt_status qlt_stream_destroy(void *stream) {
  QLT_REQUIRE_INIT();
  if (stream != nullptr) QLT_HIP(hipStreamDestroy(as_stream(stream)));
  return QLT_OK;
}

qlt_status qlt_stream_sync(void *stream) {
  QLT_REQUIRE_INIT();
  QLT_HIP(hipStreamSynchronize(as_stream(stream)));
  return QLT_OK;
}

qlt_status qlt_malloc(void **ptr, size_t bytes) {
  QLT_REQUIRE_INIT();
  QLT_CHECK_ARG(ptr != nullptr, "qlt_malloc: null output");
  *ptr = nullptr;
  if (bytes == 0) return QLT_OK;
  QLT_HIP(hipMalloc(ptr, bytes));
  return QLT_OK;
}

qlt_status qlt_free(void *ptr) {
  QLT_REQUIRE_INIT();
  if (ptr != nullptr) QLT_HIP(hipFree(ptr));
  return QLT_OK;
}

qlt_status qlt_memcpy_h2d(void *dst, const void *src, size_t bytes) {
  QLT_REQUIRE_INIT();
  if (bytes == 0) return QLT_OK;
  QLT_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  QLT_HIP(hipStreamSynchronize(nullptr));   // complete on return, for callers on other streams
  return QLT_OK;
}

qlt_status qlt_memcpy_d2h(void *dst, const void *src, size_t bytes) {
  QLT_REQUIRE_INIT();
  if (bytes == 0) return QLT_OK;
  QLT_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  QLT_HIP(hipStreamSynchronize(nullptr));   // complete on return, for callers on other streams
  return QLT_OK;
}

qlt_status qlt_memcpy_d2d(void *dst, const void *src, size_t bytes) {
  QLT_REQUIRE_INIT();
  if (bytes == 0) return QLT_OK;
  QLT_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
  QLT_HIP(hipStreamSynchronize(nullptr));   // complete on return, for callers on other streams
  return QLT_OK;
}

qlt_status qlt_memset_zero(void *dst, size_t bytes, void *stream) {
  QLT_REQUIRE_INIT();
  if (bytes == 0) return QLT_OK;
  QLT_CHECK_ARG(dst != nullptr, "qlt_memset_zero: null buffer");
  // A fill kernel rather than hipMemsetAsync, so the call is safe inside a
  // captured graph.
  const int64_t n = static_cast<int64_t>(bytes);
  hipLaunchKernelGGL(k_zero, dim3(grid_for(n / 8 + 1)), dim3(256), 0, as_stream(stream),
                     static_cast<uint8_t *>(dst), n);
  QLT_HIP(hipGetLastError());
  return QLT_OK;
}

struct qlt_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

qlt_status qlt_graph_begin(void *stream) {
  QLT_REQUIRE_INIT();
  QLT_CHECK_ARG(stream != nullptr, "qlt_graph_begin: capture needs a non-null stream");
  QLT_HIP(hipStreamBeginCapture(as_stream(stream), hipStreamCaptureModeThreadLocal));
  return QLT_OK;
}

qlt_status qlt_graph_end(void *stream, qlt_graph **graph) {
  QLT_REQUIRE_INIT();
  QLT_CHECK_ARG(graph != nullptr, "qlt_graph_end: null output");
  hipGraph_t g = nullptr;
  QLT_HIP(hipStreamEndCapture(as_stream(stream), &g));
  hipGraphExec_t ge = nullptr;
  hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    (void)hipGraphDestroy(g);
    return hip_fail(e, "hipGraphInstantiate", __FILE__, __LINE__);
  }
  auto *out = new qlt_graph;
  out->graph = g;
  out->exec = ge;
  *graph = out;
  return QLT_OK;
}

qlt_status qlt_graph_replay(qlt_graph *graph, void *stream) {
  QLT_REQUIRE_INIT();
  QLT_CHECK_ARG(graph != nullptr && graph->exec != nullptr, "qlt_graph_replay: null graph");
  QLT_HIP(hipGraphLaunch(graph->exec, as_stream(stream)));
  return QLT_OK;
}

qlt_status qlt_graph_destroy(qlt_graph *graph) {
  if (graph == nullptr) return QLT_OK;
  if (graph->exec != nullptr) (void)hipGraphExecDestroy(graph->exec);
  if (graph->graph != nullptr) (void)hipGraphDestroy(graph->graph);
  delete graph;
  return QLT_OK;
}

qlt_status qlt_fill_uniform(void *x, int64_t n_doubles, uint64_t seed, int64_t first_index,
                            double lo, double hi, void *stream) {
  QLT_REQUIRE_INIT();
  QLT_CHECK_ARG(n_doubles >= 0 && first_index >= 0, "qlt_fill_uniform: negative size");
  if (n_doubles == 0) return QLT_OK;
  QLT_CHECK_ARG(x != nullptr, "qlt_fill_uniform: null buffer");
  hipLaunchKernelGGL(k_fill_uniform, dim3(grid_for(n_doubles)), dim3(256), 0, as_stream(stream),
                     static_cast<double *>(x), n_doubles, seed, first_index, lo, hi - lo);
  QLT_HIP(hipGetLastError());
  return QLT_OK;
}

qlt_status qlt_axpby(int32_t dtype, int64_t n, const double *a, const void *x, const double *b,
                     void *y, void *stream) {
  QLT_REQUIRE_INIT();
  QLT_CHECK_ARG(valid_dtype(dtype), "qlt_axpby: bad dtype");
  QLT_CHECK_ARG(a != nullptr && b != nullptr, "qlt_axpby: null scalar");
  if (n <= 0) return QLT_OK;
  const dim3 g(grid_for(n));
  hipStream_t st = as_stream(stream);
  if (dtype == QLT_F64)
    hipLaunchKernelGGL(k_axpby_real<double>, g, dim3(256), 0, st, n, a[0], static_cast<const double *>(x), b[0],
                       static_cast<double *>(y));
  else if (dtype == QLT_C128)
    hipLaunchKernelGGL(k_axpby_cplx<double>, g, dim3(256), 0, st, n, a[0], a[1], static_cast<const double *>(x), b[0],
                       b[1], static_cast<double *>(y));
  else if (dtype == QLT_F32)
    hipLaunchKernelGGL(k_axpby_real<float>, g, dim3(256), 0, st, n, a[0], static_cast<const float *>(x), b[0],
                       static_cast<float *>(y));
  else
    hipLaunchKernelGGL(k_axpby_cplx<float>, g, dim3(256), 0, st, n, a[0], a[1], static_cast<const float *>(x), b[0],
                       b[1], static_cast<float *>(y));
  QLT_HIP(hipGetLastError());
  return QLT_OK;
}

qlt_status qlt_dot(int32_t dtype, int64_t n, const void *x, const void *y, double *out,
                   void *stream) {
  QLT_REQUIRE_INIT();
  QLT_CHECK_ARG(valid_dtype(dtype), "qlt_dot: bad dtype");
  QLT_CHECK_ARG(out != nullptr, "qlt_dot: null output");
  out[0] = out[1] = 0.0;
  if (n <= 0) return QLT_OK;
  hipStream_t s = as_stream(stream);
  const unsigned g = std::min<unsigned>(grid_for(n), kReduceBlocks);
  if (dtype == QLT_F64)
    hipLaunchKernelGGL((k_dot_partial<false, double>), dim3(g), dim3(256), 0, s, n,
                       static_cast<const double *>(x), static_cast<const double *>(y), g_reduce_buf);
  else if (dtype == QLT_C128)
    hipLaunchKernelGGL((k_dot_partial<true, double>), dim3(g), dim3(256), 0, s, n,
                       static_cast<const double *>(x), static_cast<const double *>(y), g_reduce_buf);
  else if (dtype == QLT_F32)
    hipLaunchKernelGGL((k_dot_partial<false, float>), dim3(g), dim3(256), 0, s, n,
                       static_cast<const float *>(x), static_cast<const float *>(y), g_reduce_buf);
  else
    hipLaunchKernelGGL((k_dot_partial<true, float>), dim3(g), dim3(256), 0, s, n,
                       static_cast<const float *>(x), static_cast<const float *>(y), g_reduce_buf);
  hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(256), 0, s, static_cast<int>(g), g_reduce_buf);
  QLT_HIP(hipGetLastError());
  QLT_HIP(hipMemcpyAsync(out, g_reduce_buf + 2 * kReduceBlocks, 2 * sizeof(double),
                         hipMemcpyDeviceToHost, s));
  QLT_HIP(hipStreamSynchronize(s));
  return QLT_OK;
}

qlt_status qlt_real_to_complex(int64_t n, const double *x, double *z, void *stream) {
  QLT_REQUIRE_INIT();
  if (n <= 0) return QLT_OK;
  QLT_CHECK_ARG(x != nullptr && z != nullptr, "qlt_real_to_complex: null buffer");
  hipLaunchKernelGGL(k_real_to_complex, dim3(grid_for(n)), dim3(256), 0, as_stream(stream), n, x, z);
  QLT_HIP(hipGetLastError());
  return QLT_OK;
}

qlt_status qlt_conjugate(int64_t n, double *z, void *stream) {
  QLT_REQUIRE_INIT();
  if (n <= 0) return QLT_OK;
  QLT_CHECK_ARG(z != nullptr, "qlt_conjugate: null buffer");
  hipLaunchKernelGGL(k_conjugate<double>, dim3(grid_for(n)), dim3(256), 0, as_stream(stream), n, z);
  QLT_HIP(hipGetLastError());
  return QLT_OK;
}

qlt_status qlt_conjugate_c64(int64_t n, float *z, void *stream) {
  QLT_REQUIRE_INIT();
  if (n <= 0) return QLT_OK;
  QLT_CHECK_ARG(z != nullptr, "qlt_conjugate_c64: null buffer");
  hipLaunchKernelGGL(k_conjugate<float>, dim3(grid_for(n)), dim3(256), 0, as_stream(stream), n, z);
  QLT_HIP(hipGetLastError());
  return QLT_OK;
}

qlt_status qlt_fill_uniform_f32(void *x, int64_t n_floats, uint64_t seed, int64_t first_index, double lo,
                                double hi, void *stream) {
  QLT_REQUIRE_INIT();
  if (n_floats <= 0) return QLT_OK;
  QLT_CHECK_ARG(x != nullptr, "qlt_fill_uniform_f32: null buffer");
  hipLaunchKernelGGL(k_fill_uniform_f32, dim3(grid_for(n_floats)), dim3(256), 0, as_stream(stream),
                     static_cast<float *>(x), n_floats, seed, first_index, lo, hi - lo);
  QLT_HIP(hipGetLastError());
  return QLT_OK;
}

qlt_status qlt_convert(int32_t src_dtype, int32_t dst_dtype, int64_t n, const void *src, void *dst, void *stream) {
  QLT_REQUIRE_INIT();
  QLT_CHECK_ARG(valid_dtype(src_dtype) && valid_dtype(dst_dtype), "qlt_convert: bad dtype");
  QLT_CHECK_ARG(!dtype_is_complex(src_dtype) || dtype_is_complex(dst_dtype),
                "qlt_convert: complex to real is not a conversion");
  if (n <= 0) return QLT_OK;
  QLT_CHECK_ARG(src != nullptr && dst != nullptr, "qlt_convert: null buffer");
  const dim3 g(grid_for(n));
  hipStream_t st = as_stream(stream);
  const bool cs = dtype_is_complex(src_dtype), cd = dtype_is_complex(dst_dtype);
  const bool ss = dtype_is_single(src_dtype), sd = dtype_is_single(dst_dtype);
  auto launch = [&](auto fs, auto cs_tag, auto fd, auto cd_tag) {
    using FS = decltype(fs);
    using FD = decltype(fd);
    hipLaunchKernelGGL((k_convert<FS, decltype(cs_tag)::value, FD, decltype(cd_tag)::value>), g, dim3(256), 0, st, n,
                       static_cast<const FS *>(src), static_cast<FD *>(dst));
  };
  auto by_dst = [&](auto fs, auto cs_tag) {
    if (cd) {
      if (sd) launch(fs, cs_tag, float{}, std::true_type{});
      else launch(fs, cs_tag, double{}, std::true_type{});
    } else {
      if (sd) launch(fs, cs_tag, float{}, std::false_type{});
      else launch(fs, cs_tag, double{}, std::false_type{});
    }
  };
  if (cs) {
    if (ss) by_dst(float{}, std::true_type{});
    else by_dst(double{}, std::true_type{});
  } else {
    if (ss) by_dst(float{}, std::false_type{});
    else by_dst(double{}, std::false_type{});
  }
  QLT_HIP(hipGetLastError());
  return QLT_OK;
}

}  // extern "C"

namespace {

struct Copy2D {
  int64_t rows, cols, src_off, src_ld, dst_off, dst_ld;
};

// One workgroup per rectangle; W words of type T per element (one double
// for F64 and C64, two for C128, one float for F32).
template <int W, typename T = double>
__global__ void k_copy_2d(const Copy2D *items, const T *src, T *dst) {
  const Copy2D it = items[blockIdx.x];
  const int64_t n = it.rows * it.cols;
  for (int64_t e = threadIdx.x; e < n; e += blockDim.x) {
    const int64_t r = e / it.cols, c = e - r * it.cols;
    const int64_t si = (it.src_off + r * it.src_ld + c) * W, di = (it.dst_off + r * it.dst_ld + c) * W;
#pragma unroll
    for (int w = 0; w < W; ++w) dst[di + w] = src[si + w];
  }
}

}  // namespace

extern "C" {

qlt_status qlt_copy_blocks_2d(int32_t dtype, int64_t n, const int64_t *rows, const int64_t *cols,
                              const void *src, const int64_t *src_offsets, const int64_t *src_ld,
                              void *dst, const int64_t *dst_offsets, const int64_t *dst_ld, void *stream) {
  QLT_REQUIRE_INIT();
  QLT_CHECK_ARG(valid_dtype(dtype), "qlt_copy_blocks_2d: bad dtype");
  QLT_CHECK_ARG(n >= 0 && n < (int64_t(1) << 31), "qlt_copy_blocks_2d: bad rectangle count");
  if (n == 0) return QLT_OK;
  QLT_CHECK_ARG(rows && cols && src_offsets && src_ld && dst_offsets && dst_ld, "qlt_copy_blocks_2d: null array");
  QLT_CHECK_ARG(src != nullptr && dst != nullptr, "qlt_copy_blocks_2d: null data pointer");
  hipStream_t st = as_stream(stream);
  QLT_CHECK_ARG(!stream_capturing(st), "qlt_copy_blocks_2d: not capturable (stages host arrays)");
  std::vector<Copy2D> h(static_cast<size_t>(n));
  for (int64_t i = 0; i < n; ++i) {
    QLT_CHECK_ARG(rows[i] >= 0 && cols[i] >= 0 && src_ld[i] >= cols[i] && dst_ld[i] >= cols[i] &&
                      src_offsets[i] >= 0 && dst_offsets[i] >= 0,
                  "qlt_copy_blocks_2d: bad rectangle " + std::to_string(i));
    h[static_cast<size_t>(i)] = Copy2D{rows[i], cols[i], src_offsets[i], src_ld[i], dst_offsets[i], dst_ld[i]};
  }
  Copy2D *d = nullptr;
  QLT_HIP(hipMallocAsync(reinterpret_cast<void **>(&d), sizeof(Copy2D) * h.size(), st));
  QLT_HIP(hipMemcpyAsync(d, h.data(), sizeof(Copy2D) * h.size(), hipMemcpyHostToDevice, st));
  if (dtype == QLT_F64 || dtype == QLT_C64)
    hipLaunchKernelGGL((k_copy_2d<1, double>), dim3(static_cast<unsigned>(n)), dim3(256), 0, st, d,
                       static_cast<const double *>(src), static_cast<double *>(dst));
  else if (dtype == QLT_C128)
    hipLaunchKernelGGL((k_copy_2d<2, double>), dim3(static_cast<unsigned>(n)), dim3(256), 0, st, d,
                       static_cast<const double *>(src), static_cast<double *>(dst));
  else
    hipLaunchKernelGGL((k_copy_2d<1, float>), dim3(static_cast<unsigned>(n)), dim3(256), 0, st, d,
                       static_cast<const float *>(src), static_cast<float *>(dst));
  QLT_HIP(hipGetLastError());
  QLT_HIP(hipStreamSynchronize(st));
  QLT_HIP(hipFreeAsync(d, st));
  QLT_HIP(hipStreamSynchronize(st));
  return QLT_OK;
}

qlt_status qlt_scale_segments(int32_t dtype, int64_t n_segs, const int64_t *offsets,
                              const int64_t *lengths, const double *s, void *x, void *stream) {
  QLT_REQUIRE_INIT();
  QLT_CHECK_ARG(valid_dtype(dtype), "qlt_scale_segments: bad dtype");
  QLT_CHECK_ARG(s != nullptr, "qlt_scale_segments: null scalar");
  if (n_segs <= 0) return QLT_OK;
  QLT_CHECK_ARG(offsets != nullptr && lengths != nullptr && x != nullptr,
                "qlt_scale_segments: null array");
  hipStream_t st = as_stream(stream);
  QLT_CHECK_ARG(!stream_capturing(st), "qlt_scale_segments: not capturable (stages host arrays)");
  int64_t *d = nullptr;
  QLT_HIP(hipMallocAsync(reinterpret_cast<void **>(&d), 2 * n_segs * sizeof(int64_t), st));
  QLT_HIP(hipMemcpyAsync(d, offsets, n_segs * sizeof(int64_t), hipMemcpyHostToDevice, st));
  QLT_HIP(hipMemcpyAsync(d + n_segs, lengths, n_segs * sizeof(int64_t), hipMemcpyHostToDevice, st));
  launch_scale_segments(dtype, static_cast<unsigned>(n_segs), d, d + n_segs, s, x, st);
  QLT_HIP(hipGetLastError());
  QLT_HIP(hipStreamSynchronize(st));
  QLT_HIP(hipFreeAsync(d, st));
  QLT_HIP(hipStreamSynchronize(st));
  return QLT_OK;
}

qlt_status qlt_scale_segments_noalloc(int32_t dtype, int64_t n_segs, const int64_t *offsets,
                                      const int64_t *lengths, const double *s, void *x, void *stream) {
  QLT_REQUIRE_INIT();
  QLT_CHECK_ARG(valid_dtype(dtype), "qlt_scale_segments_noalloc: bad dtype");
  QLT_CHECK_ARG(s != nullptr, "qlt_scale_segments_noalloc: null scalar");
  if (n_segs <= 0) return QLT_OK;
  QLT_CHECK_ARG(n_segs <= kNoAllocMaxSegs,
                "qlt_scale_segments_noalloc: " + std::to_string(n_segs) + " segments exceed the maximum of " +
                    std::to_string(kNoAllocMaxSegs));
  QLT_CHECK_ARG(offsets != nullptr && lengths != nullptr && x != nullptr,
                "qlt_scale_segments_noalloc: null array");
  hipStream_t st = as_stream(stream);
  QLT_CHECK_ARG(!stream_capturing(st), "qlt_scale_segments_noalloc: not capturable (stages host arrays)");
  int dev = -1;
  QLT_HIP(hipGetDevice(&dev));
  QLT_CHECK_ARG(dev >= 0 && dev < kMaxDevices && g_scratch_ptr[dev] != nullptr,
                "qlt_scale_segments_noalloc: qlt_init has not run on the current device");
  // the device's copy of the segment list; every entry the kernel reads is
  // written here first, so nothing left by an earlier call is ever read
  int64_t *d = g_scratch_ptr[dev];
  QLT_HIP(hipMemcpyAsync(d, offsets, n_segs * sizeof(int64_t), hipMemcpyHostToDevice, st));
  QLT_HIP(hipMemcpyAsync(d + kNoAllocMaxSegs, lengths, n_segs * sizeof(int64_t), hipMemcpyHostToDevice, st));
  launch_scale_segments(dtype, static_cast<unsigned>(n_segs), d, d + kNoAllocMaxSegs, s, x, st);
  QLT_HIP(hipGetLastError());
  // the segment list is reused by the next call: return only once this one is done
  QLT_HIP(hipStreamSynchronize(st));
  return QLT_OK;
}

}  // extern "C"
