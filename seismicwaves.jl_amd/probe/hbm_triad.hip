// HBM bandwidth probe: STREAM triad a = b + s * c on float4 vectors, timed
// with HIP events.  It measures the practical read+write ceiling of the
// device that the step kernels' roofline fractions are read against; it is
// not part of the library's ABI (bench.py --hbm-triad loads it).
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdint>
#include <cstdio>

namespace {

__global__ __launch_bounds__(256) void triad(float4 *__restrict__ a, const float4 *__restrict__ b,
                                             const float4 *__restrict__ c, float s, int64_t n4) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n4; q += (int64_t)gridDim.x * blockDim.x) {
    const float4 x = b[q], y = c[q];
    a[q] = make_float4(x.x + s * y.x, x.y + s * y.y, x.z + s * y.z, x.w + s * y.w);
  }
}

// The same triad with non-temporal stores (NT & 1) and non-temporal loads
// (NT & 2), on clang vector types.
typedef float v4f __attribute__((ext_vector_type(4)));
template <int NT>
__global__ __launch_bounds__(256) void triad_nt(v4f *__restrict__ a, const v4f *__restrict__ b,
                                                const v4f *__restrict__ c, float s, int64_t n4) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n4; q += (int64_t)gridDim.x * blockDim.x) {
    const v4f x = (NT & 2) ? __builtin_nontemporal_load(b + q) : b[q];
    const v4f y = (NT & 2) ? __builtin_nontemporal_load(c + q) : c[q];
    const v4f r = x + s * y;
    if (NT & 1) __builtin_nontemporal_store(r, a + q);
    else a[q] = r;
  }
}

}  // namespace

// Runs `reps` triads over three arrays of `bytes_each` bytes with `blocks`
// blocks of 256 threads (0: 8 per CU) and writes the best and the mean
// bandwidth (GB/s, 3 * bytes_each per triad) to out[0], out[1].  `mode`:
// 0 plain loads and stores, 1 non-temporal stores, 3 non-temporal loads and
// stores.  Returns 0, or a HIP error code.
extern "C" int swx_probe_hbm_triad_mode(int64_t bytes_each, int reps, int blocks, int mode, double *out);
extern "C" int swx_probe_hbm_triad(int64_t bytes_each, int reps, int blocks, double *out) {
  return swx_probe_hbm_triad_mode(bytes_each, reps, blocks, 0, out);
}

extern "C" int swx_probe_hbm_triad_mode(int64_t bytes_each, int reps, int blocks, int mode, double *out) {
  int dev = 0, ncu = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 1;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 1;
  const int64_t n4 = bytes_each / 16;
  float4 *a = nullptr, *b = nullptr, *c = nullptr;
  if (hipMalloc(&a, n4 * 16) != hipSuccess) return 2;
  if (hipMalloc(&b, n4 * 16) != hipSuccess) { (void)hipFree(a); return 2; }
  if (hipMalloc(&c, n4 * 16) != hipSuccess) { (void)hipFree(a); (void)hipFree(b); return 2; }
  (void)hipMemset(b, 0, n4 * 16);
  (void)hipMemset(c, 0, n4 * 16);
  const unsigned g = blocks > 0 ? (unsigned)blocks : (unsigned)(8 * ncu);
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  auto launch = [&]() {
    v4f *va = reinterpret_cast<v4f *>(a);
    const v4f *vb = reinterpret_cast<const v4f *>(b), *vc = reinterpret_cast<const v4f *>(c);
    if (mode == 1) triad_nt<1><<<g, 256>>>(va, vb, vc, 3.0f, n4);
    else if (mode == 3) triad_nt<3><<<g, 256>>>(va, vb, vc, 3.0f, n4);
    else triad<<<g, 256>>>(a, b, c, 3.0f, n4);
  };
  launch();   // warm-up
  double best = 0, sum = 0;
  for (int r = 0; r < reps; r++) {
    (void)hipEventRecord(e0, nullptr);
    launch();
    (void)hipEventRecord(e1, nullptr);
    (void)hipEventSynchronize(e1);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    const double gbs = 3.0 * (double)(n4 * 16) / (ms * 1e-3) / 1e9;
    best = gbs > best ? gbs : best;
    sum += gbs;
  }
  const int err = hipGetLastError() != hipSuccess ? 3 : 0;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(b);
  (void)hipFree(c);
  out[0] = best;
  out[1] = reps > 0 ? sum / reps : 0;
  return err;
}

// Launch-gap probe: a chain of dependent three-point smoothing passes over
// `n` floats (pass i reads buffer i mod 2 and writes the other), launched
// back to back on one stream, and the same chain captured once into a HIP
// graph and replayed.  It measures what a graph of a time loop would save
// against stream launches for kernels of the step kernels' size.
namespace {
__global__ __launch_bounds__(256) void smooth3(float *__restrict__ dst, const float *__restrict__ src, int64_t n) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float l = src[i > 0 ? i - 1 : i], r = src[i + 1 < n ? i + 1 : i];
  dst[i] = 0.5f * src[i] + 0.25f * (l + r);
}
}  // namespace

// out[0]: best us per pass over `reps` stream-launched chains of `launches`
// passes; out[1]: the same for replays of the captured chain; out[2]: best
// us of one pass timed alone between events; out[3]: ms to capture and
// instantiate the graph.  Returns 0, or a HIP error code.
extern "C" int swx_probe_launch_chain(int64_t n, int launches, int reps, double *out) {
  if (n <= 0 || launches <= 0 || reps <= 0) return 4;
  float *buf[2] = {nullptr, nullptr};
  if (hipMalloc(&buf[0], n * 4) != hipSuccess) return 2;
  if (hipMalloc(&buf[1], n * 4) != hipSuccess) { (void)hipFree(buf[0]); return 2; }
  (void)hipMemset(buf[0], 0, n * 4);
  (void)hipMemset(buf[1], 0, n * 4);
  hipStream_t st;
  if (hipStreamCreate(&st) != hipSuccess) { (void)hipFree(buf[0]); (void)hipFree(buf[1]); return 1; }
  const unsigned g = (unsigned)((n + 255) / 256);
  auto chain = [&]() {
    for (int i = 0; i < launches; i++) smooth3<<<g, 256, 0, st>>>(buf[(i + 1) & 1], buf[i & 1], n);
  };
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  auto timed = [&](auto &&body) {
    (void)hipEventRecord(e0, st);
    body();
    (void)hipEventRecord(e1, st);
    (void)hipEventSynchronize(e1);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    return (double)ms;
  };
  int err = 0;
  chain();   // warm-up
  (void)hipStreamSynchronize(st);
  double one = 1e30, str = 1e30, gra = 1e30;
  for (int r = 0; r < 20; r++) {
    const double ms = timed([&] { smooth3<<<g, 256, 0, st>>>(buf[1], buf[0], n); });
    one = ms < one ? ms : one;
  }
  for (int r = 0; r < reps; r++) {
    const double ms = timed(chain);
    str = ms < str ? ms : str;
  }
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  const auto t0 = std::chrono::steady_clock::now();
  if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess) err = 5;
  if (!err) {
    chain();
    if (hipStreamEndCapture(st, &graph) != hipSuccess) err = 5;
  }
  if (!err && hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) err = 6;
  const double inst_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (!err) {
    (void)hipGraphLaunch(exec, st);   // warm-up
    (void)hipStreamSynchronize(st);
    for (int r = 0; r < reps; r++) {
      const double ms = timed([&] { (void)hipGraphLaunch(exec, st); });
      gra = ms < gra ? ms : gra;
    }
  }
  if (!err && hipGetLastError() != hipSuccess) err = 3;
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipStreamDestroy(st);
  (void)hipFree(buf[0]);
  (void)hipFree(buf[1]);
  out[0] = str * 1e3 / launches;
  out[1] = err ? 0.0 : gra * 1e3 / launches;
  out[2] = one * 1e3;
  out[3] = inst_ms;
  return err;
}
