// HBM bandwidth probe: STREAM triad a = b + s * c on float4 vectors, timed
// with HIP events.  It measures the practical read+write ceiling of the
// device that the step kernels' roofline fractions are read against; it is
// not part of the library's ABI (bench.py --hbm-triad loads it).
#include <hip/hip_runtime.h>

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
