// L2 misfit and its adjoint source on the device, for the batch of a plan.
//
// Reference: src/inversion/misfits/L2Misfit.jl:24-95 (calcmisfit and
// ∂χ_∂u with windows and a diagonal inverse covariance) and the driver's
// residual step, src/models/acoustic/acou_gradient.jl:45-51 (the adjoint
// source is -∂χ/∂u).  The residual is formed in double precision, exactly
// as the host path of the Python API does (float64 residual, window mask
// and inverse covariance, then rounded to the plan's precision), so the
// adjoint source, and with it the gradient, is bitwise the host path's.
#include "swx_device.hpp"
#include "swx_internal.hpp"

namespace swx {
namespace {

constexpr int kMisfitThreads = 256;
constexpr int kMisfitBlocks = 128;   // partial sums per shot

// res = -(syn - obs) * w over the whole batch (w == nullptr: identity).
template <typename T>
__global__ void l2_adjoint_source_kernel(T *res, const T *syn, const T *obs, const double *w, int64_t n) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
    double r = (double)syn[q] - (double)obs[q];
    if (w) r *= w[q];
    res[q] = (T)(-r);
  }
}

// Partial sums of w r^2 per shot: block b of shot s covers the samples
// off[s] + b*256 + t + k*(128*256); a fixed tree in LDS, so the sum does not
// depend on scheduling.
template <typename T>
__global__ __launch_bounds__(kMisfitThreads) void l2_partial_kernel(double *part, const T *syn, const T *obs,
                                                                     const double *w, const int64_t *off) {
  __shared__ double acc[kMisfitThreads];
  const int s = blockIdx.y;
  const int64_t q0 = off[s], q1 = off[s + 1];
  double v = 0;
  for (int64_t q = q0 + (int64_t)blockIdx.x * kMisfitThreads + threadIdx.x; q < q1;
       q += (int64_t)kMisfitBlocks * kMisfitThreads) {
    const double r = (double)syn[q] - (double)obs[q];
    v += w ? r * (r * w[q]) : r * r;
  }
  acc[threadIdx.x] = v;
  __syncthreads();
  for (int k = kMisfitThreads / 2; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) acc[threadIdx.x] += acc[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[(int64_t)s * kMisfitBlocks + blockIdx.x] = acc[0];
}

__global__ void l2_finish_kernel(double *mis, const double *part, int ns) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= ns) return;
  double v = 0;
  for (int b = 0; b < kMisfitBlocks; b++) v += part[(int64_t)s * kMisfitBlocks + b];
  mis[s] = 0.5 * v;
}

template <typename T>
void l2_run(PlanCore &c, double *misfit_host) {
  const PointTables &p = c.pts;
  const int64_t n = p.total_tr;
  const T *syn = p.d_traces.as<T>();
  const T *obs = c.d_obs.as<T>();
  const double *w = c.obs_weighted ? c.d_obs_w.as<double>() : nullptr;
  const unsigned nb = (unsigned)std::min<int64_t>(std::max<int64_t>(ceil_div(n, kMisfitThreads), 1), 4096);
  l2_adjoint_source_kernel<T><<<nb, kMisfitThreads, 0, c.stream>>>(p.d_res.as<T>(), syn, obs, w, n);
  SWX_HIP(hipGetLastError());
  if (misfit_host) {
    l2_partial_kernel<T><<<dim3(kMisfitBlocks, p.nshots), kMisfitThreads, 0, c.stream>>>(
        c.d_obs_part.as<double>(), syn, obs, w, c.d_obs_off.as<int64_t>());
    SWX_HIP(hipGetLastError());
    l2_finish_kernel<<<(unsigned)ceil_div(p.nshots, 64), 64, 0, c.stream>>>(
        c.d_obs_part.as<double>() + (int64_t)p.nshots * kMisfitBlocks, c.d_obs_part.as<double>(), p.nshots);
    SWX_HIP(hipGetLastError());
    SWX_HIP(hipMemcpyAsync(misfit_host, c.d_obs_part.as<double>() + (int64_t)p.nshots * kMisfitBlocks,
                           sizeof(double) * p.nshots, hipMemcpyDeviceToHost, c.stream));
  }
}

}  // namespace

void set_observed(PlanCore &c, const void *observed, const double *weights) {
  const PointTables &p = c.pts;
  const int64_t n = p.total_tr;
  c.obs_set = false;
  c.d_obs.alloc(std::max<int64_t>(n, 1) * c.esize);
  SWX_HIP(hipMemcpyAsync(c.d_obs.get(), observed, (size_t)n * c.esize, hipMemcpyHostToDevice, c.stream));
  c.obs_weighted = weights != nullptr;
  if (weights) {
    c.d_obs_w.alloc(std::max<int64_t>(n, 1) * sizeof(double));
    SWX_HIP(hipMemcpyAsync(c.d_obs_w.get(), weights, (size_t)n * sizeof(double), hipMemcpyHostToDevice, c.stream));
  } else {
    c.d_obs_w.release();
  }
  std::vector<int64_t> off(p.nshots + 1);
  for (int s = 0; s < p.nshots; s++) off[s] = p.off_tr[s];
  off[p.nshots] = n;
  c.d_obs_off.alloc(off.size() * sizeof(int64_t));
  SWX_HIP(hipMemcpyAsync(c.d_obs_off.get(), off.data(), off.size() * sizeof(int64_t), hipMemcpyHostToDevice,
                         c.stream));
  c.d_obs_part.alloc((size_t)p.nshots * (kMisfitBlocks + 1) * sizeof(double));
  SWX_HIP(hipStreamSynchronize(c.stream));
  c.obs_set = true;
}

void l2_adjoint_source(PlanCore &c, double *misfit_host) {
  if (c.esize == 4) l2_run<float>(c, misfit_host);
  else l2_run<double>(c, misfit_host);
}

}  // namespace swx
