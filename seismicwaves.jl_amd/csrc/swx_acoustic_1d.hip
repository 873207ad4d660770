// 1D acoustic engines, constant density (CD) and variable density (VD,
// staggered): one workgroup per shot runs a whole time loop (forward, or
// replay + adjoint + correlation) in one launch, with a workgroup barrier
// between the phases of a step.  A 1D grid holds a few thousand cells, so a
// launch per step would be all launch latency; here the fields stay in the
// CU's L1/L2 for the whole loop and the host issues one launch per loop.
//
// Reference semantics (paths relative to the reference root):
//   CD step     src/models/acoustic/backends/shared/acoustic1D_xPU.jl:1-125
//   VD step     src/models/acoustic/backends/shared/acoustic1D_VD_xPU.jl:1-141
//   drivers     src/models/acoustic/acou_forward.jl:22-125,
//               src/models/acoustic/acou_gradient.jl:4-203
//   correlate   src/models/acoustic/backends/shared/correlate_gradient_xPU.jl
//
// Device layout (per shot slabs, index s):
//   CD  p [3][ns][nx] (level k in slot k mod 3, as the 2D engine), psi
//       [ns][2h], xi [ns][2h+2]; adjoint the same; D [ns][S][nx] holds
//       p^{it} - 2 p^{it-1} + p^{it-2} of the segment being correlated
//   VD  p [ns][nx], vx [ns][nx-1], psi [ns][2h], xi [ns][2h+2]; adjoint the
//       same; H [ns][S+1][nx] holds p^{cb} .. p^{cb+S} of the segment
//   checkpoints at c = k S, k = 1..K-2, of the full state of every shot
//   grad [ns][nx] (CD, VD m0) and [ns][nx-1] (VD m1), reduced in shot order
#include <algorithm>

#include "swx_device.hpp"
#include "swx_internal.hpp"

namespace swx {
namespace {

constexpr int NT1 = 256;   // threads per workgroup (4 waves)

__device__ __forceinline__ int mod3_d(int k) { return ((k % 3) + 3) % 3; }

// Per-shot view of the point tables.
struct Pts1 {
  int nts;
  int64_t otf, otr;
  int us0, us1, ur0, ur1, r0, r1;
};
__device__ __forceinline__ Pts1 pts1(const PointArgsView &pt, int s) {
  Pts1 q;
  q.nts = pt.nt[s];
  q.otf = pt.off_tf[s];
  q.otr = pt.off_tr[s];
  q.us0 = pt.usrc_off[s]; q.us1 = pt.usrc_off[s + 1];
  q.ur0 = pt.urec_off[s]; q.ur1 = pt.urec_off[s + 1];
  q.r0 = pt.rec_off[s]; q.r1 = pt.rec_off[s + 1];
  return q;
}

// Injections of step it into f: per unique cell, the values added one by
// one in the reference's serial order (inject_sources!).  Residuals of the
// adjoint are prescaled by `pre` at the cell (prescale_residuals!).  The
// summed injection of a cell is added to hist (when non-null) as well.
template <typename T>
__device__ void inject1(T *f, const PointArgsView &pt, const Pts1 &q, bool adjoint, const T *tf, const T *pre,
                        int it, T *hist) {
  const int u0 = adjoint ? q.ur0 : q.us0, u1 = adjoint ? q.ur1 : q.us1;
  for (int u = u0 + (int)threadIdx.x; u < u1; u += NT1) {
    const int64_t c = adjoint ? pt.urec_cell[u] : pt.usrc_cell[u];
    const int k0 = adjoint ? pt.urec_ptr[u] : pt.usrc_ptr[u];
    const int k1 = adjoint ? pt.urec_ptr[u + 1] : pt.usrc_ptr[u + 1];
    const int *lst = adjoint ? pt.urec_list : pt.usrc_list;
    const T *base = tf + (adjoint ? q.otr : q.otf) + (it - 1);
    T v = f[c], inj = 0;
    for (int k = k0; k < k1; k++) {
      T x = base[(int64_t)q.nts * lst[k]];
      if (pre) x = x * pre[c];
      v += x;
      inj += x;
    }
    f[c] = v;
    if (hist) hist[c] += inj;
  }
}

template <typename T>
__device__ void record1(const T *f, const PointArgsView &pt, const Pts1 &q, T *traces, int it) {
  for (int r = q.r0 + (int)threadIdx.x; r < q.r1; r += NT1)
    traces[q.otr + (it - 1) + (int64_t)q.nts * (r - q.r0)] = f[pt.rec_cell[r]];
}

// ---------------------------------------------------------------------------
// CD 1D
// ---------------------------------------------------------------------------
template <typename T>
struct Cd1dArgs {
  int nx, h, ns, S, K;
  T idx, idx2;
  const T *fact;
  T *p;                 // [3][ns][nx]
  T *psi, *xi;          // [ns][2h], [ns][2h+2]
  T *a, *apsi, *axi;    // adjoint state
  T *grad;              // [ns][nx]
  T *dbuf;              // [ns][S][nx]
  T *ck;                // [K-2][ns][ck_one]
  int64_t ck_one;       // 2 nx + 2h + 2h + 2
  const T *cpml;        // [ns][8h+4]
  PointArgsView pt;
  const T *srctf;
  T *traces;            // nullable
  const T *res;
};

// One step of forward_onestep_CPML! / adjoint_onestep_CPML!
// (acoustic1D_xPU.jl:58-125) on the three levels (po, pc, pn).  With dst,
// stores p^{it} - 2 p^{it-1} + p^{it-2} per cell (the stencil term plus the
// injection in the interior).
template <typename T>
__device__ void cd1d_step(const Cd1dArgs<T> &A, const Pts1 &q, int s, int it, const T *po, const T *pc, T *pn,
                          T *psi, T *xi, bool adjoint, T *dst, T *traces) {
  const int nx = A.nx, h = A.h;
  const AxisCoef<T> cx = axis_coef(A.cpml + (int64_t)s * (8 * h + 4), h, 0);
  // update_psi! (:15-24): entry k of the strip belongs to grid cell ii
  for (int k = threadIdx.x; k < 2 * h; k += NT1) {
    const int i = k + 1;
    const int ii = i > h ? nx - h - 1 + (i - h) : i;   // 1-based
    const T d = (pc[ii] - pc[ii - 1]) * A.idx;
    psi[k] = cx.bh(k) * psi[k] + cx.ah(k) * d;
  }
  __syncthreads();
  // update_p_CPML! (:26-43) over 2:nx-1; the outermost cells keep level
  // it-2 (the reference's rotation binds pnew to pold's array, :89-92)
  for (int c = threadIdx.x; c < nx; c += NT1) {
    const T pov = po[c];
    if (c == 0 || c == nx - 1) {
      pn[c] = pov;
      if (dst) dst[c] = (pov - T(2) * pc[c]) + pov;
      continue;
    }
    const T p0 = pc[c];
    const T d2x = (pc[c - 1] - T(2) * p0 + pc[c + 1]) * A.idx2;
    T lx = d2x;
    if (c < h) {
      const T dpsi = (psi[c] - psi[c - 1]) * A.idx;
      const T x = cx.b(c) * xi[c] + cx.a(c) * (d2x + dpsi);
      xi[c] = x;
      lx = d2x + dpsi + x;
    } else if (c >= nx - h) {
      const int k = c - nx + 2 * h + 1;   // 0-based psi / xi entry
      const T dpsi = (psi[k] - psi[k - 1]) * A.idx;
      const T x = cx.b(k) * xi[k] + cx.a(k) * (d2x + dpsi);
      xi[k] = x;
      lx = d2x + dpsi + x;
    }
    const T flap = A.fact[c] * lx;
    pn[c] = T(2) * p0 - pov + flap;
    if (dst) dst[c] = flap;
  }
  __syncthreads();
  inject1(pn, A.pt, q, adjoint, adjoint ? A.res : A.srctf, adjoint ? A.fact : (const T *)nullptr, it, dst);
  if (traces) {
    __syncthreads();
    record1(pn, A.pt, q, traces, it);
  }
  __syncthreads();
}

template <typename T>
__device__ __forceinline__ T *cd1d_level(const Cd1dArgs<T> &A, T *base, int k, int s) {
  return base + ((int64_t)mod3_d(k) * A.ns + s) * A.nx;
}

// checkpoint of shot s at c = k S: p^{c-1} | p^c | psi | xi
template <typename T>
__device__ void cd1d_ckpt(const Cd1dArgs<T> &A, int s, int k, int cb, bool save) {
  T *ck = A.ck + ((int64_t)(k - 1) * A.ns + s) * A.ck_one;
  const int nx = A.nx, h = A.h;
  T *f0 = cd1d_level(A, A.p, cb - 1, s), *f1 = cd1d_level(A, A.p, cb, s);
  T *ps = A.psi + (int64_t)s * 2 * h, *xs = A.xi + (int64_t)s * (2 * h + 2);
  for (int c = threadIdx.x; c < A.ck_one; c += NT1) {
    T *f;
    if (c < nx) f = f0 + c;
    else if (c < 2 * nx) f = f1 + (c - nx);
    else if (c < 2 * nx + 2 * h) f = ps + (c - 2 * nx);
    else f = xs + (c - 2 * nx - 2 * h);
    if (save) ck[c] = *f;
    else *f = ck[c];
  }
  __syncthreads();
}

// Forward steps it0..it1 of every shot (blockIdx.x = shot).  With save,
// saves the checkpoints and the D history of each shot's last segment.
template <typename T>
__global__ __launch_bounds__(NT1) void cd1d_forward_kernel(Cd1dArgs<T> A, int it0, int it1, int save) {
  const int s = blockIdx.x;
  const Pts1 q = pts1(A.pt, s);
  const int last = min(it1, q.nts);
  const int seg0 = ((q.nts - 1) / A.S) * A.S;
  T *psi = A.psi + (int64_t)s * 2 * A.h, *xi = A.xi + (int64_t)s * (2 * A.h + 2);
  for (int it = it0; it <= last; it++) {
    T *dst = (save && it > seg0) ? A.dbuf + ((int64_t)s * A.S + (it - seg0 - 1)) * A.nx : nullptr;
    cd1d_step(A, q, s, it, cd1d_level(A, A.p, it - 2, s), cd1d_level(A, A.p, it - 1, s),
              cd1d_level(A, A.p, it, s), psi, xi, false, dst, A.traces);
    if (save && it % A.S == 0 && it / A.S >= 1 && it / A.S <= A.K - 2) cd1d_ckpt(A, s, it / A.S, it, true);
  }
}

// Replay (from the checkpoints), adjoint loop and correlation of every
// segment of shot blockIdx.x, last segment first.  grad holds
// sum_it adj^{it} D^{it} (scaled by 1/dt^2 by the caller).
template <typename T>
__global__ __launch_bounds__(NT1) void cd1d_adjoint_kernel(Cd1dArgs<T> A) {
  const int s = blockIdx.x;
  const Pts1 q = pts1(A.pt, s);
  const int nx = A.nx, h = A.h, S = A.S;
  T *psi = A.psi + (int64_t)s * 2 * h, *xi = A.xi + (int64_t)s * (2 * h + 2);
  T *apsi = A.apsi + (int64_t)s * 2 * h, *axi = A.axi + (int64_t)s * (2 * h + 2);
  T *g = A.grad + (int64_t)s * nx;
  T *D = A.dbuf + (int64_t)s * S * nx;
  const int Ks = (q.nts + S - 1) / S;
  for (int k = Ks - 1; k >= 0; k--) {
    const int cb = k * S;
    if (cb + S < q.nts) {
      // replay of segment k from its checkpoint (the zero state for k = 0)
      if (k >= 1) {
        cd1d_ckpt(A, s, k, cb, false);
      } else {
        for (int c = threadIdx.x; c < nx; c += NT1) {
          cd1d_level(A, A.p, -1, s)[c] = T(0);
          cd1d_level(A, A.p, 0, s)[c] = T(0);
        }
        for (int c = threadIdx.x; c < 2 * h; c += NT1) psi[c] = T(0);
        for (int c = threadIdx.x; c < 2 * h + 2; c += NT1) xi[c] = T(0);
        __syncthreads();
      }
      for (int it = cb + 1; it <= cb + S; it++)
        cd1d_step(A, q, s, it, cd1d_level(A, A.p, it - 2, s), cd1d_level(A, A.p, it - 1, s),
                  cd1d_level(A, A.p, it, s), psi, xi, false, D + (int64_t)(it - cb - 1) * nx, (T *)nullptr);
    }
    const int top = min(cb + S, q.nts);
    for (int it = top; it >= cb + 1; it--) {
      T *an = cd1d_level(A, A.a, it, s);
      cd1d_step(A, q, s, it, cd1d_level(A, A.a, it + 2, s), cd1d_level(A, A.a, it + 1, s), an, apsi, axi, true,
                (T *)nullptr, (T *)nullptr);
      // correlate_gradient! (correlate_gradient_xPU.jl:1-10)
      const T *d = D + (int64_t)(it - cb - 1) * nx;
      for (int c = threadIdx.x; c < nx; c += NT1) g[c] = g[c] + an[c] * d[c];
      __syncthreads();
    }
  }
}

// ---------------------------------------------------------------------------
// VD 1D
// ---------------------------------------------------------------------------
template <typename T>
struct Vd1dArgs {
  int nx, h, ns, S, K;
  T idx;
  const T *m0, *m1;     // [nx], [nx-1]
  T *p, *vx, *psi, *xi;       // [ns][nx], [ns][nx-1], [ns][2h], [ns][2h+2]
  T *ap, *avx, *apsi, *axi;   // adjoint state
  T *g0, *g1;           // [ns][nx], [ns][nx-1]
  T *hist;              // [ns][S+1][nx]
  T *ck;                // [K-2][ns][ck_one]
  int64_t ck_one;       // nx + nx-1 + 2h + 2h+2
  const T *cpml;
  PointArgsView pt;
  const T *srctf;
  T *traces;
  const T *res;
};

// 4th-order staggered first derivative (fdgen.jl, order 4) of the n-vector f
// with the four taps at b-1 .. b+2 (0-based), taps outside 0..n-1 dropped.
template <typename T>
__device__ __forceinline__ T d4_1(const T *f, int n, int b) {
  const T c1 = T(1.0 / 24.0), c2 = T(-9.0 / 8.0);
  if (b - 1 >= 0 && b + 2 <= n - 1) return c1 * f[b - 1] + c2 * f[b] + (-c2) * f[b + 1] + (-c1) * f[b + 2];
  T acc = (b - 1 >= 0 && b - 1 <= n - 1) ? c1 * f[b - 1] : T(0);
  acc = acc + ((b >= 0 && b <= n - 1) ? c2 * f[b] : T(0));
  acc = acc + ((b + 1 >= 0 && b + 1 <= n - 1) ? (-c2) * f[b + 1] : T(0));
  acc = acc + ((b + 2 >= 0 && b + 2 <= n - 1) ? (-c1) * f[b + 2] : T(0));
  return acc;
}

// update_p_CPML! (acoustic1D_VD_xPU.jl:15-32) over 2:nx-1
template <typename T>
__device__ void vd1d_update_p(const Vd1dArgs<T> &A, const AxisCoef<T> &cx, T *p, const T *vx, T *xi) {
  const int nx = A.nx, h = A.h;
  for (int c = 1 + threadIdx.x; c <= nx - 2; c += NT1) {
    T dv = d4_1(vx, nx - 1, c - 1) * A.idx;
    if (c <= h) {
      const T x = cx.b(c) * xi[c] + cx.a(c) * dv;
      xi[c] = x;
      dv = dv + x;
    } else if (c >= nx - h - 1) {
      const int k = c - nx + 2 * h + 2;
      const T x = cx.b(k) * xi[k] + cx.a(k) * dv;
      xi[k] = x;
      dv = dv + x;
    }
    p[c] -= A.m0[c] * dv;
  }
}

// update_vx_CPML! (acoustic1D_VD_xPU.jl:34-51) over 1:nx-1
template <typename T>
__device__ void vd1d_update_v(const Vd1dArgs<T> &A, const AxisCoef<T> &cx, const T *p, T *vx, T *psi) {
  const int nx = A.nx, h = A.h;
  for (int c = threadIdx.x; c <= nx - 2; c += NT1) {
    T d = d4_1(p, nx, c) * A.idx;
    if (c < h) {
      const T x = cx.bh(c) * psi[c] + cx.ah(c) * d;
      psi[c] = x;
      d = d + x;
    } else if (c >= nx - h - 1) {
      const int k = c - nx + 2 * h + 1;
      const T x = cx.bh(k) * psi[k] + cx.ah(k) * d;
      psi[k] = x;
      d = d + x;
    }
    vx[c] -= A.m1[c] * d;
  }
}

// forward_onestep_CPML! (acoustic1D_VD_xPU.jl:66-99); with hp, stores p^{it}
template <typename T>
__device__ void vd1d_fwd_step(const Vd1dArgs<T> &A, const Pts1 &q, int s, int it, T *p, T *vx, T *psi, T *xi,
                              T *hp, T *traces) {
  const AxisCoef<T> cx = axis_coef(A.cpml + (int64_t)s * (8 * A.h + 4), A.h, 0);
  vd1d_update_p(A, cx, p, vx, xi);
  __syncthreads();
  inject1(p, A.pt, q, false, A.srctf, (const T *)nullptr, it, (T *)nullptr);
  __syncthreads();
  vd1d_update_v(A, cx, p, vx, psi);
  if (traces) record1(p, A.pt, q, traces, it);
  if (hp)
    for (int c = threadIdx.x; c < A.nx; c += NT1) hp[c] = p[c];
  __syncthreads();
}

// the full state of shot s <-> checkpoint k
template <typename T>
__device__ void vd1d_ckpt(const Vd1dArgs<T> &A, int s, int k, bool save) {
  const int nx = A.nx, h = A.h;
  T *ck = A.ck + ((int64_t)(k - 1) * A.ns + s) * A.ck_one;
  T *p = A.p + (int64_t)s * nx, *vx = A.vx + (int64_t)s * (nx - 1);
  T *ps = A.psi + (int64_t)s * 2 * h, *xs = A.xi + (int64_t)s * (2 * h + 2);
  for (int c = threadIdx.x; c < A.ck_one; c += NT1) {
    T *f;
    if (c < nx) f = p + c;
    else if (c < 2 * nx - 1) f = vx + (c - nx);
    else if (c < 2 * nx - 1 + 2 * h) f = ps + (c - (2 * nx - 1));
    else f = xs + (c - (2 * nx - 1 + 2 * h));
    if (save) ck[c] = *f;
    else *f = ck[c];
  }
  __syncthreads();
}

template <typename T>
__global__ __launch_bounds__(NT1) void vd1d_forward_kernel(Vd1dArgs<T> A, int it0, int it1, int save) {
  const int s = blockIdx.x;
  const Pts1 q = pts1(A.pt, s);
  const int nx = A.nx, h = A.h;
  const int last = min(it1, q.nts);
  const int seg0 = ((q.nts - 1) / A.S) * A.S;
  T *p = A.p + (int64_t)s * nx, *vx = A.vx + (int64_t)s * (nx - 1);
  T *psi = A.psi + (int64_t)s * 2 * h, *xi = A.xi + (int64_t)s * (2 * h + 2);
  T *H = A.hist + (int64_t)s * (A.S + 1) * nx;
  // level seg0 opens the history of the last segment (the zero initial
  // state when seg0 = 0)
  if (save && seg0 == 0 && it0 == 1)
    for (int c = threadIdx.x; c < nx; c += NT1) H[c] = T(0);
  for (int it = it0; it <= last; it++) {
    T *hp = (save && it > seg0) ? H + (int64_t)(it - seg0) * nx : nullptr;
    vd1d_fwd_step(A, q, s, it, p, vx, psi, xi, hp, A.traces);
    if (save && it == seg0) {
      for (int c = threadIdx.x; c < nx; c += NT1) H[c] = p[c];
      __syncthreads();
    }
    if (save && it % A.S == 0 && it / A.S >= 1 && it / A.S <= A.K - 2) vd1d_ckpt(A, s, it / A.S, true);
  }
}

// Replay, adjoint loop and the m0 / m1 correlations of shot blockIdx.x.
// g0 holds -sum adj_p (p^{it} - p^{it-1}) (scaled by 1/dt by the caller),
// g1 sum adj_vx d~p^{it}/dx over 2:nx-2 (acoustic1D_VD_xPU.jl:130-141).
template <typename T>
__global__ __launch_bounds__(NT1) void vd1d_adjoint_kernel(Vd1dArgs<T> A) {
  const int s = blockIdx.x;
  const Pts1 q = pts1(A.pt, s);
  const int nx = A.nx, h = A.h, S = A.S;
  const AxisCoef<T> cx = axis_coef(A.cpml + (int64_t)s * (8 * h + 4), h, 0);
  T *p = A.p + (int64_t)s * nx, *vx = A.vx + (int64_t)s * (nx - 1);
  T *psi = A.psi + (int64_t)s * 2 * h, *xi = A.xi + (int64_t)s * (2 * h + 2);
  T *ap = A.ap + (int64_t)s * nx, *avx = A.avx + (int64_t)s * (nx - 1);
  T *apsi = A.apsi + (int64_t)s * 2 * h, *axi = A.axi + (int64_t)s * (2 * h + 2);
  T *g0 = A.g0 + (int64_t)s * nx, *g1 = A.g1 + (int64_t)s * (nx - 1);
  T *H = A.hist + (int64_t)s * (S + 1) * nx;
  const int Ks = (q.nts + S - 1) / S;
  for (int k = Ks - 1; k >= 0; k--) {
    const int cb = k * S;
    if (cb + S < q.nts) {
      if (k >= 1) {
        vd1d_ckpt(A, s, k, false);
      } else {
        for (int c = threadIdx.x; c < nx; c += NT1) p[c] = T(0);
        for (int c = threadIdx.x; c < nx - 1; c += NT1) vx[c] = T(0);
        for (int c = threadIdx.x; c < 2 * h; c += NT1) psi[c] = T(0);
        for (int c = threadIdx.x; c < 2 * h + 2; c += NT1) xi[c] = T(0);
        __syncthreads();
      }
      for (int c = threadIdx.x; c < nx; c += NT1) H[c] = p[c];
      __syncthreads();
      for (int it = cb + 1; it <= cb + S; it++)
        vd1d_fwd_step(A, q, s, it, p, vx, psi, xi, H + (int64_t)(it - cb) * nx, (T *)nullptr);
    }
    const int top = min(cb + S, q.nts);
    for (int it = top; it >= cb + 1; it--) {
      // adjoint_onestep_CPML! (acoustic1D_VD_xPU.jl:101-128)
      vd1d_update_v(A, cx, ap, avx, apsi);
      __syncthreads();
      vd1d_update_p(A, cx, ap, avx, axi);
      __syncthreads();
      inject1(ap, A.pt, q, true, A.res, A.m0, it, (T *)nullptr);
      __syncthreads();
      // correlate_gradient_m0! and correlate_gradient_m1!
      const T *pc = H + (int64_t)(it - cb) * nx, *po = pc - nx;
      for (int c = threadIdx.x; c < nx; c += NT1) g0[c] = g0[c] - ap[c] * (pc[c] - po[c]);
      for (int c = 1 + threadIdx.x; c <= nx - 3; c += NT1) g1[c] = g1[c] + avx[c] * (d4_1(pc, nx, c) * A.idx);
      __syncthreads();
    }
  }
}

// ---------------------------------------------------------------------------
// Engines
// ---------------------------------------------------------------------------
template <typename T>
class Acoustic1dEngine final : public Engine {
 public:
  explicit Acoustic1dEngine(PlanCore &core) : c_(core) {
    const swx_plan_desc &d = c_.d;
    vd_ = d.physics == SWX_ACOUSTIC_VD;
    nx_ = (int)d.n[0];
    h_ = d.halo;
    ns_ = d.max_shots;
    cstride_ = 8 * h_ + 4;
    const int64_t e = sizeof(T);
    const int64_t nsp = std::max(2 * h_, 1), nxi = 2 * h_ + 2;
    state_bytes_ = 0;
    auto al = [&](DevBuf &b, int64_t elems) {
      b.alloc(std::max<int64_t>(elems, 1) * e);
      state_bytes_ += b.bytes();
    };
    al(m0_, nx_);
    if (vd_) al(m1_, nx_ - 1);
    al(p_, (vd_ ? 1 : 3) * (int64_t)ns_ * nx_);
    if (vd_) al(vx_, (int64_t)ns_ * (nx_ - 1));
    al(psi_, ns_ * nsp);
    al(xi_, ns_ * nxi);
    cpml_.alloc((int64_t)ns_ * cstride_ * e);
    if (d.gradient) {
      al(a_, (vd_ ? 1 : 3) * (int64_t)ns_ * nx_);
      if (vd_) al(avx_, (int64_t)ns_ * (nx_ - 1));
      al(apsi_, ns_ * nsp);
      al(axi_, ns_ * nxi);
      al(g0_, (int64_t)ns_ * nx_);
      al(gout0_, (int64_t)ns_ * nx_);
      if (vd_) {
        al(g1_, (int64_t)ns_ * (nx_ - 1));
        al(gout1_, (int64_t)ns_ * (nx_ - 1));
      }
      ck_one_ = vd_ ? (2 * (int64_t)nx_ - 1 + 4 * h_ + 2) : (2 * (int64_t)nx_ + 4 * h_ + 2);
      const int64_t slot = (int64_t)ns_ * nx_ * e;
      const int64_t avail = (int64_t)c_.budget - state_bytes_ - (vd_ ? slot : 0);
      S_ = choose_segment(d.nt_max, d.check_freq, slot, (int64_t)ns_ * ck_one_ * e, avail);
      K_ = (int)ceil_div(d.nt_max, S_);
      hist_.alloc((int64_t)(vd_ ? S_ + 1 : S_) * slot);
      ck_.alloc(std::max<int64_t>(std::max(K_ - 2, 0) * (int64_t)ns_ * ck_one_ * e, 1));
    }
  }

  void set_model(int nfields, const void *const *fields) override {
    SWX_REQUIRE(nfields == (vd_ ? 2 : 1) && fields[0] && (!vd_ || fields[1]), SWX_EINVAL,
                vd_ ? "acoustic VD 1D expects two model fields (fact_m0, fact_m1_x)"
                    : "acoustic CD expects one model field (fact)");
    SWX_HIP(hipMemcpyAsync(m0_.get(), fields[0], nx_ * sizeof(T), hipMemcpyHostToDevice, c_.stream));
    if (vd_) SWX_HIP(hipMemcpyAsync(m1_.get(), fields[1], (nx_ - 1) * sizeof(T), hipMemcpyHostToDevice, c_.stream));
    SWX_HIP(hipStreamSynchronize(c_.stream));
    model_set_ = true;
  }

  void on_shots_set(int nshots, const swx_shot_desc *shots) override {
    for (int s = 0; s < nshots; s++)
      SWX_REQUIRE(shots[s].nt <= c_.d.nt_max, SWX_EINVAL, "shot nt exceeds nt_max");
    const std::vector<T> hc = pack_cpml<T>(nshots, shots, 1, h_);
    SWX_HIP(hipMemcpyAsync(cpml_.get(), hc.data(), hc.size() * sizeof(T), hipMemcpyHostToDevice, c_.stream));
    SWX_HIP(hipStreamSynchronize(c_.stream));
    nsh_ = nshots;
    ntmax_ = *std::max_element(c_.pts.nt.begin(), c_.pts.nt.end());
  }

  void forward(bool save) override {
    SWX_REQUIRE(model_set_, SWX_ESTATE, "swx_set_model must be called before running");
    zero(p_); zero(vx_); zero(psi_); zero(xi_);
    // one launch for the whole loop; with snapshots, one per snapshot interval
    const int chunk = c_.snap_every > 0 ? c_.snap_every : ntmax_;
    for (int it0 = 1; it0 <= ntmax_; it0 += chunk) {
      const int it1 = std::min(it0 + chunk - 1, ntmax_);
      progress("forward step", it1, ntmax_);
      c_.ktimer.begin(c_.stream);
      if (vd_) vd1d_forward_kernel<T><<<nsh_, NT1, 0, c_.stream>>>(vd_args(), it0, it1, save && c_.d.gradient);
      else cd1d_forward_kernel<T><<<nsh_, NT1, 0, c_.stream>>>(cd_args(), it0, it1, save && c_.d.gradient);
      c_.ktimer.end(c_.stream, 0);
      SWX_HIP(hipGetLastError());
      if (c_.snap_every > 0) take_snapshots(c_, *this, it1);
    }
  }

  // "pcur" (CD: level it in slot it mod 3), "pcur" and "vcur" (VD)
  int snap_fields(int it, int s, SnapPiece *out) const override {
    const int64_t e = sizeof(T);
    if (!vd_) {
      out[0] = {p_.as<T>() + ((int64_t)mod3(it) * ns_ + s) * nx_, nx_ * e, nx_ * e, 1};
      return 1;
    }
    out[0] = {p_.as<T>() + (int64_t)s * nx_, nx_ * e, nx_ * e, 1};
    out[1] = {vx_.as<T>() + (int64_t)s * (nx_ - 1), (nx_ - 1) * e, (nx_ - 1) * e, 1};
    return 2;
  }

  void adjoint(bool per_shot) override {
    zero(a_); zero(avx_); zero(apsi_); zero(axi_); zero(g0_); zero(g1_);
    progress("adjoint", 1, 1);
    c_.ktimer.begin(c_.stream);
    if (vd_) vd1d_adjoint_kernel<T><<<nsh_, NT1, 0, c_.stream>>>(vd_args());
    else cd1d_adjoint_kernel<T><<<nsh_, NT1, 0, c_.stream>>>(cd_args());
    c_.ktimer.end(c_.stream, 1);
    SWX_HIP(hipGetLastError());
    const int ps = per_shot ? 1 : 0;
    dim3 g((unsigned)std::max<int64_t>(std::min<int64_t>(ceil_div(nx_, 256), 1024), 1));
    if (vd_) {
      reduce_shots<T><<<g, 256, 0, c_.stream>>>(gout0_.as<T>(), g0_.as<T>(), nx_, nsh_, T(1.0 / c_.d.dt), ps);
      reduce_shots<T><<<g, 256, 0, c_.stream>>>(gout1_.as<T>(), g1_.as<T>(), nx_ - 1, nsh_, T(1), ps);
    } else {
      reduce_shots<T><<<g, 256, 0, c_.stream>>>(gout0_.as<T>(), g0_.as<T>(), nx_, nsh_,
                                                T(1.0 / (c_.d.dt * c_.d.dt)), ps);
    }
    SWX_HIP(hipGetLastError());
  }

  // CD: grad (nx) per shot; VD: grad_m0 (nx) | grad_m1_x (nx-1) per shot
  void copy_grad(void *host, bool per_shot) override {
    const int n = per_shot ? nsh_ : 1;
    char *dst = static_cast<char *>(host);
    const int64_t e = sizeof(T);
    for (int s = 0; s < n; s++) {
      SWX_HIP(hipMemcpy(dst, gout0_.as<T>() + (int64_t)s * nx_, nx_ * e, hipMemcpyDeviceToHost));
      dst += nx_ * e;
      if (vd_) {
        SWX_HIP(hipMemcpy(dst, gout1_.as<T>() + (int64_t)s * (nx_ - 1), (nx_ - 1) * e, hipMemcpyDeviceToHost));
        dst += (nx_ - 1) * e;
      }
    }
  }

  int64_t bytes_state() const override { return state_bytes_; }
  int64_t bytes_checkpoints() const override { return ck_.bytes(); }
  int64_t bytes_replay() const override { return hist_.bytes(); }
  int segment() const override { return S_; }

 private:
  void zero(DevBuf &b) {
    if (b.bytes()) SWX_HIP(hipMemsetAsync(b.get(), 0, b.bytes(), c_.stream));
  }

  Cd1dArgs<T> cd_args() const {
    Cd1dArgs<T> a{};
    a.nx = nx_; a.h = h_; a.ns = ns_; a.S = S_; a.K = K_;
    a.idx = T(1.0 / c_.d.h[0]);
    a.idx2 = a.idx * a.idx;
    a.fact = m0_.as<T>();
    a.p = p_.as<T>(); a.psi = psi_.as<T>(); a.xi = xi_.as<T>();
    a.a = a_.as<T>(); a.apsi = apsi_.as<T>(); a.axi = axi_.as<T>();
    a.grad = g0_.as<T>();
    a.dbuf = hist_.as<T>();
    a.ck = ck_.as<T>(); a.ck_one = ck_one_;
    a.cpml = cpml_.as<T>();
    a.pt = view(c_.pts);
    a.srctf = c_.pts.d_srctf.as<T>();
    a.traces = c_.pts.d_traces.as<T>();
    a.res = c_.pts.d_res.as<T>();
    return a;
  }

  Vd1dArgs<T> vd_args() const {
    Vd1dArgs<T> a{};
    a.nx = nx_; a.h = h_; a.ns = ns_; a.S = S_; a.K = K_;
    a.idx = T(1.0 / c_.d.h[0]);
    a.m0 = m0_.as<T>(); a.m1 = m1_.as<T>();
    a.p = p_.as<T>(); a.vx = vx_.as<T>(); a.psi = psi_.as<T>(); a.xi = xi_.as<T>();
    a.ap = a_.as<T>(); a.avx = avx_.as<T>(); a.apsi = apsi_.as<T>(); a.axi = axi_.as<T>();
    a.g0 = g0_.as<T>(); a.g1 = g1_.as<T>();
    a.hist = hist_.as<T>();
    a.ck = ck_.as<T>(); a.ck_one = ck_one_;
    a.cpml = cpml_.as<T>();
    a.pt = view(c_.pts);
    a.srctf = c_.pts.d_srctf.as<T>();
    a.traces = c_.pts.d_traces.as<T>();
    a.res = c_.pts.d_res.as<T>();
    return a;
  }

  PlanCore &c_;
  bool vd_ = false;
  int nx_, h_, ns_, cstride_;
  DevBuf m0_, m1_, p_, vx_, psi_, xi_, cpml_;
  DevBuf a_, avx_, apsi_, axi_, g0_, g1_, gout0_, gout1_, hist_, ck_;
  int64_t ck_one_ = 0, state_bytes_ = 0;
  int S_ = 1, K_ = 1;
  int nsh_ = 0, ntmax_ = 0;
  bool model_set_ = false;
};

}  // namespace

std::unique_ptr<Engine> make_acoustic_1d(PlanCore &core) {
  if (core.d.dtype == SWX_F32) return std::make_unique<Acoustic1dEngine<float>>(core);
  return std::make_unique<Acoustic1dEngine<double>>(core);
}

}  // namespace swx
