// 2D acoustic variable-density engine (staggered pressure / velocity,
// 4th-order space, C-PML) for a batch of shots: ONE fused launch per time
// step for the forward/replay and one for the adjoint (+ gradient
// correlation).
//
// Reference semantics (paths relative to the reference root):
//   step       src/models/acoustic/backends/shared/acoustic2D_VD_xPU.jl:1-178
//   stencils   src/utils/fdgen.jl:65-161 (order 4, zero outside the array)
//   driver     src/models/acoustic/acou_gradient.jl:222-327
//   correlate  correlate_gradient_xPU.jl:12-21, acoustic2D_VD_xPU.jl:180-198
//
// Fusion.  The forward step is p <- p - m0 div~ v, inject, v <- v - m1 grad~ p.
// The velocity update needs the *new* pressure one cell left / two right of
// each velocity point, so a block recomputes the new pressure on its tile
// plus that halo in LDS (redundant work on ~20% of the cells, no extra HBM
// traffic), then updates the velocities of its tile.  Because neighbouring
// blocks read the old state around their tiles, the state is double
// buffered by step parity (read parity it-1, write parity it), including
// the C-PML memory variables.  The adjoint step (v first, then p) is fused
// the same way around the old adjoint pressure, and the correlation of
// grad_m0 / grad_m1 with the replayed pressure is done in the same launch.
//
// Layout: every field is an nx*nz slab per shot with x fastest.  vx has
// nx-1 valid columns and vz nx*(nz-1) valid rows; the extra column/row is
// kept at zero, which is exactly the zero extension of the reference's
// truncated stencils.
#include <algorithm>

#include "swx_device.hpp"
#include "swx_internal.hpp"

namespace swx {
namespace {

constexpr int TX = 64;   // tile width  (one wavefront)
constexpr int TZ = 16;   // tile height (4 rows of threads x 4)
constexpr int BY = 4;    // thread rows per block
constexpr int NT = TX * BY;

template <typename T>
struct Vd2dArgs {
  int nx, nz, h;
  int64_t N;
  T idx, idz;
  const T *m0, *m1x, *m1z;
  const T *p_o, *vx_o, *vz_o;   // read parity (shot 0 slab)
  T *p_n, *vx_n, *vz_n;         // write parity
  const T *psx_o; T *psx_n; int64_t spx;
  const T *psz_o; T *psz_n; int64_t spz;
  const T *xix_o; T *xix_n; int64_t sxx;
  const T *xiz_o; T *xiz_n; int64_t sxz;
  const T *cpml; int cstride;
  PointArgsView pt;
  const T *srctf;
  T *traces;                     // forward recording (nullable)
  const T *res;                  // adjoint residuals
  const int *nt_shot;
  int it, act_min;
  // pressure history of the segment: slot k holds p^{base+k}
  T *pbuf; int64_t pslot; int smode; int S; int cbase;
  // gradient correlation (adjoint)
  T *g0, *g1x, *g1z; T idt;
};

template <typename T>
__device__ __forceinline__ T d4(T a, T b, T c, T d) {
  // (1/24) a - (9/8) b + (9/8) c - (1/24) d  (fdgen.jl order 4, deriv 1)
  const T c1 = T(1.0 / 24.0), c2 = T(-9.0 / 8.0);
  return c1 * a + c2 * b + (-c2) * c + (-c1) * d;
}

// base of the segment a step belongs to, or -1 if the step stores nothing
__device__ __forceinline__ int store_base(int smode, int it, int nts, int S, int cbase) {
  if (smode == 1) {
    const int seg0 = ((nts - 1) / S) * S;
    return it > seg0 ? seg0 : -1;
  }
  if (smode == 2) return cbase;
  return -1;
}

// Forward / replay step: (p, vx, vz)^{it-1} -> (p, vx, vz)^{it}.
template <typename T>
__global__ __launch_bounds__(NT) void vd2d_fwd_kernel(Vd2dArgs<T> a) {
  const int s = blockIdx.z;
  const int nts = a.nt_shot[s];
  if (a.it > nts || nts <= a.act_min) return;
  const int nx = a.nx, nz = a.nz, h = a.h;
  const int I0 = blockIdx.x * TX, J0 = blockIdx.y * TZ;
  const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * TX + tx;
  const int64_t so = (int64_t)s * a.N;
  const T *cb = a.cpml + (int64_t)s * a.cstride;
  const AxisCoef<T> cx = axis_coef(cb, h, 0), cz = axis_coef(cb, h, 1);

  __shared__ T s_vx[TZ + 3][TX + 6];   // vx^{it-1}: x in [I0-3, I0+TX+3), z in [J0-1, J0+TZ+2)
  __shared__ T s_vz[TZ + 6][TX + 3];   // vz^{it-1}: x in [I0-1, I0+TX+2), z in [J0-3, J0+TZ+3)
  __shared__ T s_p[TZ + 3][TX + 3];    // p^{it}:    x in [I0-1, I0+TX+2), z in [J0-1, J0+TZ+2)

  for (int q = tid; q < (TZ + 3) * (TX + 6); q += NT) {
    const int lz = q / (TX + 6), lx = q % (TX + 6);
    const int gi = I0 - 3 + lx, gj = J0 - 1 + lz;
    T v = 0;
    if (gi >= 0 && gi <= nx - 2 && gj >= 0 && gj <= nz - 1) v = a.vx_o[so + (int64_t)gj * nx + gi];
    s_vx[lz][lx] = v;
  }
  for (int q = tid; q < (TZ + 6) * (TX + 3); q += NT) {
    const int lz = q / (TX + 3), lx = q % (TX + 3);
    const int gi = I0 - 1 + lx, gj = J0 - 3 + lz;
    T v = 0;
    if (gi >= 0 && gi <= nx - 1 && gj >= 0 && gj <= nz - 2) v = a.vz_o[so + (int64_t)gj * nx + gi];
    s_vz[lz][lx] = v;
  }
  __syncthreads();

  const int sb = store_base(a.smode, a.it, nts, a.S, a.cbase);
  // new pressure on the tile and the cross-shaped halo the velocities need
  for (int q = tid; q < (TZ + 3) * (TX + 3); q += NT) {
    const int lz = q / (TX + 3), lx = q % (TX + 3);
    const bool xarm = lz >= 1 && lz <= TZ, zarm = lx >= 1 && lx <= TX;
    T val = 0;
    const int gi = I0 - 1 + lx, gj = J0 - 1 + lz;
    if ((xarm || zarm) && gi >= 0 && gi <= nx - 1 && gj >= 0 && gj <= nz - 1) {
      const int64_t c = (int64_t)gj * nx + gi;
      const bool owned = xarm && zarm;
      const T pold = a.p_o[so + c];
      T pv = pold;
      if (gi >= 1 && gi <= nx - 2 && gj >= 1 && gj <= nz - 2) {
        // update_p_CPML! (acoustic2D_VD_xPU.jl:17-37)
        T dvx = d4(s_vx[lz][lx], s_vx[lz][lx + 1], s_vx[lz][lx + 2], s_vx[lz][lx + 3]) * a.idx;
        if (gi <= h || gi >= nx - h - 1) {
          const int k = gi <= h ? gi : gi - nx + 2 * h + 2;
          const int64_t qx = (int64_t)s * a.sxx + (int64_t)gj * (2 * h + 2) + k;
          const T xi = cx.b[k] * a.xix_o[qx] + cx.a[k] * dvx;
          if (owned) a.xix_n[qx] = xi;
          dvx = dvx + xi;
        }
        T dvz = d4(s_vz[lz][lx], s_vz[lz + 1][lx], s_vz[lz + 2][lx], s_vz[lz + 3][lx]) * a.idz;
        if (gj <= h || gj >= nz - h - 1) {
          const int k = gj <= h ? gj : gj - nz + 2 * h + 2;
          const int64_t qz = (int64_t)s * a.sxz + (int64_t)k * nx + gi;
          const T xi = cz.b[k] * a.xiz_o[qz] + cz.a[k] * dvz;
          if (owned) a.xiz_n[qz] = xi;
          dvz = dvz + xi;
        }
        pv = pv - a.m0[c] * (dvx + dvz);
      }
      // inject_sources! (acoustic2D_VD_xPU.jl:1-7), serial order per cell
      const int u = find_point(a.pt.srow, a.pt.usrc_cell, a.pt.nrows, s, gj, c);
      if (u >= 0) {
        const T *tf = a.srctf + a.pt.off_tf[s] + (a.it - 1);
        for (int k = a.pt.usrc_ptr[u]; k < a.pt.usrc_ptr[u + 1]; k++) pv += tf[(int64_t)nts * a.pt.usrc_list[k]];
      }
      val = pv;
      if (owned) {
        a.p_n[so + c] = pv;
        if (sb >= 0) {
          a.pbuf[(int64_t)(a.it - sb) * a.pslot + so + c] = pv;
          if (a.it == sb + 1) a.pbuf[so + c] = pold;
        }
        if (a.traces) {
          // record_receivers! (acoustic2D_VD_xPU.jl:9-15)
          const int r = find_point(a.pt.rrow, a.pt.urec_cell, a.pt.nrows, s, gj, c);
          if (r >= 0) {
            T *tr = a.traces + a.pt.off_tr[s] + (a.it - 1);
            for (int k = a.pt.urec_ptr[r]; k < a.pt.urec_ptr[r + 1]; k++) tr[(int64_t)nts * a.pt.urec_list[k]] = pv;
          }
        }
      }
    }
    s_p[lz][lx] = val;
  }
  __syncthreads();

  // update_vx_CPML! / update_vy_CPML! (acoustic2D_VD_xPU.jl:39-75)
  const int gi = I0 + tx;
  for (int r = 0; r < TZ / BY; r++) {
    const int lz = ty + BY * r;
    const int gj = J0 + lz;
    if (gi >= nx || gj >= nz) continue;
    const int64_t c = (int64_t)gj * nx + gi;
    if (gi <= nx - 2) {
      T d = d4(s_p[lz + 1][tx], s_p[lz + 1][tx + 1], s_p[lz + 1][tx + 2], s_p[lz + 1][tx + 3]) * a.idx;
      if (gi <= h - 1 || gi >= nx - h - 1) {
        const int k = gi <= h - 1 ? gi : gi + 1 - nx + 2 * h;
        const int64_t q = (int64_t)s * a.spx + (int64_t)gj * (2 * h) + k;
        const T ps = cx.bh[k] * a.psx_o[q] + cx.ah[k] * d;
        a.psx_n[q] = ps;
        d = d + ps;
      }
      a.vx_n[so + c] = s_vx[lz + 1][tx + 3] - a.m1x[c] * d;
    }
    if (gj <= nz - 2) {
      T d = d4(s_p[lz][tx + 1], s_p[lz + 1][tx + 1], s_p[lz + 2][tx + 1], s_p[lz + 3][tx + 1]) * a.idz;
      if (gj <= h - 1 || gj >= nz - h - 1) {
        const int k = gj <= h - 1 ? gj : gj + 1 - nz + 2 * h;
        const int64_t q = (int64_t)s * a.spz + (int64_t)k * nx + gi;
        const T ps = cz.bh[k] * a.psz_o[q] + cz.ah[k] * d;
        a.psz_n[q] = ps;
        d = d + ps;
      }
      a.vz_n[so + c] = s_vz[lz + 3][tx + 1] - a.m1z[c] * d;
    }
  }
}

// Adjoint step it: (p~, v~)^{it+1} -> (p~, v~)^{it}, then the correlation
// with p^{it}, p^{it-1} of the replayed segment.
template <typename T>
__global__ __launch_bounds__(NT) void vd2d_adj_kernel(Vd2dArgs<T> a) {
  const int s = blockIdx.z;
  const int nts = a.nt_shot[s];
  if (a.it > nts) return;
  const int nx = a.nx, nz = a.nz, h = a.h;
  const int I0 = blockIdx.x * TX, J0 = blockIdx.y * TZ;
  const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * TX + tx;
  const int64_t so = (int64_t)s * a.N;
  const T *cb = a.cpml + (int64_t)s * a.cstride;
  const AxisCoef<T> cx = axis_coef(cb, h, 0), cz = axis_coef(cb, h, 1);

  __shared__ T s_p[TZ + 6][TX + 6];   // p~^{old}: x in [I0-3, I0+TX+3), z in [J0-3, J0+TZ+3) (cross)
  __shared__ T s_vx[TZ][TX + 3];      // v~x^{new}: x in [I0-2, I0+TX+1), z in [J0, J0+TZ)
  __shared__ T s_vz[TZ + 3][TX];      // v~z^{new}: x in [I0, I0+TX), z in [J0-2, J0+TZ+1)

  for (int q = tid; q < (TZ + 6) * (TX + 6); q += NT) {
    const int lz = q / (TX + 6), lx = q % (TX + 6);
    const bool xarm = lz >= 3 && lz < TZ + 3, zarm = lx >= 3 && lx < TX + 3;
    const int gi = I0 - 3 + lx, gj = J0 - 3 + lz;
    T v = 0;
    if ((xarm || zarm) && gi >= 0 && gi <= nx - 1 && gj >= 0 && gj <= nz - 1) v = a.p_o[so + (int64_t)gj * nx + gi];
    s_p[lz][lx] = v;
  }
  __syncthreads();

  // velocities first (adjoint_onestep_CPML!, acoustic2D_VD_xPU.jl:163-171)
  for (int q = tid; q < TZ * (TX + 3); q += NT) {
    const int lz = q / (TX + 3), lx = q % (TX + 3);
    const int gi = I0 - 2 + lx, gj = J0 + lz;
    T v = 0;
    if (gi >= 0 && gi <= nx - 2 && gj <= nz - 1) {
      const int64_t c = (int64_t)gj * nx + gi;
      const bool owned = lx >= 2 && lx < TX + 2;
      T d = d4(s_p[lz + 3][lx], s_p[lz + 3][lx + 1], s_p[lz + 3][lx + 2], s_p[lz + 3][lx + 3]) * a.idx;
      if (gi <= h - 1 || gi >= nx - h - 1) {
        const int k = gi <= h - 1 ? gi : gi + 1 - nx + 2 * h;
        const int64_t qq = (int64_t)s * a.spx + (int64_t)gj * (2 * h) + k;
        const T ps = cx.bh[k] * a.psx_o[qq] + cx.ah[k] * d;
        if (owned) a.psx_n[qq] = ps;
        d = d + ps;
      }
      v = a.vx_o[so + c] - a.m1x[c] * d;
      if (owned) a.vx_n[so + c] = v;
    }
    s_vx[lz][lx] = v;
  }
  for (int q = tid; q < (TZ + 3) * TX; q += NT) {
    const int lz = q / TX, lx = q % TX;
    const int gi = I0 + lx, gj = J0 - 2 + lz;
    T v = 0;
    if (gi <= nx - 1 && gj >= 0 && gj <= nz - 2) {
      const int64_t c = (int64_t)gj * nx + gi;
      const bool owned = lz >= 2 && lz < TZ + 2;
      T d = d4(s_p[lz][lx + 3], s_p[lz + 1][lx + 3], s_p[lz + 2][lx + 3], s_p[lz + 3][lx + 3]) * a.idz;
      if (gj <= h - 1 || gj >= nz - h - 1) {
        const int k = gj <= h - 1 ? gj : gj + 1 - nz + 2 * h;
        const int64_t qq = (int64_t)s * a.spz + (int64_t)k * nx + gi;
        const T ps = cz.bh[k] * a.psz_o[qq] + cz.ah[k] * d;
        if (owned) a.psz_n[qq] = ps;
        d = d + ps;
      }
      v = a.vz_o[so + c] - a.m1z[c] * d;
      if (owned) a.vz_n[so + c] = v;
    }
    s_vz[lz][lx] = v;
  }
  __syncthreads();

  // pressure (update_p_CPML!), residual injection and correlation
  const int gi = I0 + tx;
  const int slot = a.it - a.cbase;
  const T *pc_ = a.pbuf + (int64_t)slot * a.pslot + so;
  const T *po_ = a.pbuf + (int64_t)(slot - 1) * a.pslot + so;
  for (int r = 0; r < TZ / BY; r++) {
    const int lz = ty + BY * r;
    const int gj = J0 + lz;
    if (gi >= nx || gj >= nz) continue;
    const int64_t c = (int64_t)gj * nx + gi;
    T pv = s_p[lz + 3][tx + 3];
    if (gi >= 1 && gi <= nx - 2 && gj >= 1 && gj <= nz - 2) {
      T dvx = d4(s_vx[lz][tx], s_vx[lz][tx + 1], s_vx[lz][tx + 2], s_vx[lz][tx + 3]) * a.idx;
      if (gi <= h || gi >= nx - h - 1) {
        const int k = gi <= h ? gi : gi - nx + 2 * h + 2;
        const int64_t qx = (int64_t)s * a.sxx + (int64_t)gj * (2 * h + 2) + k;
        const T xi = cx.b[k] * a.xix_o[qx] + cx.a[k] * dvx;
        a.xix_n[qx] = xi;
        dvx = dvx + xi;
      }
      T dvz = d4(s_vz[lz][tx], s_vz[lz + 1][tx], s_vz[lz + 2][tx], s_vz[lz + 3][tx]) * a.idz;
      if (gj <= h || gj >= nz - h - 1) {
        const int k = gj <= h ? gj : gj - nz + 2 * h + 2;
        const int64_t qz = (int64_t)s * a.sxz + (int64_t)k * nx + gi;
        const T xi = cz.b[k] * a.xiz_o[qz] + cz.a[k] * dvz;
        a.xiz_n[qz] = xi;
        dvz = dvz + xi;
      }
      pv = pv - a.m0[c] * (dvx + dvz);
    }
    // residual injection with prescale_residuals! by fact_m0 (:77-89)
    const int u = find_point(a.pt.rrow, a.pt.urec_cell, a.pt.nrows, s, gj, c);
    if (u >= 0) {
      const T *rs = a.res + a.pt.off_tr[s] + (a.it - 1);
      const T f = a.m0[c];
      for (int k = a.pt.urec_ptr[u]; k < a.pt.urec_ptr[u + 1]; k++) pv += rs[(int64_t)nts * a.pt.urec_list[k]] * f;
    }
    a.p_n[so + c] = pv;
    // correlate_gradient_m0! (correlate_gradient_xPU.jl:12-21)
    const T pcv = pc_[c];
    a.g0[so + c] = a.g0[so + c] - (pv * (pcv - po_[c]) * a.idt);
    // correlate_gradient_m1! (acoustic2D_VD_xPU.jl:180-198), 4th order, no C-PML
    if (gi <= nx - 2) {
      const T pm = gi >= 1 ? pc_[c - 1] : T(0);
      const T pp2 = gi + 2 <= nx - 1 ? pc_[c + 2] : T(0);
      const T dpx = d4(pm, pcv, pc_[c + 1], pp2) * a.idx;
      a.g1x[so + c] = a.g1x[so + c] + s_vx[lz][tx + 2] * dpx;
    }
    if (gj <= nz - 2) {
      const T pm = gj >= 1 ? pc_[c - nx] : T(0);
      const T pp2 = gj + 2 <= nz - 1 ? pc_[c + 2 * (int64_t)nx] : T(0);
      const T dpz = d4(pm, pcv, pc_[c + nx], pp2) * a.idz;
      a.g1z[so + c] = a.g1z[so + c] + s_vz[lz + 2][tx] * dpz;
    }
  }
}

template <typename T>
class Vd2dEngine final : public Engine {
 public:
  explicit Vd2dEngine(PlanCore &core) : c_(core) {
    const swx_plan_desc &d = c_.d;
    nx_ = (int)d.n[0];
    nz_ = (int)d.n[1];
    h_ = d.halo;
    ns_ = d.max_shots;
    N_ = (int64_t)nx_ * nz_;
    spx_ = (int64_t)2 * h_ * nz_;
    spz_ = (int64_t)nx_ * 2 * h_;
    sxx_ = (int64_t)(2 * h_ + 2) * nz_;
    sxz_ = (int64_t)nx_ * (2 * h_ + 2);
    cstride_ = 2 * (8 * h_ + 4);
    const int64_t e = sizeof(T);
    const int64_t F = ns_ * N_;
    m_.alloc(3 * N_ * e);
    st_[0].alloc(2 * 3 * F * e);  // [parity][p, vx, vz][ns][N]
    st_[1].alloc(2 * ns_ * std::max<int64_t>(spx_ + spz_, 1) * e);
    st_[2].alloc(2 * ns_ * (sxx_ + sxz_) * e);
    cpml_.alloc((int64_t)ns_ * cstride_ * e);
    state_bytes_ = m_.bytes() + st_[0].bytes() + st_[1].bytes() + st_[2].bytes() + cpml_.bytes();
    if (d.gradient) {
      ad_[0].alloc(st_[0].bytes());
      ad_[1].alloc(st_[1].bytes());
      ad_[2].alloc(st_[2].bytes());
      grad_.alloc(3 * F * e);
      gout_.alloc(3 * F * e);
      state_bytes_ += ad_[0].bytes() + ad_[1].bytes() + ad_[2].bytes() + grad_.bytes() + gout_.bytes();
      const int64_t slot = F * e;
      ck_one_ = (3 * N_ + spx_ + spz_ + sxx_ + sxz_) * ns_ * e;
      const int64_t avail = (int64_t)c_.budget - state_bytes_;
      // the replay buffer holds S+1 pressure slots
      S_ = choose_segment(d.nt_max, d.check_freq, slot, ck_one_, avail - slot);
      K_ = (int)ceil_div(d.nt_max, S_);
      pbuf_.alloc((int64_t)(S_ + 1) * slot);
      nck_ = std::max(K_ - 2, 0);
      ck_.alloc(std::max<int64_t>(nck_ * ck_one_, 1));
    }
    zero(st_);
  }

  void set_model(int nfields, const void *const *fields) override {
    SWX_REQUIRE(nfields == 3 && fields[0] && fields[1] && fields[2], SWX_EINVAL,
                "acoustic VD expects three model fields (fact_m0, fact_m1_stag x, z)");
    const size_t e = sizeof(T);
    SWX_HIP(hipMemsetAsync(m_.get(), 0, m_.bytes(), c_.stream));
    SWX_HIP(hipMemcpyAsync(M(0), fields[0], N_ * e, hipMemcpyHostToDevice, c_.stream));
    // fact_m1_stag[1] is (nx-1) x nz: pitched into the nx x nz slab
    SWX_HIP(hipMemcpy2DAsync(M(1), nx_ * e, fields[1], (nx_ - 1) * e, (nx_ - 1) * e, nz_, hipMemcpyHostToDevice,
                             c_.stream));
    SWX_HIP(hipMemcpyAsync(M(2), fields[2], (int64_t)nx_ * (nz_ - 1) * e, hipMemcpyHostToDevice, c_.stream));
    SWX_HIP(hipStreamSynchronize(c_.stream));
    model_set_ = true;
  }

  void on_shots_set(int nshots, const swx_shot_desc *shots) override {
    std::vector<T> hc = pack_cpml<T>(nshots, shots, 2, h_);
    SWX_HIP(hipMemcpyAsync(cpml_.get(), hc.data(), hc.size() * sizeof(T), hipMemcpyHostToDevice, c_.stream));
    SWX_HIP(hipStreamSynchronize(c_.stream));
    nsh_ = nshots;
    nt_ = c_.pts.nt;
    ntmax_ = *std::max_element(nt_.begin(), nt_.end());
  }

  void forward(bool save) override {
    SWX_REQUIRE(model_set_, SWX_ESTATE, "swx_set_model must be called before running");
    zero(st_);
    for (int it = 1; it <= ntmax_; it++) {
      fwd_step(it, 0, save && c_.d.gradient ? 1 : 0, 0, c_.pts.d_traces.as<T>());
      if (save && it % S_ == 0) {
        const int k = it / S_;
        if (k >= 1 && k <= K_ - 2) ckpt_copy(k, it, true);
      }
    }
  }

  void adjoint(bool per_shot) override {
    zero(ad_);
    SWX_HIP(hipMemsetAsync(grad_.get(), 0, grad_.bytes(), c_.stream));
    const int Kmax = (int)ceil_div(ntmax_, S_);
    for (int k = Kmax - 1; k >= 0; k--) {
      const int cb = k * S_;
      bool need = false;
      for (int s = 0; s < nsh_; s++) need |= nt_[s] > cb + S_;
      if (need) {
        if (k == 0) zero(st_);
        else ckpt_copy(k, cb, false);
        for (int it = cb + 1; it <= cb + S_; it++) fwd_step(it, cb + S_, 2, cb, nullptr);
      }
      const int top = std::min(cb + S_, ntmax_);
      for (int it = top; it >= cb + 1; it--) adj_step(it, cb);
    }
    const int64_t F = (int64_t)ns_ * N_;
    dim3 g((unsigned)std::min<int64_t>(ceil_div(N_, 256), 4096));
    for (int f = 0; f < 3; f++)
      reduce_shots<T><<<g, 256, 0, c_.stream>>>(gout_.as<T>() + f * F, grad_.as<T>() + f * F, N_, nsh_, T(1),
                                                per_shot ? 1 : 0);
    SWX_HIP(hipGetLastError());
  }

  void copy_grad(void *host, bool per_shot) override {
    // per copy: grad_m0 (nx*nz) | grad_m1 x ((nx-1)*nz) | grad_m1 z (nx*(nz-1))
    const size_t e = sizeof(T);
    const int nc = per_shot ? nsh_ : 1;
    const int64_t F = (int64_t)ns_ * N_;
    const int64_t per = N_ + (int64_t)(nx_ - 1) * nz_ + (int64_t)nx_ * (nz_ - 1);
    char *dst = static_cast<char *>(host);
    for (int s = 0; s < nc; s++) {
      char *o = dst + (size_t)s * per * e;
      const T *b0 = gout_.as<T>() + (int64_t)s * N_;
      SWX_HIP(hipMemcpy(o, b0, N_ * e, hipMemcpyDeviceToHost));
      o += N_ * e;
      SWX_HIP(hipMemcpy2D(o, (nx_ - 1) * e, b0 + F, nx_ * e, (nx_ - 1) * e, nz_, hipMemcpyDeviceToHost));
      o += (size_t)(nx_ - 1) * nz_ * e;
      SWX_HIP(hipMemcpy(o, b0 + 2 * F, (size_t)nx_ * (nz_ - 1) * e, hipMemcpyDeviceToHost));
    }
  }

  int64_t bytes_state() const override { return state_bytes_; }
  int64_t bytes_checkpoints() const override { return ck_.bytes(); }
  int64_t bytes_replay() const override { return pbuf_.bytes(); }
  int segment() const override { return S_; }

 private:
  T *M(int f) const { return m_.as<T>() + (int64_t)f * N_; }
  // field f (0 p, 1 vx, 2 vz) of parity par in state block b
  T *F3(const DevBuf *b, int par, int f) const { return b[0].as<T>() + ((int64_t)par * 3 + f) * ns_ * N_; }
  T *PSX(const DevBuf *b, int par) const { return b[1].as<T>() + (int64_t)par * ns_ * (spx_ + spz_); }
  T *PSZ(const DevBuf *b, int par) const { return PSX(b, par) + ns_ * spx_; }
  T *XIX(const DevBuf *b, int par) const { return b[2].as<T>() + (int64_t)par * ns_ * (sxx_ + sxz_); }
  T *XIZ(const DevBuf *b, int par) const { return XIX(b, par) + ns_ * sxx_; }

  void zero(DevBuf *b) {
    for (int k = 0; k < 3; k++) SWX_HIP(hipMemsetAsync(b[k].get(), 0, b[k].bytes(), c_.stream));
  }

  Vd2dArgs<T> args(const DevBuf *b, int it, int rpar, int wpar) const {
    Vd2dArgs<T> a{};
    a.nx = nx_; a.nz = nz_; a.h = h_; a.N = N_;
    a.idx = T(1.0 / c_.d.h[0]);
    a.idz = T(1.0 / c_.d.h[1]);
    a.m0 = M(0); a.m1x = M(1); a.m1z = M(2);
    a.p_o = F3(b, rpar, 0); a.vx_o = F3(b, rpar, 1); a.vz_o = F3(b, rpar, 2);
    a.p_n = F3(b, wpar, 0); a.vx_n = F3(b, wpar, 1); a.vz_n = F3(b, wpar, 2);
    a.psx_o = PSX(b, rpar); a.psx_n = PSX(b, wpar); a.spx = spx_;
    a.psz_o = PSZ(b, rpar); a.psz_n = PSZ(b, wpar); a.spz = spz_;
    a.xix_o = XIX(b, rpar); a.xix_n = XIX(b, wpar); a.sxx = sxx_;
    a.xiz_o = XIZ(b, rpar); a.xiz_n = XIZ(b, wpar); a.sxz = sxz_;
    a.cpml = cpml_.as<T>(); a.cstride = cstride_;
    a.pt = view(c_.pts);
    a.srctf = c_.pts.d_srctf.as<T>();
    a.res = c_.pts.d_res.as<T>();
    a.nt_shot = c_.pts.d_nt.as<int>();
    a.it = it;
    a.pbuf = pbuf_.as<T>(); a.pslot = (int64_t)ns_ * N_; a.S = S_ > 0 ? S_ : 1;
    const int64_t F = (int64_t)ns_ * N_;
    a.g0 = grad_.as<T>(); a.g1x = a.g0 ? a.g0 + F : nullptr; a.g1z = a.g0 ? a.g0 + 2 * F : nullptr;
    a.idt = T(1.0 / c_.d.dt);
    return a;
  }

  dim3 grid() const { return dim3((unsigned)ceil_div(nx_, TX), (unsigned)ceil_div(nz_, TZ), (unsigned)nsh_); }

  void fwd_step(int it, int act_min, int smode, int cbase, T *traces) {
    Vd2dArgs<T> a = args(st_, it, (it - 1) & 1, it & 1);
    a.act_min = act_min;
    a.smode = smode;
    a.cbase = cbase;
    a.traces = traces;
    c_.ktimer.begin(c_.stream);
    vd2d_fwd_kernel<T><<<grid(), dim3(TX, BY), 0, c_.stream>>>(a);
    c_.ktimer.end(c_.stream, 0);
    SWX_HIP(hipGetLastError());
  }

  void adj_step(int it, int cbase) {
    Vd2dArgs<T> a = args(ad_, it, (it + 1) & 1, it & 1);
    a.cbase = cbase;
    c_.ktimer.begin(c_.stream);
    vd2d_adj_kernel<T><<<grid(), dim3(TX, BY), 0, c_.stream>>>(a);
    c_.ktimer.end(c_.stream, 1);
    SWX_HIP(hipGetLastError());
  }

  // checkpoint k (all shots): p, vx, vz, psi_x, psi_z, xi_x, xi_z at parity cb&1
  void ckpt_copy(int k, int cb, bool save) {
    const int64_t e = sizeof(T);
    const int par = cb & 1;
    char *base = ck_.as<char>() + (int64_t)(k - 1) * ck_one_;
    struct Part { void *dev; int64_t bytes; };
    const Part parts[5] = {
        {F3(st_, par, 0), 3 * ns_ * N_ * e},
        {PSX(st_, par), ns_ * spx_ * e},
        {PSZ(st_, par), ns_ * spz_ * e},
        {XIX(st_, par), ns_ * sxx_ * e},
        {XIZ(st_, par), ns_ * sxz_ * e},
    };
    int64_t off = 0;
    for (const Part &pt : parts) {
      if (pt.bytes) {
        if (save) SWX_HIP(hipMemcpyAsync(base + off, pt.dev, pt.bytes, hipMemcpyDeviceToDevice, c_.stream));
        else SWX_HIP(hipMemcpyAsync(pt.dev, base + off, pt.bytes, hipMemcpyDeviceToDevice, c_.stream));
      }
      off += pt.bytes;
    }
  }

  PlanCore &c_;
  int nx_, nz_, h_, ns_;
  int64_t N_, spx_, spz_, sxx_, sxz_;
  int cstride_;
  DevBuf m_, cpml_;
  DevBuf st_[3], ad_[3];   // [0] p/vx/vz, [1] psi, [2] xi; both parities
  DevBuf grad_, gout_, pbuf_, ck_;
  int64_t ck_one_ = 0, state_bytes_ = 0;
  int S_ = 1, K_ = 1, nck_ = 0;
  int nsh_ = 0, ntmax_ = 0;
  std::vector<int> nt_;
  bool model_set_ = false;
};

}  // namespace

std::unique_ptr<Engine> make_acoustic_vd2d(PlanCore &core) {
  if (core.d.dtype == SWX_F32) return std::make_unique<Vd2dEngine<float>>(core);
  return std::make_unique<Vd2dEngine<double>>(core);
}

}  // namespace swx
