// 2D acoustic constant-density engine: fused C-PML time step, point
// (source/receiver) kernel, checkpoint/replay driver and gradient
// correlation for a batch of shots.
//
// Reference semantics (paths relative to the reference root):
//   step       src/models/acoustic/backends/shared/acoustic2D_xPU.jl:1-169
//   stencils   src/utils/fdgen.jl:137-193 (order 2)
//   driver     src/models/acoustic/acou_gradient.jl:4-95
//   correlate  src/models/acoustic/backends/shared/correlate_gradient_xPU.jl:1-10
//
// Device layout (x fastest, exactly the Julia layout), one slab per shot:
//   p      [4][ns][nx*nz]  time levels, p^k lives in slot k mod 4 (the
//                          reference swaps pold/pcur/pnew, :122-124; a
//                          fourth slot lets a launch step two levels)
//   psi_x  [3][ns][2h*nz]  level k in buffer k mod 3, so a cell can
//   psi_z  [3][ns][nx*2h]  recompute its left neighbour's new value
//   xi_x   [3][ns][(2h+2)*nz], xi_z [3][ns][nx*(2h+2)]
//   (the adjoint state is laid out the same way)
//   D      [S][ns][nx*nz]  p^{it-2} - 2 p^{it-1} + p^{it} of the segment
//                          being replayed (what correlate_gradient! needs)
//   ckpt   [K-2][...]      p^{c-1}, p^c, psi, xi at c = k*S
//   grad   [ns][nx*nz]     per-shot raw gradient (reduced in shot order)
#include <algorithm>
#include <cstdlib>
#include <mutex>

#include <hip/hip_ext.h>

#include "swx_device.hpp"
#include "swx_internal.hpp"

namespace swx {
namespace {

constexpr int TX = 64;  // one wavefront along the fast axis
constexpr int TZ = 4;   // waves per block; each thread steps R rows TZ apart

template <typename T>
struct Cd2dStepArgs {
  int nx, nz, h;
  int64_t N;                 // nx*nz
  T idx, idz, idx2, idz2;
  const T *__restrict__ fact;
  const T *pold;             // shot 0 slab of the three time levels
  const T *pcur;
  T *pnew;
  const T *psix_o; T *psix_n; int64_t spx;   // psi_x old/new, per-shot stride
  const T *psiz_o; T *psiz_n; int64_t spz;
  const T *xix_o; T *xix; int64_t sxx;   // xi_x of the previous / this level (may be one array)
  const T *xiz_o; T *xiz; int64_t sxz;
  const T *cpml; int cstride;                // per-shot packed coefficients
  const int *nt_shot;
  int it;                    // time step index (1-based)
  int act_min;               // shot active iff it <= nt_s and nt_s > act_min
  // D store: dmode 0 none, 1 forward last segment, 2 replay slot it-cbase-1
  T *dbuf; int64_t dslot; int dmode; int S; int cbase;
  // correlation of the current (adjoint) field with D slot it-cbase
  T *grad; int corr; int corr_lim;
  // sources/receivers: forward injects srctf and records traces (nullable);
  // the adjoint (res != nullptr) injects the residuals prescaled by fact
  PointArgsView pt;
  const T *srctf;
  T *traces;
  const T *res;
  int nbx, nby;              // tiles of TX x (TZ * R) cells along x, z
  int lean;                  // interior tiles take cd2d_lean_cells
};

// Block -> tile map giving each of the 8 XCDs a contiguous run of tiles, so
// the halo lines neighbouring tiles share are found in one L2.
__device__ __forceinline__ int xcd_tile2(int b, int nb) {
  const int per = nb >> 3;
  if (b >= per * 8) return b;
  return (b & 7) * per + (b >> 3);
}

// Interior tile: every cell of rows J0..J1 and columns I0..I0+TX-1 takes the
// plain Laplacian branch (inside the grid, outside the C-PML strips, no
// memory variable owned) and no point of the launch's families lies on
// those rows.
template <typename T>
__device__ __forceinline__ bool cd2d_is_lean(const Cd2dStepArgs<T> &a, int s, int I0, int J0, int J1) {
  const int nx = a.nx, nz = a.nz, h = a.h;
  if (!(I0 >= max(h, 1) && I0 + TX - 1 <= nx - h - 2 && J0 >= max(h, 1) && J1 <= nz - h - 2)) return false;
  const int64_t o = (int64_t)s * (a.pt.nrows + 1);
  const int *rowp = a.res == nullptr ? a.pt.srow : a.pt.rrow;
  if (rowp[o + J0] != rowp[o + J1 + 1]) return false;
  if (a.traces && a.pt.rrow[o + J0] != a.pt.rrow[o + J1 + 1]) return false;
  return true;
}

// R cells (i, J0 + r * TZ) of an interior tile: every load of the R cells is
// issued before the first store; the expressions are those of cd2d_cell's
// interior branch, so results are bitwise equal.
template <typename T, int R>
__device__ __forceinline__ void cd2d_lean_cells(const Cd2dStepArgs<T> &a, int s, int nts, int i, int J0) {
  const int nx = a.nx;
  const int64_t so = (int64_t)s * a.N;
  const T *__restrict__ pc_ = a.pcur + so;
  const T *__restrict__ po_ = a.pold + so;
  T *__restrict__ pn_ = a.pnew + so;
  const bool corr = a.corr && a.it + 1 <= nts && a.it + 1 <= a.corr_lim;
  int slot = -1;
  if (a.dmode == 1) {
    const int seg0 = ((nts - 1) / a.S) * a.S;
    if (a.it > seg0) slot = a.it - seg0 - 1;
  } else if (a.dmode == 2) {
    slot = a.it - a.cbase - 1;
  }
  const T *dsrc = corr ? a.dbuf + (int64_t)(a.it - a.cbase) * a.dslot + so : nullptr;
  T pc[R], pl[R], pr[R], pd[R], pu[R], pov[R], fa[R], gv[R], dv[R];
#pragma unroll
  for (int r = 0; r < R; r++) {
    const int c = (J0 + r * TZ) * nx + i;
    pc[r] = pc_[c];
    pl[r] = pc_[c - 1];
    pr[r] = pc_[c + 1];
    pd[r] = pc_[c - nx];
    pu[r] = pc_[c + nx];
    pov[r] = po_[c];
    fa[r] = a.fact[c];
    gv[r] = corr ? a.grad[so + c] : T(0);
    dv[r] = corr ? dsrc[c] : T(0);
  }
#pragma unroll
  for (int r = 0; r < R; r++) {
    const int c = (J0 + r * TZ) * nx + i;
    const T d2x = (pl[r] - T(2) * pc[r] + pr[r]) * a.idx2;
    const T d2z = (pd[r] - T(2) * pc[r] + pu[r]) * a.idz2;
    const T lx = d2x, lz = d2z;
    const T flap = fa[r] * (lx + lz);
    const T pnv = T(2) * pc[r] - pov[r] + flap;
    const T inj = 0;
    pn_[c] = pnv;
    if (slot >= 0) a.dbuf[(int64_t)slot * a.dslot + so + c] = flap + inj;
    if (corr) a.grad[so + c] = gv[r] + pc[r] * dv[r];
  }
}

// One cell of the step: C-PML memory variables, Laplacian, points, replay
// history and correlation.
template <typename T>
__device__ __forceinline__ void cd2d_cell(const Cd2dStepArgs<T> &a, int s, int nts, int i, int j) {
  const int nx = a.nx, nz = a.nz, h = a.h;
  if (i >= nx || j >= nz) return;
  const int64_t so = (int64_t)s * a.N;
  const int64_t c = (int64_t)j * nx + i;
  const T *pc_ = a.pcur + so;
  const T *po_ = a.pold + so;
  T *pn_ = a.pnew + so;
  const bool interior = (i >= 1 && i <= nx - 2 && j >= 1 && j <= nz - 2);
  const bool corr = a.corr && a.it + 1 <= nts && a.it + 1 <= a.corr_lim;
  // the loads that do not depend on a branch's outcome are issued together
  const T pc = pc_[c];
  const T pl = interior ? pc_[c - 1] : T(0), pr = interior ? pc_[c + 1] : T(0);
  const T pd = interior ? pc_[c - nx] : T(0), pu = interior ? pc_[c + nx] : T(0);
  const T pov = po_[c], fa = a.fact[c];
  const T gv = corr ? a.grad[so + c] : T(0);
  const T dv = corr ? a.dbuf[(int64_t)(a.it - a.cbase) * a.dslot + so + c] : T(0);
  const T *cb = a.cpml + (int64_t)s * a.cstride;
  const AxisCoef<T> cx = axis_coef(cb, h, 0);
  const AxisCoef<T> cz = axis_coef(cb, h, 1);

  // psi_x: cell i owns strip index kx (update_psi_x!, acoustic2D_xPU.jl:1-12)
  int kx = -1;
  if (i < h) kx = i;
  else if (i >= nx - h - 1 && i <= nx - 2) kx = i + 1 - nx + 2 * h;
  T psx = 0;
  if (kx >= 0) {
    const int64_t q = (int64_t)s * a.spx + (int64_t)j * (2 * h) + kx;
    const T d = (pc_[c + 1] - pc) * a.idx;
    psx = cx.bh(kx) * a.psix_o[q] + cx.ah(kx) * d;
    a.psix_n[q] = psx;
  }
  // psi_z (update_psi_y!, acoustic2D_xPU.jl:14-25)
  int kz = -1;
  if (j < h) kz = j;
  else if (j >= nz - h - 1 && j <= nz - 2) kz = j + 1 - nz + 2 * h;
  T psz = 0;
  if (kz >= 0) {
    const int64_t q = (int64_t)s * a.spz + (int64_t)kz * nx + i;
    const T d = (pc_[c + nx] - pc) * a.idz;
    psz = cz.bh(kz) * a.psiz_o[q] + cz.ah(kz) * d;
    a.psiz_n[q] = psz;
  }

  T pnv = 0, flap = 0, inj = 0;
  if (interior) {
    // update_p_CPML! (acoustic2D_xPU.jl:27-46)
    const T d2x = (pl - T(2) * pc + pr) * a.idx2;
    const T d2z = (pd - T(2) * pc + pu) * a.idz2;
    T lx = d2x, lz = d2z;
    if (i < h || i >= nx - h) {
      // left neighbour's new psi, recomputed from the old buffer
      const int kn = kx - 1;
      const int64_t qn = (int64_t)s * a.spx + (int64_t)j * (2 * h) + kn;
      const T psn = cx.bh(kn) * a.psix_o[qn] + cx.ah(kn) * ((pc - pl) * a.idx);
      const T dpsi = (psx - psn) * a.idx;
      const int64_t qx = (int64_t)s * a.sxx + (int64_t)j * (2 * h + 2) + kx;
      const T xi = cx.b(kx) * a.xix_o[qx] + cx.a(kx) * (d2x + dpsi);
      a.xix[qx] = xi;
      lx = d2x + dpsi + xi;
    }
    if (j < h || j >= nz - h) {
      const int kn = kz - 1;
      const int64_t qn = (int64_t)s * a.spz + (int64_t)kn * nx + i;
      const T psn = cz.bh(kn) * a.psiz_o[qn] + cz.ah(kn) * ((pc - pd) * a.idz);
      const T dpsi = (psz - psn) * a.idz;
      const int64_t qz = (int64_t)s * a.sxz + (int64_t)kz * nx + i;
      const T xi = cz.b(kz) * a.xiz_o[qz] + cz.a(kz) * (d2z + dpsi);
      a.xiz[qz] = xi;
      lz = d2z + dpsi + xi;
    }
    flap = fa * (lx + lz);
    pnv = T(2) * pc - pov + flap;
  }
  // Edge cells are outside the stencil range.  The reference's rotation
  // (acoustic2D_xPU.jl:122-124) binds pnew to the array of the new pold, so
  // an edge cell of level it holds level it-2 plus what is injected there.
  if (!interior) pnv = pov;
  // Points of this cell: inject_sources! / record_receivers! (forward,
  // acoustic2D_xPU.jl:48-62) or the prescaled residual injection of the
  // adjoint (:64-76, :161-162), in the reference's serial order per cell.
  if (a.res == nullptr) {
    const int u = find_point(a.pt.srow, a.pt.usrc_cell, a.pt.nrows, s, j, c);
    if (u >= 0) {
      const T *tf = a.srctf + a.pt.off_tf[s] + (a.it - 1);
      for (int k = a.pt.usrc_ptr[u]; k < a.pt.usrc_ptr[u + 1]; k++) {
        const T v = tf[(int64_t)nts * a.pt.usrc_list[k]];
        pnv += v;
        inj += v;
      }
    }
  } else {
    const int u = find_point(a.pt.rrow, a.pt.urec_cell, a.pt.nrows, s, j, c);
    if (u >= 0) {
      const T *rs = a.res + a.pt.off_tr[s] + (a.it - 1);
      const T f = fa;
      for (int k = a.pt.urec_ptr[u]; k < a.pt.urec_ptr[u + 1]; k++) pnv += rs[(int64_t)nts * a.pt.urec_list[k]] * f;
    }
  }
  pn_[c] = pnv;
  if (a.traces) {
    const int r = find_point(a.pt.rrow, a.pt.urec_cell, a.pt.nrows, s, j, c);
    if (r >= 0) {
      T *tr = a.traces + a.pt.off_tr[s] + (a.it - 1);
      for (int k = a.pt.urec_ptr[r]; k < a.pt.urec_ptr[r + 1]; k++) tr[(int64_t)nts * a.pt.urec_list[k]] = pnv;
    }
  }
  if (a.dmode) {
    int slot = -1;
    if (a.dmode == 1) {
      const int seg0 = ((nts - 1) / a.S) * a.S;
      if (a.it > seg0) slot = a.it - seg0 - 1;
    } else {
      slot = a.it - a.cbase - 1;
    }
    if (slot >= 0) {
      // p^{it} - 2 p^{it-1} + p^{it-2}: in the interior exactly the stencil
      // term plus the injection, formed without cancellation
      a.dbuf[(int64_t)slot * a.dslot + so + c] = interior ? flap + inj : (pov - T(2) * pc) + pnv;
    }
  }
  if (corr) {
    // correlate_gradient! of the previous adjoint step (adjcur = pcur here)
    a.grad[so + c] = gv + pc * dv;
  }
}

// One time step on a tile of TX x (TZ * R) cells per block (a 1D grid of
// nbx * nby * shots blocks, dealt to the XCDs in contiguous runs).
template <typename T, int R>
__global__ __launch_bounds__(TX *TZ) void cd2d_step_kernel(Cd2dStepArgs<T> a) {
  const int b = xcd_tile2(blockIdx.x, gridDim.x);
  const int t = b / a.nbx;
  const int bx = b - t * a.nbx, by = t % a.nby, s = t / a.nby;
  const int nts = a.nt_shot[s];
  if (a.it > nts || nts <= a.act_min) return;
  const int i = bx * TX + threadIdx.x;
  const int J0 = by * TZ * R;
  if (a.lean && cd2d_is_lean(a, s, bx * TX, J0, J0 + TZ * R - 1)) {
    cd2d_lean_cells<T, R>(a, s, nts, i, J0 + threadIdx.y);
    return;
  }
  for (int r = 0; r < R; r++) cd2d_cell(a, s, nts, i, J0 + threadIdx.y + r * TZ);
}

// ---------------------------------------------------------------------------
// Two forward / replay levels per launch (cd2d_step2_kernel).
//
// A launch at it steps the levels it and it+1.  A block owns a TX x H tile
// (H = TZ * 2 rows) and computes level it on the tile plus a one-cell ring
// (its "region", kept in LDS with the level-it C-PML memory variables),
// then level it+1 on the tile from the region.  The ring cells belong to
// the neighbouring tiles, whose blocks compute them as well; only the
// owning block writes a cell.  Because a block reads levels it-1 and it-2
// of the ring while its neighbours write levels it and it+1, the forward
// state keeps four pressure levels (slot k mod 4) and three levels of the
// C-PML memory variables (buffer k mod 3).  Every cell update is the
// expression of cd2d_cell / cd2d_lean_cells in the same operand order, so
// the results are bitwise those of two one-level launches.

constexpr int S2H = TZ * 2;          // tile rows
constexpr int S2W = TX + 2;          // region columns (tile + ring)
constexpr int S2R = S2H + 2;         // region rows
constexpr int S2N = S2W * S2R;       // region cells
constexpr int S2Q = (S2N + TX * TZ - 1) / (TX * TZ);   // region cells per thread

// Replay-history slot of level itl for shot s (or -1).
template <typename T>
__device__ __forceinline__ int cd2d_dslot(const Cd2dStepArgs<T> &a, int nts, int itl) {
  if (a.dmode == 1) {
    const int seg0 = ((nts - 1) / a.S) * a.S;
    return itl > seg0 ? itl - seg0 - 1 : -1;
  }
  return a.dmode == 2 ? itl - a.cbase - 1 : -1;
}

// Results of one cell update (cd2d_cell's, without the stores).
template <typename T>
struct Cd2dCell {
  T pnv, dval, psx, psz, xix, xiz;
  int kx, kz;
  bool xup, zup;   // xi_x / xi_z updated
};

// cd2d_cell's update of cell (i, j) at level itl (forward: sources;
// ADJ: the prescaled residuals),
// reading the current level through pc(di, dj), the old C-PML memory
// variables through psxo(cell i, strip k) / pszo(cell j, strip k) and
// xixo() / xizo() (the cell's own xi).
template <typename T, bool ADJ, typename PC, typename PSX, typename PSZ, typename XIX, typename XIZ>
__device__ __forceinline__ Cd2dCell<T> cd2d_cell_eval(const Cd2dStepArgs<T> &a, int s, int nts, int itl, int i,
                                                      int j, PC pc_at, T pov, T fa, PSX psxo, PSZ pszo, XIX xixo,
                                                      XIZ xizo) {
  const int nx = a.nx, nz = a.nz, h = a.h;
  const int64_t c = (int64_t)j * nx + i;
  const bool interior = (i >= 1 && i <= nx - 2 && j >= 1 && j <= nz - 2);
  const T pc = pc_at(0, 0);
  const T pl = interior ? pc_at(-1, 0) : T(0), pr = interior ? pc_at(1, 0) : T(0);
  const T pd = interior ? pc_at(0, -1) : T(0), pu = interior ? pc_at(0, 1) : T(0);
  const T *cb = a.cpml + (int64_t)s * a.cstride;
  const AxisCoef<T> cx = axis_coef(cb, h, 0);
  const AxisCoef<T> cz = axis_coef(cb, h, 1);
  Cd2dCell<T> o;
  o.kx = -1;
  if (i < h) o.kx = i;
  else if (i >= nx - h - 1 && i <= nx - 2) o.kx = i + 1 - nx + 2 * h;
  o.psx = 0;
  if (o.kx >= 0) {
    const T d = (pc_at(1, 0) - pc) * a.idx;
    o.psx = cx.bh(o.kx) * psxo(i, o.kx) + cx.ah(o.kx) * d;
  }
  o.kz = -1;
  if (j < h) o.kz = j;
  else if (j >= nz - h - 1 && j <= nz - 2) o.kz = j + 1 - nz + 2 * h;
  o.psz = 0;
  if (o.kz >= 0) {
    const T d = (pc_at(0, 1) - pc) * a.idz;
    o.psz = cz.bh(o.kz) * pszo(j, o.kz) + cz.ah(o.kz) * d;
  }
  o.xup = o.zup = false;
  o.xix = o.xiz = 0;
  T pnv = 0, flap = 0, inj = 0;
  if (interior) {
    const T d2x = (pl - T(2) * pc + pr) * a.idx2;
    const T d2z = (pd - T(2) * pc + pu) * a.idz2;
    T lx = d2x, lz = d2z;
    if (i < h || i >= nx - h) {
      const int kn = o.kx - 1;
      const T psn = cx.bh(kn) * psxo(i - 1, kn) + cx.ah(kn) * ((pc - pl) * a.idx);
      const T dpsi = (o.psx - psn) * a.idx;
      const T xi = cx.b(o.kx) * xixo() + cx.a(o.kx) * (d2x + dpsi);
      o.xix = xi;
      o.xup = true;
      lx = d2x + dpsi + xi;
    }
    if (j < h || j >= nz - h) {
      const int kn = o.kz - 1;
      const T psn = cz.bh(kn) * pszo(j - 1, kn) + cz.ah(kn) * ((pc - pd) * a.idz);
      const T dpsi = (o.psz - psn) * a.idz;
      const T xi = cz.b(o.kz) * xizo() + cz.a(o.kz) * (d2z + dpsi);
      o.xiz = xi;
      o.zup = true;
      lz = d2z + dpsi + xi;
    }
    flap = fa * (lx + lz);
    pnv = T(2) * pc - pov + flap;
  }
  if (!interior) pnv = pov;
  if (!ADJ) {
    const int u = find_point(a.pt.srow, a.pt.usrc_cell, a.pt.nrows, s, j, c);
    if (u >= 0) {
      const T *tf = a.srctf + a.pt.off_tf[s] + (itl - 1);
      for (int k = a.pt.usrc_ptr[u]; k < a.pt.usrc_ptr[u + 1]; k++) {
        const T v = tf[(int64_t)nts * a.pt.usrc_list[k]];
        pnv += v;
        inj += v;
      }
    }
  } else {
    const int u = find_point(a.pt.rrow, a.pt.urec_cell, a.pt.nrows, s, j, c);
    if (u >= 0) {
      const T *rs = a.res + a.pt.off_tr[s] + (itl - 1);
      const T f = fa;
      for (int k = a.pt.urec_ptr[u]; k < a.pt.urec_ptr[u + 1]; k++) pnv += rs[(int64_t)nts * a.pt.urec_list[k]] * f;
    }
  }
  o.pnv = pnv;
  o.dval = interior ? flap + inj : (pov - T(2) * pc) + pnv;
  return o;
}

// Stores of an owned cell at level itl: pressure, memory variables,
// replay history and receiver traces.
template <typename T>
__device__ __forceinline__ void cd2d_cell_store(const Cd2dStepArgs<T> &a, const Cd2dCell<T> &o, int s, int nts,
                                                int itl, int i, int j, T *pn, T *psxn, T *pszn, T *xixn, T *xizn) {
  const int nx = a.nx, h = a.h;
  const int64_t so = (int64_t)s * a.N;
  const int64_t c = (int64_t)j * nx + i;
  pn[so + c] = o.pnv;
  if (o.kx >= 0) psxn[(int64_t)s * a.spx + (int64_t)j * (2 * h) + o.kx] = o.psx;
  if (o.kz >= 0) pszn[(int64_t)s * a.spz + (int64_t)o.kz * nx + i] = o.psz;
  if (o.xup) xixn[(int64_t)s * a.sxx + (int64_t)j * (2 * h + 2) + o.kx] = o.xix;
  if (o.zup) xizn[(int64_t)s * a.sxz + (int64_t)o.kz * nx + i] = o.xiz;
  if (a.traces) {
    const int r = find_point(a.pt.rrow, a.pt.urec_cell, a.pt.nrows, s, j, c);
    if (r >= 0) {
      T *tr = a.traces + a.pt.off_tr[s] + (itl - 1);
      for (int k = a.pt.urec_ptr[r]; k < a.pt.urec_ptr[r + 1]; k++) tr[(int64_t)nts * a.pt.urec_list[k]] = o.pnv;
    }
  }
  const int slot = cd2d_dslot(a, nts, itl);
  if (slot >= 0) a.dbuf[(int64_t)slot * a.dslot + so + c] = o.dval;
}

// Arguments of the second level of a two-level launch: the slots of the
// second level (it+1 forward, it-1 adjoint: pressure, memory variables)
// next to those of level it in Cd2dStepArgs.
template <typename T>
struct Cd2dStep2Args {
  Cd2dStepArgs<T> a;
  T *pnew2;
  T *psix_n2, *psiz_n2, *xix_n2, *xiz_n2;
};

// Is the region of tile (I0, J0) of shot s plain: every region cell inside
// the grid, outside the C-PML strips, and no source or receiver on the
// region's rows?
template <typename T>
__device__ __forceinline__ bool cd2d_step2_is_lean(const Cd2dStepArgs<T> &a, int s, int I0, int J0) {
  const int nx = a.nx, nz = a.nz, h = a.h;
  const int i0 = I0 - 1, i1 = I0 + TX, j0 = J0 - 1, j1 = J0 + S2H;
  if (!(i0 >= max(h, 1) && i1 <= nx - h - 2 && j0 >= max(h, 1) && j1 <= nz - h - 2)) return false;
  const int64_t o = (int64_t)s * (a.pt.nrows + 1);
  const int *rowp = a.res == nullptr ? a.pt.srow : a.pt.rrow;
  if (rowp[o + j0] != rowp[o + j1 + 1]) return false;
  if (a.traces && a.pt.rrow[o + j0] != a.pt.rrow[o + j1 + 1]) return false;
  return true;
}

// Forward (ADJ = false): levels it and it+1 from it-1, it-2.  Adjoint
// (ADJ = true): adjoint levels it and it-1 from it+1, it+2, with the
// residual injection of each level and the correlations of both launches
// the one-level kernel would make (level it correlates the adjoint level
// it+1, level it-1 the adjoint level it, each with its replay history
// slot), done by the owner of each cell in the same order.
template <typename T, bool ADJ>
__global__ __launch_bounds__(TX *TZ) void cd2d_step2_kernel(Cd2dStep2Args<T> A2) {
  const Cd2dStepArgs<T> &a = A2.a;
  __shared__ T sA[S2R][S2W];    // first level on the region
  __shared__ T sC[S2R][S2W];    // its current level (it-1 forward, it+1 adjoint) on the region
  __shared__ T sX[S2R][S2W];    // psi_x of the first level (strip cells)
  __shared__ T sZ[S2R][S2W];    // psi_z of the first level
  __shared__ T sXI[S2R][S2W];   // xi_x of the first level
  __shared__ T sZI[S2R][S2W];   // xi_z of the first level
  const int b = xcd_tile2(blockIdx.x, gridDim.x);
  const int t = b / a.nbx;
  const int bx = b - t * a.nbx, by = t % a.nby, s = t / a.nby;
  const int nts = a.nt_shot[s];
  if (nts <= a.act_min) return;
  const int itA = a.it, itB = ADJ ? a.it - 1 : a.it + 1;
  // a shot takes a level while it is within its nt (an adjoint shot whose
  // last level is it-1 starts from zero fields at level it)
  const bool actA = itA <= nts, actB = itB <= nts;
  if (!actA && !actB) return;
  const int nx = a.nx, nz = a.nz, h = a.h;
  const int I0 = bx * TX, J0 = by * S2H;
  const int tid = threadIdx.y * TX + threadIdx.x;
  const int64_t so = (int64_t)s * a.N;
  const T *pc_ = a.pcur + so;
  const T *po_ = a.pold + so;
  // correlations (adjoint): level itl correlates the adjoint level itl+1
  // with history slot itl - cbase while itl+1 is within nt and the segment
  const bool corrA = ADJ && a.corr && actA && itA + 1 <= nts && itA + 1 <= a.corr_lim;
  const bool corrB = ADJ && a.corr && actB && itB + 1 <= nts && itB + 1 <= a.corr_lim;
  T gv[2], dA[2], dB[2];
  if (ADJ && (corrA || corrB)) {
#pragma unroll
    for (int r = 0; r < 2; r++) {
      const int i = min(I0 + (int)threadIdx.x, nx - 1), j = min(J0 + (int)threadIdx.y + r * TZ, nz - 1);
      const int64_t c = (int64_t)j * nx + i;
      gv[r] = a.grad[so + c];
      dA[r] = corrA ? a.dbuf[(int64_t)(itA - a.cbase) * a.dslot + so + c] : T(0);
      dB[r] = corrB ? a.dbuf[(int64_t)(itB - a.cbase) * a.dslot + so + c] : T(0);
    }
  }
  // owner's correlation of tile cell r: the current level of each of the
  // two one-level launches (pcA, then the first level pcB)
  auto correlate = [&](int r, int64_t c, T pcA, T pcB) {
    if (!(corrA || corrB)) return;
    T g = gv[r];
    if (corrA) g = g + pcA * dA[r];
    if (corrB) g = g + pcB * dB[r];
    a.grad[so + c] = g;
  };
  if (actA && a.lean && cd2d_step2_is_lean(a, s, I0, J0)) {
    // Every load first: the current level on the tile + 2 (staged in LDS),
    // the previous level and the factor on the region, the factor on the
    // tile's cells.
    constexpr int PW = TX + 4, PH = S2H + 4, PN = PW * PH;
    constexpr int PQ = (PN + TX * TZ - 1) / (TX * TZ);
    __shared__ T sP[PH][PW];      // current level: columns I0-2 .., rows J0-2 ..
    T pp[PQ], pov[S2Q], fa[S2Q], faB[2];
#pragma unroll
    for (int k = 0; k < PQ; k++) {
      const int q = min(tid + k * TX * TZ, PN - 1);
      const int ly = q / PW, lx = q - ly * PW;
      pp[k] = pc_[(J0 - 2 + ly) * nx + (I0 - 2 + lx)];
    }
#pragma unroll
    for (int k = 0; k < S2Q; k++) {
      const int q = min(tid + k * TX * TZ, S2N - 1);
      const int ly = q / S2W, lx = q - ly * S2W;
      const int c = (J0 - 1 + ly) * nx + (I0 - 1 + lx);
      pov[k] = po_[c];
      fa[k] = a.fact[c];
    }
#pragma unroll
    for (int r = 0; r < 2; r++) faB[r] = a.fact[(J0 + threadIdx.y + r * TZ) * nx + I0 + threadIdx.x];
#pragma unroll
    for (int k = 0; k < PQ; k++) {
      const int q = tid + k * TX * TZ;
      if (q < PN) sP[q / PW][q % PW] = pp[k];
    }
    const int slotA = cd2d_dslot(a, nts, itA), slotB = cd2d_dslot(a, nts, itB);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < S2Q; k++) {
      const int q = tid + k * TX * TZ;
      if (q < S2N) {
        const int ly = q / S2W, lx = q - ly * S2W;
        const T pc = sP[ly + 1][lx + 1];
        const T d2x = (sP[ly + 1][lx] - T(2) * pc + sP[ly + 1][lx + 2]) * a.idx2;
        const T d2z = (sP[ly][lx + 1] - T(2) * pc + sP[ly + 2][lx + 1]) * a.idz2;
        const T lxx = d2x, lzz = d2z;
        const T flap = fa[k] * (lxx + lzz);
        const T pnv = T(2) * pc - pov[k] + flap;
        const T inj = 0;
        sA[ly][lx] = pnv;
        if (lx >= 1 && lx <= TX && ly >= 1 && ly <= S2H) {
          const int64_t c = (int64_t)(J0 - 1 + ly) * nx + (I0 - 1 + lx);
          a.pnew[so + c] = pnv;
          if (slotA >= 0) a.dbuf[(int64_t)slotA * a.dslot + so + c] = flap + inj;
        }
      }
    }
    if (!actB) return;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 2; r++) {
      const int lx = threadIdx.x + 1, ly = threadIdx.y + r * TZ + 1;
      const int64_t c = (int64_t)(J0 - 1 + ly) * nx + (I0 - 1 + lx);
      const T p = sA[ly][lx];
      const T d2x = (sA[ly][lx - 1] - T(2) * p + sA[ly][lx + 1]) * a.idx2;
      const T d2z = (sA[ly - 1][lx] - T(2) * p + sA[ly + 1][lx]) * a.idz2;
      const T lxx = d2x, lzz = d2z;
      const T flap = faB[r] * (lxx + lzz);
      const T pnv = T(2) * p - sP[ly + 1][lx + 1] + flap;
      const T inj = 0;
      A2.pnew2[so + c] = pnv;
      if (slotB >= 0) a.dbuf[(int64_t)slotB * a.dslot + so + c] = flap + inj;
      correlate(r, c, sP[ly + 1][lx + 1], p);
    }
    return;
  }
  // General tile: the first level on the region with cd2d_cell's branches
  // (cells outside the grid are skipped: only interior cells read their
  // neighbours, and those are inside the grid)
  T faB[2];
#pragma unroll
  for (int r = 0; r < 2; r++) {
    const int i = I0 + threadIdx.x, j = J0 + threadIdx.y + r * TZ;
    faB[r] = (i < nx && j < nz) ? a.fact[(int64_t)j * nx + i] : T(0);
  }
  const T *psxo_ = a.psix_o + (int64_t)s * a.spx, *pszo_ = a.psiz_o + (int64_t)s * a.spz;
  const T *xixo_ = a.xix_o + (int64_t)s * a.sxx, *xizo_ = a.xiz_o + (int64_t)s * a.sxz;
  for (int q = tid; q < S2N; q += TX * TZ) {
    const int ly = q / S2W, lx = q - ly * S2W;
    const int i = I0 - 1 + lx, j = J0 - 1 + ly;
    if (i < 0 || i >= nx || j < 0 || j >= nz) continue;
    if (!actA) {
      // an adjoint shot that starts at the second level: zero fields
      sA[ly][lx] = sC[ly][lx] = sX[ly][lx] = sZ[ly][lx] = sXI[ly][lx] = sZI[ly][lx] = T(0);
      continue;
    }
    const int64_t c = (int64_t)j * nx + i;
    auto pcat = [&](int di, int dj) { return pc_[c + di + (int64_t)dj * nx]; };
    auto psxo = [&](int, int k) { return psxo_[(int64_t)j * (2 * h) + k]; };
    auto pszo = [&](int, int k) { return pszo_[(int64_t)k * nx + i]; };
    const Cd2dCell<T> o = cd2d_cell_eval<T, ADJ>(
        a, s, nts, itA, i, j, pcat, po_[c], a.fact[c], psxo, pszo,
        [&]() { return xixo_[(int64_t)j * (2 * h + 2) + (i < h ? i : i + 1 - nx + 2 * h)]; },
        [&]() { return xizo_[(int64_t)(j < h ? j : j + 1 - nz + 2 * h) * nx + i]; });
    sA[ly][lx] = o.pnv;
    sC[ly][lx] = pc_[c];
    sX[ly][lx] = o.psx;
    sZ[ly][lx] = o.psz;
    sXI[ly][lx] = o.xix;
    sZI[ly][lx] = o.xiz;
    if (lx >= 1 && lx <= TX && ly >= 1 && ly <= S2H)
      cd2d_cell_store(a, o, s, nts, itA, i, j, a.pnew, a.psix_n, a.psiz_n, a.xix, a.xiz);
  }
  if (!actB) return;
  __syncthreads();
  // the second level on the tile, from the region in LDS
  for (int r = 0; r < 2; r++) {
    const int lx = threadIdx.x + 1, ly = threadIdx.y + r * TZ + 1;
    const int i = I0 - 1 + lx, j = J0 - 1 + ly;
    if (i >= nx || j >= nz) continue;
    const int64_t c = (int64_t)j * nx + i;
    auto pcat = [&](int di, int dj) { return sA[ly + dj][lx + di]; };
    auto psxo = [&](int ii, int) { return sX[ly][ii - (I0 - 1)]; };
    auto pszo = [&](int jj, int) { return sZ[jj - (J0 - 1)][lx]; };
    const Cd2dCell<T> o = cd2d_cell_eval<T, ADJ>(a, s, nts, itB, i, j, pcat, sC[ly][lx], faB[r], psxo, pszo,
                                                 [&]() { return sXI[ly][lx]; }, [&]() { return sZI[ly][lx]; });
    cd2d_cell_store(a, o, s, nts, itB, i, j, A2.pnew2, A2.psix_n2, A2.psiz_n2, A2.xix_n2, A2.xiz_n2);
    correlate(r, c, sC[ly][lx], sA[ly][lx]);
  }
}

// grad += adjcur * D for the first step of a segment (the other steps are
// fused into the adjoint stencil).
template <typename T>
__global__ void cd2d_correlate_kernel(T *grad, const T *adj, const T *d, int64_t N, const int *nt_shot, int cbase) {
  const int s = blockIdx.y;
  if (nt_shot[s] <= cbase) return;
  const int64_t so = (int64_t)s * N;
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < N; q += (int64_t)gridDim.x * blockDim.x)
    grad[so + q] += adj[so + q] * d[so + q];
}

template <typename T>
class Cd2dEngine final : public Engine {
 public:
  explicit Cd2dEngine(PlanCore &core) : c_(core) {
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
    // rows per thread (1, 2 or 4) and the interior-tile path; env overrides
    // for measurement
    if (const char *v = std::getenv("SWX_CD2D_ROWS")) {
      const int r = std::atoi(v);
      SWX_REQUIRE(r == 1 || r == 2 || r == 4, SWX_EINVAL, "SWX_CD2D_ROWS must be 1, 2 or 4");
      rows_ = r;
      tuned_ = true;
    }
    if (const char *v = std::getenv("SWX_CD2D_LEAN")) lean_ = std::atoi(v) != 0;
    // levels per forward / replay launch (1 or 2, cd2d_step2_kernel)
    if (const char *v = std::getenv("SWX_CD2D_STEPS")) {
      const int n = std::atoi(v);
      SWX_REQUIRE(n == 1 || n == 2, SWX_EINVAL, "SWX_CD2D_STEPS must be 1 or 2");
      steps_ = n;
      steps_env_ = true;
    }
    const size_t e = sizeof(T);
    fact_.alloc(N_ * e);
    // forward state: four pressure levels (slot k mod 4) and the C-PML
    // memory variables of three levels (buffer k mod 3), so that a launch
    // stepping two levels never overwrites a level a neighbouring block
    // still reads (cd2d_step2_kernel); the one-step kernel works on any
    // distinct slots
    p_.alloc(4 * ns_ * N_ * e);
    psix_.alloc(std::max<int64_t>(3 * ns_ * spx_, 1) * e);
    psiz_.alloc(std::max<int64_t>(3 * ns_ * spz_, 1) * e);
    xix_.alloc(std::max<int64_t>(3 * ns_ * sxx_, 1) * e);
    xiz_.alloc(std::max<int64_t>(3 * ns_ * sxz_, 1) * e);
    cpml_.alloc((int64_t)ns_ * cstride_ * e);
    state_bytes_ = fact_.bytes() + p_.bytes() + psix_.bytes() + psiz_.bytes() + xix_.bytes() + xiz_.bytes();
    if (d.gradient) {
      a_.alloc(4 * ns_ * N_ * e);
      // adjoint state: the same level layout as the forward state
      apsix_.alloc(psix_.bytes());
      apsiz_.alloc(psiz_.bytes());
      axix_.alloc(xix_.bytes());
      axiz_.alloc(xiz_.bytes());
      grad_.alloc(ns_ * N_ * e);
      gout_.alloc(ns_ * N_ * e);
      state_bytes_ += a_.bytes() + apsix_.bytes() + apsiz_.bytes() + axix_.bytes() +
                      axiz_.bytes() + grad_.bytes() + gout_.bytes();
      const int64_t slot = ns_ * N_ * (int64_t)e;
      ck_bytes_one_ = (2 * N_ + spx_ + spz_ + sxx_ + sxz_) * ns_ * (int64_t)e;
      int64_t avail = (int64_t)c_.budget - state_bytes_;
      S_ = choose_segment(d.nt_max, d.check_freq, slot, ck_bytes_one_, avail);
      K_ = (int)ceil_div(d.nt_max, S_);
      dbuf_.alloc((int64_t)S_ * slot);
      nck_ = std::max(K_ - 2, 0);
      ck_.alloc(std::max<int64_t>(nck_ * ck_bytes_one_, 1));
    }
    zero_forward_state();
  }

  void set_model(int nfields, const void *const *fields) override {
    SWX_REQUIRE(nfields == 1 && fields[0], SWX_EINVAL, "acoustic CD expects one model field (fact)");
    SWX_HIP(hipMemcpyAsync(fact_.get(), fields[0], N_ * sizeof(T), hipMemcpyHostToDevice, c_.stream));
    SWX_HIP(hipStreamSynchronize(c_.stream));
    model_set_ = true;
  }

  void on_shots_set(int nshots, const swx_shot_desc *shots) override {
    std::vector<T> hc((size_t)nshots * cstride_);
    const int la = 2 * h_ + 2, lh = 2 * h_;
    for (int s = 0; s < nshots; s++) {
      T *dst = hc.data() + (size_t)s * cstride_;
      for (int ax = 0; ax < 2; ax++) {
        const int len[4] = {la, lh, la, lh};
        T *p = dst + ax * (8 * h_ + 4);
        for (int k = 0; k < 4; k++) {
          const T *src = static_cast<const T *>(shots[s].cpml[ax][k]);
          SWX_REQUIRE(src || len[k] == 0, SWX_EINVAL, "missing C-PML coefficient array");
          if (len[k]) std::memcpy(p, src, sizeof(T) * len[k]);
          p += len[k];
        }
      }
      SWX_REQUIRE(shots[s].nt <= c_.d.nt_max, SWX_EINVAL, "shot nt exceeds nt_max");
    }
    SWX_HIP(hipMemcpyAsync(cpml_.get(), hc.data(), hc.size() * sizeof(T), hipMemcpyHostToDevice, c_.stream));
    SWX_HIP(hipStreamSynchronize(c_.stream));
    nsh_ = nshots;
    nt_ = c_.pts.nt;
    ntmax_ = *std::max_element(nt_.begin(), nt_.end());
  }

  void forward(bool save) override {
    SWX_REQUIRE(model_set_, SWX_ESTATE, "swx_set_model must be called before running");
    zero_forward_state();
    T *tr = c_.pts.d_traces.as<T>();
    if (!tuned_) tune_rows(save, tr);
    if (!steps_tuned_) tune_steps(save, tr);
    for (int it = 1; it <= ntmax_;) {
      progress("forward step", it, ntmax_);
      // two levels per launch: the state of both levels is in memory
      // afterwards, so snapshots and checkpoints of either are taken after it
      const int n = steps_ == 2 && it + 1 <= ntmax_ ? 2 : 1;
      if (n == 2) step2_fwd(it, /*act_min=*/0, save ? 1 : 0, /*cbase=*/0, tr);
      else step_fwd(it, /*act_min=*/0, save ? 1 : 0, /*cbase=*/0, tr);
      for (int l = it; l < it + n; l++) {
        take_snapshots(c_, *this, l);
        if (save && l % S_ == 0) {
          const int k = l / S_;
          if (k >= 1 && k <= K_ - 2) save_ckpt(k, l);
        }
      }
      it += n;
    }
  }

  // "pcur" (acou_forward.jl:55-57): level it is in slot it mod 4
  int snap_fields(int it, int s, SnapPiece *out) const override {
    out[0] = {P(it) + (int64_t)s * N_, (int64_t)(N_ * sizeof(T)), (int64_t)(N_ * sizeof(T)), 1};
    return 1;
  }

  void adjoint(bool per_shot) override {
    (void)per_shot;
    const int64_t e = sizeof(T);
    if (!adj_steps_tuned_) tune_adj_steps();
    SWX_HIP(hipMemsetAsync(a_.get(), 0, a_.bytes(), c_.stream));
    SWX_HIP(hipMemsetAsync(apsix_.get(), 0, apsix_.bytes(), c_.stream));
    SWX_HIP(hipMemsetAsync(apsiz_.get(), 0, apsiz_.bytes(), c_.stream));
    SWX_HIP(hipMemsetAsync(axix_.get(), 0, axix_.bytes(), c_.stream));
    SWX_HIP(hipMemsetAsync(axiz_.get(), 0, axiz_.bytes(), c_.stream));
    SWX_HIP(hipMemsetAsync(grad_.get(), 0, grad_.bytes(), c_.stream));
    const int Kmax = (int)ceil_div(ntmax_, S_);
    for (int k = Kmax - 1; k >= 0; k--) {
      progress("adjoint segment", Kmax - k, Kmax);
      const int cb = k * S_;
      // replay for shots whose segment k is not their last one
      bool need = false;
      for (int s = 0; s < nsh_; s++) need |= nt_[s] > cb + S_;
      if (need) {
        restore_ckpt(k, cb);
        for (int it = cb + 1; it <= cb + S_;) {
          const int n = steps_ == 2 && it + 1 <= cb + S_ ? 2 : 1;
          if (n == 2) step2_fwd(it, cb + S_, 2, cb, nullptr);
          else step_fwd(it, cb + S_, 2, cb, nullptr);
          it += n;
        }
      }
      const int top = std::min(cb + S_, ntmax_);
      for (int it = top; it >= cb + 1;) {
        const int n = adj_steps_ == 2 && it - 1 >= cb + 1 ? 2 : 1;
        if (n == 2) step2_adj(it, cb);
        else step_adj(it, cb);
        it -= n;
      }
      // correlation of the first step of the segment
      dim3 g((unsigned)std::min<int64_t>(ceil_div(N_, 256), 1024), nsh_);
      cd2d_correlate_kernel<T><<<g, 256, 0, c_.stream>>>(grad_.as<T>(), A(cb + 1),
                                                          dbuf_.as<T>(), N_, c_.pts.d_nt.as<int>(), cb);
      SWX_HIP(hipGetLastError());
    }
    const T scale = T(1.0 / (c_.d.dt * c_.d.dt));
    dim3 g((unsigned)std::min<int64_t>(ceil_div(N_, 256), 4096));
    reduce_shots<T><<<g, 256, 0, c_.stream>>>(gout_.as<T>(), grad_.as<T>(), N_, nsh_, scale, per_shot ? 1 : 0);
    SWX_HIP(hipGetLastError());
    (void)e;
  }

  void copy_grad(void *host, bool per_shot) override {
    const size_t bytes = (per_shot ? (size_t)nsh_ : 1) * N_ * sizeof(T);
    SWX_HIP(hipMemcpy(host, gout_.get(), bytes, hipMemcpyDeviceToHost));
  }

  int64_t bytes_state() const override { return state_bytes_; }
  int64_t bytes_checkpoints() const override { return ck_.bytes(); }
  int64_t bytes_replay() const override { return dbuf_.bytes(); }
  int segment() const override { return S_; }
  // rows per thread of the one-level kernel; levels per launch, 10 x
  // adjoint + forward / replay
  void variants(int32_t v[2]) const override { v[0] = rows_; v[1] = 10 * adj_steps_ + steps_; }

 private:
  T *P(int k) const { return p_.as<T>() + (int64_t)(((k % 4) + 4) % 4) * ns_ * N_; }
  // forward C-PML memory variables of level k
  T *PSX(int k) const { return psix_.as<T>() + (int64_t)mod3(k) * ns_ * spx_; }
  T *PSZ(int k) const { return psiz_.as<T>() + (int64_t)mod3(k) * ns_ * spz_; }
  T *XIX(int k) const { return xix_.as<T>() + (int64_t)mod3(k) * ns_ * sxx_; }
  T *XIZ(int k) const { return xiz_.as<T>() + (int64_t)mod3(k) * ns_ * sxz_; }
  T *A(int k) const { return a_.as<T>() + (int64_t)(((k % 4) + 4) % 4) * ns_ * N_; }
  T *APSX(int k) const { return apsix_.as<T>() + (int64_t)mod3(k) * ns_ * spx_; }
  T *APSZ(int k) const { return apsiz_.as<T>() + (int64_t)mod3(k) * ns_ * spz_; }
  T *AXIX(int k) const { return axix_.as<T>() + (int64_t)mod3(k) * ns_ * sxx_; }
  T *AXIZ(int k) const { return axiz_.as<T>() + (int64_t)mod3(k) * ns_ * sxz_; }

  void zero_forward_state() {
    SWX_HIP(hipMemsetAsync(p_.get(), 0, p_.bytes(), c_.stream));
    SWX_HIP(hipMemsetAsync(psix_.get(), 0, psix_.bytes(), c_.stream));
    SWX_HIP(hipMemsetAsync(psiz_.get(), 0, psiz_.bytes(), c_.stream));
    SWX_HIP(hipMemsetAsync(xix_.get(), 0, xix_.bytes(), c_.stream));
    SWX_HIP(hipMemsetAsync(xiz_.get(), 0, xiz_.bytes(), c_.stream));
  }

  Cd2dStepArgs<T> base_args(int it) const {
    Cd2dStepArgs<T> a{};
    a.nx = nx_; a.nz = nz_; a.h = h_; a.N = N_;
    a.idx = T(1.0 / c_.d.h[0]);
    a.idz = T(1.0 / c_.d.h[1]);
    a.idx2 = a.idx * a.idx;
    a.idz2 = a.idz * a.idz;
    a.fact = fact_.as<T>();
    a.spx = spx_; a.spz = spz_; a.sxx = sxx_; a.sxz = sxz_;
    a.cpml = cpml_.as<T>(); a.cstride = cstride_;
    a.nt_shot = c_.pts.d_nt.as<int>();
    a.it = it;
    a.dbuf = dbuf_.as<T>(); a.dslot = (int64_t)ns_ * N_; a.S = S_ > 0 ? S_ : 1;
    a.pt = view(c_.pts);
    a.srctf = c_.pts.d_srctf.as<T>();
    return a;
  }

  // e0 / e1 (nullable): events the dispatch records at the kernel's start
  // and end
  template <int R>
  void launch_rows(Cd2dStepArgs<T> a, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr) {
    a.nbx = (int)ceil_div(nx_, TX);
    a.nby = (int)ceil_div(nz_, TZ * R);
    const unsigned nb = (unsigned)((int64_t)a.nbx * a.nby * nsh_);
    hipExtLaunchKernelGGL(cd2d_step_kernel<T, R>, dim3(nb), dim3(TX, TZ), 0, c_.stream, e0, e1, 0, a);
  }

  // A launch here lasts about 9 us, so event packets around it would add a
  // fifth to the measured time: the per-launch timing uses the events the
  // dispatch itself records.
  void launch(Cd2dStepArgs<T> a) {
    a.lean = lean_ ? 1 : 0;
    hipEvent_t e0, e1;
    c_.ktimer.dispatch_events(a.corr ? 1 : 0, &e0, &e1);
    if (rows_ == 4) launch_rows<4>(a, e0, e1);
    else if (rows_ == 1) launch_rows<1>(a, e0, e1);
    else launch_rows<2>(a, e0, e1);
    SWX_HIP(hipGetLastError());
  }

  // Rows per thread (1, 2 or 4) timed on step 1 of the zero initial state,
  // which every repetition recomputes identically (all stencil inputs are
  // zero; the new level is the injection alone).  Every choice gives
  // bitwise the same results.
  void tune_rows(bool save, T *tr) {
    static const int rows[3] = {2, 1, 4};
    Cd2dStepArgs<T> a = fwd_args(1, 0, save ? 1 : 0, 0, tr);
    a.lean = lean_ ? 1 : 0;
    const int k = fastest_candidate(c_.stream, 3, [&](int c) {
      if (rows[c] == 4) launch_rows<4>(a);
      else if (rows[c] == 1) launch_rows<1>(a);
      else launch_rows<2>(a);
      SWX_HIP(hipGetLastError());
    });
    rows_ = rows[k];
    tuned_ = true;
  }

  Cd2dStepArgs<T> fwd_args(int it, int act_min, int dmode, int cbase, T *traces) const {
    Cd2dStepArgs<T> a = base_args(it);
    a.pold = P(it - 2); a.pcur = P(it - 1); a.pnew = P(it);
    a.psix_o = PSX(it - 1); a.psix_n = PSX(it);
    a.psiz_o = PSZ(it - 1); a.psiz_n = PSZ(it);
    a.xix_o = XIX(it - 1); a.xix = XIX(it);
    a.xiz_o = XIZ(it - 1); a.xiz = XIZ(it);
    a.act_min = act_min;
    a.dmode = c_.d.gradient ? dmode : 0;
    a.cbase = cbase;
    a.corr = 0;
    a.traces = traces;
    return a;
  }

  void step_fwd(int it, int act_min, int dmode, int cbase, T *traces) {
    launch(fwd_args(it, act_min, dmode, cbase, traces));
  }

  // levels it and it+1 in one launch (cd2d_step2_kernel)
  Cd2dStep2Args<T> step2_args(int it, int act_min, int dmode, int cbase, T *traces) const {
    Cd2dStep2Args<T> b;
    b.a = fwd_args(it, act_min, dmode, cbase, traces);
    b.a.lean = lean_ ? 1 : 0;
    b.a.nbx = (int)ceil_div(nx_, TX);
    b.a.nby = (int)ceil_div(nz_, S2H);
    b.pnew2 = P(it + 1);
    b.psix_n2 = PSX(it + 1);
    b.psiz_n2 = PSZ(it + 1);
    b.xix_n2 = XIX(it + 1);
    b.xiz_n2 = XIZ(it + 1);
    return b;
  }
  template <bool ADJ>
  void launch_step2(const Cd2dStep2Args<T> &b, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr) {
    const unsigned nb = (unsigned)((int64_t)b.a.nbx * b.a.nby * nsh_);
    hipExtLaunchKernelGGL(cd2d_step2_kernel<T, ADJ>, dim3(nb), dim3(TX, TZ), 0, c_.stream, e0, e1, 0, b);
  }
  void step2_fwd(int it, int act_min, int dmode, int cbase, T *traces) {
    hipEvent_t e0, e1;
    c_.ktimer.dispatch_events(0, &e0, &e1);
    launch_step2<false>(step2_args(it, act_min, dmode, cbase, traces), e0, e1);
    SWX_HIP(hipGetLastError());
  }

  // One level per launch (two launches of the tuned row count) against two
  // levels per launch, timed on levels 1 and 2 of the zero initial state,
  // which every repetition recomputes identically.  Both give bitwise the
  // same results; SWX_CD2D_STEPS pins the choice.
  void tune_steps(bool save, T *tr) {
    steps_tuned_ = true;
    if (steps_env_ || ntmax_ < 2) return;
    Cd2dStepArgs<T> a1 = fwd_args(1, 0, save ? 1 : 0, 0, tr), a2 = fwd_args(2, 0, save ? 1 : 0, 0, tr);
    a1.lean = a2.lean = lean_ ? 1 : 0;
    const Cd2dStep2Args<T> b = step2_args(1, 0, save ? 1 : 0, 0, tr);
    auto one = [&](Cd2dStepArgs<T> x) {
      if (rows_ == 4) launch_rows<4>(x);
      else if (rows_ == 1) launch_rows<1>(x);
      else launch_rows<2>(x);
    };
    const int k = fastest_candidate(c_.stream, 2, [&](int c) {
      if (c == 0) {
        one(a1);
        one(a2);
      } else {
        launch_step2<false>(b);
      }
      SWX_HIP(hipGetLastError());
    });
    steps_ = k == 0 ? 1 : 2;
  }

  Cd2dStepArgs<T> adj_args(int it, int cbase) const {
    Cd2dStepArgs<T> a = base_args(it);
    a.pold = A(it + 2); a.pcur = A(it + 1); a.pnew = A(it);
    a.psix_o = APSX(it + 1); a.psix_n = APSX(it);
    a.psiz_o = APSZ(it + 1); a.psiz_n = APSZ(it);
    a.xix_o = AXIX(it + 1); a.xix = AXIX(it);
    a.xiz_o = AXIZ(it + 1); a.xiz = AXIZ(it);
    a.act_min = 0;
    a.dmode = 0;
    a.cbase = cbase;
    a.grad = grad_.as<T>();
    a.corr = 1;
    a.corr_lim = cbase + S_;
    a.res = c_.pts.d_res.as<T>();
    return a;
  }
  void step_adj(int it, int cbase) { launch(adj_args(it, cbase)); }

  // adjoint levels it and it-1 in one launch
  Cd2dStep2Args<T> adj2_args(int it, int cbase) const {
    Cd2dStep2Args<T> b;
    b.a = adj_args(it, cbase);
    b.a.lean = lean_ ? 1 : 0;
    b.a.nbx = (int)ceil_div(nx_, TX);
    b.a.nby = (int)ceil_div(nz_, S2H);
    b.pnew2 = A(it - 1);
    b.psix_n2 = APSX(it - 1);
    b.psiz_n2 = APSZ(it - 1);
    b.xix_n2 = AXIX(it - 1);
    b.xiz_n2 = AXIZ(it - 1);
    return b;
  }
  void step2_adj(int it, int cbase) {
    hipEvent_t e0, e1;
    c_.ktimer.dispatch_events(1, &e0, &e1);
    launch_step2<true>(adj2_args(it, cbase), e0, e1);
    SWX_HIP(hipGetLastError());
  }

  // Adjoint levels per launch, timed like tune_steps on the top two levels
  // of the first segment before the adjoint state and the gradient are
  // cleared (what the timed launches write is discarded).
  void tune_adj_steps() {
    adj_steps_tuned_ = true;
    adj_steps_ = steps_;
    if (steps_env_ || ntmax_ < 2 || S_ < 2) return;
    const int it = std::min(S_, ntmax_);
    Cd2dStepArgs<T> a1 = adj_args(it, 0), a2 = adj_args(it - 1, 0);
    a1.lean = a2.lean = lean_ ? 1 : 0;
    const Cd2dStep2Args<T> b = adj2_args(it, 0);
    auto one = [&](Cd2dStepArgs<T> x) {
      if (rows_ == 4) launch_rows<4>(x);
      else if (rows_ == 1) launch_rows<1>(x);
      else launch_rows<2>(x);
    };
    const int k = fastest_candidate(c_.stream, 2, [&](int c) {
      if (c == 0) {
        one(a1);
        one(a2);
      } else {
        launch_step2<true>(b);
      }
      SWX_HIP(hipGetLastError());
    });
    adj_steps_ = k == 0 ? 1 : 2;
  }

  // checkpoint layout (all shots): p^{c-1} | p^c | psi_x | psi_z | xi_x | xi_z
  void ckpt_copy(int k, int cb, bool save) {
    const int64_t e = sizeof(T);
    char *base = ck_.as<char>() + (int64_t)(k - 1) * ck_bytes_one_;
    const int64_t fb = (int64_t)ns_ * N_ * e;
    struct Part { void *dev; int64_t bytes; };
    Part parts[6] = {
        {P(cb - 1), fb},
        {P(cb), fb},
        {PSX(cb), ns_ * spx_ * e},
        {PSZ(cb), ns_ * spz_ * e},
        {XIX(cb), ns_ * sxx_ * e},
        {XIZ(cb), ns_ * sxz_ * e},
    };
    int64_t off = 0;
    for (auto &pt : parts) {
      if (pt.bytes) {
        if (save)
          SWX_HIP(hipMemcpyAsync(base + off, pt.dev, pt.bytes, hipMemcpyDeviceToDevice, c_.stream));
        else
          SWX_HIP(hipMemcpyAsync(pt.dev, base + off, pt.bytes, hipMemcpyDeviceToDevice, c_.stream));
      }
      off += pt.bytes;
    }
  }
  void save_ckpt(int k, int it) { ckpt_copy(k, it, true); }
  void restore_ckpt(int k, int cb) {
    if (k == 0) {
      zero_forward_state();
    } else {
      ckpt_copy(k, cb, false);
    }
  }

  PlanCore &c_;
  int nx_, nz_, h_, ns_;
  int64_t N_, spx_, spz_, sxx_, sxz_;
  int cstride_;
  int rows_ = 2;
  bool tuned_ = false;   // rows_ chosen (by SWX_CD2D_ROWS or tune_rows)
  int steps_ = 1;        // levels per forward / replay launch
  int adj_steps_ = 1;    // levels per adjoint launch
  bool steps_env_ = false, steps_tuned_ = false, adj_steps_tuned_ = false;
  bool lean_ = true;
  DevBuf fact_, p_, psix_, psiz_, xix_, xiz_, cpml_;
  DevBuf a_, apsix_, apsiz_, axix_, axiz_, grad_, gout_;
  DevBuf dbuf_, ck_;
  int64_t ck_bytes_one_ = 0, state_bytes_ = 0;
  int S_ = 1, K_ = 1, nck_ = 0;
  int nsh_ = 0, ntmax_ = 0;
  std::vector<int> nt_;
  bool model_set_ = false;
};

}  // namespace

std::unique_ptr<Engine> make_acoustic_cd2d(PlanCore &core) {
  if (core.d.dtype == SWX_F32) return std::make_unique<Cd2dEngine<float>>(core);
  return std::make_unique<Cd2dEngine<double>>(core);
}

}  // namespace swx

// ---------------------------------------------------------------------------
// swx_forward_noalloc: the forward of one shot without any device allocation
// ---------------------------------------------------------------------------
// Every byte of device memory the call uses (model, three time levels, C-PML
// memory variables and coefficients, point tables, source time function and
// traces) is carved out of this array.  Each GPU the library runs on holds
// its own copy until the process ends; calls on one GPU share it.
static __device__ __align__(256) unsigned char scratch[SWX_NOALLOC_SCRATCH_BYTES];

namespace swx {
namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_dev[kMaxDevices] = {};

// Device address of `scratch` on the current device `dev`, looked up once
// per device.  The host-side name `scratch` is not a device address.
void *scratch_on(int dev) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  SWX_REQUIRE(dev >= 0 && dev < kMaxDevices, SWX_EINVAL, "invalid device ordinal");
  if (!g_scratch_dev[dev]) {
    void *ptr = nullptr;
    SWX_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    SWX_REQUIRE(ptr != nullptr, SWX_EHIP, "hipGetSymbolAddress(scratch) returned no address");
    g_scratch_dev[dev] = ptr;
  }
  return g_scratch_dev[dev];
}

__global__ void zero_words(uint32_t *p, int64_t n) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x)
    p[q] = 0u;
}

// Offsets of the pieces in the scratch array, each aligned to 256 bytes.
struct Carve {
  size_t off = 0;
  size_t take(size_t bytes) {
    off = (off + 255) & ~size_t(255);
    const size_t o = off;
    off += bytes;
    return o;
  }
};

template <typename T>
void forward_noalloc(const swx_plan_desc &d, const T *fact, const swx_shot_desc &shot, T *traces_host) {
  const int nx = (int)d.n[0], nz = (int)d.n[1], h = d.halo;
  const int64_t N = (int64_t)nx * nz;
  const int64_t spx = (int64_t)2 * h * nz, spz = (int64_t)nx * 2 * h;
  const int64_t sxx = (int64_t)(2 * h + 2) * nz, sxz = (int64_t)nx * (2 * h + 2);
  const int cstride = 2 * (8 * h + 4);
  HostPointTables t;
  host_point_tables(t, 2, d.n, 1, &shot);
  const std::vector<T> cp = pack_cpml<T>(1, &shot, 2, h);
  const int nt = shot.nt;
  // the pieces copied from the host come first, the ones cleared on the
  // device after them
  Carve cv;
  const size_t o_fact = cv.take(N * sizeof(T));
  const size_t o_cpml = cv.take(cp.size() * sizeof(T));
  const size_t o_nt = cv.take(sizeof(int));
  const size_t o_offtf = cv.take(sizeof(int64_t)), o_offtr = cv.take(sizeof(int64_t));
  const size_t o_usrc_off = cv.take(t.usrc_off.size() * sizeof(int));
  const size_t o_usrc_cell = cv.take(t.usrc_cell.size() * sizeof(int64_t));
  const size_t o_usrc_ptr = cv.take(t.usrc_ptr.size() * sizeof(int));
  const size_t o_usrc_list = cv.take(t.usrc_list.size() * sizeof(int));
  const size_t o_urec_off = cv.take(t.urec_off.size() * sizeof(int));
  const size_t o_urec_cell = cv.take(t.urec_cell.size() * sizeof(int64_t));
  const size_t o_urec_ptr = cv.take(t.urec_ptr.size() * sizeof(int));
  const size_t o_urec_list = cv.take(t.urec_list.size() * sizeof(int));
  const size_t o_rec_off = cv.take(t.rec_off.size() * sizeof(int));
  const size_t o_rec_cell = cv.take(t.rec_cell.size() * sizeof(int64_t));
  const size_t o_srow = cv.take(t.srow.size() * sizeof(int));
  const size_t o_rrow = cv.take(t.rrow.size() * sizeof(int));
  const size_t o_tf = cv.take((size_t)t.total_tf * sizeof(T));
  const size_t copy_end = cv.take(0);
  const size_t o_p = cv.take((size_t)3 * N * sizeof(T));
  const size_t o_psix = cv.take((size_t)2 * spx * sizeof(T)), o_psiz = cv.take((size_t)2 * spz * sizeof(T));
  const size_t o_xix = cv.take((size_t)sxx * sizeof(T)), o_xiz = cv.take((size_t)sxz * sizeof(T));
  const size_t o_tr = cv.take((size_t)t.total_tr * sizeof(T));
  const size_t total = cv.take(0);
  SWX_REQUIRE(total <= SWX_NOALLOC_SCRATCH_BYTES, SWX_ENOMEM,
              "swx_forward_noalloc: the shot needs " + std::to_string(total) + " bytes of device memory, more than "
              "the " + std::to_string((size_t)SWX_NOALLOC_SCRATCH_BYTES) + "-byte scratch array");

  std::vector<char> stage(copy_end);
  auto put = [&](size_t o, const void *src, size_t bytes) {
    if (bytes) std::memcpy(stage.data() + o, src, bytes);
  };
  const int64_t zero64 = 0;
  put(o_fact, fact, N * sizeof(T));
  put(o_cpml, cp.data(), cp.size() * sizeof(T));
  put(o_nt, &nt, sizeof(int));
  put(o_offtf, &zero64, sizeof(int64_t));
  put(o_offtr, &zero64, sizeof(int64_t));
  put(o_usrc_off, t.usrc_off.data(), t.usrc_off.size() * sizeof(int));
  put(o_usrc_cell, t.usrc_cell.data(), t.usrc_cell.size() * sizeof(int64_t));
  put(o_usrc_ptr, t.usrc_ptr.data(), t.usrc_ptr.size() * sizeof(int));
  put(o_usrc_list, t.usrc_list.data(), t.usrc_list.size() * sizeof(int));
  put(o_urec_off, t.urec_off.data(), t.urec_off.size() * sizeof(int));
  put(o_urec_cell, t.urec_cell.data(), t.urec_cell.size() * sizeof(int64_t));
  put(o_urec_ptr, t.urec_ptr.data(), t.urec_ptr.size() * sizeof(int));
  put(o_urec_list, t.urec_list.data(), t.urec_list.size() * sizeof(int));
  put(o_rec_off, t.rec_off.data(), t.rec_off.size() * sizeof(int));
  put(o_rec_cell, t.rec_cell.data(), t.rec_cell.size() * sizeof(int64_t));
  put(o_srow, t.srow.data(), t.srow.size() * sizeof(int));
  put(o_rrow, t.rrow.data(), t.rrow.size() * sizeof(int));
  put(o_tf, shot.srctf, (size_t)t.total_tf * sizeof(T));

  char *base = static_cast<char *>(scratch_on(d.device));
  const hipStream_t stream = nullptr;
  SWX_HIP(hipMemcpyAsync(base, stage.data(), copy_end, hipMemcpyHostToDevice, stream));
  // the time levels, memory variables and traces start at zero
  const int64_t words = (int64_t)(total - copy_end) / 4;
  zero_words<<<(unsigned)std::min<int64_t>(ceil_div(words, 256), 2048), 256, 0, stream>>>(
      reinterpret_cast<uint32_t *>(base + copy_end), words);
  SWX_HIP(hipGetLastError());

  auto at = [&](size_t o) { return base + o; };
  PointArgsView pv{reinterpret_cast<const int *>(at(o_nt)), reinterpret_cast<const int64_t *>(at(o_offtf)),
                   reinterpret_cast<const int64_t *>(at(o_offtr)), reinterpret_cast<const int *>(at(o_usrc_off)),
                   reinterpret_cast<const int64_t *>(at(o_usrc_cell)), reinterpret_cast<const int *>(at(o_usrc_ptr)),
                   reinterpret_cast<const int *>(at(o_usrc_list)), reinterpret_cast<const int *>(at(o_urec_off)),
                   reinterpret_cast<const int64_t *>(at(o_urec_cell)), reinterpret_cast<const int *>(at(o_urec_ptr)),
                   reinterpret_cast<const int *>(at(o_urec_list)), reinterpret_cast<const int *>(at(o_rec_off)),
                   reinterpret_cast<const int64_t *>(at(o_rec_cell)), reinterpret_cast<const int *>(at(o_srow)),
                   reinterpret_cast<const int *>(at(o_rrow)), t.nrows};
  T *p = reinterpret_cast<T *>(at(o_p));
  T *psix = reinterpret_cast<T *>(at(o_psix));
  T *psiz = reinterpret_cast<T *>(at(o_psiz));
  T *tr = reinterpret_cast<T *>(at(o_tr));
  constexpr int R = 2;
  Cd2dStepArgs<T> a{};
  a.nx = nx; a.nz = nz; a.h = h; a.N = N;
  a.idx = T(1.0 / d.h[0]);
  a.idz = T(1.0 / d.h[1]);
  a.idx2 = a.idx * a.idx;
  a.idz2 = a.idz * a.idz;
  a.fact = reinterpret_cast<const T *>(at(o_fact));
  a.spx = spx; a.spz = spz; a.sxx = sxx; a.sxz = sxz;
  a.xix = reinterpret_cast<T *>(at(o_xix));
  a.xiz = reinterpret_cast<T *>(at(o_xiz));
  a.xix_o = a.xix;
  a.xiz_o = a.xiz;
  a.cpml = reinterpret_cast<const T *>(at(o_cpml)); a.cstride = cstride;
  a.nt_shot = pv.nt;
  a.S = 1;
  a.pt = pv;
  a.srctf = reinterpret_cast<const T *>(at(o_tf));
  a.traces = tr;
  a.lean = 1;
  a.nbx = (int)ceil_div(nx, TX);
  a.nby = (int)ceil_div(nz, TZ * R);
  const unsigned nb = (unsigned)((int64_t)a.nbx * a.nby);
  for (int it = 1; it <= nt; it++) {
    a.it = it;
    a.pold = p + (int64_t)mod3(it - 2) * N;
    a.pcur = p + (int64_t)mod3(it - 1) * N;
    a.pnew = p + (int64_t)mod3(it) * N;
    a.psix_o = psix + (int64_t)((it - 1) & 1) * spx;
    a.psix_n = psix + (int64_t)(it & 1) * spx;
    a.psiz_o = psiz + (int64_t)((it - 1) & 1) * spz;
    a.psiz_n = psiz + (int64_t)(it & 1) * spz;
    cd2d_step_kernel<T, R><<<nb, dim3(TX, TZ), 0, stream>>>(a);
    SWX_HIP(hipGetLastError());
  }
  SWX_HIP(hipMemcpyAsync(traces_host, tr, (size_t)t.total_tr * sizeof(T), hipMemcpyDeviceToHost, stream));
  SWX_HIP(hipStreamSynchronize(stream));
}

}  // namespace
}  // namespace swx

extern "C" int swx_forward_noalloc(const swx_plan_desc *desc, const void *fact, const swx_shot_desc *shot,
                                   void *traces) {
  using namespace swx;
  return guarded([&] {
    SWX_REQUIRE(desc && fact && shot && traces, SWX_EINVAL, "null argument");
    const swx_plan_desc &d = *desc;
    SWX_REQUIRE(d.physics == SWX_ACOUSTIC_CD && d.ndim == 2, SWX_EUNSUPPORTED,
                "swx_forward_noalloc runs 2D acoustic constant-density shots");
    validate_plan_desc(d);
    SWX_REQUIRE(shot->nt >= 1 && shot->nt <= d.nt_max, SWX_EINVAL, "shot nt must be in 1..nt_max");
    SWX_HIP(hipSetDevice(d.device));
    if (d.dtype == SWX_F32)
      forward_noalloc<float>(d, static_cast<const float *>(fact), *shot, static_cast<float *>(traces));
    else
      forward_noalloc<double>(d, static_cast<const double *>(fact), *shot, static_cast<double *>(traces));
  });
}
