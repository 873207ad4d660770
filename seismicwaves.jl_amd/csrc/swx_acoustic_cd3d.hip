// 3D acoustic constant-density engine for a batch of shots: one fused
// launch per time step (C-PML memory variables, pressure update, source
// injection / receiver recording or residual injection, replay-history
// store and gradient correlation).
//
// Reference semantics (paths relative to the reference root):
//   step       src/models/acoustic/backends/shared/acoustic3D_xPU.jl:1-219
//   stencils   src/utils/fdgen.jl:137-193 (order 2)
//   driver     src/models/acoustic/acou_gradient.jl:4-95
//   correlate  src/models/acoustic/backends/shared/correlate_gradient_xPU.jl:1-10
//
// Kernel shape: a block is a 64 x 4 column tile of the (x, y) plane that
// marches along z (the slow axis) over a chunk of planes.  Each thread keeps
// p(k-1), p(k), p(k+1) in registers and rotates them, so the z neighbours
// and the current plane are read from HBM once; x/y neighbours come from
// the L1/L2 lines the neighbouring lanes load anyway.  The z C-PML memory
// variable of plane k-1 stays in a register for the xi update of plane k.
// Tiles clear of the x/y C-PML strips, the grid edges and the points take a
// march with no per-cell tests (cd3d_lean_column) in the same launch.
//
// Device layout (x fastest), one slab per shot:
//   p       [3][ns][N]            p^k lives in slot k mod 3 (pold/pcur/pnew rotation)
//   psi_d   [2][ns][strip_d(2h)]  double buffered by step parity
//   xi_d    [ns][strip_d(2h+2)]
//   D       [S][ns][N]            p^{it-2} - 2 p^{it-1} + p^{it} of the segment
//   ckpt    [K-2][...]            p^{c-1}, p^c, psi, xi at c = k*S
//   grad    [ns][N]
#include <algorithm>
#include <cstdlib>

#include "swx_device.hpp"
#include "swx_internal.hpp"

namespace swx {
namespace {

constexpr int TX = 64;

template <typename T>
struct Cd3dArgs {
  int n[3], h;
  int nbx, nby, nbk, kch;       // tiles along x, y; z chunks; planes per chunk
  int lean;                     // interior column tiles take cd3d_lean_column
  int64_t N;
  T id[3], id2[3];
  const T *fact;
  const T *pold, *pcur;
  T *pnew;
  const T *psi_o[3];
  T *psi_n[3];
  int64_t sps[3];               // per-shot strides of psi_d
  T *xi[3];
  int64_t sxi[3];
  const T *cpml; int cstride;
  const int *nt_shot;
  int it, act_min;
  T *dbuf; int64_t dslot; int dmode; int S; int cbase;
  T *grad; int corr; int corr_lim;
  PointArgsView pt;
  const T *srctf;
  T *traces;
  const T *res;
};

__device__ __forceinline__ int xcd_tile3(int b, int nb) {
  const int per = nb >> 3;
  if (b >= per * 8) return b;
  return (b & 7) * per + (b >> 3);
}

// strip index of the cell owning psi along an axis of size n (or -1)
__device__ __forceinline__ int psi_owner(int i, int n, int h) {
  if (i < h) return i;
  if (i >= n - h - 1 && i <= n - 2) return i + 1 - n + 2 * h;
  return -1;
}

// Interior column tile of the 3D step (cd3d_is_lean): no x/y C-PML strip or
// grid edge in reach of the tile, and no source / receiver on its rows in
// any plane of its z chunk.  The z-march of cd3d_step_kernel with the same
// expressions (z C-PML and the edge planes handled per plane), without the
// per-cell x/y tests and point lookups, on 32-bit in-plane offsets.
template <typename T, int TY>
__device__ __forceinline__ bool cd3d_is_lean(const Cd3dArgs<T> &a, int s, int bx, int by, int K0, int K1) {
  const int nx = a.n[0], ny = a.n[1], h = a.h;
  const int i0 = bx * TX, i1 = i0 + TX - 1, j0 = by * TY, j1 = j0 + TY - 1;
  if (!(i0 >= max(h, 1) && i1 <= nx - h - 2 && j0 >= max(h, 1) && j1 <= ny - h - 2)) return false;
  // the planes of the chunk spread over the threads of the block
  const int *rowp = (a.res == nullptr ? a.pt.srow : a.pt.rrow) + (int64_t)s * (a.pt.nrows + 1);
  const int *rowr = a.pt.rrow + (int64_t)s * (a.pt.nrows + 1);
  int hit = 0;
  for (int k = K0 + threadIdx.y * TX + threadIdx.x; k < K1; k += TX * TY) {
    const int64_t r0 = j0 + (int64_t)ny * k, r1 = j1 + 1 + (int64_t)ny * k;
    hit |= rowp[r0] != rowp[r1] || (a.traces && rowr[r0] != rowr[r1]);
  }
  return !__syncthreads_or(hit);
}

// Replay-history slot this launch stores for shot s (or -1).
template <typename T>
__device__ __forceinline__ int cd3d_dslot(const Cd3dArgs<T> &a, int nts) {
  if (a.dmode == 1) {
    const int seg0 = ((nts - 1) / a.S) * a.S;
    return a.it > seg0 ? a.it - seg0 - 1 : -1;
  }
  return a.dmode == 2 ? a.it - a.cbase - 1 : -1;
}

// Compile-time flags, so that the march's loads and stores are
// unconditional where they can be: CORR, the launch correlates this shot's
// adjoint field; DEEP, every plane of the chunk is inside the z-interior (no
// z C-PML memory variable, no edge plane), so a plane is one straight-line
// update; DST (with DEEP), the launch stores the replay history (slot dslot).
template <typename T, bool CORR, bool DEEP, bool DST>
__device__ __forceinline__ void cd3d_lean_column(const Cd3dArgs<T> &a, int s, int dslot, int i, int j, int K0,
                                                 int K1) {
  const int nx = a.n[0], ny = a.n[1], nz = a.n[2], h = a.h;
  const int64_t so = (int64_t)s * a.N;
  const int64_t sxy = (int64_t)nx * ny;
  const T *__restrict__ pc_ = a.pcur + so;
  const T *__restrict__ po_ = a.pold + so;
  T *__restrict__ pn_ = a.pnew + so;
  const T *cb = a.cpml + (int64_t)s * a.cstride;
  const AxisCoef<T> cz = axis_coef(cb, h, 2);
  const bool corr = CORR;
  const int cslot = a.it - a.cbase;
  const int cxy = j * nx + i;
  T pm = K0 >= 1 ? pc_[(int64_t)(K0 - 1) * sxy + cxy] : T(0);
  T pc = pc_[(int64_t)K0 * sxy + cxy];
  T psz_prev = 0;
  struct Plane { T pp, pl, pr, pd, pu, po, fa, g, d; };
  auto load_plane = [&](int k) -> Plane {
    Plane q;
    const int64_t c = (int64_t)k * sxy + cxy;
    const T pnext = pc_[k + 1 < nz ? c + sxy : c];   // in bounds on the last plane too
    q.pp = k + 1 < nz ? pnext : T(0);
    q.pl = pc_[c - 1];
    q.pr = pc_[c + 1];
    q.pd = pc_[c - nx];
    q.pu = pc_[c + nx];
    q.po = po_[c];
    q.fa = a.fact[c];
    q.g = CORR ? a.grad[so + c] : T(0);
    q.d = CORR ? a.dbuf[(int64_t)cslot * a.dslot + so + c] : T(0);
    return q;
  };
  // One plane of the march; cur holds the operands loaded for plane k.
  auto step = [&](int k, const Plane &cur) {
    const int64_t c = (int64_t)k * sxy + cxy;
    const T pp = cur.pp;
    if (DEEP) {
      // the interior branch below with no z C-PML term
      const T d2x = (cur.pl - T(2) * pc + cur.pr) * a.id2[0];
      const T d2y = (cur.pd - T(2) * pc + cur.pu) * a.id2[1];
      const T d2z = (pm - T(2) * pc + pp) * a.id2[2];
      const T lx = d2x, ly = d2y, lz = d2z;
      const T flap = cur.fa * (lx + ly + lz);
      const T pnv = T(2) * pc - cur.po + flap;
      pn_[c] = pnv;
      if (DST) {
        const T inj = 0;
        a.dbuf[(int64_t)dslot * a.dslot + so + c] = flap + inj;
      }
      if (CORR) a.grad[so + c] = cur.g + pc * cur.d;
      pm = pc;
      pc = pp;
      return;
    }
    T psz = 0;
    const int kz = psi_owner(k, nz, h);
    if (kz >= 0) {
      const int64_t q = (int64_t)s * a.sps[2] + (int64_t)kz * sxy + cxy;
      psz = cz.bh(kz) * a.psi_o[2][q] + cz.ah(kz) * ((pp - pc) * a.id[2]);
      a.psi_n[2][q] = psz;
    }
    const bool interior = k >= 1 && k <= nz - 2;
    T pnv = 0, flap = 0;
    if (interior) {
      const T d2x = (cur.pl - T(2) * pc + cur.pr) * a.id2[0];
      const T d2y = (cur.pd - T(2) * pc + cur.pu) * a.id2[1];
      const T d2z = (pm - T(2) * pc + pp) * a.id2[2];
      const T lx = d2x, ly = d2y;
      T lz = d2z;
      if (k < h || k >= nz - h) {
        const int kn = kz - 1;
        T psn;
        if (k > K0) {
          psn = psz_prev;
        } else {
          const int64_t qn = (int64_t)s * a.sps[2] + (int64_t)kn * sxy + cxy;
          psn = cz.bh(kn) * a.psi_o[2][qn] + cz.ah(kn) * ((pc - pm) * a.id[2]);
        }
        const T dpsi = (psz - psn) * a.id[2];
        const int64_t qz = (int64_t)s * a.sxi[2] + (int64_t)kz * sxy + cxy;
        const T xi = cz.b(kz) * a.xi[2][qz] + cz.a(kz) * (d2z + dpsi);
        a.xi[2][qz] = xi;
        lz = d2z + dpsi + xi;
      }
      flap = cur.fa * (lx + ly + lz);
      pnv = T(2) * pc - cur.po + flap;
    } else {
      // outer face: pnew is pold's array in the reference
      pnv = cur.po;
    }
    pn_[c] = pnv;
    if (dslot >= 0) {
      const T inj = 0;
      a.dbuf[(int64_t)dslot * a.dslot + so + c] = interior ? flap + inj : (cur.po - T(2) * pc) + pnv;
    }
    if (corr) a.grad[so + c] = cur.g + pc * cur.d;
    psz_prev = psz;
    pm = pc;
    pc = pp;
  };
  // Three plane buffers used in turn (the loop is unrolled by three): the
  // loads of planes k+1 and k+2 stay in flight while plane k is stepped.
  // With one buffer copied forward each plane, the copy of registers whose
  // loads are still pending makes the wave wait for them at the end of
  // every plane.
  Plane A = load_plane(K0), B = load_plane(min(K0 + 1, K1 - 1));
  for (int k = K0; k < K1; k += 3) {
    const Plane C = load_plane(min(k + 2, K1 - 1));
    step(k, A);
    if (k + 1 >= K1) break;
    A = load_plane(min(k + 3, K1 - 1));
    step(k + 1, B);
    if (k + 2 >= K1) break;
    B = load_plane(min(k + 4, K1 - 1));
    step(k + 2, C);
  }
}

// Interior z-march of cd3d_lean_column for a chunk inside the z-interior
// (DEEP) with each plane of pcur staged through LDS: the block loads the
// plane's 64 x TY tile and its one-cell x/y halo once (one load per value
// instead of the five overlapping loads of cd3d_lean_column) into one of
// two LDS planes used in turn, and takes the x/y neighbours from there.  The
// staged values, old pressure and factor of planes k+1 and k+2 are in
// flight while plane k is stepped; one barrier per plane.  The same
// expressions in the same order as cd3d_lean_column, so the results are
// bitwise equal.  Lean tiles keep every halo cell inside the grid.
template <int TY>
struct Cd3dHalo {
  static constexpr int NH = 2 * TX + 2 * TY;                 // halo cells of a plane
  static constexpr int PER = (NH + TX * TY - 1) / (TX * TY);  // per thread
};

template <typename T, int TY, bool CORR, bool DST>
__device__ __forceinline__ void cd3d_lds_column(const Cd3dArgs<T> &a, T *sp, int s, int dslot, int i0, int j0,
                                                int K0, int K1) {
  using H = Cd3dHalo<TY>;
  constexpr int W = TX + 2;
  const int nx = a.n[0], ny = a.n[1];
  const int64_t so = (int64_t)s * a.N;
  const int64_t sxy = (int64_t)nx * ny;
  const T *__restrict__ pc_ = a.pcur + so;
  const T *__restrict__ po_ = a.pold + so;
  T *__restrict__ pn_ = a.pnew + so;
  const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * TX + tx;
  const int cxy = (j0 + ty) * nx + i0 + tx;
  const int cslot = a.it - a.cbase;
  // halo cell q of the plane: rows -1 and TY over the tile columns, then
  // columns -1 and TX over the tile rows (LDS coordinates lx, lz; in-plane
  // offset off from the tile's corner cell)
  int hl[H::PER], hoff[H::PER];
#pragma unroll
  for (int r = 0; r < H::PER; r++) {
    const int q = tid + r * TX * TY;
    int lx = -1, ly = -1;
    if (q < TX) { lx = q + 1; ly = 0; }
    else if (q < 2 * TX) { lx = q - TX + 1; ly = TY + 1; }
    else if (q < 2 * TX + TY) { lx = 0; ly = q - 2 * TX + 1; }
    else if (q < H::NH) { lx = TX + 1; ly = q - 2 * TX - TY + 1; }
    hl[r] = ly >= 0 ? ly * W + lx : -1;
    hoff[r] = ly >= 0 ? (j0 + ly - 1) * nx + i0 + lx - 1 : 0;
  }
  struct Plane { T ctr, hal[H::PER], po, fa, g, d; };
  auto load_plane = [&](int k) -> Plane {
    Plane q;
    const int64_t base = (int64_t)k * sxy;
    q.ctr = pc_[base + cxy];
#pragma unroll
    for (int r = 0; r < H::PER; r++) q.hal[r] = hl[r] >= 0 ? pc_[base + hoff[r]] : T(0);
    q.po = po_[base + cxy];
    q.fa = a.fact[base + cxy];
    q.g = CORR ? a.grad[so + base + cxy] : T(0);
    q.d = CORR ? a.dbuf[(int64_t)cslot * a.dslot + so + base + cxy] : T(0);
    return q;
  };
  auto stage = [&](int buf, const Plane &q) {
    T *s_ = sp + buf * (TY + 2) * W;
    s_[(ty + 1) * W + tx + 1] = q.ctr;
#pragma unroll
    for (int r = 0; r < H::PER; r++)
      if (hl[r] >= 0) s_[hl[r]] = q.hal[r];
  };
  T pm = pc_[(int64_t)(K0 - 1) * sxy + cxy];
  T pc;
  // plane k sits in LDS plane (k - K0) & 1; cur holds its operands, nxt
  // the staged values of plane k+1 (pp) and its operands
  auto step = [&](int k, const Plane &cur, const Plane &nxt) {
    const T *s_ = sp + ((k - K0) & 1) * (TY + 2) * W;
    const T pl = s_[(ty + 1) * W + tx], pr = s_[(ty + 1) * W + tx + 2];
    const T pd = s_[ty * W + tx + 1], pu = s_[(ty + 2) * W + tx + 1];
    const T pp = nxt.ctr;
    const int64_t c = (int64_t)k * sxy + cxy;
    const T d2x = (pl - T(2) * pc + pr) * a.id2[0];
    const T d2y = (pd - T(2) * pc + pu) * a.id2[1];
    const T d2z = (pm - T(2) * pc + pp) * a.id2[2];
    const T lx = d2x, ly = d2y, lz = d2z;
    const T flap = cur.fa * (lx + ly + lz);
    const T pnv = T(2) * pc - cur.po + flap;
    pn_[c] = pnv;
    if (DST) {
      const T inj = 0;
      a.dbuf[(int64_t)dslot * a.dslot + so + c] = flap + inj;
    }
    if (CORR) a.grad[so + c] = cur.g + pc * cur.d;
    pm = pc;
    pc = pp;
  };
  // Four plane buffers used in turn (the loop is unrolled by four): while
  // plane k is stepped, the loads of planes k+1 .. k+3 are in flight (the
  // kernel's register budget is set by the general march, so the deeper
  // queue costs no occupancy).  Plane K1 (the last plane's z neighbour) is
  // loaded too, which the DEEP bound keeps inside the grid.
  Plane A = load_plane(K0);
  Plane B = load_plane(K0 + 1);
  Plane C = load_plane(min(K0 + 2, K1));
  Plane D;
  stage(0, A);
  pc = A.ctr;
  __syncthreads();
  for (int k = K0; k < K1; k += 4) {
    D = load_plane(min(k + 3, K1));
    stage((k + 1 - K0) & 1, B);
    step(k, A, B);
    __syncthreads();
    if (k + 1 >= K1) break;
    A = load_plane(min(k + 4, K1));
    stage((k + 2 - K0) & 1, C);
    step(k + 1, B, C);
    __syncthreads();
    if (k + 2 >= K1) break;
    B = load_plane(min(k + 5, K1));
    stage((k + 3 - K0) & 1, D);
    step(k + 2, C, D);
    __syncthreads();
    if (k + 3 >= K1) break;
    C = load_plane(min(k + 6, K1));
    stage((k + 4 - K0) & 1, A);
    step(k + 3, D, A);
    __syncthreads();
  }
}

// TY: tile height (rows of the block); MODE 0 takes interior tiles through
// cd3d_lean_column and the others through the general march in one launch,
// MODE 1 runs only the interior tiles and MODE 2 only the others (the two
// launched one after the other, so each gets its own register budget);
// MODE 3 is MODE 0 with the chunks of interior tiles inside the z-interior
// stepped by cd3d_lds_column.
template <typename T, int TY, int MODE>
__global__ __launch_bounds__(TX *TY) void cd3d_step_kernel(Cd3dArgs<T> a) {
  const int b = xcd_tile3(blockIdx.x, gridDim.x);
  int t = b;
  const int bx = t % a.nbx; t /= a.nbx;
  const int by = t % a.nby; t /= a.nby;
  const int bk = t % a.nbk;
  const int s = t / a.nbk;
  const int nts = a.nt_shot[s];
  if (a.it > nts || nts <= a.act_min) return;
  const int nx = a.n[0], ny = a.n[1], nz = a.n[2], h = a.h;
  const int i = bx * TX + threadIdx.x, j = by * TY + threadIdx.y;
  const int K0 = bk * a.kch, K1 = min(K0 + a.kch, nz);
  if (MODE == 1 || (a.lean && (MODE == 0 || MODE == 3))) {
    if (cd3d_is_lean<T, TY>(a, s, bx, by, K0, K1)) {
      if (MODE != 2) {
        const bool corr = a.corr && a.it + 1 <= nts && a.it + 1 <= a.corr_lim;
        const int dslot = cd3d_dslot(a, nts);
        const bool deep = K0 >= max(h, 1) && K1 - 1 <= nz - h - 2;
        if (MODE == 3 && deep) {
          // two LDS planes of the tile and its x/y halo
          __shared__ T sp[2 * (TY + 2) * (TX + 2)];
          const int i0 = bx * TX, j0 = by * TY;
          if (corr) {
            if (dslot >= 0) cd3d_lds_column<T, TY, true, true>(a, sp, s, dslot, i0, j0, K0, K1);
            else cd3d_lds_column<T, TY, true, false>(a, sp, s, dslot, i0, j0, K0, K1);
          } else {
            if (dslot >= 0) cd3d_lds_column<T, TY, false, true>(a, sp, s, dslot, i0, j0, K0, K1);
            else cd3d_lds_column<T, TY, false, false>(a, sp, s, dslot, i0, j0, K0, K1);
          }
        } else if (deep) {
          if (corr) {
            if (dslot >= 0) cd3d_lean_column<T, true, true, true>(a, s, dslot, i, j, K0, K1);
            else cd3d_lean_column<T, true, true, false>(a, s, dslot, i, j, K0, K1);
          } else {
            if (dslot >= 0) cd3d_lean_column<T, false, true, true>(a, s, dslot, i, j, K0, K1);
            else cd3d_lean_column<T, false, true, false>(a, s, dslot, i, j, K0, K1);
          }
        } else {
          if (corr) cd3d_lean_column<T, true, false, false>(a, s, dslot, i, j, K0, K1);
          else cd3d_lean_column<T, false, false, false>(a, s, dslot, i, j, K0, K1);
        }
      }
      return;
    }
    if (MODE == 1) return;
  } else if (MODE == 2) {
    if (cd3d_is_lean<T, TY>(a, s, bx, by, K0, K1)) return;
  }
  const bool col_ok = i < nx && j < ny;
  const int64_t so = (int64_t)s * a.N;
  const int64_t sxy = (int64_t)nx * ny;
  const T *pc_ = a.pcur + so;
  const T *po_ = a.pold + so;
  T *pn_ = a.pnew + so;
  const T *cb = a.cpml + (int64_t)s * a.cstride;
  const AxisCoef<T> cx = axis_coef(cb, h, 0), cy = axis_coef(cb, h, 1), cz = axis_coef(cb, h, 2);
  const int kx = col_ok ? psi_owner(i, nx, h) : -1;
  const int ky = col_ok ? psi_owner(j, ny, h) : -1;
  const bool xin = i >= 1 && i <= nx - 2, yin = j >= 1 && j <= ny - 2;
  const bool xcpml = xin && (i < h || i >= nx - h), ycpml = yin && (j < h || j >= ny - h);

  int dslot = -1;
  if (a.dmode == 1) {
    const int seg0 = ((nts - 1) / a.S) * a.S;
    if (a.it > seg0) dslot = a.it - seg0 - 1;
  } else if (a.dmode == 2) {
    dslot = a.it - a.cbase - 1;
  }
  const bool corr = a.corr && a.it + 1 <= nts && a.it + 1 <= a.corr_lim;
  const int *rowp = a.res == nullptr ? a.pt.srow : a.pt.rrow;
  const int64_t *rcell = a.res == nullptr ? a.pt.usrc_cell : a.pt.urec_cell;
  const int *rowr = a.pt.rrow + (int64_t)s * (a.pt.nrows + 1);
  const int *rowi = rowp + (int64_t)s * (a.pt.nrows + 1);
  const int jy0 = by * TY, jy1 = min(jy0 + TY, ny);

  // rolling z window
  const int64_t cxy = (int64_t)j * nx + i;
  T pm = (col_ok && K0 >= 1) ? pc_[(int64_t)(K0 - 1) * sxy + cxy] : T(0);
  T pc = col_ok ? pc_[(int64_t)K0 * sxy + cxy] : T(0);
  T psz_prev = 0;        // new psi_z of plane k-1 (valid if k-1 >= K0)
  // Operands of plane k that do not depend on plane k-1's results are
  // loaded one plane ahead, so their latency overlaps the previous plane's
  // work (the loop would otherwise wait on every plane's loads).
  const int cslot = a.it - a.cbase;
  struct Plane { T pp, pl, pr, pd, pu, po, fa, g, d; };
  auto load_plane = [&](int k) -> Plane {
    Plane q{0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (!col_ok) return q;
    const int64_t c = (int64_t)k * sxy + cxy;
    q.pp = k + 1 < nz ? pc_[c + sxy] : T(0);
    q.pl = i >= 1 ? pc_[c - 1] : T(0);
    q.pr = i + 1 < nx ? pc_[c + 1] : T(0);
    q.pd = j >= 1 ? pc_[c - nx] : T(0);
    q.pu = j + 1 < ny ? pc_[c + nx] : T(0);
    q.po = po_[c];
    q.fa = a.fact[c];
    if (corr) {
      q.g = a.grad[so + c];
      q.d = a.dbuf[(int64_t)cslot * a.dslot + so + c];
    }
    return q;
  };
  Plane nxt = load_plane(K0);
  for (int k = K0; k < K1; k++) {
    const int64_t c = (int64_t)k * sxy + cxy;
    const Plane cur = nxt;
    if (k + 1 < K1) nxt = load_plane(k + 1);
    const T pp = cur.pp;
    // wave-uniform: any point on this plane's rows of the block?
    const bool any_pt = rowi[jy0 + (int64_t)ny * k] < rowi[jy1 + (int64_t)ny * k];
    const bool any_rec = a.traces && rowr[jy0 + (int64_t)ny * k] < rowr[jy1 + (int64_t)ny * k];
    if (col_ok) {
      const T pl = cur.pl, pr = cur.pr;
      const T pd = cur.pd, pu = cur.pu;
      // update_psi_{x,y,z}! (acoustic3D_xPU.jl:1-38)
      T psx = 0, psy = 0, psz = 0;
      if (kx >= 0) {
        const int64_t q = (int64_t)s * a.sps[0] + ((int64_t)k * ny + j) * (2 * h) + kx;
        psx = cx.bh(kx) * a.psi_o[0][q] + cx.ah(kx) * ((pr - pc) * a.id[0]);
        a.psi_n[0][q] = psx;
      }
      if (ky >= 0) {
        const int64_t q = (int64_t)s * a.sps[1] + ((int64_t)k * (2 * h) + ky) * nx + i;
        psy = cy.bh(ky) * a.psi_o[1][q] + cy.ah(ky) * ((pu - pc) * a.id[1]);
        a.psi_n[1][q] = psy;
      }
      const int kz = psi_owner(k, nz, h);
      if (kz >= 0) {
        const int64_t q = (int64_t)s * a.sps[2] + (int64_t)kz * sxy + cxy;
        psz = cz.bh(kz) * a.psi_o[2][q] + cz.ah(kz) * ((pp - pc) * a.id[2]);
        a.psi_n[2][q] = psz;
      }
      const bool interior = xin && yin && k >= 1 && k <= nz - 2;
      T pnv = 0, flap = 0, inj = 0;
      if (interior) {
        // update_p_CPML! (acoustic3D_xPU.jl:40-63): sum over x, y, z
        const T d2x = (pl - T(2) * pc + pr) * a.id2[0];
        const T d2y = (pd - T(2) * pc + pu) * a.id2[1];
        const T d2z = (pm - T(2) * pc + pp) * a.id2[2];
        T lx = d2x, ly = d2y, lz = d2z;
        if (xcpml) {
          const int kn = kx - 1;
          const int64_t qn = (int64_t)s * a.sps[0] + ((int64_t)k * ny + j) * (2 * h) + kn;
          const T psn = cx.bh(kn) * a.psi_o[0][qn] + cx.ah(kn) * ((pc - pl) * a.id[0]);
          const T dpsi = (psx - psn) * a.id[0];
          const int64_t qx = (int64_t)s * a.sxi[0] + ((int64_t)k * ny + j) * (2 * h + 2) + kx;
          const T xi = cx.b(kx) * a.xi[0][qx] + cx.a(kx) * (d2x + dpsi);
          a.xi[0][qx] = xi;
          lx = d2x + dpsi + xi;
        }
        if (ycpml) {
          const int kn = ky - 1;
          const int64_t qn = (int64_t)s * a.sps[1] + ((int64_t)k * (2 * h) + kn) * nx + i;
          const T psn = cy.bh(kn) * a.psi_o[1][qn] + cy.ah(kn) * ((pc - pd) * a.id[1]);
          const T dpsi = (psy - psn) * a.id[1];
          const int64_t qy = (int64_t)s * a.sxi[1] + ((int64_t)k * (2 * h + 2) + ky) * nx + i;
          const T xi = cy.b(ky) * a.xi[1][qy] + cy.a(ky) * (d2y + dpsi);
          a.xi[1][qy] = xi;
          ly = d2y + dpsi + xi;
        }
        if (k < h || k >= nz - h) {
          const int kn = kz - 1;
          T psn;
          if (k > K0) {
            psn = psz_prev;
          } else {
            const int64_t qn = (int64_t)s * a.sps[2] + (int64_t)kn * sxy + cxy;
            psn = cz.bh(kn) * a.psi_o[2][qn] + cz.ah(kn) * ((pc - pm) * a.id[2]);
          }
          const T dpsi = (psz - psn) * a.id[2];
          const int64_t qz = (int64_t)s * a.sxi[2] + (int64_t)kz * sxy + cxy;
          const T xi = cz.b(kz) * a.xi[2][qz] + cz.a(kz) * (d2z + dpsi);
          a.xi[2][qz] = xi;
          lz = d2z + dpsi + xi;
        }
        flap = cur.fa * (lx + ly + lz);
        pnv = T(2) * pc - cur.po + flap;
      }
      // Cells on the outer faces are outside the stencil range.  The
      // reference's rotation (acoustic3D_xPU.jl:155-157) binds pnew to the
      // array of the new pold, so they hold level it-2 plus any injection.
      if (!interior) pnv = cur.po;
      if (any_pt) {
        const int64_t row = j + (int64_t)ny * k;
        const int u = find_point(rowp, rcell, a.pt.nrows, s, row, c);
        if (u >= 0) {
          if (a.res == nullptr) {
            const T *tf = a.srctf + a.pt.off_tf[s] + (a.it - 1);
            for (int q = a.pt.usrc_ptr[u]; q < a.pt.usrc_ptr[u + 1]; q++) {
              const T v = tf[(int64_t)nts * a.pt.usrc_list[q]];
              pnv += v;
              inj += v;
            }
          } else {
            const T *rs = a.res + a.pt.off_tr[s] + (a.it - 1);
            const T f = cur.fa;
            for (int q = a.pt.urec_ptr[u]; q < a.pt.urec_ptr[u + 1]; q++) pnv += rs[(int64_t)nts * a.pt.urec_list[q]] * f;
          }
        }
      }
      pn_[c] = pnv;
      if (any_rec) {
        const int r = find_point(a.pt.rrow, a.pt.urec_cell, a.pt.nrows, s, j + (int64_t)ny * k, c);
        if (r >= 0) {
          T *tr = a.traces + a.pt.off_tr[s] + (a.it - 1);
          for (int q = a.pt.urec_ptr[r]; q < a.pt.urec_ptr[r + 1]; q++) tr[(int64_t)nts * a.pt.urec_list[q]] = pnv;
        }
      }
      if (dslot >= 0) {
        // p^{it} - 2 p^{it-1} + p^{it-2}: in the interior this is exactly the
        // stencil term plus the injection, formed without cancellation
        a.dbuf[(int64_t)dslot * a.dslot + so + c] = interior ? flap + inj : (cur.po - T(2) * pc) + pnv;
      }
      if (corr) a.grad[so + c] = cur.g + pc * cur.d;
      psz_prev = psz;
    }
    pm = pc;
    pc = pp;
  }
}

// grad += adjcur * D for the first step of a segment
template <typename T>
__global__ void cd3d_correlate_kernel(T *grad, const T *adj, const T *d, int64_t N, const int *nt_shot, int cbase) {
  const int s = blockIdx.y;
  if (nt_shot[s] <= cbase) return;
  const int64_t so = (int64_t)s * N;
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < N; q += (int64_t)gridDim.x * blockDim.x)
    grad[so + q] += adj[so + q] * d[so + q];
}

template <typename T>
class Cd3dEngine final : public Engine {
 public:
  explicit Cd3dEngine(PlanCore &core) : c_(core) {
    const swx_plan_desc &d = c_.d;
    for (int k = 0; k < 3; k++) n_[k] = (int)d.n[k];
    h_ = d.halo;
    ns_ = d.max_shots;
    N_ = (int64_t)n_[0] * n_[1] * n_[2];
    for (int ax = 0; ax < 3; ax++) {
      sps_[ax] = strip(ax, 2 * h_);
      sxi_[ax] = strip(ax, 2 * h_ + 2);
    }
    cstride_ = 3 * (8 * h_ + 4);
    const int64_t e = sizeof(T);
    fact_.alloc(N_ * e);
    p_.alloc(3 * ns_ * N_ * e);
    for (int ax = 0; ax < 3; ax++) {
      psi_[ax].alloc(std::max<int64_t>(2 * ns_ * sps_[ax], 1) * e);
      xi_[ax].alloc(ns_ * sxi_[ax] * e);
    }
    cpml_.alloc((int64_t)ns_ * cstride_ * e);
    state_bytes_ = fact_.bytes() + p_.bytes() + cpml_.bytes();
    for (int ax = 0; ax < 3; ax++) state_bytes_ += psi_[ax].bytes() + xi_[ax].bytes();
    if (d.gradient) {
      a_.alloc(p_.bytes());
      for (int ax = 0; ax < 3; ax++) {
        apsi_[ax].alloc(psi_[ax].bytes());
        axi_[ax].alloc(xi_[ax].bytes());
        state_bytes_ += apsi_[ax].bytes() + axi_[ax].bytes();
      }
      grad_.alloc(ns_ * N_ * e);
      gout_.alloc(ns_ * N_ * e);
      state_bytes_ += a_.bytes() + grad_.bytes() + gout_.bytes();
      const int64_t slot = ns_ * N_ * e;
      ck_one_ = (2 * N_ + sps_[0] + sps_[1] + sps_[2] + sxi_[0] + sxi_[1] + sxi_[2]) * ns_ * e;
      const int64_t avail = (int64_t)c_.budget - state_bytes_;
      S_ = choose_segment(d.nt_max, d.check_freq, slot, ck_one_, avail);
      K_ = (int)ceil_div(d.nt_max, S_);
      dbuf_.alloc((int64_t)S_ * slot);
      nck_ = std::max(K_ - 2, 0);
      ck_.alloc(std::max<int64_t>(nck_ * ck_one_, 1));
    }
    // 64 x 4 tiles in one launch, 24 planes per chunk: measured best of
    // 64 x {8, 4, 2} tiles, one or two launches and 8 to 64 planes on
    // 600 x 600 x 400 (the smaller blocks let more of them share a CU than
    // the general path's register count would with 64 x 8)
    kch_ = 24;
    if (const char *v = std::getenv("SWX_CD3D_ZCHUNK")) {
      kch_ = std::max(1, std::atoi(v));
      kch_env_ = true;
    }
    if (const char *v = std::getenv("SWX_CD3D_LEAN")) lean_ = std::atoi(v) != 0;
    // launch shape: 8 / 4 / 2 = one launch with 64 x 8 / 64 x 4 / 64 x 2
    // tiles, 108 / 104 / 102 = interior tiles and the others in two launches,
    // 208 / 204 / 202 = one launch with the LDS-staged interior march
    if (const char *v = std::getenv("SWX_CD3D_VARIANT")) {
      const int c = std::atoi(v);
      SWX_REQUIRE((c % 100 == 8 || c % 100 == 4 || c % 100 == 2) && c > 0 && c < 300, SWX_EINVAL,
                  "SWX_CD3D_VARIANT must be 8, 4, 2, 108, 104, 102, 208, 204 or 202");
      variant_ = c;
      variant_env_ = true;
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
    std::vector<T> hc = pack_cpml<T>(nshots, shots, 3, h_);
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
    if (!tuned_) tune_shape(save, tr);
    for (int it = 1; it <= ntmax_; it++) {
      progress("forward step", it, ntmax_);
      step_fwd(it, 0, save ? 1 : 0, 0, tr);
      take_snapshots(c_, *this, it);
      if (save && it % S_ == 0) {
        const int k = it / S_;
        if (k >= 1 && k <= K_ - 2) ckpt_copy(k, it, true);
      }
    }
  }

  // "pcur" (acou_forward.jl:55-57): level it is in slot it mod 3
  int snap_fields(int it, int s, SnapPiece *out) const override {
    out[0] = {P(it) + (int64_t)s * N_, (int64_t)(N_ * sizeof(T)), (int64_t)(N_ * sizeof(T)), 1};
    return 1;
  }

  void adjoint(bool per_shot) override {
    SWX_HIP(hipMemsetAsync(a_.get(), 0, a_.bytes(), c_.stream));
    for (int ax = 0; ax < 3; ax++) {
      SWX_HIP(hipMemsetAsync(apsi_[ax].get(), 0, apsi_[ax].bytes(), c_.stream));
      SWX_HIP(hipMemsetAsync(axi_[ax].get(), 0, axi_[ax].bytes(), c_.stream));
    }
    SWX_HIP(hipMemsetAsync(grad_.get(), 0, grad_.bytes(), c_.stream));
    const int Kmax = (int)ceil_div(ntmax_, S_);
    for (int k = Kmax - 1; k >= 0; k--) {
      progress("adjoint segment", Kmax - k, Kmax);
      const int cb = k * S_;
      bool need = false;
      for (int s = 0; s < nsh_; s++) need |= nt_[s] > cb + S_;
      if (need) {
        if (k == 0) zero_forward_state();
        else ckpt_copy(k, cb, false);
        for (int it = cb + 1; it <= cb + S_; it++) step_fwd(it, cb + S_, 2, cb, nullptr);
      }
      const int top = std::min(cb + S_, ntmax_);
      for (int it = top; it >= cb + 1; it--) step_adj(it, cb);
      dim3 g((unsigned)std::min<int64_t>(ceil_div(N_, 256), 2048), nsh_);
      cd3d_correlate_kernel<T><<<g, 256, 0, c_.stream>>>(grad_.as<T>(), A(cb + 1), dbuf_.as<T>(), N_,
                                                          c_.pts.d_nt.as<int>(), cb);
      SWX_HIP(hipGetLastError());
    }
    const T scale = T(1.0 / (c_.d.dt * c_.d.dt));
    dim3 g((unsigned)std::min<int64_t>(ceil_div(N_, 256), 8192));
    reduce_shots<T><<<g, 256, 0, c_.stream>>>(gout_.as<T>(), grad_.as<T>(), N_, nsh_, scale, per_shot ? 1 : 0);
    SWX_HIP(hipGetLastError());
  }

  void copy_grad(void *host, bool per_shot) override {
    const size_t bytes = (per_shot ? (size_t)nsh_ : 1) * N_ * sizeof(T);
    SWX_HIP(hipMemcpy(host, gout_.get(), bytes, hipMemcpyDeviceToHost));
  }

  int64_t bytes_state() const override { return state_bytes_; }
  int64_t bytes_checkpoints() const override { return ck_.bytes(); }
  int64_t bytes_replay() const override { return dbuf_.bytes(); }
  int segment() const override { return S_; }
  void variants(int32_t v[2]) const override { v[0] = variant_; v[1] = kch_; }

 private:
  int64_t strip(int ax, int w) const {
    int64_t v = 1;
    for (int d = 0; d < 3; d++) v *= d == ax ? w : n_[d];
    return v;
  }
  T *P(int k) const { return p_.as<T>() + (int64_t)mod3(k) * ns_ * N_; }
  T *A(int k) const { return a_.as<T>() + (int64_t)mod3(k) * ns_ * N_; }

  void zero_forward_state() {
    SWX_HIP(hipMemsetAsync(p_.get(), 0, p_.bytes(), c_.stream));
    for (int ax = 0; ax < 3; ax++) {
      SWX_HIP(hipMemsetAsync(psi_[ax].get(), 0, psi_[ax].bytes(), c_.stream));
      SWX_HIP(hipMemsetAsync(xi_[ax].get(), 0, xi_[ax].bytes(), c_.stream));
    }
  }

  Cd3dArgs<T> base_args(int it) const {
    Cd3dArgs<T> a{};
    for (int k = 0; k < 3; k++) {
      a.n[k] = n_[k];
      a.id[k] = T(1.0 / c_.d.h[k]);
      a.id2[k] = a.id[k] * a.id[k];
      a.sps[k] = sps_[k];
      a.sxi[k] = sxi_[k];
    }
    a.h = h_;
    a.N = N_;
    a.nbx = (int)ceil_div(n_[0], TX);
    a.nby = 0;  // set per launch shape
    a.kch = kch_;
    a.nbk = (int)ceil_div(n_[2], kch_);
    a.lean = lean_ ? 1 : 0;
    a.fact = fact_.as<T>();
    a.cpml = cpml_.as<T>();
    a.cstride = cstride_;
    a.nt_shot = c_.pts.d_nt.as<int>();
    a.it = it;
    a.dbuf = dbuf_.as<T>();
    a.dslot = (int64_t)ns_ * N_;
    a.S = S_ > 0 ? S_ : 1;
    a.pt = view(c_.pts);
    a.srctf = c_.pts.d_srctf.as<T>();
    return a;
  }

  // kind 0: one launch; 1: interior tiles and the others in two launches;
  // 2: one launch with the LDS-staged march (cd3d_lds_column)
  template <int TY>
  void launch_shape(Cd3dArgs<T> a, int kind) {
    a.nby = (int)ceil_div(n_[1], TY);
    const unsigned nb = (unsigned)((int64_t)a.nbx * a.nby * a.nbk * nsh_);
    if (kind == 1 && a.lean) {
      cd3d_step_kernel<T, TY, 1><<<nb, dim3(TX, TY), 0, c_.stream>>>(a);
      cd3d_step_kernel<T, TY, 2><<<nb, dim3(TX, TY), 0, c_.stream>>>(a);
    } else if (kind == 2) {
      cd3d_step_kernel<T, TY, 3><<<nb, dim3(TX, TY), 0, c_.stream>>>(a);
    } else {
      cd3d_step_kernel<T, TY, 0><<<nb, dim3(TX, TY), 0, c_.stream>>>(a);
    }
  }

  void launch_variant(const Cd3dArgs<T> &a, int variant) {
    const int ty = variant % 100, kind = variant / 100;
    if (ty == 4) launch_shape<4>(a, kind);
    else if (ty == 2) launch_shape<2>(a, kind);
    else launch_shape<8>(a, kind);
  }

  void launch(const Cd3dArgs<T> &a, int kind) {
    c_.ktimer.begin(c_.stream);
    launch_variant(a, variant_);
    c_.ktimer.end(c_.stream, kind);
    SWX_HIP(hipGetLastError());
  }

  // Launch shape (tile height 8 / 4 / 2; one launch, two launches or the
  // LDS-staged march) and then the planes per chunk, timed on step 1 of the zero initial state, which every
  // repetition recomputes identically.  Every choice gives bitwise the same
  // results; SWX_CD3D_VARIANT / SWX_CD3D_ZCHUNK pin them instead.
  void tune_shape(bool save, T *tr) {
    auto run = [&](int variant, int kch) {
      variant_ = variant;
      kch_ = kch;
      launch_variant(fwd_args(1, 0, save ? 1 : 0, 0, tr), variant);
      SWX_HIP(hipGetLastError());
    };
    if (!variant_env_) {
      static const int vars[9] = {4, 8, 2, 104, 108, 102, 204, 208, 202};
      const int kch = kch_;
      variant_ = vars[fastest_candidate(c_.stream, 9, [&](int c) { run(vars[c], kch); })];
    }
    if (!kch_env_) {
      static const int chunks[3] = {24, 12, 48};
      const int var = variant_;
      kch_ = chunks[fastest_candidate(c_.stream, 3, [&](int c) { run(var, chunks[c]); })];
    }
    tuned_ = true;
  }

  Cd3dArgs<T> fwd_args(int it, int act_min, int dmode, int cbase, T *traces) const {
    Cd3dArgs<T> a = base_args(it);
    a.pold = P(it - 2); a.pcur = P(it - 1); a.pnew = P(it);
    for (int ax = 0; ax < 3; ax++) {
      a.psi_o[ax] = psi_[ax].as<T>() + (int64_t)((it - 1) & 1) * ns_ * sps_[ax];
      a.psi_n[ax] = psi_[ax].as<T>() + (int64_t)(it & 1) * ns_ * sps_[ax];
      a.xi[ax] = xi_[ax].as<T>();
    }
    a.act_min = act_min;
    a.dmode = c_.d.gradient ? dmode : 0;
    a.cbase = cbase;
    a.traces = traces;
    return a;
  }

  void step_fwd(int it, int act_min, int dmode, int cbase, T *traces) {
    launch(fwd_args(it, act_min, dmode, cbase, traces), 0);
  }

  void step_adj(int it, int cbase) {
    Cd3dArgs<T> a = base_args(it);
    a.pold = A(it + 2); a.pcur = A(it + 1); a.pnew = A(it);
    for (int ax = 0; ax < 3; ax++) {
      a.psi_o[ax] = apsi_[ax].as<T>() + (int64_t)((it + 1) & 1) * ns_ * sps_[ax];
      a.psi_n[ax] = apsi_[ax].as<T>() + (int64_t)(it & 1) * ns_ * sps_[ax];
      a.xi[ax] = axi_[ax].as<T>();
    }
    a.cbase = cbase;
    a.grad = grad_.as<T>();
    a.corr = 1;
    a.corr_lim = cbase + S_;
    a.res = c_.pts.d_res.as<T>();
    launch(a, 1);
  }

  // checkpoint k (all shots): p^{c-1} | p^c | psi x, y, z | xi x, y, z
  void ckpt_copy(int k, int cb, bool save) {
    const int64_t e = sizeof(T);
    char *base = ck_.as<char>() + (int64_t)(k - 1) * ck_one_;
    const int64_t fb = (int64_t)ns_ * N_ * e;
    struct Part { void *dev; int64_t bytes; };
    std::vector<Part> parts = {{P(cb - 1), fb}, {P(cb), fb}};
    for (int ax = 0; ax < 3; ax++)
      parts.push_back({psi_[ax].as<T>() + (int64_t)(cb & 1) * ns_ * sps_[ax], ns_ * sps_[ax] * e});
    for (int ax = 0; ax < 3; ax++) parts.push_back({xi_[ax].get(), ns_ * sxi_[ax] * e});
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
  int n_[3], h_, ns_;
  int64_t N_, sps_[3], sxi_[3];
  int cstride_, kch_ = 24;
  bool lean_ = true;
  int variant_ = 4;
  bool variant_env_ = false, kch_env_ = false, tuned_ = false;
  DevBuf fact_, p_, psi_[3], xi_[3], cpml_;
  DevBuf a_, apsi_[3], axi_[3], grad_, gout_;
  DevBuf dbuf_, ck_;
  int64_t ck_one_ = 0, state_bytes_ = 0;
  int S_ = 1, K_ = 1, nck_ = 0;
  int nsh_ = 0, ntmax_ = 0;
  std::vector<int> nt_;
  bool model_set_ = false;
};

}  // namespace

std::unique_ptr<Engine> make_acoustic_cd3d(PlanCore &core) {
  if (core.d.dtype == SWX_F32) return std::make_unique<Cd3dEngine<float>>(core);
  return std::make_unique<Cd3dEngine<double>>(core);
}

}  // namespace swx
