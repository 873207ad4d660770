// 2D isotropic elastic engine (displacement-stress, 4th-order staggered
// space, 2nd-order time, C-PML, optional free surface on the top row) for a
// batch of shots: two launches per time step (stresses from ucur, then
// displacements), a receiver gather when traces are recorded, and the
// gradient correlation fused into the adjoint stress launch.
//
// Reference semantics (paths relative to the reference root):
//   step       src/models/elastic/backends/shared/elastic2D_iso_xPU.jl:1-448
//   derivs     src/models/elastic/backends/shared/freesurface_derivatives_4th_mirror.jl:1-233
//   C-PML      src/utils/fdgenerated.jl:178-199 (index maps of d~x4th/d~y4th)
//   driver     src/models/elastic/ela_gradient.jl:4-362
//   correlate  src/models/elastic/backends/shared/correlate_gradient_xPU.jl:1-82
//
// Layout: every field is an nx*nz slab per shot with x fastest (the Julia
// column-major layout); ux has nx-1 valid columns, uz nz-1 valid rows and
// sxz both, the spare column/row is never read as data.  The displacement
// update reads uold and ucur only at its own cell, so unew overwrites uold in
// place: the time level u^k lives in slot k mod 2.  The C-PML memory
// variables keep the reference's own shapes (8 arrays per shot).
//
// Adjoint.  The correlation of adjoint step c needs spatial derivatives of
// the new adjoint field a^c, which the next adjoint stress launch computes
// anyway (it reads a^c as its ucur), so that launch correlates step c with
// the replayed forward history u^{c-2}, u^{c-1}, u^c; the first step of each
// replay segment is correlated by a correlate-only launch.
#include <algorithm>
#include <cstdlib>
#include <map>

#include "swx_device.hpp"
#include "swx_internal.hpp"

namespace swx {
namespace {

constexpr int TX = 64;  // one wavefront along x
constexpr int TZ = 4;

// Point tables of one tap family for a batch (injection view: taps grouped
// by cell, in point order inside a cell; gather view: taps per point).
template <typename T>
struct TapView {
  const int64_t *cell;  // unique cells, ascending per shot
  const int *eptr;      // entries of unique cell u: [eptr[u], eptr[u+1])
  const int *epoint;    // point index (within its shot) of each entry
  const T *ecoef;
  const int *rows;      // row index of the unique cells (find_point)
  const int *poff;      // first point of each shot (gather view)
  const int *pptr;      // taps of global point p: [pptr[p], pptr[p+1])
  const int64_t *pcell;
  const T *pcoef;
};

template <typename T>
struct ElArgs {
  int nx, nz, h, freetop;
  int nbx, nbz, ch;             // z-marching launch: strips, chunks, rows per chunk
  int tbx, tbz, xcd;            // block launches: tiles along x, z; XCD-aware tile map
  int order;                    // 1: the tiles outside [ibx0, ibx1] x [iby0, iby1] first
  int ibx0, ibx1, iby0, iby1;   // tiles clear of the C-PML strips and grid edges
  int64_t N, nrows;
  T idx, idz, idt2;
  const T *lam, *mu, *muhh, *rdx, *rdz;  // rdx = dt^2/rho_ihalf, rdz = dt^2/rho_jhalf
  const T *cpml; int cstride;
  const int *nt_shot;
  int it, act_min;
  const T *ucur;          // [ns][2][N]: ux | uz at level it-1
  T *unew;                // level it (= slot of it-2, updated in place)
  T *sig;                 // [ns][3][N]: sxx | szz | sxz
  const T *psi_o; T *psi_n; int64_t psi_stride;   // C-PML memory variables, by step parity
  int64_t psioff[8];
  // points: kind 0 none, 1 moment tensor (stress), 2 force or residual (displacement)
  int kind;
  TapView<T> f0, f1;
  const T *tf; const int64_t *off_tf; int ncomp;   // per shot (nt, ncomp, npts)
  const T *mom; const int *off_mom;                // moment tensors (3 per source)
  // forward history for the gradient: slot k holds u^{base+k-1}
  T *hist; int64_t hslot; int hmode; int S; int cbase;
  // correlation (stress launch on the adjoint field)
  T *grad; int64_t gstride; int update; int corr_step; int corr_lim;
};

// Block -> (strip, chunk, shot) map giving each of the 8 XCDs a contiguous
// run, so neighbouring strips share halo lines in one L2.
__device__ __forceinline__ int xcd_tile_el(int b, int nb) {
  const int per = nb >> 3;
  if (b >= per * 8) return b;
  return (b & 7) * per + (b >> 3);
}

// Tile (bx, by) and shot s of a block launch (a 1D grid of tbx * tbz * shots
// blocks).  With xcd set, each XCD steps a contiguous run of tiles, so the
// x-halo lines a tile shares with its neighbours are found in one L2.
struct ElBlock {
  int bx, by, s;
};
//
// With order set, the first blocks of the launch take the frame of tiles
// around the interior rectangle [ibx0, ibx1] x [iby0, iby1] (C-PML strips
// and grid edges: slower, latency-bound blocks) and the rest the interior
// tiles, each range dealt to the XCDs in contiguous runs.  The slow blocks
// then start first and overlap the interior ones instead of forming the
// launch's tail.
template <typename T>
__device__ __forceinline__ ElBlock el_block(const ElArgs<T> &a) {
  if (!a.order) {
    const int b = a.xcd ? xcd_tile_el(blockIdx.x, gridDim.x) : (int)blockIdx.x;
    const int t = b / a.tbx;
    return ElBlock{b - t * a.tbx, t % a.tbz, t / a.tbz};
  }
  const int nIx = a.ibx1 - a.ibx0 + 1, nIy = a.iby1 - a.iby0 + 1;   // 0 x 0 when empty
  const int nI = nIx * nIy, F = a.tbx * a.tbz - nI;
  const int ns = gridDim.x / (a.tbx * a.tbz);
  const int G = F * ns;
  int b = blockIdx.x;
  if (b < G) {
    b = xcd_tile_el(b, G);
    const int s = b / F;
    int f = b - s * F;
    const int top = a.iby0 * a.tbx;             // rows above the interior
    if (f < top) return ElBlock{f % a.tbx, f / a.tbx, s};
    f -= top;
    const int side = a.tbx - nIx;                // frame tiles of an interior row
    if (f < nIy * side) {
      const int r = f / side, c = f - r * side;
      return ElBlock{c < a.ibx0 ? c : c + nIx, a.iby0 + r, s};
    }
    f -= nIy * side;                             // rows below the interior
    return ElBlock{f % a.tbx, a.iby0 + nIy + f / a.tbx, s};
  }
  b = xcd_tile_el(b - G, nI * ns);
  const int s = b / nI, r = b - s * nI;
  return ElBlock{a.ibx0 + r % nIx, a.iby0 + r / nIx, s};
}

// 4th-order staggered difference, freesurface_derivatives_4th_mirror.jl:2-7
template <typename T>
__device__ __forceinline__ T d4(T f1, T f2, T f3, T f4, T inv) {
  return (T(1.0 / 24.0) * f1 - T(27.0 / 24.0) * f2 + T(27.0 / 24.0) * f3 - T(1.0 / 24.0) * f4) * inv;
}

// Field accessor with the reference's 1-based indices.
template <typename T>
struct F2 {
  const T *p; int nx;
  __device__ __forceinline__ T operator()(int i, int j) const {
    SWX_DASSERT(i >= 1 && i <= nx && j >= 1);
    return p[(i - 1) + (int64_t)(j - 1) * nx];
  }
};

// Derivative wrappers with the boundary rules of
// freesurface_derivatives_4th_mirror.jl:72-233.
template <typename T, typename A>
__device__ __forceinline__ T d_sxx_dx(A s, int i, int j, int nx, T id) {
  if (i == 1) return d4<T>(0, s(i, j), s(i + 1, j), s(i + 2, j), id);
  if (i == nx - 1) return d4<T>(s(i - 1, j), s(i, j), s(i + 1, j), 0, id);
  return d4<T>(s(i - 1, j), s(i, j), s(i + 1, j), s(i + 2, j), id);
}
template <typename T, typename A>
__device__ __forceinline__ T d_szz_dz(A s, int i, int j, int nz, bool ft, T id) {
  if (j == 1) return d4<T>(ft ? -s(i, j + 1) : T(0), s(i, j), s(i, j + 1), s(i, j + 2), id);
  if (j == nz - 1) return d4<T>(s(i, j - 1), s(i, j), s(i, j + 1), 0, id);
  return d4<T>(s(i, j - 1), s(i, j), s(i, j + 1), s(i, j + 2), id);
}
template <typename T, typename A>
__device__ __forceinline__ T d_sxz_dx(A s, int i, int j, int nx, T id) {
  if (i == 1) return d4<T>(0, 0, s(i, j), s(i + 1, j), id);
  if (i == 2) return d4<T>(0, s(i - 1, j), s(i, j), s(i + 1, j), id);
  if (i == nx - 1) return d4<T>(s(i - 2, j), s(i - 1, j), s(i, j), 0, id);
  if (i == nx) return d4<T>(s(i - 2, j), s(i - 1, j), 0, 0, id);
  return d4<T>(s(i - 2, j), s(i - 1, j), s(i, j), s(i + 1, j), id);
}
template <typename T, typename A>
__device__ __forceinline__ T d_sxz_dz(A s, int i, int j, int nz, bool ft, T id) {
  if (j == 1) {
    if (ft) return d4<T>(-s(i, j + 1), -s(i, j), s(i, j), s(i, j + 1), id);
    return d4<T>(0, 0, s(i, j), s(i, j + 1), id);
  }
  if (j == 2) return d4<T>(ft ? -s(i, j - 1) : T(0), s(i, j - 1), s(i, j), s(i, j + 1), id);
  if (j == nz - 1) return d4<T>(s(i, j - 2), s(i, j - 1), s(i, j), 0, id);
  if (j == nz) return d4<T>(s(i, j - 2), s(i, j - 1), 0, 0, id);
  return d4<T>(s(i, j - 2), s(i, j - 1), s(i, j), s(i, j + 1), id);
}
template <typename T, typename A>
__device__ __forceinline__ T d_ux_dx(A u, int i, int j, int nx, T id) {
  if (i == 1) return d4<T>(0, 0, u(i, j), u(i + 1, j), id);
  if (i == 2) return d4<T>(0, u(i - 1, j), u(i, j), u(i + 1, j), id);
  if (i == nx - 1) return d4<T>(u(i - 2, j), u(i - 1, j), u(i, j), 0, id);
  if (i == nx) return d4<T>(u(i - 2, j), u(i - 1, j), 0, 0, id);
  return d4<T>(u(i - 2, j), u(i - 1, j), u(i, j), u(i + 1, j), id);
}
// duz/dz; at the free surface from sigma_zz = 0: -lam/(lam+2mu) dux/dx
template <typename T, typename A, typename B>
__device__ __forceinline__ T d_uz_dz(A ux, B uz, T l, T m, int i, int j, int nx, int nz, bool ft, T idx, T idz) {
  if (j == 1) {
    if (ft) return (-l / (l + T(2) * m)) * d_ux_dx<T>(ux, i, j, nx, idx);
    return d4<T>(0, 0, uz(i, j), uz(i, j + 1), idz);
  }
  if (j == 2) return d4<T>(ft ? uz(i, j - 1) : T(0), uz(i, j - 1), uz(i, j), uz(i, j + 1), idz);
  if (j == nz - 1) return d4<T>(uz(i, j - 2), uz(i, j - 1), uz(i, j), 0, idz);
  if (j == nz) return d4<T>(uz(i, j - 2), uz(i, j - 1), 0, 0, idz);
  return d4<T>(uz(i, j - 2), uz(i, j - 1), uz(i, j), uz(i, j + 1), idz);
}
template <typename T, typename A>
__device__ __forceinline__ T d_ux_dz(A u, int i, int j, int nz, bool ft, T id) {
  if (j == 1) return d4<T>(ft ? u(i, j + 1) : T(0), u(i, j), u(i, j + 1), u(i, j + 2), id);
  if (j == nz - 1) return d4<T>(u(i, j - 1), u(i, j), u(i, j + 1), 0, id);
  return d4<T>(u(i, j - 1), u(i, j), u(i, j + 1), u(i, j + 2), id);
}
template <typename T, typename A>
__device__ __forceinline__ T d_uz_dx(A u, int i, int j, int nx, T id) {
  if (i == 1) return d4<T>(0, u(i, j), u(i + 1, j), u(i + 2, j), id);
  if (i == nx - 1) return d4<T>(u(i - 1, j), u(i, j), u(i + 1, j), 0, id);
  return d4<T>(u(i - 1, j), u(i, j), u(i + 1, j), u(i + 2, j), id);
}

// C-PML strip index (1-based, 0 = none) of d~4th (fdgenerated.jl:178-195) for
// base index I along an axis on which the differentiated field has n points.
__device__ __forceinline__ int strip_half(int I, int n, int h) {  // half = true
  const int idim = I + 1;
  if (idim <= h + 1) return idim;
  if (idim >= n - h) return I - (n - h) + 2 + h;
  return 0;
}
__device__ __forceinline__ int strip_full(int I, int n, int h) {  // half = false
  if (I <= h) return I;
  if (I >= n - h) return I - (n - h) + 1 + h;
  return 0;
}

// Is block (bx, by) of shot s an interior block of the stress (stress =
// true) or displacement step: every cell of the block takes the plain
// 4th-order branch of each derivative helper (3 <= i <= n-2 along both
// axes), lies outside every C-PML strip (h+2 <= i <= n-h-1), and no tap of
// the launch's point family lies in its rows?
template <typename T>
__device__ __forceinline__ bool el_is_lean(const ElArgs<T> &a, int s, int bx, int by, bool stress) {
  const int i0 = bx * TX + 1, i1 = i0 + TX - 1, j0 = by * TZ + 1, j1 = j0 + TZ - 1;
  const int lo = max(3, a.h + 2);
  if (!(i0 >= lo && i1 <= min(a.nx - 2, a.nx - a.h - 1) && j0 >= lo && j1 <= min(a.nz - 2, a.nz - a.h - 1)))
    return false;
  if (a.kind == (stress ? 1 : 2)) {
    const int64_t o = (int64_t)s * (a.nrows + 1);
    if (a.f0.rows[o + j0 - 1] != a.f0.rows[o + j1]) return false;
    if (a.f1.rows[o + j0 - 1] != a.f1.rows[o + j1]) return false;
  }
  return true;
}

// Stress step (and, on the adjoint field, the correlation) of an interior
// block: the expressions of el_stress_kernel, in the same operand order,
// with every global load issued first.
template <typename T>
__device__ __forceinline__ void el_stress_lean_block(const ElArgs<T> &a, ElBlock B, bool upd, bool corr) {
  const int nx = a.nx, s = B.s;
  const int i = B.bx * TX + threadIdx.x + 1;
  const int j = B.by * TZ + threadIdx.y + 1;
  const int64_t N = a.N;
  const int c = (i - 1) + (j - 1) * nx;
  const T *__restrict__ ux = a.ucur + (int64_t)s * 2 * N;
  const T *__restrict__ uz = ux + N;
  // ux along x (i-2..i+1) and z (j-1..j+2), uz along z (j-2..j+1) and x (i-1..i+2)
  const T x_m2 = ux[c - 2], x_m1 = ux[c - 1], x_0 = ux[c], x_p1 = ux[c + 1];
  const T x_zm1 = ux[c - nx], x_zp1 = ux[c + nx], x_zp2 = ux[c + 2 * nx];
  const T z_m2 = uz[c - 2 * nx], z_m1 = uz[c - nx], z_0 = uz[c], z_p1 = uz[c + nx];
  const T z_xm1 = uz[c - 1], z_xp1 = uz[c + 1], z_xp2 = uz[c + 2];
  const T l = a.lam[c], m = a.mu[c], mhh = a.muhh[c];
  T h0x = 0, h1x = 0, h2x = 0, h0z = 0, h1z = 0, h2z = 0;
  T fx_m2 = 0, fx_m1 = 0, fx_p1 = 0, fx_zm1 = 0, fx_zp1 = 0, fx_zp2 = 0;
  T fz_m2 = 0, fz_m1 = 0, fz_p1 = 0, fz_xm1 = 0, fz_xp1 = 0, fz_xp2 = 0;
  T g0 = 0, g1 = 0, g2 = 0, g3 = 0, g4 = 0;
  T *g = a.grad + (int64_t)s * a.gstride;
  if (corr) {
    const int sl = a.corr_step - a.cbase + 1;   // slot of u^{corr_step}
    const T *h0 = a.hist + (int64_t)(sl - 2) * a.hslot + (int64_t)s * 2 * N;
    const T *h1 = h0 + a.hslot;
    const T *h2 = h1 + a.hslot;
    h0x = h0[c]; h1x = h1[c]; h2x = h2[c];
    h0z = h0[N + c]; h1z = h1[N + c]; h2z = h2[N + c];
    fx_m2 = h1[c - 2]; fx_m1 = h1[c - 1]; fx_p1 = h1[c + 1];
    fx_zm1 = h1[c - nx]; fx_zp1 = h1[c + nx]; fx_zp2 = h1[c + 2 * nx];
    fz_m2 = h1[N + c - 2 * nx]; fz_m1 = h1[N + c - nx]; fz_p1 = h1[N + c + nx];
    fz_xm1 = h1[N + c - 1]; fz_xp1 = h1[N + c + 1]; fz_xp2 = h1[N + c + 2];
    g0 = g[c]; g1 = g[N + c]; g2 = g[2 * N + c]; g3 = g[3 * N + c]; g4 = g[4 * N + c];
  }
  const T dux = d4<T>(x_m2, x_m1, x_0, x_p1, a.idx);
  const T duz = d4<T>(z_m2, z_m1, z_0, z_p1, a.idz);
  const T duzx = d4<T>(z_xm1, z_0, z_xp1, z_xp2, a.idx);
  const T duxz = d4<T>(x_zm1, x_0, x_zp1, x_zp2, a.idz);
  if (upd) {
    T *sg = a.sig + (int64_t)s * 3 * N;
    const T dxc = dux, dzc = duz;
    sg[c] = (l + T(2) * m) * dxc + l * dzc;
    sg[N + c] = l * dxc + (l + T(2) * m) * dzc;
    const T c1 = duzx, c2 = duxz;
    sg[2 * N + c] = mhh * (c1 + c2);
  }
  if (corr) {
    // correlate_gradients! of adjoint step corr_step (adjcur = ucur here)
    const T v = x_0 * (h0x - T(2) * h1x + h2x) * a.idt2;
    g[c] = g0 + v;
    g[N + c] = g1 + z_0 * (h0z - T(2) * h1z + h2z) * a.idt2;
    const T exx = d4<T>(fx_m2, fx_m1, h1x, fx_p1, a.idx);
    const T ezz = d4<T>(fz_m2, fz_m1, h1z, fz_p1, a.idz);
    const T dv = exx + ezz, dva = dux + duz;
    g[2 * N + c] = g2 + dv * dva;
    g[3 * N + c] = g3 + T(2) * (exx * dux + ezz * duz);
    const T exz = (d4<T>(fz_xm1, h1z, fz_xp1, fz_xp2, a.idx) + d4<T>(fx_zm1, h1x, fx_zp1, fx_zp2, a.idz)) / T(2);
    const T exza = (duzx + duxz) / T(2);
    g[4 * N + c] = g4 + T(2) * (exz * exza + exz * exza);
  }
}

// Displacement step of an interior block: the expressions of
// el_disp_kernel, in the same operand order, with every global load issued
// first, and the history store.
template <typename T>
__device__ __forceinline__ void el_disp_lean_block(const ElArgs<T> &a, ElBlock B, int nts) {
  const int nx = a.nx, s = B.s;
  const int i = B.bx * TX + threadIdx.x + 1;
  const int j = B.by * TZ + threadIdx.y + 1;
  const int64_t N = a.N;
  const int c = (i - 1) + (j - 1) * nx;
  const T *__restrict__ uc = a.ucur + (int64_t)s * 2 * N;
  T *__restrict__ un = a.unew + (int64_t)s * 2 * N;
  const T *__restrict__ sxx = a.sig + (int64_t)s * 3 * N;
  const T *__restrict__ szz = sxx + N;
  const T *__restrict__ sxz = sxx + 2 * N;
  const T a_m1 = sxx[c - 1], a_0 = sxx[c], a_p1 = sxx[c + 1], a_p2 = sxx[c + 2];
  const T b_zm2 = sxz[c - 2 * nx], b_zm1 = sxz[c - nx], b_0 = sxz[c], b_zp1 = sxz[c + nx];
  const T b_xm2 = sxz[c - 2], b_xm1 = sxz[c - 1], b_xp1 = sxz[c + 1];
  const T z_m1 = szz[c - nx], z_0 = szz[c], z_p1 = szz[c + nx], z_p2 = szz[c + 2 * nx];
  const T oldx = un[c], oldz = un[N + c], curx = uc[c], curz = uc[N + c];
  const T rx = a.rdx[c], rz = a.rdz[c];
  const T d1x = d4<T>(a_m1, a_0, a_p1, a_p2, a.idx);
  const T d2x = d4<T>(b_zm2, b_zm1, b_0, b_zp1, a.idz);
  const T nux = T(2) * curx - oldx + rx * (d1x + d2x);
  const T d1z = d4<T>(b_xm2, b_xm1, b_0, b_xp1, a.idx);
  const T d2z = d4<T>(z_m1, z_0, z_p1, z_p2, a.idz);
  const T nuz = T(2) * curz - oldz + rz * (d1z + d2z);
  un[c] = nux;
  un[N + c] = nuz;
  if (a.hmode) {
    const int base = a.hmode == 1 ? (nts / a.S) * a.S : a.cbase;
    if (a.it > base) {
      T *hs = a.hist + (int64_t)s * 2 * N + c;
      const int sl = a.it - base + 1;
      hs[(int64_t)sl * a.hslot] = nux;
      hs[(int64_t)sl * a.hslot + N] = nuz;
      if (a.hmode == 1 && a.it == base + 1) {
        hs[0] = oldx;
        hs[N] = oldz;
        hs[a.hslot] = curx;
        hs[a.hslot + N] = curz;
      }
    }
  }
}

// Interior blocks of a stress / displacement step; other blocks return at
// once and el_stress_kernel / el_disp_kernel<T, false>, launched right
// after on the same grid, step them.
template <typename T>
__global__ __launch_bounds__(TX *TZ) void el_stress_lean(ElArgs<T> a) {
  const ElBlock B = el_block(a);
  const int s = B.s;
  const int nts = a.nt_shot[s];
  const bool upd = a.update && a.it <= nts && nts > a.act_min;
  const bool corr = a.corr_step > 0 && a.corr_step <= nts && a.corr_step <= a.corr_lim;
  if (!upd && !corr) return;
  if (!el_is_lean(a, s, B.bx, B.by, true)) return;
  el_stress_lean_block(a, B, upd, corr);
}
template <typename T>
__global__ __launch_bounds__(TX *TZ) void el_disp_lean(ElArgs<T> a) {
  const ElBlock B = el_block(a);
  const int s = B.s;
  const int nts = a.nt_shot[s];
  if (a.it > nts || nts <= a.act_min) return;
  if (!el_is_lean(a, s, B.bx, B.by, false)) return;
  el_disp_lean_block(a, B, nts);
}

// Stress update of one step (update_sxx_szz! :39-60, update_sxz! :62-79),
// moment-tensor injection (:81-94) and, on the adjoint field, the fused
// correlation of adjoint step corr_step (correlate_gradient_xPU.jl:1-82).
template <typename T, bool FOLD>
__global__ __launch_bounds__(TX *TZ) void el_stress_kernel(ElArgs<T> a) {
  const ElBlock B = el_block(a);
  const int s = B.s;
  const int nts = a.nt_shot[s];
  const bool upd = a.update && a.it <= nts && nts > a.act_min;
  const bool corr = a.corr_step > 0 && a.corr_step <= nts && a.corr_step <= a.corr_lim;
  if (!upd && !corr) return;
  if (el_is_lean(a, s, B.bx, B.by, true)) {   // interior block
    if (FOLD) el_stress_lean_block(a, B, upd, corr);
    return;
  }
  const int i = B.bx * TX + threadIdx.x + 1;
  const int j = B.by * TZ + threadIdx.y + 1;
  const int nx = a.nx, nz = a.nz, h = a.h;
  if (i > nx || j > nz) return;
  const bool ft = a.freetop != 0;
  const int64_t N = a.N;
  const int64_t c = (int64_t)(i - 1) + (int64_t)(j - 1) * nx;
  const F2<T> ux{a.ucur + (int64_t)s * 2 * N, nx}, uz{a.ucur + (int64_t)s * 2 * N + N, nx};
  T *sg = a.sig + (int64_t)s * 3 * N;
  const T *pso = a.psi_o + (int64_t)s * a.psi_stride;   // C-PML memory, parity of the read state
  T *psn = a.psi_n + (int64_t)s * a.psi_stride;         // parity of the written state
  const T *cb = a.cpml + (int64_t)s * a.cstride;
  const AxisCoef<T> cx = axis_coef(cb, h, 0), cz = axis_coef(cb, h, 1);

  const bool nrm = i >= 2 && i <= nx - 1 && j >= (ft ? 1 : 2) && j <= nz - 1;
  const bool shr = i <= nx - 1 && j <= nz - 1;
  T dux = 0, duz = 0, duzx = 0, duxz = 0;
  T vxx = 0, vzz = 0, vxz = 0;
  bool wn = false, ws = false;
  if (nrm) {
    const T l = a.lam[c], m = a.mu[c];
    dux = d_ux_dx(ux, i, j, nx, a.idx);
    duz = d_uz_dz(ux, uz, l, m, i, j, nx, nz, ft, a.idx, a.idz);
    if (upd) {
      T dxc = dux, dzc = duz;
      int k = strip_half(i - 1, nx - 1, h);
      if (k) {
        const int64_t q = a.psioff[4] + (k - 1) + (int64_t)(j - 1) * (2 * h + 2);
        const T pn = cx.b(k - 1) * pso[q] + cx.a(k - 1) * dux;
        psn[q] = pn;
        dxc = dux + pn;
      }
      k = strip_half(j - 1, nz - 1, h);
      if (k) {
        const int64_t q = a.psioff[7] + (i - 1) + (int64_t)(k - 1) * nx;
        const T pn = cz.b(k - 1) * pso[q] + cz.a(k - 1) * duz;
        psn[q] = pn;
        dzc = duz + pn;
      }
      vxx = (l + T(2) * m) * dxc + l * dzc;
      vzz = j == 1 ? T(0) : l * dxc + (l + T(2) * m) * dzc;
      wn = true;
    }
  }
  if (shr) {
    duzx = d_uz_dx(uz, i, j, nx, a.idx);
    duxz = d_ux_dz(ux, i, j, nz, ft, a.idz);
    if (upd) {
      T c1 = duzx, c2 = duxz;
      int k = strip_full(i, nx, h);
      if (k) {
        const int64_t q = a.psioff[5] + (k - 1) + (int64_t)(j - 1) * (2 * h);
        const T pn = cx.bh(k - 1) * pso[q] + cx.ah(k - 1) * duzx;
        psn[q] = pn;
        c1 = duzx + pn;
      }
      k = strip_full(j, nz, h);
      if (k) {
        const int64_t q = a.psioff[6] + (i - 1) + (int64_t)(k - 1) * (nx - 1);
        const T pn = cz.bh(k - 1) * pso[q] + cz.ah(k - 1) * duxz;
        psn[q] = pn;
        c2 = duxz + pn;
      }
      vxz = a.muhh[c] * (c1 + c2);
      ws = true;
    }
  }
  if (upd && a.kind == 1) {
    // moment-tensor sources in index order; cells the stencil does not
    // update keep their previous value plus the injection
    const T *tf = a.tf + a.off_tf[s] + (a.it - 1);
    const T *M = a.mom + 3 * (int64_t)a.off_mom[s];
    int u = find_point(a.f0.rows, a.f0.cell, a.nrows, s, j - 1, c);
    if (u >= 0) {
      if (!wn) { vxx = sg[c]; vzz = sg[N + c]; }
      for (int e = a.f0.eptr[u]; e < a.f0.eptr[u + 1]; e++) {
        const int p = a.f0.epoint[e];
        const T f = tf[(int64_t)nts * p];
        vxx += M[3 * p + 0] * a.f0.ecoef[e] * f;
        vzz += M[3 * p + 1] * a.f0.ecoef[e] * f;
      }
      wn = true;
    }
    u = find_point(a.f1.rows, a.f1.cell, a.nrows, s, j - 1, c);
    if (u >= 0) {
      if (!ws) vxz = sg[2 * N + c];
      for (int e = a.f1.eptr[u]; e < a.f1.eptr[u + 1]; e++) {
        const int p = a.f1.epoint[e];
        vxz += M[3 * p + 2] * a.f1.ecoef[e] * tf[(int64_t)nts * p];
      }
      ws = true;
    }
  }
  if (wn) { sg[c] = vxx; sg[N + c] = vzz; }
  if (ws) sg[2 * N + c] = vxz;

  if (corr) {
    // correlate_gradients! of adjoint step corr_step (adjcur = ucur here)
    const int sl = a.corr_step - a.cbase + 1;   // slot of u^{corr_step}
    const T *h0 = a.hist + (int64_t)(sl - 2) * a.hslot + (int64_t)s * 2 * N;
    const T *h1 = h0 + a.hslot;
    const T *h2 = h1 + a.hslot;
    T *g = a.grad + (int64_t)s * a.gstride;
    if (i <= nx - 1) {
      const T v = ux(i, j) * (h0[c] - T(2) * h1[c] + h2[c]) * a.idt2;
      g[c] += (j == 1 && ft) ? v / T(2) : v;
    }
    if (j <= nz - 1) g[N + c] += uz(i, j) * (h0[N + c] - T(2) * h1[N + c] + h2[N + c]) * a.idt2;
    const F2<T> fx{h1, nx}, fz{h1 + N, nx};
    if (nrm) {
      const T exx = d_ux_dx(fx, i, j, nx, a.idx);
      const T ezz = d_uz_dz(fx, fz, a.lam[c], a.mu[c], i, j, nx, nz, ft, a.idx, a.idz);
      const T dv = exx + ezz, dva = dux + duz;
      if (j == 1 && ft) {
        g[2 * N + c] += dv * dva / T(2);
        g[3 * N + c] += exx * dux + ezz * duz;
      } else {
        g[2 * N + c] += dv * dva;
        g[3 * N + c] += T(2) * (exx * dux + ezz * duz);
      }
    }
    if (shr) {
      const T exz = (d_uz_dx(fz, i, j, nx, a.idx) + d_ux_dz(fx, i, j, nz, ft, a.idz)) / T(2);
      const T exza = (duzx + duxz) / T(2);
      g[4 * N + c] += T(2) * (exz * exza + exz * exza);
    }
  }
}

// Displacement update (update_ux! :1-18, update_uz! :20-37) into the slot of
// uold, force / adjoint-residual injection (:96-106) and the forward history.
template <typename T, bool FOLD>
__global__ __launch_bounds__(TX *TZ) void el_disp_kernel(ElArgs<T> a) {
  const ElBlock B = el_block(a);
  const int s = B.s;
  const int nts = a.nt_shot[s];
  if (a.it > nts || nts <= a.act_min) return;
  if (el_is_lean(a, s, B.bx, B.by, false)) {   // interior block
    if (FOLD) el_disp_lean_block(a, B, nts);
    return;
  }
  const int i = B.bx * TX + threadIdx.x + 1;
  const int j = B.by * TZ + threadIdx.y + 1;
  const int nx = a.nx, nz = a.nz, h = a.h;
  if (i > nx || j > nz) return;
  const bool ft = a.freetop != 0;
  const int64_t N = a.N;
  const int64_t c = (int64_t)(i - 1) + (int64_t)(j - 1) * nx;
  const T *uc = a.ucur + (int64_t)s * 2 * N;
  T *un = a.unew + (int64_t)s * 2 * N;
  const T *sg = a.sig + (int64_t)s * 3 * N;
  const F2<T> sxx{sg, nx}, szz{sg + N, nx}, sxz{sg + 2 * N, nx};
  const T *pso = a.psi_o + (int64_t)s * a.psi_stride;   // C-PML memory, parity of the read state
  T *psn = a.psi_n + (int64_t)s * a.psi_stride;         // parity of the written state
  const T *cb = a.cpml + (int64_t)s * a.cstride;
  const AxisCoef<T> cx = axis_coef(cb, h, 0), cz = axis_coef(cb, h, 1);
  const bool hx = i <= nx - 1, hz = j <= nz - 1;
  T oldx = 0, oldz = 0, curx = 0, curz = 0, nux = 0, nuz = 0;
  if (hx) {
    const T d1 = d_sxx_dx(sxx, i, j, nx, a.idx);
    const T d2 = d_sxz_dz(sxz, i, j, nz, ft, a.idz);
    T c1 = d1, c2 = d2;
    int k = strip_full(i, nx, h);
    if (k) {
      const int64_t q = a.psioff[0] + (k - 1) + (int64_t)(j - 1) * (2 * h);
      const T pn = cx.bh(k - 1) * pso[q] + cx.ah(k - 1) * d1;
      psn[q] = pn;
      c1 = d1 + pn;
    }
    k = strip_half(j - 1, nz - 1, h);
    if (k) {
      const int64_t q = a.psioff[3] + (i - 1) + (int64_t)(k - 1) * (nx - 1);
      const T pn = cz.b(k - 1) * pso[q] + cz.a(k - 1) * d2;
      psn[q] = pn;
      c2 = d2 + pn;
    }
    oldx = un[c];
    curx = uc[c];
    nux = T(2) * curx - oldx + a.rdx[c] * (c1 + c2);
  }
  if (hz) {
    const T d1 = d_sxz_dx(sxz, i, j, nx, a.idx);
    const T d2 = d_szz_dz(szz, i, j, nz, ft, a.idz);
    T c1 = d1, c2 = d2;
    int k = strip_half(i - 1, nx - 1, h);
    if (k) {
      const int64_t q = a.psioff[1] + (k - 1) + (int64_t)(j - 1) * (2 * h + 2);
      const T pn = cx.b(k - 1) * pso[q] + cx.a(k - 1) * d1;
      psn[q] = pn;
      c1 = d1 + pn;
    }
    k = strip_full(j, nz, h);
    if (k) {
      const int64_t q = a.psioff[2] + (i - 1) + (int64_t)(k - 1) * nx;
      const T pn = cz.bh(k - 1) * pso[q] + cz.ah(k - 1) * d2;
      psn[q] = pn;
      c2 = d2 + pn;
    }
    oldz = un[N + c];
    curz = uc[N + c];
    nuz = T(2) * curz - oldz + a.rdz[c] * (c1 + c2);
  }
  if (a.kind == 2) {
    // point forces (forward) or adjoint residuals, in point order
    const T *tf = a.tf + a.off_tf[s] + (a.it - 1);
    if (hx) {
      const int u = find_point(a.f0.rows, a.f0.cell, a.nrows, s, j - 1, c);
      if (u >= 0)
        for (int e = a.f0.eptr[u]; e < a.f0.eptr[u + 1]; e++)
          nux += a.f0.ecoef[e] * tf[(int64_t)nts * (a.ncomp * a.f0.epoint[e])] * a.rdx[c];
    }
    if (hz) {
      const int u = find_point(a.f1.rows, a.f1.cell, a.nrows, s, j - 1, c);
      if (u >= 0)
        for (int e = a.f1.eptr[u]; e < a.f1.eptr[u + 1]; e++)
          nuz += a.f1.ecoef[e] * tf[(int64_t)nts * (a.ncomp * a.f1.epoint[e] + 1)] * a.rdz[c];
    }
  }
  if (hx) un[c] = nux;
  if (hz) un[N + c] = nuz;
  if (a.hmode) {
    // forward: the final segment of this shot, from its last checkpoint
    // floor(nt/S)*S on (LinearCheckpointer, checkpointers.jl); replay: the
    // segment being replayed (its first two levels come from the checkpoint)
    const int base = a.hmode == 1 ? (nts / a.S) * a.S : a.cbase;
    if (a.it > base) {
      T *hs = a.hist + (int64_t)s * 2 * N + c;
      const int sl = a.it - base + 1;
      hs[(int64_t)sl * a.hslot] = nux;
      hs[(int64_t)sl * a.hslot + N] = nuz;
      if (a.hmode == 1 && a.it == base + 1) {
        hs[0] = oldx;
        hs[N] = oldz;
        hs[a.hslot] = curx;
        hs[a.hslot + N] = curz;
      }
    }
  }
}

// Copies one displacement level (ux | uz of every shot) into a history slot
// for the shots being replayed (nt_s > act_min).
template <typename T>
__global__ void el_hist_copy(T *dst, const T *src, int64_t n2, const int *nt_shot, int act_min) {
  const int s = blockIdx.y;
  if (nt_shot[s] <= act_min) return;
  const int64_t o = (int64_t)s * n2;
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n2; q += (int64_t)gridDim.x * blockDim.x)
    dst[o + q] = src[o + q];
}


// Row ring in LDS seen through the reference's 1-based (i, j): row j lives
// in slot j mod NR, column i at index i - i0.
template <typename T, int NR, int PITCH>
struct RingA {
  const T *b;
  int i0;
  __device__ __forceinline__ T operator()(int i, int j) const {
    return b[(((j % NR) + NR) % NR) * PITCH + (i - i0)];
  }
};

// Moment-tensor injection into the normal-stress cells the stencil never
// updates (columns 1 and nx, row nz, row 1 without free surface): there the
// reference's stresses keep accumulating (elastic2D_iso_xPU.jl:81-94 on a
// field update_sxx_szz! does not touch).  Run before el_march_kernel, which
// reads these cells from HBM.
template <typename T>
__global__ void el_edge_inject_kernel(ElArgs<T> a) {
  const int s = blockIdx.y;
  const int nts = a.nt_shot[s];
  if (a.it > nts || nts <= a.act_min) return;
  const int *rows = a.f0.rows + (int64_t)s * (a.nrows + 1);
  const int u = rows[0] + blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= rows[a.nrows]) return;
  const int64_t c = a.f0.cell[u];
  const int i = (int)(c % a.nx) + 1, j = (int)(c / a.nx) + 1;
  const bool nrm = i >= 2 && i <= a.nx - 1 && j >= (a.freetop ? 1 : 2) && j <= a.nz - 1;
  if (nrm) return;
  T *sg = a.sig + (int64_t)s * 3 * a.N;
  const T *tf = a.tf + a.off_tf[s] + (a.it - 1);
  const T *M = a.mom + 3 * (int64_t)a.off_mom[s];
  T vxx = sg[c], vzz = sg[a.N + c];
  for (int e = a.f0.eptr[u]; e < a.f0.eptr[u + 1]; e++) {
    const int p = a.f0.epoint[e];
    const T f = tf[(int64_t)nts * p];
    vxx += M[3 * p + 0] * a.f0.ecoef[e] * f;
    vzz += M[3 * p + 1] * a.f0.ecoef[e] * f;
  }
  sg[c] = vxx;
  sg[a.N + c] = vzz;
}

// One time step marching down z: stresses of row j (own columns plus the
// halo columns the displacement stencils need) from LDS rings of ucur rows,
// then the displacements of row j-2, whose stress rows j-4..j are in the
// stress rings.  The stresses never go to HBM (except the never-updated
// edge cells, see el_edge_inject_kernel).  A block owns a strip of 64*W
// columns and a chunk of rows.  Forward (ADJ = false): moment-tensor and
// force injection, history store.  Adjoint (ADJ = true): residual injection
// and the correlation of adjoint step corr_step on the own stress cells,
// with rings of the replayed level corr_step-1.  Two barriers per row.
template <typename T, int W, bool ADJ>
__global__ __launch_bounds__(64 * W) void el_march_kernel(ElArgs<T> a) {
  constexpr int TXM = 64 * W;
  constexpr int PU = TXM + 6;   // ux, uz rings: columns [I1-3, I1+TXM+2]
  constexpr int PS = TXM + 3;   // sxx ring [I1-1, I1+TXM+1]; sxz ring [I1-2, I1+TXM]; history rings
  constexpr int NH = ADJ ? 5 : 1;
  const int b = xcd_tile_el(blockIdx.x, gridDim.x);
  const int bx = b % a.nbx, bzs = b / a.nbx, bz = bzs % a.nbz, s = bzs / a.nbz;
  const int nts = a.nt_shot[s];
  if (ADJ ? a.it > nts : (a.it > nts || nts <= a.act_min)) return;
  const int nx = a.nx, nz = a.nz, h = a.h;
  const bool ft = a.freetop != 0;
  const int64_t N = a.N;
  const int I1 = bx * TXM + 1;
  const int Zd0 = bz * a.ch + 1, Zd1 = min(bz * a.ch + a.ch, nz) + 1;   // owned rows [Zd0, Zd1)
  const int tx = threadIdx.x;
  const int i = I1 + tx;
  const T *ucx = a.ucur + (int64_t)s * 2 * N, *ucz = ucx + N;
  T *un = a.unew + (int64_t)s * 2 * N;
  const T *pso = a.psi_o + (int64_t)s * a.psi_stride;
  T *psn = a.psi_n + (int64_t)s * a.psi_stride;
  const T *cb = a.cpml + (int64_t)s * a.cstride;
  const AxisCoef<T> cx = axis_coef(cb, h, 0), cz = axis_coef(cb, h, 1);
  const bool corr = ADJ && a.corr_step > 0 && a.corr_step <= nts && a.corr_step <= a.corr_lim;

  __shared__ T rux[6][PU], ruz[5][PU];
  __shared__ T rsxx[3][PS], rszz[4][TXM], rsxz[5][PS];
  __shared__ T rhx[NH][PS], rhz[NH][PS];   // level corr_step-1: ux [I1-2, I1+TXM], uz [I1-1, I1+TXM+1]
  const RingA<T, 6, PU> UX{&rux[0][0], I1 - 3};
  const RingA<T, 5, PU> UZ{&ruz[0][0], I1 - 3};
  const RingA<T, 3, PS> SXX{&rsxx[0][0], I1 - 1};
  const RingA<T, 4, TXM> SZZ{&rszz[0][0], I1};
  const RingA<T, 5, PS> SXZ{&rsxz[0][0], I1 - 2};
  const RingA<T, NH, PS> HX{&rhx[0][0], I1 - 2};
  const RingA<T, NH, PS> HZ{&rhz[0][0], I1 - 1};
  auto slot = [](int j, int nr) { return ((j % nr) + nr) % nr; };

  // column duties: u halo (tx < 6), sxx halo / sxz halo / history halos (tx < 3)
  const int hu_lx = tx < 3 ? tx : TXM + tx, hu_x = I1 - 3 + hu_lx;
  const int hxx_x = tx == 0 ? I1 - 1 : I1 + TXM + tx - 1;
  const int hxz_x = tx < 2 ? I1 - 2 + tx : I1 + TXM;
  const int hhx_x = hxz_x;                       // history ux halo: I1-2, I1-1, I1+TXM
  const int hhz_x = tx == 0 ? I1 - 1 : I1 + TXM + tx - 1;   // history uz halo: I1-1, I1+TXM, I1+TXM+1
  auto ldx = [&](const T *f, int x, int j) -> T {   // ux-grid field (nx-1) x nz, 1-based
    return (x >= 1 && x <= nx - 1 && j >= 1 && j <= nz) ? f[(int64_t)(j - 1) * nx + (x - 1)] : T(0);
  };
  auto ldz = [&](const T *f, int x, int j) -> T {   // uz-grid field nx x (nz-1)
    return (x >= 1 && x <= nx && j >= 1 && j <= nz - 1) ? f[(int64_t)(j - 1) * nx + (x - 1)] : T(0);
  };
  auto ldn = [&](const T *f, int x, int j) -> T {   // node field nx x nz
    return (x >= 1 && x <= nx && j >= 1 && j <= nz) ? f[(int64_t)(j - 1) * nx + (x - 1)] : T(0);
  };
  auto ldh = [&](const T *f, int x, int j) -> T {   // sxz-grid field (nx-1) x (nz-1)
    return (x >= 1 && x <= nx - 1 && j >= 1 && j <= nz - 1) ? f[(int64_t)(j - 1) * nx + (x - 1)] : T(0);
  };
  const T *h0 = nullptr, *h1 = nullptr, *h2 = nullptr;
  if (ADJ && corr) {
    const int sl = a.corr_step - a.cbase + 1;
    h0 = a.hist + (int64_t)(sl - 2) * a.hslot + (int64_t)s * 2 * N;
    h1 = h0 + a.hslot;
    h2 = h1 + a.hslot;
  }

  // preload: ucur rows Zd0-3 .. Zd0-1 (ux), Zd0-4 .. Zd0-2 (uz); history alike
  for (int j = Zd0 - 3; j <= Zd0 - 1; j++) {
    rux[slot(j, 6)][tx + 3] = ldx(ucx, i, j);
    if (tx < 6) rux[slot(j, 6)][hu_lx] = ldx(ucx, hu_x, j);
    if (ADJ && corr) {
      rhx[slot(j, NH)][tx + 2] = ldx(h1, i, j);
      if (tx < 3) rhx[slot(j, NH)][hhx_x - (I1 - 2)] = ldx(h1, hhx_x, j);
    }
  }
  for (int j = Zd0 - 4; j <= Zd0 - 2; j++) {
    ruz[slot(j, 5)][tx + 3] = ldz(ucz, i, j);
    if (tx < 6) ruz[slot(j, 5)][hu_lx] = ldz(ucz, hu_x, j);
    if (ADJ && corr) {
      rhz[slot(j, NH)][tx + 1] = ldz(h1 + N, i, j);
      if (tx < 3) rhz[slot(j, NH)][hhz_x - (I1 - 1)] = ldz(h1 + N, hhz_x, j);
    }
  }

  // prefetch registers of iteration j (loaded during iteration j-1)
  T n_ux = 0, n_uxh = 0, n_uz = 0, n_uzh = 0, n_l = 0, n_m = 0, n_lh = 0, n_mh = 0, n_mhh = 0, n_mhhh = 0;
  T n_ox = 0, n_oz = 0, n_rx = 0, n_rz = 0;
  T n_hx = 0, n_hxh = 0, n_hz = 0, n_hzh = 0, n_h0x = 0, n_h0z = 0, n_h2x = 0, n_h2z = 0;
  T n_g[5] = {0, 0, 0, 0, 0};
  auto issue = [&](int j) {
    n_ux = ldx(ucx, i, j + 2);
    n_uz = ldz(ucz, i, j + 1);
    n_l = ldn(a.lam, i, j);
    n_m = ldn(a.mu, i, j);
    n_mhh = ldh(a.muhh, i, j);
    n_ox = ldx(un, i, j - 2);
    n_oz = ldz(un + N, i, j - 2);
    n_rx = ldx(a.rdx, i, j - 2);
    n_rz = ldz(a.rdz, i, j - 2);
    if (tx < 6) {
      n_uxh = ldx(ucx, hu_x, j + 2);
      n_uzh = ldz(ucz, hu_x, j + 1);
    }
    if (tx < 3) {
      n_lh = ldn(a.lam, hxx_x, j);
      n_mh = ldn(a.mu, hxx_x, j);
      n_mhhh = ldh(a.muhh, hxz_x, j);
    }
    if (ADJ && corr) {
      n_hx = ldx(h1, i, j + 2);
      n_hz = ldz(h1 + N, i, j + 1);
      if (tx < 3) {
        n_hxh = ldx(h1, hhx_x, j + 2);
        n_hzh = ldz(h1 + N, hhz_x, j + 1);
      }
      n_h0x = ldx(h0, i, j);
      n_h2x = ldx(h2, i, j);
      n_h0z = ldz(h0 + N, i, j);
      n_h2z = ldz(h2 + N, i, j);
      T *g = a.grad + (int64_t)s * a.gstride;
      n_g[0] = ldx(g, i, j);
      n_g[1] = ldz(g + N, i, j);
      n_g[2] = ldn(g + 2 * N, i, j);
      n_g[3] = ldn(g + 3 * N, i, j);
      n_g[4] = ldh(g + 4 * N, i, j);
    }
  };
  issue(Zd0 - 2);
  __syncthreads();

  // normal stresses at (x, j); raw derivatives returned for the correlation
  auto normal = [&](int x, int j, T l, T m, bool owned, T &vxx, T &vzz, T &dux, T &duz) {
    vxx = vzz = dux = duz = 0;
    if (x < 1 || x > nx || j < 1 || j > nz) return;
    const int64_t c = (int64_t)(j - 1) * nx + (x - 1);
    const bool nrm = x >= 2 && x <= nx - 1 && j >= (ft ? 1 : 2) && j <= nz - 1;
    if (!nrm) {
      // never updated: zero, or the accumulated injections (forward)
      if (!ADJ && a.kind == 1) {
        const T *sg = a.sig + (int64_t)s * 3 * N;
        vxx = sg[c];
        vzz = sg[N + c];
      }
      return;
    }
    dux = d_ux_dx<T>(UX, x, j, nx, a.idx);
    duz = d_uz_dz<T>(UX, UZ, l, m, x, j, nx, nz, ft, a.idx, a.idz);
    T dxc = dux, dzc = duz;
    int k = strip_half(x - 1, nx - 1, h);
    if (k) {
      const int64_t q = a.psioff[4] + (k - 1) + (int64_t)(j - 1) * (2 * h + 2);
      const T pn = cx.b(k - 1) * pso[q] + cx.a(k - 1) * dux;
      if (owned) psn[q] = pn;
      dxc = dux + pn;
    }
    k = strip_half(j - 1, nz - 1, h);
    if (k) {
      const int64_t q = a.psioff[7] + (x - 1) + (int64_t)(k - 1) * nx;
      const T pn = cz.b(k - 1) * pso[q] + cz.a(k - 1) * duz;
      if (owned) psn[q] = pn;
      dzc = duz + pn;
    }
    vxx = (l + T(2) * m) * dxc + l * dzc;
    vzz = j == 1 ? T(0) : l * dxc + (l + T(2) * m) * dzc;
    if (!ADJ && a.kind == 1) {
      const int u = find_point(a.f0.rows, a.f0.cell, a.nrows, s, j - 1, c);
      if (u >= 0) {
        const T *tf = a.tf + a.off_tf[s] + (a.it - 1);
        const T *M = a.mom + 3 * (int64_t)a.off_mom[s];
        for (int e = a.f0.eptr[u]; e < a.f0.eptr[u + 1]; e++) {
          const int p = a.f0.epoint[e];
          const T f = tf[(int64_t)nts * p];
          vxx += M[3 * p + 0] * a.f0.ecoef[e] * f;
          vzz += M[3 * p + 1] * a.f0.ecoef[e] * f;
        }
      }
    }
  };
  // shear stress at (x, j)
  auto shear = [&](int x, int j, T mhh, bool owned, T &vxz, T &duzx, T &duxz) {
    vxz = duzx = duxz = 0;
    if (x < 1 || x > nx - 1 || j < 1 || j > nz - 1) return;
    const int64_t c = (int64_t)(j - 1) * nx + (x - 1);
    duzx = d_uz_dx<T>(UZ, x, j, nx, a.idx);
    duxz = d_ux_dz<T>(UX, x, j, nz, ft, a.idz);
    T c1 = duzx, c2 = duxz;
    int k = strip_full(x, nx, h);
    if (k) {
      const int64_t q = a.psioff[5] + (k - 1) + (int64_t)(j - 1) * (2 * h);
      const T pn = cx.bh(k - 1) * pso[q] + cx.ah(k - 1) * duzx;
      if (owned) psn[q] = pn;
      c1 = duzx + pn;
    }
    k = strip_full(j, nz, h);
    if (k) {
      const int64_t q = a.psioff[6] + (x - 1) + (int64_t)(k - 1) * (nx - 1);
      const T pn = cz.bh(k - 1) * pso[q] + cz.ah(k - 1) * duxz;
      if (owned) psn[q] = pn;
      c2 = duxz + pn;
    }
    vxz = mhh * (c1 + c2);
    if (!ADJ && a.kind == 1) {
      const int u = find_point(a.f1.rows, a.f1.cell, a.nrows, s, j - 1, c);
      if (u >= 0) {
        const T *tf = a.tf + a.off_tf[s] + (a.it - 1);
        const T *M = a.mom + 3 * (int64_t)a.off_mom[s];
        for (int e = a.f1.eptr[u]; e < a.f1.eptr[u + 1]; e++) {
          const int p = a.f1.epoint[e];
          vxz += M[3 * p + 2] * a.f1.ecoef[e] * tf[(int64_t)nts * p];
        }
      }
    }
  };

  const int sb = a.hmode == 1 ? (nts / a.S) * a.S : a.cbase;
  for (int j = Zd0 - 2; j <= Zd1 + 1; j++) {
    // commit ucur rows j+2 (ux), j+1 (uz) and the history rows
    rux[slot(j + 2, 6)][tx + 3] = n_ux;
    ruz[slot(j + 1, 5)][tx + 3] = n_uz;
    if (tx < 6) {
      rux[slot(j + 2, 6)][hu_lx] = n_uxh;
      ruz[slot(j + 1, 5)][hu_lx] = n_uzh;
    }
    if (ADJ && corr) {
      rhx[slot(j + 2, NH)][tx + 2] = n_hx;
      rhz[slot(j + 1, NH)][tx + 1] = n_hz;
      if (tx < 3) {
        rhx[slot(j + 2, NH)][hhx_x - (I1 - 2)] = n_hxh;
        rhz[slot(j + 1, NH)][hhz_x - (I1 - 1)] = n_hzh;
      }
    }
    const T c_l = n_l, c_m = n_m, c_lh = n_lh, c_mh = n_mh, c_mhh = n_mhh, c_mhhh = n_mhhh;
    const T c_ox = n_ox, c_oz = n_oz, c_rx = n_rx, c_rz = n_rz;
    const T c_h0x = n_h0x, c_h0z = n_h0z, c_h2x = n_h2x, c_h2z = n_h2z;
    T c_g[5];
#pragma unroll
    for (int q = 0; q < 5; q++) c_g[q] = n_g[q];
    if (j + 1 <= Zd1 + 1) issue(j + 1);
    __syncthreads();

    // ---- stresses of row j --------------------------------------------
    const bool own_row = j >= Zd0 && j < Zd1;
    const bool own = own_row && i <= nx;
    {
      T vxx, vzz, dux, duz, vxz, duzx, duxz;
      normal(i, j, c_l, c_m, own, vxx, vzz, dux, duz);
      shear(i, j, c_mhh, own, vxz, duzx, duxz);
      rsxx[slot(j, 3)][tx + 1] = vxx;
      rszz[slot(j, 4)][tx] = vzz;
      rsxz[slot(j, 5)][tx + 2] = vxz;
      if (ADJ && corr && own) {
        // correlate_gradients! (correlate_gradient_xPU.jl:1-82) of step corr_step
        T *g = a.grad + (int64_t)s * a.gstride;
        const int64_t c = (int64_t)(j - 1) * nx + (i - 1);
        if (i <= nx - 1) {
          const T v = UX(i, j) * (c_h0x - T(2) * HX(i, j) + c_h2x) * a.idt2;
          g[c] = c_g[0] + ((j == 1 && ft) ? v / T(2) : v);
        }
        if (j <= nz - 1) g[N + c] = c_g[1] + UZ(i, j) * (c_h0z - T(2) * HZ(i, j) + c_h2z) * a.idt2;
        if (i >= 2 && i <= nx - 1 && j >= (ft ? 1 : 2) && j <= nz - 1) {
          const T exx = d_ux_dx<T>(HX, i, j, nx, a.idx);
          const T ezz = d_uz_dz<T>(HX, HZ, c_l, c_m, i, j, nx, nz, ft, a.idx, a.idz);
          const T dv = exx + ezz, dva = dux + duz;
          if (j == 1 && ft) {
            g[2 * N + c] = c_g[2] + dv * dva / T(2);
            g[3 * N + c] = c_g[3] + (exx * dux + ezz * duz);
          } else {
            g[2 * N + c] = c_g[2] + dv * dva;
            g[3 * N + c] = c_g[3] + T(2) * (exx * dux + ezz * duz);
          }
        }
        if (i <= nx - 1 && j <= nz - 1) {
          const T exz = (d_uz_dx<T>(HZ, i, j, nx, a.idx) + d_ux_dz<T>(HX, i, j, nz, ft, a.idz)) / T(2);
          const T exza = (duzx + duxz) / T(2);
          g[4 * N + c] = c_g[4] + T(2) * (exz * exza + exz * exza);
        }
      }
    }
    if (tx < 3) {
      T vxx, vzz, d0, d1;
      normal(hxx_x, j, c_lh, c_mh, false, vxx, vzz, d0, d1);
      rsxx[slot(j, 3)][hxx_x - (I1 - 1)] = vxx;
      T vxz, d2, d3;
      shear(hxz_x, j, c_mhhh, false, vxz, d2, d3);
      rsxz[slot(j, 5)][hxz_x - (I1 - 2)] = vxz;
    }
    __syncthreads();

    // ---- displacements of row jd = j-2 (update_ux!/update_uz!, :1-37) ---
    const int jd = j - 2;
    if (jd >= Zd0 && jd < Zd1 && i <= nx) {
      const int64_t c = (int64_t)(jd - 1) * nx + (i - 1);
      const bool hx = i <= nx - 1, hz = jd <= nz - 1;
      T curx = 0, curz = 0, nux = 0, nuz = 0;
      if (hx) {
        const T d1 = d_sxx_dx<T>(SXX, i, jd, nx, a.idx);
        const T d2 = d_sxz_dz<T>(SXZ, i, jd, nz, ft, a.idz);
        T c1 = d1, c2 = d2;
        int k = strip_full(i, nx, h);
        if (k) {
          const int64_t q = a.psioff[0] + (k - 1) + (int64_t)(jd - 1) * (2 * h);
          const T pn = cx.bh(k - 1) * pso[q] + cx.ah(k - 1) * d1;
          psn[q] = pn;
          c1 = d1 + pn;
        }
        k = strip_half(jd - 1, nz - 1, h);
        if (k) {
          const int64_t q = a.psioff[3] + (i - 1) + (int64_t)(k - 1) * (nx - 1);
          const T pn = cz.b(k - 1) * pso[q] + cz.a(k - 1) * d2;
          psn[q] = pn;
          c2 = d2 + pn;
        }
        curx = UX(i, jd);
        nux = T(2) * curx - c_ox + c_rx * (c1 + c2);
      }
      if (hz) {
        const T d1 = d_sxz_dx<T>(SXZ, i, jd, nx, a.idx);
        const T d2 = d_szz_dz<T>(SZZ, i, jd, nz, ft, a.idz);
        T c1 = d1, c2 = d2;
        int k = strip_half(i - 1, nx - 1, h);
        if (k) {
          const int64_t q = a.psioff[1] + (k - 1) + (int64_t)(jd - 1) * (2 * h + 2);
          const T pn = cx.b(k - 1) * pso[q] + cx.a(k - 1) * d1;
          psn[q] = pn;
          c1 = d1 + pn;
        }
        k = strip_full(jd, nz, h);
        if (k) {
          const int64_t q = a.psioff[2] + (i - 1) + (int64_t)(k - 1) * nx;
          const T pn = cz.bh(k - 1) * pso[q] + cz.ah(k - 1) * d2;
          psn[q] = pn;
          c2 = d2 + pn;
        }
        curz = UZ(i, jd);
        nuz = T(2) * curz - c_oz + c_rz * (c1 + c2);
      }
      if (a.kind == 2) {
        // point forces (forward) or adjoint residuals, in point order
        const T *tf = a.tf + a.off_tf[s] + (a.it - 1);
        if (hx) {
          const int u = find_point(a.f0.rows, a.f0.cell, a.nrows, s, jd - 1, c);
          if (u >= 0)
            for (int e = a.f0.eptr[u]; e < a.f0.eptr[u + 1]; e++)
              nux += a.f0.ecoef[e] * tf[(int64_t)nts * (a.ncomp * a.f0.epoint[e])] * c_rx;
        }
        if (hz) {
          const int u = find_point(a.f1.rows, a.f1.cell, a.nrows, s, jd - 1, c);
          if (u >= 0)
            for (int e = a.f1.eptr[u]; e < a.f1.eptr[u + 1]; e++)
              nuz += a.f1.ecoef[e] * tf[(int64_t)nts * (a.ncomp * a.f1.epoint[e] + 1)] * c_rz;
        }
      }
      if (hx) un[c] = nux;
      if (hz) un[N + c] = nuz;
      if (!ADJ && a.hmode && a.it > sb) {
        T *hs = a.hist + (int64_t)s * 2 * N + c;
        const int sl = a.it - sb + 1;
        hs[(int64_t)sl * a.hslot] = nux;
        hs[(int64_t)sl * a.hslot + N] = nuz;
        if (a.hmode == 1 && a.it == sb + 1) {
          hs[0] = c_ox;
          hs[N] = c_oz;
          hs[a.hslot] = curx;
          hs[a.hslot + N] = curz;
        }
      }
    }
  }
}

// Rows of displacement cells per block of the fused tile step (64 columns).
template <typename T> struct ElTile;
template <> struct ElTile<float> { static constexpr int rows = 16; };
template <> struct ElTile<double> { static constexpr int rows = 8; };

// Region of a tile held in LDS and read through the reference's 1-based
// (i, j): element (i, j) at [(j - j0) * W + (i - i0)].
template <typename T, int W>
struct TileA {
  const T *b;
  int i0, j0;
  __device__ __forceinline__ T operator()(int i, int j) const {
    SWX_DASSERT(i - i0 >= 0 && i - i0 < W && j - j0 >= 0);
    return b[(j - j0) * W + (i - i0)];
  }
};

// correlate_gradients! of adjoint step corr_step (correlate_gradient_xPU.jl:
// 1-82) at stress cell (x, j) of shot s, as el_stress_kernel: ux, uz the
// adjoint field there, dux .. duxz its raw derivatives.
template <typename T>
__device__ __forceinline__ void el_tile_correlate(const ElArgs<T> &a, int s, int x, int j, T ux, T uz, bool nrm,
                                                  bool shr, T dux, T duz, T duzx, T duxz) {
  const int nx = a.nx, nz = a.nz;
  const bool ft = a.freetop != 0;
  const int64_t N = a.N;
  const int64_t c = (int64_t)(j - 1) * nx + (x - 1);
  const int sl = a.corr_step - a.cbase + 1;   // slot of u^{corr_step}
  const T *h0 = a.hist + (int64_t)(sl - 2) * a.hslot + (int64_t)s * 2 * N;
  const T *h1 = h0 + a.hslot;
  const T *h2 = h1 + a.hslot;
  T *g = a.grad + (int64_t)s * a.gstride;
  if (x <= nx - 1) {
    const T v = ux * (h0[c] - T(2) * h1[c] + h2[c]) * a.idt2;
    g[c] += (j == 1 && ft) ? v / T(2) : v;
  }
  if (j <= nz - 1) g[N + c] += uz * (h0[N + c] - T(2) * h1[N + c] + h2[N + c]) * a.idt2;
  const F2<T> fx{h1, nx}, fz{h1 + N, nx};
  if (nrm) {
    const T exx = d_ux_dx(fx, x, j, nx, a.idx);
    const T ezz = d_uz_dz(fx, fz, a.lam[c], a.mu[c], x, j, nx, nz, ft, a.idx, a.idz);
    const T dv = exx + ezz, dva = dux + duz;
    if (j == 1 && ft) {
      g[2 * N + c] += dv * dva / T(2);
      g[3 * N + c] += exx * dux + ezz * duz;
    } else {
      g[2 * N + c] += dv * dva;
      g[3 * N + c] += T(2) * (exx * dux + ezz * duz);
    }
  }
  if (shr) {
    const T exz = (d_uz_dx(fz, x, j, nx, a.idx) + d_ux_dz(fx, x, j, nz, ft, a.idz)) / T(2);
    const T exza = (duzx + duxz) / T(2);
    g[4 * N + c] += T(2) * (exz * exza + exz * exza);
  }
}

// The step of an interior tile (no boundary branch of any derivative
// helper, no C-PML strip and no point in reach): every global load of the
// tile (model factors of the stress cells, uold and dt^2/rho of the
// displacement cells, ucur) is issued before the first barrier, then the
// stresses and the displacements are formed with the helpers' interior
// expressions in the same operand order (bitwise the general path).
template <typename T, bool ADJ>
__device__ __forceinline__ void el_tile_lean(const ElArgs<T> &a, int s, int nts, int I1, int J1, bool corr, T *lux,
                                             T *luz, T *lsxx, T *lszz, T *lsxz) {
  constexpr int R = ElTile<T>::rows;
  constexpr int UW = TX + 8, UH = R + 8;
  constexpr int SW = TX + 4, SH = R + 4;
  constexpr int NT = TX * TZ;
  constexpr int NSQ = (SW * SH + NT - 1) / NT;   // stress cells per thread
  constexpr int NR = R / TZ;                     // displacement rows per thread
  const int nx = a.nx;
  const int64_t N = a.N;
  const int tid = threadIdx.y * TX + threadIdx.x;
  const T *ucx = a.ucur + (int64_t)s * 2 * N, *ucz = ucx + N;
  T *un = a.unew + (int64_t)s * 2 * N;
  constexpr int NU = (UW * UH + NT - 1) / NT;    // staged ucur cells per thread
  // Every load below is unconditional (indexes past the last cell of a
  // region are clamped onto it and their values not used), so they are all
  // in flight together before the first LDS store.
  T ml[NSQ], mm[NSQ], mh[NSQ];
#pragma unroll
  for (int k = 0; k < NSQ; k++) {
    const int q = min(tid + k * NT, SW * SH - 1);
    const int r = q / SW, x = I1 - 2 + (q - r * SW), j = J1 - 2 + r;
    const int64_t c = (int64_t)(j - 1) * nx + (x - 1);
    ml[k] = a.lam[c];
    mm[k] = a.mu[c];
    mh[k] = a.muhh[c];
  }
  const int i = I1 + threadIdx.x;
  T ox[NR], oz[NR], rx[NR], rz[NR];
#pragma unroll
  for (int k = 0; k < NR; k++) {
    const int j = J1 + threadIdx.y + k * TZ;
    const int64_t c = (int64_t)(i - 1) + (int64_t)(j - 1) * nx;
    ox[k] = un[c];
    oz[k] = un[N + c];
    rx[k] = a.rdx[c];
    rz[k] = a.rdz[c];
  }
  // ucur on the tile + 4: inside the grid for an interior tile (column nx,
  // outside the ux grid, only when h = 0: loaded, stored as zero)
  T vx[NU], vz[NU];
#pragma unroll
  for (int k = 0; k < NU; k++) {
    const int q = min(tid + k * NT, UW * UH - 1);
    const int r = q / UW, x = I1 - 4 + (q - r * UW), j = J1 - 4 + r;
    const int64_t c = (int64_t)(j - 1) * nx + (x - 1);
    vx[k] = ucx[c];
    vz[k] = ucz[c];
  }
  // adjoint with the correlation: the forward field of step corr_step
  // (history level h1) on the stress box, staged in LDS after the
  // displacements, and the history levels h0, h2 and the five gradient
  // fields at the tile's own cells (loaded after the first barrier)
  const T *h0p = nullptr;
  if (ADJ && corr) h0p = a.hist + (int64_t)(a.corr_step - a.cbase - 1) * a.hslot + (int64_t)s * 2 * N;
  T hx[ADJ ? NSQ : 1], hz[ADJ ? NSQ : 1];
  if (ADJ && corr) {
    const T *h1p = h0p + a.hslot;
#pragma unroll
    for (int k = 0; k < NSQ; k++) {
      const int q = min(tid + k * NT, SW * SH - 1);
      const int r = q / SW, x = I1 - 2 + (q - r * SW), j = J1 - 2 + r;
      const int64_t c = (int64_t)(j - 1) * nx + (x - 1);
      hx[k] = h1p[c];
      hz[k] = h1p[N + c];
    }
  }
#pragma unroll
  for (int k = 0; k < NU; k++) {
    const int q = tid + k * NT;
    if (q < UW * UH) {
      const int x = I1 - 4 + q % UW;
      lux[q] = x <= nx - 1 ? vx[k] : T(0);
      luz[q] = vz[k];
    }
  }
  __syncthreads();
  const TileA<T, UW> UX{lux, I1 - 4, J1 - 4}, UZ{luz, I1 - 4, J1 - 4};
  T c0x[ADJ ? NR : 1], c0z[ADJ ? NR : 1], c2x[ADJ ? NR : 1], c2z[ADJ ? NR : 1], cg[ADJ ? NR : 1][5];
  if (ADJ && corr) {
    const T *h2p = h0p + 2 * a.hslot;
    const T *g = a.grad + (int64_t)s * a.gstride;
#pragma unroll
    for (int k = 0; k < NR; k++) {
      const int j = J1 + threadIdx.y + k * TZ;
      const int64_t c = (int64_t)(i - 1) + (int64_t)(j - 1) * nx;
      c0x[k] = h0p[c];
      c0z[k] = h0p[N + c];
      c2x[k] = h2p[c];
      c2z[k] = h2p[N + c];
#pragma unroll
      for (int f = 0; f < 5; f++) cg[k][f] = g[f * N + c];
    }
  }
#pragma unroll
  for (int k = 0; k < NSQ; k++) {
    const int q = tid + k * NT;
    if (q < SW * SH) {
      const int r = q / SW, x = I1 - 2 + (q - r * SW), j = J1 - 2 + r;
      const T l = ml[k], m = mm[k], mhh = mh[k];
      const T dux = d4<T>(UX(x - 2, j), UX(x - 1, j), UX(x, j), UX(x + 1, j), a.idx);
      const T duz = d4<T>(UZ(x, j - 2), UZ(x, j - 1), UZ(x, j), UZ(x, j + 1), a.idz);
      const T dxc = dux, dzc = duz;
      lsxx[q] = (l + T(2) * m) * dxc + l * dzc;
      lszz[q] = l * dxc + (l + T(2) * m) * dzc;
      const T duzx = d4<T>(UZ(x - 1, j), UZ(x, j), UZ(x + 1, j), UZ(x + 2, j), a.idx);
      const T duxz = d4<T>(UX(x, j - 1), UX(x, j), UX(x, j + 1), UX(x, j + 2), a.idz);
      const T c1 = duzx, c2 = duxz;
      lsxz[q] = mhh * (c1 + c2);
    }
  }
  __syncthreads();
  const TileA<T, SW> SXX{lsxx, I1 - 2, J1 - 2}, SZZ{lszz, I1 - 2, J1 - 2}, SXZ{lsxz, I1 - 2, J1 - 2};
  const int sb = a.hmode == 1 ? (nts / a.S) * a.S : a.cbase;
#pragma unroll
  for (int k = 0; k < NR; k++) {
    const int j = J1 + threadIdx.y + k * TZ;
    const int64_t c = (int64_t)(i - 1) + (int64_t)(j - 1) * nx;
    const T curx = UX(i, j), curz = UZ(i, j);
    const T d1x = d4<T>(SXX(i - 1, j), SXX(i, j), SXX(i + 1, j), SXX(i + 2, j), a.idx);
    const T d2x = d4<T>(SXZ(i, j - 2), SXZ(i, j - 1), SXZ(i, j), SXZ(i, j + 1), a.idz);
    const T c1x = d1x, c2x = d2x;
    const T nux = T(2) * curx - ox[k] + rx[k] * (c1x + c2x);
    const T d1z = d4<T>(SXZ(i - 2, j), SXZ(i - 1, j), SXZ(i, j), SXZ(i + 1, j), a.idx);
    const T d2z = d4<T>(SZZ(i, j - 1), SZZ(i, j), SZZ(i, j + 1), SZZ(i, j + 2), a.idz);
    const T c1z = d1z, c2z = d2z;
    const T nuz = T(2) * curz - oz[k] + rz[k] * (c1z + c2z);
    un[c] = nux;
    un[N + c] = nuz;
    if (a.hmode && a.it > sb) {
      T *hs = a.hist + (int64_t)s * 2 * N + c;
      const int sl = a.it - sb + 1;
      hs[(int64_t)sl * a.hslot] = nux;
      hs[(int64_t)sl * a.hslot + N] = nuz;
      if (a.hmode == 1 && a.it == sb + 1) {
        hs[0] = ox[k];
        hs[N] = oz[k];
        hs[a.hslot] = curx;
        hs[a.hslot + N] = curz;
      }
    }
  }
  if (!(ADJ && corr)) return;
  // correlate_gradients! of step corr_step at the tile's own stress cells
  // (el_tile_correlate's expressions for an interior cell, same operand
  // order): h1 goes to the stress planes of LDS, which the displacements
  // have consumed
  __syncthreads();
#pragma unroll
  for (int k = 0; k < (ADJ ? NSQ : 0); k++) {
    const int q = tid + k * NT;
    if (q < SW * SH) {
      lsxx[q] = hx[k];
      lszz[q] = hz[k];
    }
  }
  __syncthreads();
  const TileA<T, SW> FX{lsxx, I1 - 2, J1 - 2}, FZ{lszz, I1 - 2, J1 - 2};
  T *g = a.grad + (int64_t)s * a.gstride;
#pragma unroll
  for (int k = 0; k < (ADJ ? NR : 0); k++) {
    const int x = i, j = J1 + threadIdx.y + k * TZ;
    const int64_t c = (int64_t)(j - 1) * nx + (x - 1);
    const T ux = UX(x, j), uz = UZ(x, j);
    const T dux = d4<T>(UX(x - 2, j), UX(x - 1, j), UX(x, j), UX(x + 1, j), a.idx);
    const T duz = d4<T>(UZ(x, j - 2), UZ(x, j - 1), UZ(x, j), UZ(x, j + 1), a.idz);
    const T duzx = d4<T>(UZ(x - 1, j), UZ(x, j), UZ(x + 1, j), UZ(x + 2, j), a.idx);
    const T duxz = d4<T>(UX(x, j - 1), UX(x, j), UX(x, j + 1), UX(x, j + 2), a.idz);
    // the statements of el_tile_correlate (the x term is formed on its own
    // before the sum, the z term inside it), so that contraction into FMAs
    // happens at the same places
    const T v = ux * (c0x[k] - T(2) * FX(x, j) + c2x[k]) * a.idt2;
    g[c] = cg[k][0] + v;
    g[N + c] = cg[k][1] + uz * (c0z[k] - T(2) * FZ(x, j) + c2z[k]) * a.idt2;
    const T exx = d4<T>(FX(x - 2, j), FX(x - 1, j), FX(x, j), FX(x + 1, j), a.idx);
    const T ezz = d4<T>(FZ(x, j - 2), FZ(x, j - 1), FZ(x, j), FZ(x, j + 1), a.idz);
    const T dv = exx + ezz, dva = dux + duz;
    g[2 * N + c] = cg[k][2] + dv * dva;
    g[3 * N + c] = cg[k][3] + T(2) * (exx * dux + ezz * duz);
    const T exz = (d4<T>(FZ(x - 1, j), FZ(x, j), FZ(x + 1, j), FZ(x + 2, j), a.idx) +
                   d4<T>(FX(x, j - 1), FX(x, j), FX(x, j + 1), FX(x, j + 2), a.idz)) / T(2);
    const T exza = (duzx + duxz) / T(2);
    g[4 * N + c] = cg[k][4] + T(2) * (exz * exza + exz * exza);
  }
}

// One step, stresses and displacements in one launch, on a tile of 64 x R
// displacement cells: forward or replay (ADJ = false), or adjoint (ADJ =
// true: residual injection, and the correlation of adjoint step corr_step
// on the tile's own stress cells, as el_stress_kernel does).  ucur is staged in LDS on the
// tile plus 4 cells on every side; the stresses are computed into LDS on
// the tile plus the 2 cells the displacement stencils reach on every side
// (recomputed by neighbouring tiles, whose C-PML memory variables only the
// owning tile writes), then the displacements of the tile are stepped.  The
// stresses never go to HBM, except the never-updated edge cells that
// accumulate moment-tensor injections (el_edge_inject_kernel, run first).
// The expressions are those of el_stress_kernel and el_disp_kernel in the
// same operand order, so the results are bitwise equal.
template <typename T, bool ADJ>
__global__ __launch_bounds__(TX *TZ) void el_tile_kernel(ElArgs<T> a) {
  constexpr int R = ElTile<T>::rows;
  constexpr int UW = TX + 8, UH = R + 8;   // ucur: columns I1-4 .. I1+TX+3, rows J1-4 .. J1+R+3
  constexpr int SW = TX + 4, SH = R + 4;   // stresses: columns I1-2 .. I1+TX+1, rows J1-2 .. J1+R+1
  __shared__ T lux[UH * UW], luz[UH * UW];
  __shared__ T lsxx[SH * SW], lszz[SH * SW], lsxz[SH * SW];
  const int b = xcd_tile_el(blockIdx.x, gridDim.x);
  const int bx = b % a.tbx, t = b / a.tbx, by = t % a.tbz, s = t / a.tbz;
  const int nts = a.nt_shot[s];
  if (a.it > nts || (!ADJ && nts <= a.act_min)) return;
  const bool corr = ADJ && a.corr_step > 0 && a.corr_step <= nts && a.corr_step <= a.corr_lim;
  const int nx = a.nx, nz = a.nz, h = a.h;
  const bool ft = a.freetop != 0;
  const int64_t N = a.N;
  const int I1 = bx * TX + 1, J1 = by * R + 1;
  const int tid = threadIdx.y * TX + threadIdx.x;
  const T *ucx = a.ucur + (int64_t)s * 2 * N, *ucz = ucx + N;
  T *un = a.unew + (int64_t)s * 2 * N;
  const T *pso = a.psi_o + (int64_t)s * a.psi_stride;
  T *psn = a.psi_n + (int64_t)s * a.psi_stride;
  const T *cb = a.cpml + (int64_t)s * a.cstride;
  const AxisCoef<T> cx = axis_coef(cb, h, 0), cz = axis_coef(cb, h, 1);
  // Interior tile: every stress cell of the tile + 2 and every displacement
  // cell takes the plain 4th-order branch of each derivative helper, lies
  // outside every C-PML strip, and no point of the shot lies on its rows.
  bool lean = I1 - 2 >= max(3, h + 2) && I1 + TX + 1 <= min(nx - 2, nx - h - 2) && J1 - 2 >= max(3, h + 2) &&
              J1 + R + 1 <= min(nz - 2, nz - h - 2);
  if (lean && a.kind != 0) {
    const int64_t o = (int64_t)s * (a.nrows + 1);
    const int r0 = J1 - 3, r1 = J1 + R + 1;   // rows J1-2 .. J1+R+1 (0-based r0 .. r1-1)
    lean = a.f0.rows[o + r0] == a.f0.rows[o + r1] && a.f1.rows[o + r0] == a.f1.rows[o + r1];
  }
  if (lean) {
    el_tile_lean<T, ADJ>(a, s, nts, I1, J1, corr, lux, luz, lsxx, lszz, lsxz);
    return;
  }

  // ucur on the tile + 4 (zero outside the ux / uz grids); the loads read
  // a clamped in-grid cell where the value is not used, so they are all in
  // flight before the first LDS store
  {
    constexpr int NU = (UW * UH + TX * TZ - 1) / (TX * TZ);
    T vx[NU], vz[NU];
#pragma unroll
    for (int k = 0; k < NU; k++) {
      const int q = min(tid + k * TX * TZ, UW * UH - 1);
      const int r = q / UW, x = I1 - 4 + (q - r * UW), j = J1 - 4 + r;
      const int64_t c = (int64_t)(min(max(j, 1), nz) - 1) * nx + (min(max(x, 1), nx) - 1);
      vx[k] = ucx[c];
      vz[k] = ucz[c];
    }
#pragma unroll
    for (int k = 0; k < NU; k++) {
      const int q = tid + k * TX * TZ;
      if (q < UW * UH) {
        const int r = q / UW, x = I1 - 4 + (q - r * UW), j = J1 - 4 + r;
        lux[q] = (x >= 1 && x <= nx - 1 && j >= 1 && j <= nz) ? vx[k] : T(0);
        luz[q] = (x >= 1 && x <= nx && j >= 1 && j <= nz - 1) ? vz[k] : T(0);
      }
    }
  }
  __syncthreads();
  const TileA<T, UW> UX{lux, I1 - 4, J1 - 4}, UZ{luz, I1 - 4, J1 - 4};

  // stresses on the tile + 2 (update_sxx_szz!, update_sxz!, :39-79, and the
  // moment-tensor injection, :81-94)
  for (int q = tid; q < SW * SH; q += TX * TZ) {
    const int r = q / SW, x = I1 - 2 + (q - r * SW), j = J1 - 2 + r;
    const bool own = x >= I1 && x < I1 + TX && j >= J1 && j < J1 + R;
    T vxx = 0, vzz = 0, vxz = 0;
    T dux = 0, duz = 0, duzx = 0, duxz = 0;   // raw derivatives (adjoint correlation)
    bool nrm = false, shr = false;
    if (x >= 1 && x <= nx && j >= 1 && j <= nz) {
      const int64_t c = (int64_t)(j - 1) * nx + (x - 1);
      nrm = x >= 2 && x <= nx - 1 && j >= (ft ? 1 : 2) && j <= nz - 1;
      if (nrm) {
        const T l = a.lam[c], m = a.mu[c];
        dux = d_ux_dx<T>(UX, x, j, nx, a.idx);
        duz = d_uz_dz<T>(UX, UZ, l, m, x, j, nx, nz, ft, a.idx, a.idz);
        T dxc = dux, dzc = duz;
        int k = strip_half(x - 1, nx - 1, h);
        if (k) {
          const int64_t qq = a.psioff[4] + (k - 1) + (int64_t)(j - 1) * (2 * h + 2);
          const T pn = cx.b(k - 1) * pso[qq] + cx.a(k - 1) * dux;
          if (own) psn[qq] = pn;
          dxc = dux + pn;
        }
        k = strip_half(j - 1, nz - 1, h);
        if (k) {
          const int64_t qq = a.psioff[7] + (x - 1) + (int64_t)(k - 1) * nx;
          const T pn = cz.b(k - 1) * pso[qq] + cz.a(k - 1) * duz;
          if (own) psn[qq] = pn;
          dzc = duz + pn;
        }
        vxx = (l + T(2) * m) * dxc + l * dzc;
        vzz = j == 1 ? T(0) : l * dxc + (l + T(2) * m) * dzc;
        if (a.kind == 1) {
          const int u = find_point(a.f0.rows, a.f0.cell, a.nrows, s, j - 1, c);
          if (u >= 0) {
            const T *tf = a.tf + a.off_tf[s] + (a.it - 1);
            const T *M = a.mom + 3 * (int64_t)a.off_mom[s];
            for (int e = a.f0.eptr[u]; e < a.f0.eptr[u + 1]; e++) {
              const int p = a.f0.epoint[e];
              const T f = tf[(int64_t)nts * p];
              vxx += M[3 * p + 0] * a.f0.ecoef[e] * f;
              vzz += M[3 * p + 1] * a.f0.ecoef[e] * f;
            }
          }
        }
      } else if (a.kind == 1) {
        // never updated: the injections accumulated so far (this step's
        // included, el_edge_inject_kernel)
        const T *sg = a.sig + (int64_t)s * 3 * N;
        vxx = sg[c];
        vzz = sg[N + c];
      }
      shr = x <= nx - 1 && j <= nz - 1;
      if (shr) {
        duzx = d_uz_dx<T>(UZ, x, j, nx, a.idx);
        duxz = d_ux_dz<T>(UX, x, j, nz, ft, a.idz);
        T c1 = duzx, c2 = duxz;
        int k = strip_full(x, nx, h);
        if (k) {
          const int64_t qq = a.psioff[5] + (k - 1) + (int64_t)(j - 1) * (2 * h);
          const T pn = cx.bh(k - 1) * pso[qq] + cx.ah(k - 1) * duzx;
          if (own) psn[qq] = pn;
          c1 = duzx + pn;
        }
        k = strip_full(j, nz, h);
        if (k) {
          const int64_t qq = a.psioff[6] + (x - 1) + (int64_t)(k - 1) * (nx - 1);
          const T pn = cz.bh(k - 1) * pso[qq] + cz.ah(k - 1) * duxz;
          if (own) psn[qq] = pn;
          c2 = duxz + pn;
        }
        vxz = a.muhh[c] * (c1 + c2);
        if (a.kind == 1) {
          const int u = find_point(a.f1.rows, a.f1.cell, a.nrows, s, j - 1, c);
          if (u >= 0) {
            const T *tf = a.tf + a.off_tf[s] + (a.it - 1);
            const T *M = a.mom + 3 * (int64_t)a.off_mom[s];
            for (int e = a.f1.eptr[u]; e < a.f1.eptr[u + 1]; e++) {
              const int p = a.f1.epoint[e];
              vxz += M[3 * p + 2] * a.f1.ecoef[e] * tf[(int64_t)nts * p];
            }
          }
        }
      }
    }
    lsxx[q] = vxx;
    lszz[q] = vzz;
    lsxz[q] = vxz;
    if (ADJ && corr && own && x <= nx && j <= nz)
      el_tile_correlate<T>(a, s, x, j, UX(x, j), UZ(x, j), nrm, shr, dux, duz, duzx, duxz);
  }
  __syncthreads();
  const TileA<T, SW> SXX{lsxx, I1 - 2, J1 - 2}, SZZ{lszz, I1 - 2, J1 - 2}, SXZ{lsxz, I1 - 2, J1 - 2};

  // displacements of the tile (update_ux!, update_uz!, :1-37; force
  // injection :96-106; forward history)
  const int i = I1 + threadIdx.x;
  if (i > nx) return;
  const int sb = a.hmode == 1 ? (nts / a.S) * a.S : a.cbase;
  for (int rr = threadIdx.y; rr < R; rr += TZ) {
    const int j = J1 + rr;
    if (j > nz) break;
    const int64_t c = (int64_t)(i - 1) + (int64_t)(j - 1) * nx;
    const bool hx = i <= nx - 1, hz = j <= nz - 1;
    T oldx = 0, oldz = 0, curx = 0, curz = 0, nux = 0, nuz = 0;
    if (hx) {
      const T d1 = d_sxx_dx<T>(SXX, i, j, nx, a.idx);
      const T d2 = d_sxz_dz<T>(SXZ, i, j, nz, ft, a.idz);
      T c1 = d1, c2 = d2;
      int k = strip_full(i, nx, h);
      if (k) {
        const int64_t q = a.psioff[0] + (k - 1) + (int64_t)(j - 1) * (2 * h);
        const T pn = cx.bh(k - 1) * pso[q] + cx.ah(k - 1) * d1;
        psn[q] = pn;
        c1 = d1 + pn;
      }
      k = strip_half(j - 1, nz - 1, h);
      if (k) {
        const int64_t q = a.psioff[3] + (i - 1) + (int64_t)(k - 1) * (nx - 1);
        const T pn = cz.b(k - 1) * pso[q] + cz.a(k - 1) * d2;
        psn[q] = pn;
        c2 = d2 + pn;
      }
      oldx = un[c];
      curx = UX(i, j);
      nux = T(2) * curx - oldx + a.rdx[c] * (c1 + c2);
    }
    if (hz) {
      const T d1 = d_sxz_dx<T>(SXZ, i, j, nx, a.idx);
      const T d2 = d_szz_dz<T>(SZZ, i, j, nz, ft, a.idz);
      T c1 = d1, c2 = d2;
      int k = strip_half(i - 1, nx - 1, h);
      if (k) {
        const int64_t q = a.psioff[1] + (k - 1) + (int64_t)(j - 1) * (2 * h + 2);
        const T pn = cx.b(k - 1) * pso[q] + cx.a(k - 1) * d1;
        psn[q] = pn;
        c1 = d1 + pn;
      }
      k = strip_full(j, nz, h);
      if (k) {
        const int64_t q = a.psioff[2] + (i - 1) + (int64_t)(k - 1) * nx;
        const T pn = cz.bh(k - 1) * pso[q] + cz.ah(k - 1) * d2;
        psn[q] = pn;
        c2 = d2 + pn;
      }
      oldz = un[N + c];
      curz = UZ(i, j);
      nuz = T(2) * curz - oldz + a.rdz[c] * (c1 + c2);
    }
    if (a.kind == 2) {
      const T *tf = a.tf + a.off_tf[s] + (a.it - 1);
      if (hx) {
        const int u = find_point(a.f0.rows, a.f0.cell, a.nrows, s, j - 1, c);
        if (u >= 0)
          for (int e = a.f0.eptr[u]; e < a.f0.eptr[u + 1]; e++)
            nux += a.f0.ecoef[e] * tf[(int64_t)nts * (a.ncomp * a.f0.epoint[e])] * a.rdx[c];
      }
      if (hz) {
        const int u = find_point(a.f1.rows, a.f1.cell, a.nrows, s, j - 1, c);
        if (u >= 0)
          for (int e = a.f1.eptr[u]; e < a.f1.eptr[u + 1]; e++)
            nuz += a.f1.ecoef[e] * tf[(int64_t)nts * (a.ncomp * a.f1.epoint[e] + 1)] * a.rdz[c];
      }
    }
    if (hx) un[c] = nux;
    if (hz) un[N + c] = nuz;
    if (a.hmode && a.it > sb) {
      T *hs = a.hist + (int64_t)s * 2 * N + c;
      const int sl = a.it - sb + 1;
      hs[(int64_t)sl * a.hslot] = nux;
      hs[(int64_t)sl * a.hslot + N] = nuz;
      if (a.hmode == 1 && a.it == sb + 1) {
        hs[0] = oldx;
        hs[N] = oldz;
        hs[a.hslot] = curx;
        hs[a.hslot + N] = curz;
      }
    }
  }
}

// traces[it, comp, r] = sum over taps of coef * u (record_receivers2D_* and
// the per-receiver reduction, elastic2D_iso_xPU.jl:108-118, :216-231).
template <typename T>
__global__ void el_record_kernel(const T *u, int64_t N, TapView<T> r0, TapView<T> r1, const int *nt_shot,
                                 const int64_t *off_tr, T *traces, int it, int maxrec) {
  const int s = blockIdx.y;
  const int nts = nt_shot[s];
  if (it > nts) return;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  const int comp = q & 1, r = q >> 1;
  const int p0 = r0.poff[s], np = r0.poff[s + 1] - p0;
  if (r >= np || r >= maxrec) return;
  const TapView<T> &t = comp ? r1 : r0;
  const T *f = u + (int64_t)s * 2 * N + comp * N;
  T acc = 0;
  for (int k = t.pptr[p0 + r]; k < t.pptr[p0 + r + 1]; k++) acc += t.pcoef[k] * f[t.pcell[k]];
  traces[off_tr[s] + (it - 1) + (int64_t)nts * (comp + 2 * r)] = acc;
}

// dt^2 / rho on a padded slab (the spare column or row holds rho = 0 and
// stays 0; it is never read as data)
template <typename T>
__global__ void el_rdt_kernel(T *out, const T *rho, int64_t n, T dt2) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x)
    out[q] = rho[q] != T(0) ? dt2 / rho[q] : T(0);
}

// Host-side tap table of one family for a batch of shots.
struct TapTable {
  DevBuf cell, eptr, epoint, ecoef, rows, poff, pptr, pcell, pcoef;
  template <typename T>
  TapView<T> view() const {
    return TapView<T>{cell.as<int64_t>(), eptr.as<int>(),   epoint.as<int>(), ecoef.as<T>(), rows.as<int>(),
                      poff.as<int>(),     pptr.as<int>(),   pcell.as<int64_t>(), pcoef.as<T>()};
  }
};

template <typename T>
void build_taps(TapTable &t, int nshots, const int *npts, const swx_taps *const *fam, int nx, int nz, int ni, int nj,
                const char *what, hipStream_t stream) {
  std::vector<int64_t> cell, pcell;
  std::vector<int> eptr(1, 0), epoint, rows, poff(1, 0), pptr(1, 0);
  std::vector<T> ecoef, pcoef;
  for (int s = 0; s < nshots; s++) {
    const swx_taps &f = *fam[s];
    SWX_REQUIRE(f.ptr && (f.ptr[npts[s]] == 0 || (f.ij && f.coef)), SWX_EINVAL, "null tap array");
    std::map<int64_t, std::vector<std::pair<int, T>>> g;
    const T *coef = static_cast<const T *>(f.coef);
    for (int p = 0; p < npts[s]; p++) {
      SWX_REQUIRE(f.ptr[p] <= f.ptr[p + 1], SWX_EINVAL, "tap offsets must be non-decreasing");
      for (int k = f.ptr[p]; k < f.ptr[p + 1]; k++) {
        const int64_t i = f.ij[2 * (int64_t)k], j = f.ij[2 * (int64_t)k + 1];
        if (i < 1 || i > ni || j < 1 || j > nj)
          throw Error(SWX_EINVAL, std::string(what) + " position of shot " + std::to_string(s + 1) +
                                      " is not inside the grid!");
        const int64_t c = (i - 1) + (j - 1) * (int64_t)nx;
        g[c].push_back({p, coef[k]});
        pcell.push_back(c);
        pcoef.push_back(coef[k]);
      }
      pptr.push_back((int)pcell.size());
    }
    poff.push_back(poff.back() + npts[s]);
    const int u0 = (int)cell.size();
    for (auto &kv : g) {
      cell.push_back(kv.first);
      for (auto &pe : kv.second) {
        epoint.push_back(pe.first);
        ecoef.push_back(pe.second);
      }
      eptr.push_back((int)epoint.size());
    }
    int u = u0;
    for (int64_t r = 0; r <= nz; r++) {
      while (u < (int)cell.size() && cell[u] / nx < r) u++;
      rows.push_back(u);
    }
  }
  auto up = [&](DevBuf &b, const void *src, size_t bytes) {
    b.alloc(std::max<size_t>(bytes, 16));
    if (bytes) SWX_HIP(hipMemcpyAsync(b.get(), src, bytes, hipMemcpyHostToDevice, stream));
  };
  up(t.cell, cell.data(), cell.size() * sizeof(int64_t));
  up(t.eptr, eptr.data(), eptr.size() * sizeof(int));
  up(t.epoint, epoint.data(), epoint.size() * sizeof(int));
  up(t.ecoef, ecoef.data(), ecoef.size() * sizeof(T));
  up(t.rows, rows.data(), rows.size() * sizeof(int));
  up(t.poff, poff.data(), poff.size() * sizeof(int));
  up(t.pptr, pptr.data(), pptr.size() * sizeof(int));
  up(t.pcell, pcell.data(), pcell.size() * sizeof(int64_t));
  up(t.pcoef, pcoef.data(), pcoef.size() * sizeof(T));
  SWX_HIP(hipStreamSynchronize(stream));
}

template <typename T>
class El2dEngine final : public Engine {
 public:
  explicit El2dEngine(PlanCore &core) : c_(core) {
    const swx_plan_desc &d = c_.d;
    nx_ = (int)d.n[0];
    nz_ = (int)d.n[1];
    h_ = d.halo;
    ns_ = d.max_shots;
    N_ = (int64_t)nx_ * nz_;
    const int64_t h = h_, nx = nx_, nz = nz_;
    const int64_t sz[8] = {2 * h * nz,       (2 * h + 2) * (nz - 1), nx * 2 * h,       (nx - 1) * (2 * h + 2),
                           (2 * h + 2) * nz, 2 * h * (nz - 1),       (nx - 1) * 2 * h, nx * (2 * h + 2)};
    int64_t off = 0;
    for (int k = 0; k < 8; k++) {
      psioff_[k] = off;
      off += sz[k];
    }
    psi_stride_ = std::max<int64_t>(off, 1);
    cstride_ = 2 * (8 * h_ + 4);
    const size_t e = sizeof(T);
    model_.alloc(5 * N_ * e);
    u_.alloc(2 * ns_ * 2 * N_ * e);
    sig_.alloc(ns_ * 3 * N_ * e);
    psi_.alloc(2 * ns_ * psi_stride_ * e);   // two step parities
    cpml_.alloc((int64_t)ns_ * cstride_ * e);
    state_bytes_ = model_.bytes() + u_.bytes() + sig_.bytes() + psi_.bytes();
    if (d.gradient) {
      au_.alloc(u_.bytes());
      asig_.alloc(sig_.bytes());
      apsi_.alloc(psi_.bytes());
      grad_.alloc(ns_ * 5 * N_ * e);
      gout_.alloc(ns_ * 5 * N_ * e);
      state_bytes_ += au_.bytes() + asig_.bytes() + apsi_.bytes() + grad_.bytes() + gout_.bytes();
      const int64_t slot = ns_ * 2 * N_ * (int64_t)e;
      ck_bytes_one_ = (int64_t)u_.bytes() + (int64_t)psi_.bytes() / 2;
      const int64_t avail = (int64_t)c_.budget - state_bytes_ - 2 * slot - 2 * ck_bytes_one_;
      S_ = choose_segment(d.nt_max, d.check_freq, slot, ck_bytes_one_, avail);
      // check_freq = 1 keeps every step (no replay); otherwise a checkpoint
      // at every multiple of S, as the reference's LinearCheckpointer
      store_all_ = d.check_freq <= 1 && S_ == d.nt_max;
      K_ = (int)ceil_div(d.nt_max, S_);
      nck_ = store_all_ ? 0 : d.nt_max / S_;
      hist_.alloc((int64_t)(S_ + 2) * slot);
      ck_.alloc(std::max<int64_t>(nck_ * ck_bytes_one_, 1));
    }
    zero_forward_state(true);
  }

  void set_model(int nfields, const void *const *fields) override {
    SWX_REQUIRE(nfields == 5, SWX_EINVAL, "elastic expects five model fields (lambda, mu, rho_ihalf, rho_jhalf, "
                                          "mu_ihalf_jhalf)");
    for (int k = 0; k < 5; k++) SWX_REQUIRE(fields[k], SWX_EINVAL, "null model field");
    const size_t e = sizeof(T);
    SWX_HIP(hipMemsetAsync(model_.get(), 0, model_.bytes(), c_.stream));
    SWX_HIP(hipMemcpyAsync(M(0), fields[0], N_ * e, hipMemcpyHostToDevice, c_.stream));
    SWX_HIP(hipMemcpyAsync(M(1), fields[1], N_ * e, hipMemcpyHostToDevice, c_.stream));
    DevBuf tmp(2 * N_ * e);
    T *rx = tmp.as<T>(), *rz = rx + N_;
    SWX_HIP(hipMemsetAsync(tmp.get(), 0, tmp.bytes(), c_.stream));
    // rho_ihalf (nx-1, nz) and mu_ihalf_jhalf (nx-1, nz-1) into nx-pitched slabs
    SWX_HIP(hipMemcpy2DAsync(rx, nx_ * e, fields[2], (nx_ - 1) * e, (nx_ - 1) * e, nz_, hipMemcpyHostToDevice,
                             c_.stream));
    SWX_HIP(hipMemcpyAsync(rz, fields[3], (int64_t)nx_ * (nz_ - 1) * e, hipMemcpyHostToDevice, c_.stream));
    SWX_HIP(hipMemcpy2DAsync(M(2), nx_ * e, fields[4], (nx_ - 1) * e, (nx_ - 1) * e, nz_ - 1, hipMemcpyHostToDevice,
                             c_.stream));
    const T dt = T(c_.d.dt);
    dim3 g((unsigned)std::min<int64_t>(ceil_div(N_, 256), 2048));
    el_rdt_kernel<T><<<g, 256, 0, c_.stream>>>(M(3), rx, N_, dt * dt);
    el_rdt_kernel<T><<<g, 256, 0, c_.stream>>>(M(4), rz, N_, dt * dt);
    SWX_HIP(hipGetLastError());
    SWX_HIP(hipStreamSynchronize(c_.stream));
    model_set_ = true;
  }

  void on_shots_set(int, const swx_shot_desc *) override {
    throw Error(SWX_EINVAL, "elastic plans take swx_set_shots_elastic");
  }

  void on_elastic_shots(int nshots, const swx_elastic_shot_desc *shots) override {
    const int esz = sizeof(T);
    PointTables &pt = c_.pts;
    pt.nshots = nshots;
    pt.nt.assign(nshots, 0);
    pt.nsrc.assign(nshots, 0);
    pt.nrec.assign(nshots, 0);
    pt.off_tf.assign(nshots, 0);
    pt.off_tr.assign(nshots, 0);
    std::vector<int> nsrc(nshots), nrec(nshots), offm(nshots + 1, 0);
    int64_t otf = 0, otr = 0;
    kind_ = shots[0].source_kind;
    for (int s = 0; s < nshots; s++) {
      const swx_elastic_shot_desc &sh = shots[s];
      SWX_REQUIRE(sh.source_kind == SWX_SRC_MOMENT_TENSOR || sh.source_kind == SWX_SRC_EXTERNAL_FORCE, SWX_EINVAL,
                  "source_kind must be SWX_SRC_MOMENT_TENSOR or SWX_SRC_EXTERNAL_FORCE");
      SWX_REQUIRE(sh.source_kind == kind_, SWX_EINVAL, "Types of shots are inconsistent.");
      SWX_REQUIRE(sh.nsrc > 0, SWX_EINVAL, "There must be at least one source!");
      SWX_REQUIRE(sh.nrec > 0, SWX_EINVAL, "There must be at least one receiver!");
      SWX_REQUIRE(sh.srctf && (kind_ == SWX_SRC_EXTERNAL_FORCE || sh.moments), SWX_EINVAL,
                  "null source array");
      pt.nt[s] = sh.nt;
      pt.nsrc[s] = nsrc[s] = sh.nsrc;
      pt.nrec[s] = nrec[s] = sh.nrec;
      pt.off_tf[s] = otf;
      pt.off_tr[s] = otr;
      offm[s + 1] = offm[s] + sh.nsrc;
      otf += (int64_t)sh.nt * sh.nsrc * (kind_ == SWX_SRC_EXTERNAL_FORCE ? 2 : 1);
      otr += (int64_t)sh.nt * 2 * sh.nrec;
    }
    pt.total_tf = otf;
    pt.total_tr = otr;
    std::vector<const swx_taps *> f[4];
    for (int s = 0; s < nshots; s++) {
      f[0].push_back(&shots[s].src[0]);
      f[1].push_back(&shots[s].src[1]);
      f[2].push_back(&shots[s].rec[0]);
      f[3].push_back(&shots[s].rec[1]);
    }
    const bool mt = kind_ == SWX_SRC_MOMENT_TENSOR;
    // grids of the families: sxx/szz nx x nz, sxz (nx-1) x (nz-1), ux (nx-1) x nz, uz nx x (nz-1)
    build_taps<T>(taps_[0], nshots, nsrc.data(), f[0].data(), nx_, nz_, mt ? nx_ : nx_ - 1, mt ? nz_ : nz_, "Source",
                  c_.stream);
    build_taps<T>(taps_[1], nshots, nsrc.data(), f[1].data(), nx_, nz_, mt ? nx_ - 1 : nx_, nz_ - 1, "Source",
                  c_.stream);
    build_taps<T>(taps_[2], nshots, nrec.data(), f[2].data(), nx_, nz_, nx_ - 1, nz_, "Receiver", c_.stream);
    build_taps<T>(taps_[3], nshots, nrec.data(), f[3].data(), nx_, nz_, nx_, nz_ - 1, "Receiver", c_.stream);
    maxrec_ = *std::max_element(nrec.begin(), nrec.end());
    // moment-tensor taps on normal-stress cells the stencil never updates
    edge_taps_ = false;
    max_ucell0_ = 0;
    if (mt) {
      for (int s = 0; s < nshots; s++) {
        std::vector<int64_t> cells;
        const swx_taps &t0 = shots[s].src[0];
        for (int k = 0; k < t0.ptr[nsrc[s]]; k++) {
          const int64_t i = t0.ij[2 * (int64_t)k], j = t0.ij[2 * (int64_t)k + 1];
          cells.push_back((i - 1) + (j - 1) * (int64_t)nx_);
          const bool nrm = i >= 2 && i <= nx_ - 1 && j >= (c_.d.freetop ? 1 : 2) && j <= nz_ - 1;
          edge_taps_ |= !nrm;
        }
        std::sort(cells.begin(), cells.end());
        cells.erase(std::unique(cells.begin(), cells.end()), cells.end());
        max_ucell0_ = std::max<int>(max_ucell0_, (int)cells.size());
      }
    }
    auto up = [&](DevBuf &b, const void *src, size_t bytes) {
      b.alloc(std::max<size_t>(bytes, 16));
      if (bytes) SWX_HIP(hipMemcpyAsync(b.get(), src, bytes, hipMemcpyHostToDevice, c_.stream));
    };
    up(pt.d_nt, pt.nt.data(), sizeof(int) * nshots);
    up(pt.d_off_tf, pt.off_tf.data(), sizeof(int64_t) * nshots);
    up(pt.d_off_tr, pt.off_tr.data(), sizeof(int64_t) * nshots);
    up(offm_, offm.data(), sizeof(int) * offm.size());
    std::vector<char> tf((size_t)otf * esz), mom((size_t)offm[nshots] * 3 * esz);
    for (int s = 0; s < nshots; s++) {
      const int comps = mt ? 1 : 2;
      std::memcpy(tf.data() + (size_t)pt.off_tf[s] * esz, shots[s].srctf,
                  (size_t)shots[s].nt * comps * shots[s].nsrc * esz);
      if (mt) std::memcpy(mom.data() + (size_t)offm[s] * 3 * esz, shots[s].moments, (size_t)shots[s].nsrc * 3 * esz);
    }
    up(pt.d_srctf, tf.data(), tf.size());
    up(mom_, mom.data(), mom.size());
    pt.d_traces.alloc(std::max<size_t>((size_t)otr * esz, 16));
    pt.d_res.alloc(std::max<size_t>((size_t)otr * esz, 16));
    SWX_HIP(hipMemsetAsync(pt.d_traces.get(), 0, (size_t)otr * esz, c_.stream));
    // C-PML coefficients per shot
    std::vector<swx_shot_desc> tmp(nshots);
    for (int s = 0; s < nshots; s++)
      for (int ax = 0; ax < 3; ax++)
        for (int k = 0; k < 4; k++) tmp[s].cpml[ax][k] = shots[s].cpml[ax][k];
    std::vector<T> hc = pack_cpml<T>(nshots, tmp.data(), 2, h_);
    SWX_HIP(hipMemcpyAsync(cpml_.get(), hc.data(), hc.size() * sizeof(T), hipMemcpyHostToDevice, c_.stream));
    SWX_HIP(hipStreamSynchronize(c_.stream));
    nsh_ = nshots;
    nt_ = pt.nt;
    ntmax_ = *std::max_element(nt_.begin(), nt_.end());
  }

  void forward(bool save) override {
    SWX_REQUIRE(model_set_, SWX_ESTATE, "swx_set_model must be called before running");
    if (var_fwd_ < 0) {
      // tuning launches run on the state before it is cleared below
      var_fwd_ = tune("SWX_EL2D_VARIANT_FWD", [&](int v) { step_fwd_v(v, 1, 0, 0, 0, nullptr); }, true);
    }
    zero_forward_state(true);
    if (save) SWX_HIP(hipMemsetAsync(hist_.get(), 0, hist_.bytes(), c_.stream));
    T *tr = c_.pts.d_traces.as<T>();
    for (int it = 1; it <= ntmax_; it++) {
      progress("forward step", it, ntmax_);
      step_fwd(it, 0, save ? 1 : 0, 0, tr);
      take_snapshots(c_, *this, it);
      if (save && it % S_ == 0) {
        const int k = it / S_;
        if (k >= 1 && k <= nck_) ckpt_copy(k, true);
      }
    }
  }

  // "ucur" = [ux, uz] and "σ" = [σxx, σzz, σxz] (ela_forward.jl:74-77)
  // right after step it; the stresses are in HBM for the variants with a
  // stress launch (the z-marching ones keep them on chip, so a forward with
  // snapshots runs variant 1)
  int snap_fields(int it, int s, SnapPiece *out) const override {
    const int64_t e = sizeof(T), nx = nx_, nz = nz_;
    const T *u = U(u_, it) + (int64_t)s * 2 * N_;
    const T *sg = sig_.as<T>() + (int64_t)s * 3 * N_;
    out[0] = {u, (nx - 1) * e, nx * e, nz};
    out[1] = {u + N_, nx * e, nx * e, nz - 1};
    out[2] = {sg, nx * e, nx * e, nz};
    out[3] = {sg + N_, nx * e, nx * e, nz};
    out[4] = {sg + 2 * N_, (nx - 1) * e, nx * e, nz - 1};
    return 5;
  }

  void adjoint(bool per_shot) override {
    if (var_adj_ < 0) {
      // timed on a step of the last segment that correlates (its stress
      // launch correlates the step after it), so the candidates are timed
      // with the correlation's loads and stores; the history slots it reads
      // are in range, and the adjoint state and gradients are cleared below
      const int cbt = (int)(ceil_div(ntmax_, S_) - 1) * S_;
      const int itt = ntmax_ - 1 > cbt ? ntmax_ - 1 : ntmax_;
      var_adj_ = tune("SWX_EL2D_VARIANT_ADJ", [&](int v) { step_adj_v(v, itt, cbt); }, false);
    }
    SWX_HIP(hipMemsetAsync(au_.get(), 0, au_.bytes(), c_.stream));
    SWX_HIP(hipMemsetAsync(asig_.get(), 0, asig_.bytes(), c_.stream));
    SWX_HIP(hipMemsetAsync(apsi_.get(), 0, apsi_.bytes(), c_.stream));
    SWX_HIP(hipMemsetAsync(grad_.get(), 0, grad_.bytes(), c_.stream));
    const int Kmax = (int)ceil_div(ntmax_, S_);
    const int64_t n2 = 2 * N_;
    for (int k = Kmax - 1; k >= 0; k--) {
      progress("adjoint segment", Kmax - k, Kmax);
      const int cb = k * S_;
      // shots whose checkpoint (k+1)S exists replay segment k: levels cb-1
      // and cb from checkpoint k, cb+1 .. cb+S-1 recomputed, cb+S-1 and
      // cb+S from checkpoint k+1 (initrecover!/recover!/getsaved,
      // checkpointers.jl:80-105)
      const int act = cb + S_ - 1;
      bool need = false;
      for (int s = 0; s < nsh_; s++) need |= !store_all_ && nt_[s] > act;
      if (need) {
        if (k == 0) zero_forward_state(false);
        else ckpt_copy(k, false);
        T *hs = hist_.as<T>();
        const int64_t hslot = (int64_t)ns_ * n2;
        dim3 g((unsigned)std::min<int64_t>(ceil_div(n2, 256), 1024), (unsigned)nsh_);
        el_hist_copy<T><<<g, 256, 0, c_.stream>>>(hs, U(u_, cb - 1), n2, c_.pts.d_nt.as<int>(), act);
        el_hist_copy<T><<<g, 256, 0, c_.stream>>>(hs + hslot, U(u_, cb), n2, c_.pts.d_nt.as<int>(), act);
        SWX_HIP(hipGetLastError());
        for (int it = cb + 1; it <= cb + S_ - 1; it++) step_fwd(it, act, 2, cb, nullptr);
        // u^{cb+S-1} and u^{cb+S} are read from checkpoint k+1 (getsaved
        // prefers checkpointed levels over the replayed buffer)
        const T *ck1 = reinterpret_cast<const T *>(ck_.as<char>() + (int64_t)k * ck_bytes_one_);
        for (int t = cb + S_ - 1; t <= cb + S_; t++)
          el_hist_copy<T><<<g, 256, 0, c_.stream>>>(hs + (int64_t)(t - cb + 1) * hslot, ck1 + (int64_t)(t % 2) * ns_ * n2,
                                                    n2, c_.pts.d_nt.as<int>(), act);
        SWX_HIP(hipGetLastError());
      }
      const int top = std::min(cb + S_, ntmax_);
      for (int it = top; it >= cb + 1; it--) step_adj(it, cb);
      // correlation of the first step of the segment
      ElArgs<T> a = base_args(cb);
      a.ucur = U(au_, cb + 1);
      a.sig = asig_.as<T>();
      a.update = 0;
      a.corr_step = cb + 1;
      a.corr_lim = cb + S_;
      a.cbase = cb;
      launch_stress(a, 1, var_adj_ == 1);
    }
    const int64_t G = 5 * N_;
    dim3 g((unsigned)std::min<int64_t>(ceil_div(G, 256), 4096));
    reduce_shots<T><<<g, 256, 0, c_.stream>>>(gout_.as<T>(), grad_.as<T>(), G, nsh_, T(1), per_shot ? 1 : 0);
    SWX_HIP(hipGetLastError());
  }

  void copy_grad(void *host, bool per_shot) override {
    // unpad: g_rho_ihalf (nx-1, nz) | g_rho_jhalf (nx, nz-1) | g_lambda | g_mu | g_mu_ihalf_jhalf (nx-1, nz-1)
    const size_t e = sizeof(T);
    char *o = static_cast<char *>(host);
    const int copies = per_shot ? nsh_ : 1;
    for (int s = 0; s < copies; s++) {
      const char *b = gout_.as<char>() + (size_t)s * 5 * N_ * e;
      SWX_HIP(hipMemcpy2D(o, (nx_ - 1) * e, b, nx_ * e, (nx_ - 1) * e, nz_, hipMemcpyDeviceToHost));
      o += (size_t)(nx_ - 1) * nz_ * e;
      SWX_HIP(hipMemcpy(o, b + N_ * e, (size_t)nx_ * (nz_ - 1) * e, hipMemcpyDeviceToHost));
      o += (size_t)nx_ * (nz_ - 1) * e;
      SWX_HIP(hipMemcpy(o, b + 2 * N_ * e, (size_t)2 * N_ * e, hipMemcpyDeviceToHost));
      o += (size_t)2 * N_ * e;
      SWX_HIP(hipMemcpy2D(o, (nx_ - 1) * e, b + 4 * N_ * e, nx_ * e, (nx_ - 1) * e, nz_ - 1, hipMemcpyDeviceToHost));
      o += (size_t)(nx_ - 1) * (nz_ - 1) * e;
    }
  }

  void variants(int32_t v[2]) const override { v[0] = var_fwd_; v[1] = var_adj_; }
  int64_t bytes_state() const override { return state_bytes_; }
  int64_t bytes_checkpoints() const override { return ck_.bytes(); }
  int64_t bytes_replay() const override { return hist_.bytes(); }
  int segment() const override { return S_; }

 private:
  T *M(int k) const { return model_.as<T>() + (int64_t)k * N_; }
  // C-PML memory variables of step parity p (all shots)
  T *PS(const DevBuf &b, int k) const { return b.as<T>() + (int64_t)(k & 1) * ns_ * psi_stride_; }
  // time level k of a displacement buffer (slot k mod 2)
  T *U(const DevBuf &b, int k) const { return b.as<T>() + (int64_t)(((k % 2) + 2) % 2) * ns_ * 2 * N_; }

  void zero_forward_state(bool with_stress) {
    SWX_HIP(hipMemsetAsync(u_.get(), 0, u_.bytes(), c_.stream));
    SWX_HIP(hipMemsetAsync(psi_.get(), 0, psi_.bytes(), c_.stream));
    // the stresses are not part of the reference's checkpoints: a replay
    // from the start keeps them (only never-updated edge cells matter)
    if (with_stress) SWX_HIP(hipMemsetAsync(sig_.get(), 0, sig_.bytes(), c_.stream));
  }

  ElArgs<T> base_args(int it) const {
    ElArgs<T> a{};
    a.nx = nx_; a.nz = nz_; a.h = h_; a.freetop = c_.d.freetop ? 1 : 0;
    a.N = N_; a.nrows = nz_;
    a.idx = T(1.0 / c_.d.h[0]);
    a.idz = T(1.0 / c_.d.h[1]);
    a.idt2 = T(1.0 / (c_.d.dt * c_.d.dt));
    a.lam = M(0); a.mu = M(1); a.muhh = M(2); a.rdx = M(3); a.rdz = M(4);
    a.cpml = cpml_.as<T>(); a.cstride = cstride_;
    a.nt_shot = c_.pts.d_nt.as<int>();
    a.it = it;
    a.psi_stride = psi_stride_;
    for (int k = 0; k < 8; k++) a.psioff[k] = psioff_[k];
    a.off_tf = c_.pts.d_off_tf.as<int64_t>();
    a.mom = mom_.as<T>(); a.off_mom = offm_.as<int>();
    a.hist = hist_.as<T>(); a.hslot = (int64_t)ns_ * 2 * N_;
    a.S = store_all_ ? (1 << 30) : S_;   // store-all: the final segment is the whole run
    a.grad = grad_.as<T>(); a.gstride = 5 * N_;
    a.update = 1;
    return a;
  }

  // Interior tile rectangle for the frame-first block order: the tiles whose
  // cells satisfy the geometric part of el_is_lean (lo <= i <= hi along x
  // and z); the point rows are left to el_is_lean.
  void set_order(ElArgs<T> &a) const {
    a.order = order_ ? 1 : 0;
    const int lo = std::max(3, h_ + 2);
    const int hx = std::min(nx_ - 2, nx_ - h_ - 1), hz = std::min(nz_ - 2, nz_ - h_ - 1);
    a.ibx0 = (int)ceil_div(lo - 1, TX);
    a.ibx1 = hx >= TX ? (hx - TX) / TX : -1;
    a.iby0 = (int)ceil_div(lo - 1, TZ);
    a.iby1 = hz >= TZ ? (hz - TZ) / TZ : -1;
    a.ibx1 = std::min(a.ibx1, a.tbx - 1);
    a.iby1 = std::min(a.iby1, a.tbz - 1);
    if (a.ibx1 < a.ibx0 || a.iby1 < a.iby0) {   // no interior tile: all frame
      a.ibx0 = 0; a.ibx1 = -1;
      a.iby0 = a.tbz; a.iby1 = a.tbz - 1;
    }
  }

  // fold: interior and other blocks in one launch; otherwise an interior
  // launch followed by an other-block launch
  void launch_stress(ElArgs<T> a, int kind, bool fold) {
    a.tbx = (int)ceil_div(nx_, TX);
    a.tbz = (int)ceil_div(nz_, TZ);
    a.xcd = xcd_ ? 1 : 0;
    set_order(a);
    dim3 blk(TX, TZ), grd((unsigned)((int64_t)a.tbx * a.tbz * nsh_));
    c_.ktimer.begin(c_.stream);
    if (fold) {
      el_stress_kernel<T, true><<<grd, blk, 0, c_.stream>>>(a);
    } else {
      el_stress_lean<T><<<grd, blk, 0, c_.stream>>>(a);
      el_stress_kernel<T, false><<<grd, blk, 0, c_.stream>>>(a);
    }
    c_.ktimer.end(c_.stream, kind);
    SWX_HIP(hipGetLastError());
  }
  void launch_disp(ElArgs<T> a, int kind, bool fold) {
    a.tbx = (int)ceil_div(nx_, TX);
    a.tbz = (int)ceil_div(nz_, TZ);
    a.xcd = xcd_ ? 1 : 0;
    set_order(a);
    dim3 blk(TX, TZ), grd((unsigned)((int64_t)a.tbx * a.tbz * nsh_));
    c_.ktimer.begin(c_.stream);
    if (fold) {
      el_disp_kernel<T, true><<<grd, blk, 0, c_.stream>>>(a);
    } else {
      el_disp_lean<T><<<grd, blk, 0, c_.stream>>>(a);
      el_disp_kernel<T, false><<<grd, blk, 0, c_.stream>>>(a);
    }
    c_.ktimer.end(c_.stream, kind);
    SWX_HIP(hipGetLastError());
  }

  void step_fwd(int it, int act_min, int hmode, int cbase, T *traces) {
    // with snapshots the stresses must be in HBM: no fused variant
    const bool fused = var_fwd_ == 2 || var_fwd_ > 100;
    step_fwd_v(c_.snap_every > 0 && fused ? 1 : var_fwd_, it, act_min, hmode, cbase, traces);
  }
  void step_adj(int it, int cbase) { step_adj_v(var_adj_, it, cbase); }

  // z-marching launch shape: strips of 64*W columns, chunks sized for
  // about 2048 blocks (8 per CU) when the grid allows it
  unsigned grid_march(ElArgs<T> &a, int w) const {
    a.nbx = (int)ceil_div(nx_, 64 * w);
    const int64_t per = (int64_t)a.nbx * nsh_;
    const int nbz = (int)std::max<int64_t>(1, std::min<int64_t>(ceil_div(2048, per), ceil_div(nz_, 32)));
    a.ch = (int)ceil_div(nz_, nbz);
    a.nbz = (int)ceil_div(nz_, a.ch);
    return (unsigned)((int64_t)a.nbx * a.nbz * nsh_);
  }
  template <bool ADJ>
  void launch_march(ElArgs<T> a, int v, int kind) {
    const unsigned g = grid_march(a, v - 100);
    c_.ktimer.begin(c_.stream);
    if (v == 102) el_march_kernel<T, 2, ADJ><<<g, 128, 0, c_.stream>>>(a);
    else el_march_kernel<T, 4, ADJ><<<g, 256, 0, c_.stream>>>(a);
    c_.ktimer.end(c_.stream, kind);
    SWX_HIP(hipGetLastError());
  }

  // Variants: 0 = stress and displacement steps each as an interior-block
  // launch plus an other-block launch, 1 = the same steps with both kinds of
  // blocks in one launch each, 2 = both steps in one launch on 64 x R tiles
  // with the stresses in LDS, 102/104 = one z-marching launch on 128/256-
  // column strips.  The tuner times 0, 1 and 2, and the z-marching variants
  // when kMarchInTuner is set; the environment can pin any.
  static constexpr bool kMarchInTuner = false;
  template <typename F>
  int tune(const char *env, F step, bool forward) {
    (void)forward;
    std::vector<int> cands = {0, 1, 2};
    if (kMarchInTuner) {
      cands.push_back(102);
      cands.push_back(104);
    }
    if (const char *v = std::getenv(env)) {
      const int t = std::atoi(v);
      if (t == 0 || t == 1 || t == 2 || t == 102 || t == 104) return t;
    }
    return cands[fastest_candidate(c_.stream, (int)cands.size(), [&](int k) { step(cands[k]); })];
  }

  void step_fwd_v(int v, int it, int act_min, int hmode, int cbase, T *traces) {
    if (v == 2 || v > 100) {
      ElArgs<T> a = base_args(it);
      a.ucur = U(u_, it - 1);
      a.unew = U(u_, it);
      a.sig = sig_.as<T>();
      a.psi_o = PS(psi_, it - 1);
      a.psi_n = PS(psi_, it);
      a.act_min = act_min;
      a.f0 = taps_[0].view<T>();
      a.f1 = taps_[1].view<T>();
      a.tf = c_.pts.d_srctf.as<T>();
      a.ncomp = 2;
      a.kind = kind_ == SWX_SRC_MOMENT_TENSOR ? 1 : 2;
      a.hmode = c_.d.gradient ? hmode : 0;
      a.cbase = cbase;
      if (a.kind == 1 && edge_taps_) {
        dim3 g((unsigned)ceil_div(std::max(max_ucell0_, 1), 64), (unsigned)nsh_);
        el_edge_inject_kernel<T><<<g, 64, 0, c_.stream>>>(a);
        SWX_HIP(hipGetLastError());
      }
      if (v == 2) {
        a.tbx = (int)ceil_div(nx_, TX);
        a.tbz = (int)ceil_div(nz_, ElTile<T>::rows);
        const unsigned g = (unsigned)((int64_t)a.tbx * a.tbz * nsh_);
        c_.ktimer.begin(c_.stream);
        el_tile_kernel<T, false><<<g, dim3(TX, TZ), 0, c_.stream>>>(a);
        c_.ktimer.end(c_.stream, 0);
        SWX_HIP(hipGetLastError());
      } else {
        launch_march<false>(a, v, 0);
      }
      record(it, traces);
      return;
    }
    ElArgs<T> a = base_args(it);
    a.ucur = U(u_, it - 1);
    a.unew = U(u_, it);
    a.sig = sig_.as<T>();
    a.psi_o = PS(psi_, it - 1);
    a.psi_n = PS(psi_, it);
    a.act_min = act_min;
    a.f0 = taps_[0].view<T>();
    a.f1 = taps_[1].view<T>();
    a.tf = c_.pts.d_srctf.as<T>();
    a.ncomp = 2;
    a.kind = kind_ == SWX_SRC_MOMENT_TENSOR ? 1 : 0;
    launch_stress(a, 0, v == 1);
    a.kind = kind_ == SWX_SRC_EXTERNAL_FORCE ? 2 : 0;
    a.hmode = c_.d.gradient ? hmode : 0;
    a.cbase = cbase;
    launch_disp(a, 0, v == 1);
    record(it, traces);
  }

  void record(int it, T *traces) {
    if (!traces) return;
    const int thr = 2 * maxrec_;
    dim3 g((unsigned)ceil_div(thr, 128), (unsigned)nsh_);
    el_record_kernel<T><<<g, 128, 0, c_.stream>>>(U(u_, it), N_, taps_[2].view<T>(), taps_[3].view<T>(),
                                                   c_.pts.d_nt.as<int>(), c_.pts.d_off_tr.as<int64_t>(), traces, it,
                                                   maxrec_);
    SWX_HIP(hipGetLastError());
  }

  void step_adj_v(int v, int it, int cbase) {
    if (v == 2 || v > 100) {
      ElArgs<T> a = base_args(it);
      a.ucur = U(au_, it + 1);
      a.unew = U(au_, it);
      a.sig = asig_.as<T>();
      a.psi_o = PS(apsi_, it + 1);
      a.psi_n = PS(apsi_, it);
      a.cbase = cbase;
      a.corr_step = it + 1;
      a.corr_lim = cbase + S_;
      a.kind = 2;
      a.f0 = taps_[2].view<T>();
      a.f1 = taps_[3].view<T>();
      a.tf = c_.pts.d_res.as<T>();
      a.off_tf = c_.pts.d_off_tr.as<int64_t>();
      a.ncomp = 2;
      if (v == 2) {
        a.tbx = (int)ceil_div(nx_, TX);
        a.tbz = (int)ceil_div(nz_, ElTile<T>::rows);
        const unsigned g = (unsigned)((int64_t)a.tbx * a.tbz * nsh_);
        c_.ktimer.begin(c_.stream);
        el_tile_kernel<T, true><<<g, dim3(TX, TZ), 0, c_.stream>>>(a);
        c_.ktimer.end(c_.stream, 1);
        SWX_HIP(hipGetLastError());
      } else {
        launch_march<true>(a, v, 1);
      }
      return;
    }
    ElArgs<T> a = base_args(it);
    a.ucur = U(au_, it + 1);
    a.unew = U(au_, it);
    a.sig = asig_.as<T>();
    a.psi_o = PS(apsi_, it + 1);
    a.psi_n = PS(apsi_, it);
    a.cbase = cbase;
    a.corr_step = it + 1;
    a.corr_lim = cbase + S_;
    a.kind = 0;
    launch_stress(a, 1, v == 1);
    a.corr_step = 0;
    a.kind = 2;
    a.f0 = taps_[2].view<T>();
    a.f1 = taps_[3].view<T>();
    a.tf = c_.pts.d_res.as<T>();
    a.off_tf = c_.pts.d_off_tr.as<int64_t>();
    a.ncomp = 2;
    launch_disp(a, 1, v == 1);
  }

  // checkpoint k (k = 1 .. floor(nt_max/S)): both displacement levels
  // (u^{kS-1}, u^{kS}) and the C-PML memory variables of every shot
  void ckpt_copy(int k, bool save) {
    char *base = ck_.as<char>() + (int64_t)(k - 1) * ck_bytes_one_;
    struct Part { void *dev; int64_t bytes; };
    Part parts[2] = {{u_.get(), (int64_t)u_.bytes()},
                     {PS(psi_, k * S_), (int64_t)psi_.bytes() / 2}};
    int64_t off = 0;
    for (auto &p : parts) {
      if (save)
        SWX_HIP(hipMemcpyAsync(base + off, p.dev, p.bytes, hipMemcpyDeviceToDevice, c_.stream));
      else
        SWX_HIP(hipMemcpyAsync(p.dev, base + off, p.bytes, hipMemcpyDeviceToDevice, c_.stream));
      off += p.bytes;
    }
  }

  PlanCore &c_;
  int nx_, nz_, h_, ns_;
  int64_t N_, psi_stride_ = 1;
  int64_t psioff_[8];
  int cstride_;
  DevBuf model_, u_, sig_, psi_, cpml_;
  DevBuf au_, asig_, apsi_, grad_, gout_, hist_, ck_;
  DevBuf mom_, offm_;
  TapTable taps_[4];
  int kind_ = SWX_SRC_MOMENT_TENSOR, maxrec_ = 0;
  bool edge_taps_ = false;
  int max_ucell0_ = 0;
  int var_fwd_ = -1, var_adj_ = -1;
  // XCD-aware tile map of the block launches (env SWX_EL2D_XCD=0 turns it off)
  bool xcd_ = [] {
    const char *v = std::getenv("SWX_EL2D_XCD");
    return v == nullptr || std::atoi(v) != 0;
  }();
  // frame-first block order (env SWX_EL2D_ORDER=0 turns it off)
  bool order_ = [] {
    const char *v = std::getenv("SWX_EL2D_ORDER");
    return v == nullptr || std::atoi(v) != 0;
  }();
  int64_t ck_bytes_one_ = 0, state_bytes_ = 0;
  int S_ = 1, K_ = 1, nck_ = 0;
  bool store_all_ = true;
  int nsh_ = 0, ntmax_ = 0;
  std::vector<int> nt_;
  bool model_set_ = false;
};

}  // namespace

std::unique_ptr<Engine> make_elastic2d(PlanCore &core) {
  if (core.d.dtype == SWX_F32) return std::make_unique<El2dEngine<float>>(core);
  return std::make_unique<El2dEngine<double>>(core);
}

}  // namespace swx
