/*
 * CPU restatement of the 3D acoustic constant-density path of SeismicWaves.jl
 * (TEST INFRASTRUCTURE: the checker, never the product path).  Included by
 * swx_oracle.c after oracle_impl.h, once per precision.
 *
 * Reference: src/models/acoustic/backends/shared/acoustic3D_xPU.jl:1-219,
 * stencils src/utils/fdgen.jl:137-193 (order 2), driver
 * src/models/acoustic/acou_gradient.jl:4-95, checkpointer
 * src/utils/checkpointers.jl.  Loops use the reference's 1-based indices.
 */

#define A3(arr, n1, n2, i, j, k) (arr)[((i) - 1) + (size_t)(n1) * (((j) - 1) + (size_t)(n2) * ((k) - 1))]

typedef struct {
  real *p[3];                 /* pold, pcur, pnew                          */
  real *psi[3], *xi[3];       /* x: 2h x ny x nz, y: nx x 2h x nz, z: nx x ny x 2h (xi: 2h+2) */
} SFX(cd3d_state);

static size_t SFX(cd3d_strip)(const oracle_geom *g, int axis, int w) {
  size_t n = 1;
  for (int d = 0; d < 3; d++) n *= (d == axis) ? (size_t)w : (size_t)g->n[d];
  return n;
}

static void SFX(cd3d_state_alloc)(const oracle_geom *g, SFX(cd3d_state) *s) {
  size_t N = (size_t)g->n[0] * g->n[1] * g->n[2];
  for (int k = 0; k < 3; k++) s->p[k] = (real *)calloc(N, sizeof(real));
  for (int d = 0; d < 3; d++) {
    s->psi[d] = (real *)calloc(SFX(cd3d_strip)(g, d, 2 * g->halo) + 1, sizeof(real));
    s->xi[d] = (real *)calloc(SFX(cd3d_strip)(g, d, 2 * g->halo + 2), sizeof(real));
  }
}
static void SFX(cd3d_state_free)(SFX(cd3d_state) *s) {
  for (int k = 0; k < 3; k++) free(s->p[k]);
  for (int d = 0; d < 3; d++) { free(s->psi[d]); free(s->xi[d]); }
}
static void SFX(cd3d_rotate)(SFX(cd3d_state) *s) {
  real *old = s->p[0];
  s->p[0] = s->p[1];
  s->p[1] = s->p[2];
  s->p[2] = old;
}

/* forward_onestep_CPML! of acoustic3D_xPU.jl (update_psi_x/y/z!, then
 * update_p_CPML!, inject_sources!, record_receivers!).  As in 2D, the
 * reference's rotation (acoustic3D_xPU.jl:155-157) binds pnew to the array of
 * the new pold, so cells on the six outer faces keep level it-2 plus any
 * injection: their pold values are copied into pnew before the injection. */
static void SFX(cd3d_onestep)(const oracle_geom *g, const real *fact, SFX(cd3d_state) *st,
                              const SFX(cpml) *c, int nsrc, const int64_t *src, const real *srctf, int nt,
                              int it, int nrec, const int64_t *rec, real *traces) {
  const int64_t nx = g->n[0], ny = g->n[1], nz = g->n[2];
  const int halo = g->halo;
  const real _d[3] = {(real)(1.0 / g->h[0]), (real)(1.0 / g->h[1]), (real)(1.0 / g->h[2])};
  const real *pold = st->p[0], *pcur = st->p[1];
  real *pnew = st->p[2];
  const int64_t n[3] = {nx, ny, nz};
  /* update_psi_{x,y,z}! (acoustic3D_xPU.jl:1-38): strip index i in 1:2halo
   * along the axis, shifted to the right side beyond halo */
  for (int ax = 0; ax < 3; ax++) {
    int64_t m[3] = {nx, ny, nz};
    m[ax] = 2 * halo;
    real *psi = st->psi[ax];
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t k = 1; k <= m[2]; k++)
      for (int64_t j = 1; j <= m[1]; j++)
        for (int64_t i = 1; i <= m[0]; i++) {
          int64_t I[3] = {i, j, k};
          int64_t t = I[ax];
          int64_t tt = t > halo ? n[ax] - halo - 1 + (t - halo) : t;
          int64_t P[3] = {i, j, k}, Q[3] = {i, j, k};
          P[ax] = tt;
          Q[ax] = tt + 1;
          real d = (A3(pcur, nx, ny, Q[0], Q[1], Q[2]) - A3(pcur, nx, ny, P[0], P[1], P[2])) * _d[ax];
          int64_t kk;
          if (tt <= halo) kk = tt;
          else if (tt >= n[ax] - halo) kk = tt - (n[ax] - halo) + 1 + halo;
          else continue;
          int64_t S[3] = {i, j, k};
          S[ax] = kk;
          const real *bh = c[ax].b_h, *ah = c[ax].a_h;
          A3(psi, m[0], m[1], S[0], S[1], S[2]) = bh[kk - 1] * A3(psi, m[0], m[1], S[0], S[1], S[2]) + ah[kk - 1] * d;
        }
  }
  /* update_p_CPML! (acoustic3D_xPU.jl:40-63) */
#pragma omp parallel for collapse(2) schedule(static)
  for (int64_t k = 2; k <= nz - 1; k++)
    for (int64_t j = 2; j <= ny - 1; j++)
      for (int64_t i = 2; i <= nx - 1; i++) {
        const real pc = A3(pcur, nx, ny, i, j, k);
        real lap = 0;
        for (int ax = 0; ax < 3; ax++) {
          int64_t I[3] = {i, j, k};
          int64_t L[3] = {i, j, k}, R[3] = {i, j, k};
          L[ax] -= 1;
          R[ax] += 1;
          real d2 = (A3(pcur, nx, ny, L[0], L[1], L[2]) - 2 * pc + A3(pcur, nx, ny, R[0], R[1], R[2])) * (_d[ax] * _d[ax]);
          int64_t t = I[ax], kk = -1;
          if (t <= halo) kk = t;
          else if (t >= n[ax] - halo + 1) kk = t - (n[ax] - halo) + 1 + halo;
          if (kk > 0) {
            int64_t m[3] = {nx, ny, nz}, mx[3] = {nx, ny, nz};
            m[ax] = 2 * halo;
            mx[ax] = 2 * halo + 2;
            int64_t A[3] = {i, j, k}, B[3] = {i, j, k};
            A[ax] = kk;
            B[ax] = kk - 1;
            const real *psi = st->psi[ax];
            real dpsi = (A3(psi, m[0], m[1], A[0], A[1], A[2]) - A3(psi, m[0], m[1], B[0], B[1], B[2])) * _d[ax];
            real *xip = &A3(st->xi[ax], mx[0], mx[1], A[0], A[1], A[2]);
            *xip = c[ax].b[kk - 1] * *xip + c[ax].a[kk - 1] * (d2 + dpsi);
            lap += d2 + dpsi + *xip;
          } else {
            lap += d2;
          }
        }
        A3(pnew, nx, ny, i, j, k) = 2 * pc - A3(pold, nx, ny, i, j, k) + A3(fact, nx, ny, i, j, k) * lap;
      }
#pragma omp parallel for schedule(static)
  for (int64_t k = 1; k <= nz; k++)
    for (int64_t j = 1; j <= ny; j++) {
      const int face = (k == 1 || k == nz || j == 1 || j == ny);
      for (int64_t i = 1; i <= nx; i += (face || i == nx) ? 1 : nx - 1)
        A3(pnew, nx, ny, i, j, k) = A3(pold, nx, ny, i, j, k);
    }
  for (int s = 0; s < nsrc; s++)
    A3(pnew, nx, ny, src[s], src[s + nsrc], src[s + 2 * nsrc]) += srctf[(it - 1) + (size_t)nt * s];
  if (traces)
    for (int r = 0; r < nrec; r++)
      traces[(it - 1) + (size_t)nt * r] = A3(pnew, nx, ny, rec[r], rec[r + nrec], rec[r + 2 * nrec]);
}

int SFX(oracle_cd3d_forward)(const oracle_geom *g, const real *fact, const SFX(cpml) *c, int nt, int nsrc,
                             const int64_t *src, const real *srctf, int nrec, const int64_t *rec, real *traces) {
  SFX(cd3d_state) s;
  SFX(cd3d_state_alloc)(g, &s);
  for (int it = 1; it <= nt; it++) {
    SFX(cd3d_onestep)(g, fact, &s, c, nsrc, src, srctf, nt, it, nrec, rec, traces);
    SFX(cd3d_rotate)(&s);
    if (snap_due(it)) snap_put(s.p[1], (size_t)g->n[0] * g->n[1] * g->n[2] * sizeof(real));   /* "pcur" */
    oracle_progress("cd3d forward", it, nt);
  }
  SFX(cd3d_state_free)(&s);
  return 0;
}

/* swgradient_1shot! (acou_gradient.jl:4-95) for 3D with the LinearCheckpointer
 * (pcur width 2, psi, xi; buffered pcur).  Raw grad_vp. */
int SFX(oracle_cd3d_gradient)(const oracle_geom *g, const real *fact, const SFX(cpml) *c, int nt, int nsrc,
                              const int64_t *src, const real *srctf, int nrec, const int64_t *rec,
                              const real *adjsrc, int cf, real *grad, real *traces) {
  if (cf < 1 || cf >= nt) return -1;
  const int64_t nx = g->n[0], ny = g->n[1];
  const size_t N = (size_t)g->n[0] * g->n[1] * g->n[2];
  SFX(cd3d_state) s, a;
  SFX(cd3d_state_alloc)(g, &s);
  SFX(cd3d_state_alloc)(g, &a);
  size_t npsi[3], nxi[3];
  for (int d = 0; d < 3; d++) {
    npsi[d] = SFX(cd3d_strip)(g, d, 2 * g->halo) + 1;
    nxi[d] = SFX(cd3d_strip)(g, d, 2 * g->halo + 2);
  }
  const int last = (nt / cf) * cf;
  int curr = last;
  /* pcur checkpoints for it in -1..nt+1 where it%cf==0 or (it+1)%cf==0 */
  real **ckp = (real **)calloc(nt + 3, sizeof(real *));
#define HAS_P(t) ((t) >= -1 && (t) <= nt + 1 && ((((t) >= 0) && (t) % cf == 0) || (((t) + 1) >= 0 && ((t) + 1) % cf == 0)))
  for (int t = -1; t <= nt + 1; t++)
    if (HAS_P(t)) ckp[t + 1] = (real *)calloc(N, sizeof(real));
  const int nk = (nt + 1) / cf + 1;
  real ***ckpsi = (real ***)calloc(nk, sizeof(real **));
  for (int k = 0; k < nk; k++) {
    ckpsi[k] = (real **)calloc(6, sizeof(real *));
    for (int d = 0; d < 3; d++) {
      ckpsi[k][d] = (real *)calloc(npsi[d], sizeof(real));
      ckpsi[k][3 + d] = (real *)calloc(nxi[d], sizeof(real));
    }
  }
  real **buf = (real **)calloc(cf + 1, sizeof(real *));
  for (int k = 0; k <= cf; k++) buf[k] = (real *)calloc(N, sizeof(real));
  real *own_tr = traces ? traces : (real *)calloc((size_t)nt * nrec, sizeof(real));
  for (int it = 1; it <= nt; it++) {
    SFX(cd3d_onestep)(g, fact, &s, c, nsrc, src, srctf, nt, it, nrec, rec, own_tr);
    SFX(cd3d_rotate)(&s);
    for (int itw = it; itw <= it + 1; itw++)
      if (itw % cf == 0) memcpy(ckp[it + 1], s.p[1], N * sizeof(real));
    if (it % cf == 0)
      for (int d = 0; d < 3; d++) {
        memcpy(ckpsi[it / cf][d], s.psi[d], npsi[d] * sizeof(real));
        memcpy(ckpsi[it / cf][3 + d], s.xi[d], nxi[d] * sizeof(real));
      }
    if (it >= last) memcpy(buf[it - last], s.p[1], N * sizeof(real));
    oracle_progress("cd3d gradient, forward", it, nt);
  }
  real *res = (real *)malloc((size_t)nt * nrec * sizeof(real));
  for (int r = 0; r < nrec; r++) {
    real f = A3(fact, nx, ny, rec[r], rec[r + nrec], rec[r + 2 * nrec]);
    for (int it = 0; it < nt; it++) res[it + (size_t)nt * r] = adjsrc[it + (size_t)nt * r] * f;
  }
  memset(grad, 0, N * sizeof(real));
  const real _dt2 = (real)(1.0 / (g->dt * g->dt));
#define SAVED(t) (HAS_P(t) ? ckp[(t) + 1] : buf[(t) - curr])
  for (int it = nt; it >= 1; it--) {
    SFX(cd3d_onestep)(g, fact, &a, c, nrec, rec, res, nt, it, 0, NULL, NULL);
    SFX(cd3d_rotate)(&a);
    oracle_progress("cd3d gradient, adjoint", nt - it + 1, nt);
    int t = it - 2;
    if (!((curr <= t && t <= curr + cf) || HAS_P(t))) {
      int old = curr;
      curr -= cf;
      memcpy(buf[0], ckp[curr + 1], N * sizeof(real));
      memcpy(buf[cf], ckp[old + 1], N * sizeof(real));
      memcpy(s.p[0], ckp[curr], N * sizeof(real));
      memcpy(s.p[1], ckp[curr + 1], N * sizeof(real));
      for (int d = 0; d < 3; d++) {
        memcpy(s.psi[d], ckpsi[curr / cf][d], npsi[d] * sizeof(real));
        memcpy(s.xi[d], ckpsi[curr / cf][3 + d], nxi[d] * sizeof(real));
      }
      for (int rit = curr + 1; rit <= curr + cf - 1; rit++) {
        SFX(cd3d_onestep)(g, fact, &s, c, nsrc, src, srctf, nt, rit, 0, NULL, NULL);
        SFX(cd3d_rotate)(&s);
        memcpy(buf[rit - curr], s.p[1], N * sizeof(real));
      }
    }
    const real *pa = SAVED(it - 2), *pb = SAVED(it - 1), *pc = SAVED(it);
    const real *adj = a.p[1];
#pragma omp parallel for schedule(static)
    for (size_t q = 0; q < N; q++) grad[q] = grad[q] + adj[q] * (pa[q] - 2 * pb[q] + pc[q]) * _dt2;
  }
#undef SAVED
#undef HAS_P
  free(res);
  if (!traces) free(own_tr);
  for (int k = 0; k <= cf; k++) free(buf[k]);
  free(buf);
  for (int k = 0; k < nk; k++) {
    for (int d = 0; d < 6; d++) free(ckpsi[k][d]);
    free(ckpsi[k]);
  }
  free(ckpsi);
  for (int t = -1; t <= nt + 1; t++) free(ckp[t + 1]);
  free(ckp);
  SFX(cd3d_state_free)(&s);
  SFX(cd3d_state_free)(&a);
  return 0;
}
