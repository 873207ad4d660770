/*
 * 1D acoustic constant density (CD) and variable density (VD) time stepping
 * and adjoint gradients.  TEST INFRASTRUCTURE ONLY (see oracle_impl.h).
 * Included by swx_oracle.c after oracle_impl.h (which defines A2), once per
 * precision.  Indices below are the reference's 1-based ones.
 */

#define A1(arr, i) (arr)[(i) - 1]

/* ------------------------------------------------------------------------ */
/* CD 1D                                                                    */
/* ------------------------------------------------------------------------ */

typedef struct {
  real *p[3];        /* pold, pcur, pnew (rotated by pointer swaps) */
  real *psi, *xi;    /* 2halo (+1 spare) and 2(halo+1)              */
} SFX(cd1d_state);

static void SFX(cd1d_state_alloc)(int64_t nx, int halo, SFX(cd1d_state) *s) {
  for (int k = 0; k < 3; k++) s->p[k] = (real *)calloc((size_t)nx, sizeof(real));
  s->psi = (real *)calloc((size_t)(2 * halo) + 1, sizeof(real));
  s->xi = (real *)calloc((size_t)(2 * halo + 2), sizeof(real));
}
static void SFX(cd1d_state_free)(SFX(cd1d_state) *s) {
  for (int k = 0; k < 3; k++) free(s->p[k]);
  free(s->psi);
  free(s->xi);
}
static void SFX(cd1d_rotate)(SFX(cd1d_state) *s) {
  /* acoustic1D_xPU.jl:89-92 (see cd2d_onestep for the edge consequence) */
  real *old = s->p[0];
  s->p[0] = s->p[1];
  s->p[1] = s->p[2];
  s->p[2] = old;
}

/* forward_onestep_CPML! / adjoint_onestep_CPML! of acoustic1D_xPU.jl:58-125:
 * update_psi! over 1:2halo (:15-24), update_p_CPML! over 2:nx-1 (:26-43),
 * inject_sources! (:1-6), record_receivers! (:8-13).  The outermost cells
 * keep level it-2 (pnew shares pold's array after the rotation). */
static void SFX(cd1d_onestep)(const oracle_geom *g, const real *fact, SFX(cd1d_state) *st,
                              const SFX(cpml) *cx, int nsrc, const int64_t *src, const real *srctf, int nt,
                              int it, int nrec, const int64_t *rec, real *traces) {
  const int64_t nx = g->n[0];
  const int halo = g->halo;
  const real _dx = (real)(1.0 / g->h[0]);
  const real *pold = st->p[0], *pcur = st->p[1];
  real *pnew = st->p[2];
  for (int64_t i = 1; i <= 2 * halo; i++) {
    int64_t ii = i > halo ? nx - halo - 1 + (i - halo) : i;
    real d = (A1(pcur, ii + 1) - A1(pcur, ii)) * _dx;
    if (ii <= halo) {
      A1(st->psi, ii) = cx->b_h[ii - 1] * A1(st->psi, ii) + cx->a_h[ii - 1] * d;
    } else if (ii >= nx - halo) {
      int64_t k = ii - (nx - halo) + 1 + halo;
      A1(st->psi, k) = cx->b_h[k - 1] * A1(st->psi, k) + cx->a_h[k - 1] * d;
    }
  }
  for (int64_t i = 2; i <= nx - 1; i++) {
    real pc = A1(pcur, i);
    real d2x = (A1(pcur, i - 1) - 2 * pc + A1(pcur, i + 1)) * (_dx * _dx);
    real lx;
    if (i <= halo) {
      real dpsi = (A1(st->psi, i) - A1(st->psi, i - 1)) * _dx;
      A1(st->xi, i) = cx->b[i - 1] * A1(st->xi, i) + cx->a[i - 1] * (d2x + dpsi);
      lx = d2x + dpsi + A1(st->xi, i);
    } else if (i >= nx - halo + 1) {
      int64_t k = i - (nx - halo) + 1 + halo;
      real dpsi = (A1(st->psi, k) - A1(st->psi, k - 1)) * _dx;
      A1(st->xi, k) = cx->b[k - 1] * A1(st->xi, k) + cx->a[k - 1] * (d2x + dpsi);
      lx = d2x + dpsi + A1(st->xi, k);
    } else {
      lx = d2x;
    }
    A1(pnew, i) = 2 * pc - A1(pold, i) + A1(fact, i) * lx;
  }
  A1(pnew, 1) = A1(pold, 1);
  A1(pnew, nx) = A1(pold, nx);
  for (int s = 0; s < nsrc; s++) A1(pnew, src[s]) += srctf[(it - 1) + (size_t)nt * s];
  if (traces)
    for (int r = 0; r < nrec; r++) traces[(it - 1) + (size_t)nt * r] = A1(pnew, rec[r]);
}

/* swforward_1shot! for CD 1D (acou_forward.jl:22-62) */
int SFX(oracle_cd1d_forward)(const oracle_geom *g, const real *fact, const SFX(cpml) *cx, int nt, int nsrc,
                             const int64_t *src, const real *srctf, int nrec, const int64_t *rec, real *traces) {
  SFX(cd1d_state) s;
  SFX(cd1d_state_alloc)(g->n[0], g->halo, &s);
  for (int it = 1; it <= nt; it++) {
    SFX(cd1d_onestep)(g, fact, &s, cx, nsrc, src, srctf, nt, it, nrec, rec, traces);
    SFX(cd1d_rotate)(&s);
    if (snap_due(it)) snap_put(s.p[1], (size_t)g->n[0] * sizeof(real));   /* "pcur" */
  }
  SFX(cd1d_state_free)(&s);
  return 0;
}

/* swgradient_1shot! for CD 1D (acou_gradient.jl:4-95) up to the raw
 * grad_vp field, with the LinearCheckpointer of acou_models.jl:159-173
 * (checkpointers.jl:1-107): pcur checkpointed with width 2, psi and xi at
 * multiples of cf, pcur buffered from floor(nt/cf)*cf. */
int SFX(oracle_cd1d_gradient)(const oracle_geom *g, const real *fact, const SFX(cpml) *cx, int nt, int nsrc,
                              const int64_t *src, const real *srctf, int nrec, const int64_t *rec,
                              const real *adjsrc, int cf, real *grad, real *traces) {
  if (cf < 1 || cf >= nt) return -1;
  const int64_t nx = g->n[0];
  const int halo = g->halo;
  const size_t N = (size_t)nx, npsi = (size_t)(2 * halo) + 1, nxi = (size_t)(2 * halo + 2);
  SFX(cd1d_state) s, a;
  SFX(cd1d_state_alloc)(nx, halo, &s);
  SFX(cd1d_state_alloc)(nx, halo, &a);
  const int last = (nt / cf) * cf;
  int curr = last;
  /* ck_p[it + 1] for it in -1..nt+1 with it % cf == 0 or (it + 1) % cf == 0 */
  real **ck_p = (real **)calloc(nt + 3, sizeof(real *));
#define HAS_P(t) ((t) >= -1 && (t) <= nt + 1 && (((t) >= 0 && (t) % cf == 0) || ((t) + 1 >= 0 && ((t) + 1) % cf == 0)))
  for (int t = -1; t <= nt + 1; t++)
    if (HAS_P(t)) ck_p[t + 1] = (real *)calloc(N, sizeof(real));
  const int nk = (nt + 1) / cf + 1;
  real **ck_psi = (real **)calloc(nk, sizeof(real *)), **ck_xi = (real **)calloc(nk, sizeof(real *));
  for (int k = 0; k < nk; k++) {
    ck_psi[k] = (real *)calloc(npsi, sizeof(real));
    ck_xi[k] = (real *)calloc(nxi, sizeof(real));
  }
  real **buf = (real **)calloc(cf + 1, sizeof(real *));
  for (int k = 0; k <= cf; k++) buf[k] = (real *)calloc(N, sizeof(real));
  /* savecheckpoint! (checkpointers.jl:44-56) of the current state at it */
#define SAVE(it_, pc_)                                                                 \
  do {                                                                                 \
    for (int itw = (it_); itw <= (it_) + 1; itw++)                                     \
      if (itw % cf == 0) memcpy(ck_p[(it_) + 1], (pc_), N * sizeof(real));             \
    if ((it_) >= 0 && (it_) % cf == 0) {                                               \
      memcpy(ck_psi[(it_) / cf], s.psi, npsi * sizeof(real));                          \
      memcpy(ck_xi[(it_) / cf], s.xi, nxi * sizeof(real));                             \
    }                                                                                  \
    if ((it_) >= last) memcpy(buf[(it_) - last], (pc_), N * sizeof(real));             \
  } while (0)
#define GETSAVED(t) (HAS_P(t) ? ck_p[(t) + 1] : buf[(t) - curr])
  SAVE(-1, s.p[0]);   /* acou_models.jl:169-172 */
  SAVE(0, s.p[1]);
  real *own_tr = traces ? traces : (real *)calloc((size_t)nt * nrec, sizeof(real));
  for (int it = 1; it <= nt; it++) {
    SFX(cd1d_onestep)(g, fact, &s, cx, nsrc, src, srctf, nt, it, nrec, rec, own_tr);
    SFX(cd1d_rotate)(&s);
    SAVE(it, s.p[1]);
  }
  /* prescale_residuals! (acoustic1D_xPU.jl:45-56) */
  real *res = (real *)malloc((size_t)nt * nrec * sizeof(real));
  for (int r = 0; r < nrec; r++)
    for (int it = 0; it < nt; it++) res[it + (size_t)nt * r] = adjsrc[it + (size_t)nt * r] * A1(fact, rec[r]);
  memset(grad, 0, N * sizeof(real));
  const real _dt2 = (real)(1.0 / (g->dt * g->dt));
  for (int it = nt; it >= 1; it--) {
    SFX(cd1d_onestep)(g, fact, &a, cx, nrec, rec, res, nt, it, 0, NULL, NULL);
    SFX(cd1d_rotate)(&a);
    int t = it - 2;
    int saved = (curr <= t && t <= curr + cf) || HAS_P(t);
    if (!saved) {
      /* initrecover! + recover! (checkpointers.jl:87-105) */
      int old = curr;
      curr -= cf;
      /* levels curr - 1 and curr are checkpointed (pcur has width 2) */
      memcpy(buf[0], ck_p[curr + 1], N * sizeof(real));
      memcpy(buf[cf], ck_p[old + 1], N * sizeof(real));
      memcpy(s.p[0], ck_p[curr], N * sizeof(real));
      memcpy(s.p[1], ck_p[curr + 1], N * sizeof(real));
      memcpy(s.psi, ck_psi[curr / cf], npsi * sizeof(real));
      memcpy(s.xi, ck_xi[curr / cf], nxi * sizeof(real));
      for (int rit = curr + 1; rit <= curr + cf - 1; rit++) {
        SFX(cd1d_onestep)(g, fact, &s, cx, nsrc, src, srctf, nt, rit, 0, NULL, NULL);
        SFX(cd1d_rotate)(&s);
        memcpy(buf[rit - curr], s.p[1], N * sizeof(real));
      }
    }
    const real *pa = GETSAVED(it - 2), *pb = GETSAVED(it - 1), *pc = GETSAVED(it);
    /* correlate_gradient! (correlate_gradient_xPU.jl:1-10) */
    for (size_t q = 0; q < N; q++) grad[q] = grad[q] + a.p[1][q] * (pa[q] - 2 * pb[q] + pc[q]) * _dt2;
  }
#undef SAVE
#undef GETSAVED
#undef HAS_P
  free(res);
  if (!traces) free(own_tr);
  for (int t = -1; t <= nt + 1; t++) free(ck_p[t + 1]);
  free(ck_p);
  for (int k = 0; k < nk; k++) { free(ck_psi[k]); free(ck_xi[k]); }
  free(ck_psi); free(ck_xi);
  for (int k = 0; k <= cf; k++) free(buf[k]);
  free(buf);
  SFX(cd1d_state_free)(&s);
  SFX(cd1d_state_free)(&a);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* VD 1D (staggered pressure / velocity)                                    */
/* ------------------------------------------------------------------------ */

/* 4th-order staggered first derivative (fdgen.jl:65-135, order=4) of the
 * n-element vector A at base index b; terms outside 1..n dropped. */
static real SFX(d4_1d)(const real *A, int64_t n, int64_t b) {
#define at(k) A1(A, (k))
  const real c1 = (real)(1.0 / 24.0), c2 = (real)(-9.0 / 8.0);
  if (b - 1 >= 1 && b + 2 <= n) return c1 * at(b - 1) + c2 * at(b) + (-c2) * at(b + 1) + (-c1) * at(b + 2);
  return D4_TERM(b - 1, c1) + D4_TERM(b, c2) + D4_TERM(b + 1, -c2) + D4_TERM(b + 2, -c1);
#undef at
}

typedef struct {
  real *p, *vx;      /* nx and nx-1 */
  real *psi, *xi;    /* 2halo (+1 spare) and 2(halo+1) */
} SFX(vd1d_state);

static void SFX(vd1d_state_alloc)(int64_t nx, int halo, SFX(vd1d_state) *s) {
  s->p = (real *)calloc((size_t)nx, sizeof(real));
  s->vx = (real *)calloc((size_t)(nx - 1), sizeof(real));
  s->psi = (real *)calloc((size_t)(2 * halo) + 1, sizeof(real));
  s->xi = (real *)calloc((size_t)(2 * halo + 2), sizeof(real));
}
static void SFX(vd1d_state_free)(SFX(vd1d_state) *s) {
  free(s->p); free(s->vx); free(s->psi); free(s->xi);
}
static void SFX(vd1d_state_copy)(int64_t nx, int halo, SFX(vd1d_state) *dst, const SFX(vd1d_state) *src) {
  memcpy(dst->p, src->p, (size_t)nx * sizeof(real));
  memcpy(dst->vx, src->vx, (size_t)(nx - 1) * sizeof(real));
  memcpy(dst->psi, src->psi, ((size_t)(2 * halo) + 1) * sizeof(real));
  memcpy(dst->xi, src->xi, (size_t)(2 * halo + 2) * sizeof(real));
}

/* update_p_CPML! (acoustic1D_VD_xPU.jl:15-32), range 2:nx-1 */
static void SFX(vd1d_update_p)(const oracle_geom *g, const real *m0, SFX(vd1d_state) *s, const SFX(cpml) *cx) {
  const int64_t nx = g->n[0];
  const int halo = g->halo;
  const real _dx = (real)(1.0 / g->h[0]);
  for (int64_t i = 2; i <= nx - 1; i++) {
    /* @d~x(vx, a_x, b_x, xi_x, order=4, I=(i-1,), halfgrid=false) */
    real dvx = SFX(d4_1d)(s->vx, nx - 1, i - 1) * _dx;
    if (i <= halo + 1) {
      A1(s->xi, i) = cx->b[i - 1] * A1(s->xi, i) + cx->a[i - 1] * dvx;
      dvx = dvx + A1(s->xi, i);
    } else if (i >= nx - halo) {
      int64_t k = i - (nx - halo) + 1 + (halo + 1);
      A1(s->xi, k) = cx->b[k - 1] * A1(s->xi, k) + cx->a[k - 1] * dvx;
      dvx = dvx + A1(s->xi, k);
    }
    A1(s->p, i) -= A1(m0, i) * dvx;
  }
}

/* update_vx_CPML! (acoustic1D_VD_xPU.jl:34-51), range 1:nx-1 */
static void SFX(vd1d_update_v)(const oracle_geom *g, const real *m1, SFX(vd1d_state) *s, const SFX(cpml) *cx) {
  const int64_t nx = g->n[0];
  const int halo = g->halo;
  const real _dx = (real)(1.0 / g->h[0]);
  for (int64_t i = 1; i <= nx - 1; i++) {
    real d = SFX(d4_1d)(s->p, nx, i) * _dx;
    if (i <= halo) {
      A1(s->psi, i) = cx->b_h[i - 1] * A1(s->psi, i) + cx->a_h[i - 1] * d;
      d = d + A1(s->psi, i);
    } else if (i >= nx - halo) {
      int64_t k = i - (nx - halo) + 1 + halo;
      A1(s->psi, k) = cx->b_h[k - 1] * A1(s->psi, k) + cx->a_h[k - 1] * d;
      d = d + A1(s->psi, k);
    }
    A1(s->vx, i) -= A1(m1, i) * d;
  }
}

/* forward_onestep_CPML! (acoustic1D_VD_xPU.jl:66-99) */
static void SFX(vd1d_fwd_step)(const oracle_geom *g, const real *m0, const real *m1, SFX(vd1d_state) *s,
                               const SFX(cpml) *cx, int nsrc, const int64_t *src, const real *srctf, int nt,
                               int it, int nrec, const int64_t *rec, real *traces) {
  SFX(vd1d_update_p)(g, m0, s, cx);
  for (int k = 0; k < nsrc; k++) A1(s->p, src[k]) += srctf[(it - 1) + (size_t)nt * k];
  SFX(vd1d_update_v)(g, m1, s, cx);
  if (traces)
    for (int r = 0; r < nrec; r++) traces[(it - 1) + (size_t)nt * r] = A1(s->p, rec[r]);
}

/* adjoint_onestep_CPML! (acoustic1D_VD_xPU.jl:101-128) */
static void SFX(vd1d_adj_step)(const oracle_geom *g, const real *m0, const real *m1, SFX(vd1d_state) *s,
                               const SFX(cpml) *cx, int nrec, const int64_t *rec, const real *res, int nt, int it) {
  SFX(vd1d_update_v)(g, m1, s, cx);
  SFX(vd1d_update_p)(g, m0, s, cx);
  for (int r = 0; r < nrec; r++) A1(s->p, rec[r]) += res[(it - 1) + (size_t)nt * r];
}

/* swforward_1shot! for VD 1D (acou_forward.jl:83-125) */
int SFX(oracle_vd1d_forward)(const oracle_geom *g, const real *m0, const real *m1, const SFX(cpml) *cx, int nt,
                             int nsrc, const int64_t *src, const real *srctf, int nrec, const int64_t *rec,
                             real *traces) {
  const int64_t nx = g->n[0];
  SFX(vd1d_state) s;
  SFX(vd1d_state_alloc)(nx, g->halo, &s);
  for (int it = 1; it <= nt; it++) {
    SFX(vd1d_fwd_step)(g, m0, m1, &s, cx, nsrc, src, srctf, nt, it, nrec, rec, traces);
    if (snap_due(it)) {   /* "pcur", then "vcur" = [vx] */
      snap_put(s.p, (size_t)nx * sizeof(real));
      snap_put(s.vx, (size_t)(nx - 1) * sizeof(real));
    }
  }
  SFX(vd1d_state_free)(&s);
  return 0;
}

/* swgradient_1shot! for VD 1D (acou_gradient.jl:97-184) up to the raw
 * grad_m0 and grad_m1_stag fields; checkpointer as oracle_vd2d_gradient.
 * correlate_gradient_m1! in 1D covers 2:nx-2 only (acoustic1D_VD_xPU.jl:
 * 130-141), unlike the 2D one. */
int SFX(oracle_vd1d_gradient)(const oracle_geom *g, const real *m0, const real *m1, const SFX(cpml) *cx, int nt,
                              int nsrc, const int64_t *src, const real *srctf, int nrec, const int64_t *rec,
                              const real *adjsrc, int cf, real *g_m0, real *g_m1, real *traces) {
  if (cf < 1 || cf >= nt) return -1;
  const int64_t nx = g->n[0];
  const int halo = g->halo;
  const size_t N = (size_t)nx;
  SFX(vd1d_state) s, a;
  SFX(vd1d_state_alloc)(nx, halo, &s);
  SFX(vd1d_state_alloc)(nx, halo, &a);
  const int last = (nt / cf) * cf;
  int curr = last;
  const int nk = (nt + 1) / cf + 1;
  SFX(vd1d_state) *ck = (SFX(vd1d_state) *)calloc(nk, sizeof(SFX(vd1d_state)));
  for (int k = 0; k < nk; k++) SFX(vd1d_state_alloc)(nx, halo, &ck[k]);   /* ck[0] = zero state */
  real **buf = (real **)calloc(cf + 1, sizeof(real *));
  for (int k = 0; k <= cf; k++) buf[k] = (real *)calloc(N, sizeof(real));
  real *own_tr = traces ? traces : (real *)calloc((size_t)nt * nrec, sizeof(real));
  for (int it = 1; it <= nt; it++) {
    SFX(vd1d_fwd_step)(g, m0, m1, &s, cx, nsrc, src, srctf, nt, it, nrec, rec, own_tr);
    if (it % cf == 0) SFX(vd1d_state_copy)(nx, halo, &ck[it / cf], &s);
    if (it >= last) memcpy(buf[it - last], s.p, N * sizeof(real));
  }
  /* prescale_residuals! (acoustic1D_VD_xPU.jl:53-64) */
  real *res = (real *)malloc((size_t)nt * nrec * sizeof(real));
  for (int r = 0; r < nrec; r++)
    for (int it = 0; it < nt; it++) res[it + (size_t)nt * r] = adjsrc[it + (size_t)nt * r] * A1(m0, rec[r]);
  memset(g_m0, 0, N * sizeof(real));
  memset(g_m1, 0, (N - 1) * sizeof(real));
  const real _dt = (real)(1.0 / g->dt);
  const real _dx = (real)(1.0 / g->h[0]);
#define VD_SAVED(t) (((t) % cf == 0) ? ck[(t) / cf].p : buf[(t) - curr])
  for (int it = nt; it >= 1; it--) {
    SFX(vd1d_adj_step)(g, m0, m1, &a, cx, nrec, rec, res, nt, it);
    int t = it - 1;
    int saved = (curr <= t && t <= curr + cf) || (t % cf == 0);
    if (!saved) {
      int old = curr;
      curr -= cf;
      memcpy(buf[0], ck[curr / cf].p, N * sizeof(real));
      memcpy(buf[cf], ck[old / cf].p, N * sizeof(real));
      SFX(vd1d_state_copy)(nx, halo, &s, &ck[curr / cf]);
      for (int rit = curr + 1; rit <= curr + cf - 1; rit++) {
        SFX(vd1d_fwd_step)(g, m0, m1, &s, cx, nsrc, src, srctf, nt, rit, 0, NULL, NULL);
        memcpy(buf[rit - curr], s.p, N * sizeof(real));
      }
    }
    const real *pc = VD_SAVED(it);
    const real *po = VD_SAVED(it - 1);
    /* correlate_gradient_m0! (correlate_gradient_xPU.jl:12-21) */
    for (size_t q = 0; q < N; q++) g_m0[q] = g_m0[q] - (a.p[q] * (pc[q] - po[q]) * _dt);
    /* correlate_gradient_m1! (acoustic1D_VD_xPU.jl:130-141) */
    for (int64_t i = 2; i <= nx - 2; i++)
      A1(g_m1, i) = A1(g_m1, i) + A1(a.vx, i) * (SFX(d4_1d)(pc, nx, i) * _dx);
  }
#undef VD_SAVED
  free(res);
  if (!traces) free(own_tr);
  for (int k = 0; k <= cf; k++) free(buf[k]);
  free(buf);
  for (int k = 0; k < nk; k++) SFX(vd1d_state_free)(&ck[k]);
  free(ck);
  SFX(vd1d_state_free)(&s);
  SFX(vd1d_state_free)(&a);
  return 0;
}
