/*
 * CPU restatement of the SeismicWaves.jl finite-difference time stepping and
 * adjoint-gradient path.  TEST INFRASTRUCTURE ONLY: this file is the checker
 * (and the timed CPU baseline), never the product path.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 *
 * This header is included twice by swx_oracle.c, once with `real` = double
 * (suffix _f64, the FP64 oracle) and once with `real` = float (suffix _f32,
 * used only as the same-precision CPU baseline).
 *
 * Conventions follow the reference exactly: arrays are column-major with the
 * first (x) index fastest, all loops below are written with the reference's
 * 1-based indices and the macros A2/A3 subtract one.  Every function cites the
 * reference file:line it restates (paths relative to the reference root).
 */

#define A2(arr, n1, i, j) (arr)[((i) - 1) + (size_t)(n1) * ((j) - 1)]

/* ------------------------------------------------------------------------ */
/* 2D acoustic constant density                                             */
/* ------------------------------------------------------------------------ */

/* update_psi_x!, update_psi_y!, update_p_CPML!, inject_sources!,
 * record_receivers! and forward_onestep_CPML! of
 * src/models/acoustic/backends/shared/acoustic2D_xPU.jl:1-126 with the stencil
 * semantics of src/utils/fdgen.jl:137-193 (order 2: first derivative
 * (A[i+1]-A[i])/h on the half grid, second derivative (A[i-1]-2A[i]+A[i+1])/h^2;
 * psi <- b_h psi + a_h dA in the strips, xi <- b xi + a (d2A + dpsi)).
 * The caller rotates pold/pcur/pnew (acoustic2D_xPU.jl:122-124).  Those three
 * sequential assignments leave "pnew" bound to the same array as the new
 * "pold", so the reference runs on two arrays and writes pnew over pold in
 * place: a cell outside the stencil range (the outermost rows and columns)
 * keeps the value of level it-2 plus what is injected there.  With three
 * arrays here, the edge cells of pold are copied into pnew to match.       */
static void SFX(cd2d_onestep)(const oracle_geom *g, const real *fact,
                              const real *pold, const real *pcur, real *pnew,
                              real *psi_x, real *psi_z, real *xi_x, real *xi_z,
                              const SFX(cpml) *cx, const SFX(cpml) *cz,
                              int nsrc, const int64_t *src, const real *srctf,
                              int nt, int it, int nrec, const int64_t *rec,
                              real *traces) {
  const int64_t nx = g->n[0], nz = g->n[1];
  const int halo = g->halo;
  const real _dx = (real)(1.0 / g->h[0]), _dz = (real)(1.0 / g->h[1]);
  const int64_t npx = 2 * halo, npz = 2 * halo;         /* psi strip widths */
  const int64_t nxx = 2 * (halo + 1), nxz = 2 * (halo + 1); /* xi widths  */

  /* update_psi_x!  (acoustic2D_xPU.jl:1-12), range (1:2halo, 1:nz) */
#pragma omp parallel for schedule(static)
  for (int64_t j = 1; j <= nz; j++) {
    for (int64_t i = 1; i <= 2 * halo; i++) {
      int64_t ii = i > halo ? nx - halo - 1 + (i - halo) : i;
      real d = (A2(pcur, nx, ii + 1, j) - A2(pcur, nx, ii, j)) * _dx;
      if (ii <= halo) {
        A2(psi_x, npx, ii, j) = (real)(cx->b_h[ii - 1] * A2(psi_x, npx, ii, j) + cx->a_h[ii - 1] * d);
      } else if (ii >= nx - halo) {
        int64_t k = ii - (nx - halo) + 1 + halo;
        A2(psi_x, npx, k, j) = (real)(cx->b_h[k - 1] * A2(psi_x, npx, k, j) + cx->a_h[k - 1] * d);
      }
    }
  }
  /* update_psi_y!  (acoustic2D_xPU.jl:14-25), range (1:nx, 1:2halo) */
#pragma omp parallel for schedule(static)
  for (int64_t j = 1; j <= 2 * halo; j++) {
    int64_t jj = j > halo ? nz - halo - 1 + (j - halo) : j;
    for (int64_t i = 1; i <= nx; i++) {
      real d = (A2(pcur, nx, i, jj + 1) - A2(pcur, nx, i, jj)) * _dz;
      if (jj <= halo) {
        A2(psi_z, nx, i, jj) = (real)(cz->b_h[jj - 1] * A2(psi_z, nx, i, jj) + cz->a_h[jj - 1] * d);
      } else if (jj >= nz - halo) {
        int64_t k = jj - (nz - halo) + 1 + halo;
        A2(psi_z, nx, i, k) = (real)(cz->b_h[k - 1] * A2(psi_z, nx, i, k) + cz->a_h[k - 1] * d);
      }
    }
  }
  (void)npz;
  /* update_p_CPML!  (acoustic2D_xPU.jl:27-46), range (2:nx-1, 2:nz-1) */
#pragma omp parallel for schedule(static)
  for (int64_t j = 2; j <= nz - 1; j++) {
    for (int64_t i = 2; i <= nx - 1; i++) {
      real pc = A2(pcur, nx, i, j);
      real d2x = (A2(pcur, nx, i - 1, j) - 2 * pc + A2(pcur, nx, i + 1, j)) * (_dx * _dx);
      real d2z = (A2(pcur, nx, i, j - 1) - 2 * pc + A2(pcur, nx, i, j + 1)) * (_dz * _dz);
      real lx, lz;
      if (i <= halo) {
        real dpsi = (A2(psi_x, npx, i, j) - A2(psi_x, npx, i - 1, j)) * _dx;
        A2(xi_x, nxx, i, j) = (real)(cx->b[i - 1] * A2(xi_x, nxx, i, j) + cx->a[i - 1] * (d2x + dpsi));
        lx = d2x + dpsi + A2(xi_x, nxx, i, j);
      } else if (i >= nx - halo + 1) {
        int64_t k = i - (nx - halo) + 1 + halo;
        real dpsi = (A2(psi_x, npx, k, j) - A2(psi_x, npx, k - 1, j)) * _dx;
        A2(xi_x, nxx, k, j) = (real)(cx->b[k - 1] * A2(xi_x, nxx, k, j) + cx->a[k - 1] * (d2x + dpsi));
        lx = d2x + dpsi + A2(xi_x, nxx, k, j);
      } else {
        lx = d2x;
      }
      if (j <= halo) {
        real dpsi = (A2(psi_z, nx, i, j) - A2(psi_z, nx, i, j - 1)) * _dz;
        A2(xi_z, nx, i, j) = (real)(cz->b[j - 1] * A2(xi_z, nx, i, j) + cz->a[j - 1] * (d2z + dpsi));
        lz = d2z + dpsi + A2(xi_z, nx, i, j);
      } else if (j >= nz - halo + 1) {
        int64_t k = j - (nz - halo) + 1 + halo;
        real dpsi = (A2(psi_z, nx, i, k) - A2(psi_z, nx, i, k - 1)) * _dz;
        A2(xi_z, nx, i, k) = (real)(cz->b[k - 1] * A2(xi_z, nx, i, k) + cz->a[k - 1] * (d2z + dpsi));
        lz = d2z + dpsi + A2(xi_z, nx, i, k);
      } else {
        lz = d2z;
      }
      A2(pnew, nx, i, j) = 2 * pc - A2(pold, nx, i, j) + A2(fact, nx, i, j) * (lx + lz);
    }
  }
  (void)nxz;
  /* edge cells: pnew is pold's array in the reference (see above) */
  for (int64_t i = 1; i <= nx; i++) {
    A2(pnew, nx, i, 1) = A2(pold, nx, i, 1);
    A2(pnew, nx, i, nz) = A2(pold, nx, i, nz);
  }
  for (int64_t j = 2; j <= nz - 1; j++) {
    A2(pnew, nx, 1, j) = A2(pold, nx, 1, j);
    A2(pnew, nx, nx, j) = A2(pold, nx, nx, j);
  }
  /* inject_sources! (acoustic2D_xPU.jl:48-54), serial order */
  for (int s = 0; s < nsrc; s++) {
    int64_t is = src[s], js = src[s + nsrc];
    A2(pnew, nx, is, js) += srctf[(it - 1) + (size_t)nt * s];
  }
  /* record_receivers! (acoustic2D_xPU.jl:56-62) */
  if (traces) {
    for (int r = 0; r < nrec; r++) {
      int64_t ir = rec[r], jr = rec[r + nrec];
      traces[(it - 1) + (size_t)nt * r] = A2(pnew, nx, ir, jr);
    }
  }
}

typedef struct {
  real *p[3];      /* pold, pcur, pnew (rotated by pointer swaps)          */
  real *psi_x, *psi_z, *xi_x, *xi_z;
} SFX(cd2d_state);

static size_t SFX(cd2d_nx_psi)(const oracle_geom *g) { return (size_t)(2 * g->halo) * g->n[1]; }
static size_t SFX(cd2d_nz_psi)(const oracle_geom *g) { return (size_t)g->n[0] * (2 * g->halo); }
static size_t SFX(cd2d_nx_xi)(const oracle_geom *g) { return (size_t)(2 * (g->halo + 1)) * g->n[1]; }
static size_t SFX(cd2d_nz_xi)(const oracle_geom *g) { return (size_t)g->n[0] * (2 * (g->halo + 1)); }

static void SFX(cd2d_state_alloc)(const oracle_geom *g, SFX(cd2d_state) *s) {
  size_t N = (size_t)g->n[0] * g->n[1];
  for (int k = 0; k < 3; k++) s->p[k] = (real *)calloc(N, sizeof(real));
  s->psi_x = (real *)calloc(SFX(cd2d_nx_psi)(g) + 1, sizeof(real));
  s->psi_z = (real *)calloc(SFX(cd2d_nz_psi)(g) + 1, sizeof(real));
  s->xi_x = (real *)calloc(SFX(cd2d_nx_xi)(g), sizeof(real));
  s->xi_z = (real *)calloc(SFX(cd2d_nz_xi)(g), sizeof(real));
}
static void SFX(cd2d_state_free)(SFX(cd2d_state) *s) {
  for (int k = 0; k < 3; k++) free(s->p[k]);
  free(s->psi_x); free(s->psi_z); free(s->xi_x); free(s->xi_z);
}
static void SFX(cd2d_rotate)(SFX(cd2d_state) *s) {
  /* acoustic2D_xPU.jl:122-124: pold <- pcur, pcur <- pnew, pnew <- pold */
  real *old = s->p[0];
  s->p[0] = s->p[1];
  s->p[1] = s->p[2];
  s->p[2] = old;
}

/* swforward_1shot! for the CD model (src/models/acoustic/acou_forward.jl:22-62):
 * reset, nt forward steps, traces recorded every step.  src/rec hold 1-based
 * grid indices (nsrc x 2 and nrec x 2, column-major), srctf is the already
 * scaled source term (nt x nsrc) of possrcrec_scaletf (acou_forward.jl:6-20). */
int SFX(oracle_cd2d_forward)(const oracle_geom *g, const real *fact,
                             const SFX(cpml) *cx, const SFX(cpml) *cz,
                             int nt, int nsrc, const int64_t *src, const real *srctf,
                             int nrec, const int64_t *rec, real *traces) {
  SFX(cd2d_state) s;
  SFX(cd2d_state_alloc)(g, &s);
  for (int it = 1; it <= nt; it++) {
    SFX(cd2d_onestep)(g, fact, s.p[0], s.p[1], s.p[2], s.psi_x, s.psi_z, s.xi_x, s.xi_z,
                      cx, cz, nsrc, src, srctf, nt, it, nrec, rec, traces);
    SFX(cd2d_rotate)(&s);
    if (snap_due(it)) snap_put(s.p[1], (size_t)g->n[0] * g->n[1] * sizeof(real));   /* "pcur" */
    oracle_progress("cd2d forward", it, nt);
  }
  SFX(cd2d_state_free)(&s);
  return 0;
}

/* LinearCheckpointer (src/utils/checkpointers.jl:1-107) specialised to the CD
 * model: checkpointed fields pcur (width 2), psi and xi; buffered field pcur
 * (acou_models.jl:159-173).  Checkpoint slot for step `it` exists iff
 * it % cf == 0 or (it+1) % cf == 0 with 0 <= it (+1) <= nt+1.               */
typedef struct {
  int nt, cf, last, curr;
  size_t N, npx, npz, nxx, nxz;
  real **ck_p;       /* indexed by it+1 (it in -1..nt+1); NULL if no slot  */
  real **ck_psi_x, **ck_psi_z, **ck_xi_x, **ck_xi_z; /* indexed by it/cf   */
  real **buf;        /* cf+1 copies of pcur                                 */
} SFX(cd2d_ckpt);

static int SFX(ck_has_p)(const SFX(cd2d_ckpt) *c, int it) {
  if (it < -1 || it > c->nt + 1) return 0;
  int a = (it >= 0 && it <= c->nt + 1 && it % c->cf == 0);
  int b = (it + 1 >= 0 && it + 1 <= c->nt + 1 && (it + 1) % c->cf == 0);
  return a || b;
}

static void SFX(cd2d_ckpt_init)(SFX(cd2d_ckpt) *c, const oracle_geom *g, int nt, int cf) {
  c->nt = nt; c->cf = cf;
  c->last = (nt / cf) * cf;   /* checkpointers.jl:19 */
  c->curr = c->last;
  c->N = (size_t)g->n[0] * g->n[1];
  c->npx = SFX(cd2d_nx_psi)(g); c->npz = SFX(cd2d_nz_psi)(g);
  c->nxx = SFX(cd2d_nx_xi)(g); c->nxz = SFX(cd2d_nz_xi)(g);
  c->ck_p = (real **)calloc(nt + 3, sizeof(real *));
  for (int it = -1; it <= nt + 1; it++)
    if (SFX(ck_has_p)(c, it)) c->ck_p[it + 1] = (real *)calloc(c->N, sizeof(real));
  int nk = (nt + 1) / cf + 1;
  c->ck_psi_x = (real **)calloc(nk, sizeof(real *));
  c->ck_psi_z = (real **)calloc(nk, sizeof(real *));
  c->ck_xi_x = (real **)calloc(nk, sizeof(real *));
  c->ck_xi_z = (real **)calloc(nk, sizeof(real *));
  for (int k = 0; k < nk; k++) {
    c->ck_psi_x[k] = (real *)calloc(c->npx + 1, sizeof(real));
    c->ck_psi_z[k] = (real *)calloc(c->npz + 1, sizeof(real));
    c->ck_xi_x[k] = (real *)calloc(c->nxx, sizeof(real));
    c->ck_xi_z[k] = (real *)calloc(c->nxz, sizeof(real));
  }
  c->buf = (real **)calloc(cf + 1, sizeof(real *));
  for (int k = 0; k <= cf; k++) c->buf[k] = (real *)calloc(c->N, sizeof(real));
}

static void SFX(cd2d_ckpt_free)(SFX(cd2d_ckpt) *c) {
  for (int it = -1; it <= c->nt + 1; it++) free(c->ck_p[it + 1]);
  free(c->ck_p);
  int nk = (c->nt + 1) / c->cf + 1;
  for (int k = 0; k < nk; k++) {
    free(c->ck_psi_x[k]); free(c->ck_psi_z[k]); free(c->ck_xi_x[k]); free(c->ck_xi_z[k]);
  }
  free(c->ck_psi_x); free(c->ck_psi_z); free(c->ck_xi_x); free(c->ck_xi_z);
  for (int k = 0; k <= c->cf; k++) free(c->buf[k]);
  free(c->buf);
}

/* savecheckpoint! (checkpointers.jl:44-56) of pcur (width 2) + psi, xi.
 * `it` = -1 saves only pcur (acou_models.jl:169).                          */
static void SFX(cd2d_ckpt_save)(SFX(cd2d_ckpt) *c, const SFX(cd2d_state) *s, int it) {
  size_t N = c->N;
  for (int itw = it; itw <= it + 1; itw++)
    if (itw % c->cf == 0) memcpy(c->ck_p[it + 1], s->p[1], N * sizeof(real));
  if (it >= 0 && it % c->cf == 0) {
    int k = it / c->cf;
    memcpy(c->ck_psi_x[k], s->psi_x, c->npx * sizeof(real));
    memcpy(c->ck_psi_z[k], s->psi_z, c->npz * sizeof(real));
    memcpy(c->ck_xi_x[k], s->xi_x, c->nxx * sizeof(real));
    memcpy(c->ck_xi_z[k], s->xi_z, c->nxz * sizeof(real));
  }
  if (it >= c->last) memcpy(c->buf[it - c->last], s->p[1], N * sizeof(real));
}

static int SFX(cd2d_issaved)(const SFX(cd2d_ckpt) *c, int it) {
  /* isbuffered || ischeckpointed (checkpointers.jl:58-78) */
  if (c->curr <= it && it <= c->curr + c->cf) return 1;
  return SFX(ck_has_p)(c, it);
}
static const real *SFX(cd2d_getsaved)(const SFX(cd2d_ckpt) *c, int it) {
  if (SFX(ck_has_p)(c, it)) return c->ck_p[it + 1];
  return c->buf[it - c->curr];
}

/* swgradient_1shot! for the CD model (src/models/acoustic/acou_gradient.jl:
 * 4-95), returning the raw grad_vp field before muting and the 2/vp^3
 * rescale.  `adjsrc` is -dchi/du (nt x nrec), not yet prescaled: the
 * prescale by fact at the receivers (acoustic2D_xPU.jl:64-76) is done here.
 * Traces of the forward pass are written to `traces` when non-NULL.         */
int SFX(oracle_cd2d_gradient)(const oracle_geom *g, const real *fact,
                              const SFX(cpml) *cx, const SFX(cpml) *cz,
                              int nt, int nsrc, const int64_t *src, const real *srctf,
                              int nrec, const int64_t *rec, const real *adjsrc,
                              int check_freq, real *grad, real *traces) {
  if (check_freq < 1 || check_freq >= nt) return -1; /* checkpointers.jl:17 */
  const int64_t nx = g->n[0];
  size_t N = (size_t)g->n[0] * g->n[1];
  SFX(cd2d_state) s, a;
  SFX(cd2d_state_alloc)(g, &s);
  SFX(cd2d_state_alloc)(g, &a);
  SFX(cd2d_ckpt) ck;
  SFX(cd2d_ckpt_init)(&ck, g, nt, check_freq);
  /* acou_models.jl:169-172: save the (zero) initial time levels */
  {
    SFX(cd2d_state) tmp = s;
    tmp.p[1] = s.p[0];
    SFX(cd2d_ckpt_save)(&ck, &tmp, -1);
    SFX(cd2d_ckpt_save)(&ck, &s, 0);
  }
  real *own_traces = traces ? traces : (real *)calloc((size_t)nt * nrec, sizeof(real));
  /* forward time loop (acou_gradient.jl:33-42) */
  for (int it = 1; it <= nt; it++) {
    SFX(cd2d_onestep)(g, fact, s.p[0], s.p[1], s.p[2], s.psi_x, s.psi_z, s.xi_x, s.xi_z,
                      cx, cz, nsrc, src, srctf, nt, it, nrec, rec, own_traces);
    SFX(cd2d_rotate)(&s);
    SFX(cd2d_ckpt_save)(&ck, &s, it);
    oracle_progress("cd2d gradient, forward", it, nt);
  }
  /* prescale_residuals! (acoustic2D_xPU.jl:64-76) */
  real *res = (real *)malloc((size_t)nt * nrec * sizeof(real));
  for (int r = 0; r < nrec; r++) {
    real f = A2(fact, nx, rec[r], rec[r + nrec]);
    for (int it = 0; it < nt; it++) res[it + (size_t)nt * r] = adjsrc[it + (size_t)nt * r] * f;
  }
  memset(grad, 0, N * sizeof(real));
  const real _dt2 = (real)(1.0 / (g->dt * g->dt));
  /* adjoint time loop (acou_gradient.jl:55-82) */
  for (int it = nt; it >= 1; it--) {
    SFX(cd2d_onestep)(g, fact, a.p[0], a.p[1], a.p[2], a.psi_x, a.psi_z, a.xi_x, a.xi_z,
                      cx, cz, nrec, rec, res, nt, it, 0, NULL, NULL);
    SFX(cd2d_rotate)(&a);
    oracle_progress("cd2d gradient, adjoint", nt - it + 1, nt);
    if (!SFX(cd2d_issaved)(&ck, it - 2)) {
      /* initrecover! (checkpointers.jl:87-94) */
      int old = ck.curr;
      ck.curr -= ck.cf;
      memcpy(ck.buf[0], SFX(cd2d_getsaved)(&ck, ck.curr), N * sizeof(real));
      memcpy(ck.buf[ck.cf], ck.ck_p[old + 1], N * sizeof(real));
      int k = ck.curr / ck.cf;
      memcpy(s.p[0], SFX(cd2d_getsaved)(&ck, ck.curr - 1), N * sizeof(real));
      memcpy(s.p[1], SFX(cd2d_getsaved)(&ck, ck.curr), N * sizeof(real));
      memcpy(s.psi_x, ck.ck_psi_x[k], ck.npx * sizeof(real));
      memcpy(s.psi_z, ck.ck_psi_z[k], ck.npz * sizeof(real));
      memcpy(s.xi_x, ck.ck_xi_x[k], ck.nxx * sizeof(real));
      memcpy(s.xi_z, ck.ck_xi_z[k], ck.nxz * sizeof(real));
      /* recover! (checkpointers.jl:96-105) */
      for (int rit = ck.curr + 1; rit <= ck.curr + ck.cf - 1; rit++) {
        SFX(cd2d_onestep)(g, fact, s.p[0], s.p[1], s.p[2], s.psi_x, s.psi_z, s.xi_x, s.xi_z,
                          cx, cz, nsrc, src, srctf, nt, rit, 0, NULL, NULL);
        SFX(cd2d_rotate)(&s);
        memcpy(ck.buf[rit - ck.curr], s.p[1], N * sizeof(real));
      }
    }
    const real *pa = SFX(cd2d_getsaved)(&ck, it - 2);
    const real *pb = SFX(cd2d_getsaved)(&ck, it - 1);
    const real *pc = SFX(cd2d_getsaved)(&ck, it);
    const real *adjcur = a.p[1];
    /* correlate_gradient! (correlate_gradient_xPU.jl:1-10) */
#pragma omp parallel for schedule(static)
    for (size_t q = 0; q < N; q++) grad[q] = grad[q] + adjcur[q] * (pa[q] - 2 * pb[q] + pc[q]) * _dt2;
  }
  free(res);
  if (!traces) free(own_traces);
  SFX(cd2d_ckpt_free)(&ck);
  SFX(cd2d_state_free)(&s);
  SFX(cd2d_state_free)(&a);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* 2D acoustic variable density (staggered pressure / velocity)             */
/* ------------------------------------------------------------------------ */

/* 4th-order staggered first derivative of fdgen.jl:65-135 with order=4,
 * deriv=1 (weights 1/24, -9/8, 9/8, -1/24 at offsets -1, 0, 1, 2 from the
 * base index `b`, terms outside 1..n dropped as the generated boundary
 * branches do).  `at(k)` reads the 1-based element k along the axis.        */
#define D4_TERM(k, w) (((k) >= 1 && (k) <= n) ? (w) * at(k) : (real)0)

/* derivative along x of A (n1 x n2, column-major) at (b, j) */
static real SFX(d4x)(const real *A, int64_t n1, int64_t n, int64_t b, int64_t j) {
#define at(k) A2(A, n1, (k), j)
  const real c1 = (real)(1.0 / 24.0), c2 = (real)(-9.0 / 8.0);
  if (b - 1 >= 1 && b + 2 <= n) return c1 * at(b - 1) + c2 * at(b) + (-c2) * at(b + 1) + (-c1) * at(b + 2);
  return D4_TERM(b - 1, c1) + D4_TERM(b, c2) + D4_TERM(b + 1, -c2) + D4_TERM(b + 2, -c1);
#undef at
}
/* derivative along z of A (n1 x n) at (i, b) */
static real SFX(d4z)(const real *A, int64_t n1, int64_t n, int64_t i, int64_t b) {
#define at(k) A2(A, n1, i, (k))
  const real c1 = (real)(1.0 / 24.0), c2 = (real)(-9.0 / 8.0);
  if (b - 1 >= 1 && b + 2 <= n) return c1 * at(b - 1) + c2 * at(b) + (-c2) * at(b + 1) + (-c1) * at(b + 2);
  return D4_TERM(b - 1, c1) + D4_TERM(b, c2) + D4_TERM(b + 1, -c2) + D4_TERM(b + 2, -c1);
#undef at
}

typedef struct {
  real *p, *vx, *vz;              /* p: nx x nz, vx: (nx-1) x nz, vz: nx x (nz-1) */
  real *psi_x, *psi_z, *xi_x, *xi_z;
} SFX(vd2d_state);

/* update_p_CPML! (acoustic2D_VD_xPU.jl:17-37): p -= m0 (d~x vx + d~z vz) */
static void SFX(vd2d_update_p)(const oracle_geom *g, const real *m0, SFX(vd2d_state) *s,
                               const SFX(cpml) *cx, const SFX(cpml) *cz) {
  const int64_t nx = g->n[0], nz = g->n[1];
  const int halo = g->halo;
  const int64_t nxx = 2 * (halo + 1);
  const real _dx = (real)(1.0 / g->h[0]), _dz = (real)(1.0 / g->h[1]);
#pragma omp parallel for schedule(static)
  for (int64_t j = 2; j <= nz - 1; j++) {
    for (int64_t i = 2; i <= nx - 1; i++) {
      /* @d~x(vx, a_x, b_x, xi_x, order=4, I=(i-1,j), halfgrid=false): idim = i */
      real dvx = SFX(d4x)(s->vx, nx - 1, nx - 1, i - 1, j) * _dx;
      if (i <= halo + 1) {
        A2(s->xi_x, nxx, i, j) = cx->b[i - 1] * A2(s->xi_x, nxx, i, j) + cx->a[i - 1] * dvx;
        dvx = dvx + A2(s->xi_x, nxx, i, j);
      } else if (i >= nx - halo) {
        int64_t k = i - (nx - halo) + 1 + (halo + 1);
        A2(s->xi_x, nxx, k, j) = cx->b[k - 1] * A2(s->xi_x, nxx, k, j) + cx->a[k - 1] * dvx;
        dvx = dvx + A2(s->xi_x, nxx, k, j);
      }
      real dvz = SFX(d4z)(s->vz, nx, nz - 1, i, j - 1) * _dz;
      if (j <= halo + 1) {
        A2(s->xi_z, nx, i, j) = cz->b[j - 1] * A2(s->xi_z, nx, i, j) + cz->a[j - 1] * dvz;
        dvz = dvz + A2(s->xi_z, nx, i, j);
      } else if (j >= nz - halo) {
        int64_t k = j - (nz - halo) + 1 + (halo + 1);
        A2(s->xi_z, nx, i, k) = cz->b[k - 1] * A2(s->xi_z, nx, i, k) + cz->a[k - 1] * dvz;
        dvz = dvz + A2(s->xi_z, nx, i, k);
      }
      A2(s->p, nx, i, j) -= A2(m0, nx, i, j) * (dvx + dvz);
    }
  }
}

/* update_vx_CPML! and update_vy_CPML! (acoustic2D_VD_xPU.jl:39-75) */
static void SFX(vd2d_update_v)(const oracle_geom *g, const real *m1x, const real *m1z, SFX(vd2d_state) *s,
                               const SFX(cpml) *cx, const SFX(cpml) *cz) {
  const int64_t nx = g->n[0], nz = g->n[1];
  const int halo = g->halo;
  const int64_t npx = 2 * halo;
  const real _dx = (real)(1.0 / g->h[0]), _dz = (real)(1.0 / g->h[1]);
#pragma omp parallel for schedule(static)
  for (int64_t j = 1; j <= nz; j++) {
    for (int64_t i = 1; i <= nx - 1; i++) {
      real d = SFX(d4x)(s->p, nx, nx, i, j) * _dx;
      if (i <= halo) {
        A2(s->psi_x, npx, i, j) = cx->b_h[i - 1] * A2(s->psi_x, npx, i, j) + cx->a_h[i - 1] * d;
        d = d + A2(s->psi_x, npx, i, j);
      } else if (i >= nx - halo) {
        int64_t k = i - (nx - halo) + 1 + halo;
        A2(s->psi_x, npx, k, j) = cx->b_h[k - 1] * A2(s->psi_x, npx, k, j) + cx->a_h[k - 1] * d;
        d = d + A2(s->psi_x, npx, k, j);
      }
      A2(s->vx, nx - 1, i, j) -= A2(m1x, nx - 1, i, j) * d;
    }
  }
#pragma omp parallel for schedule(static)
  for (int64_t j = 1; j <= nz - 1; j++) {
    for (int64_t i = 1; i <= nx; i++) {
      real d = SFX(d4z)(s->p, nx, nz, i, j) * _dz;
      if (j <= halo) {
        A2(s->psi_z, nx, i, j) = cz->b_h[j - 1] * A2(s->psi_z, nx, i, j) + cz->a_h[j - 1] * d;
        d = d + A2(s->psi_z, nx, i, j);
      } else if (j >= nz - halo) {
        int64_t k = j - (nz - halo) + 1 + halo;
        A2(s->psi_z, nx, i, k) = cz->b_h[k - 1] * A2(s->psi_z, nx, i, k) + cz->a_h[k - 1] * d;
        d = d + A2(s->psi_z, nx, i, k);
      }
      A2(s->vz, nx, i, j) -= A2(m1z, nx, i, j) * d;
    }
  }
}

static void SFX(vd2d_inject)(const oracle_geom *g, real *p, int n, const int64_t *idx, const real *tf, int nt,
                             int it, const real *scale_field) {
  const int64_t nx = g->n[0];
  for (int s = 0; s < n; s++) {
    int64_t i = idx[s], j = idx[s + n];
    real v = tf[(it - 1) + (size_t)nt * s];
    if (scale_field) v = v * A2(scale_field, nx, i, j);
    A2(p, nx, i, j) += v;
  }
}

/* forward_onestep_CPML! (acoustic2D_VD_xPU.jl:91-137) */
static void SFX(vd2d_fwd_step)(const oracle_geom *g, const real *m0, const real *m1x, const real *m1z,
                               SFX(vd2d_state) *s, const SFX(cpml) *cx, const SFX(cpml) *cz, int nsrc,
                               const int64_t *src, const real *srctf, int nt, int it, int nrec,
                               const int64_t *rec, real *traces) {
  const int64_t nx = g->n[0];
  SFX(vd2d_update_p)(g, m0, s, cx, cz);
  SFX(vd2d_inject)(g, s->p, nsrc, src, srctf, nt, it, NULL);
  SFX(vd2d_update_v)(g, m1x, m1z, s, cx, cz);
  if (traces)
    for (int r = 0; r < nrec; r++) traces[(it - 1) + (size_t)nt * r] = A2(s->p, nx, rec[r], rec[r + nrec]);
}

/* adjoint_onestep_CPML! (acoustic2D_VD_xPU.jl:139-178); the residual is
 * prescaled by fact_m0 at the receivers (prescale_residuals!, :77-89) */
static void SFX(vd2d_adj_step)(const oracle_geom *g, const real *m0, const real *m1x, const real *m1z,
                               SFX(vd2d_state) *s, const SFX(cpml) *cx, const SFX(cpml) *cz, int nrec,
                               const int64_t *rec, const real *res, int nt, int it) {
  SFX(vd2d_update_v)(g, m1x, m1z, s, cx, cz);
  SFX(vd2d_update_p)(g, m0, s, cx, cz);
  SFX(vd2d_inject)(g, s->p, nrec, rec, res, nt, it, NULL);
}

static void SFX(vd2d_state_alloc)(const oracle_geom *g, SFX(vd2d_state) *s) {
  const int64_t nx = g->n[0], nz = g->n[1];
  const int h = g->halo;
  s->p = (real *)calloc((size_t)nx * nz, sizeof(real));
  s->vx = (real *)calloc((size_t)(nx - 1) * nz, sizeof(real));
  s->vz = (real *)calloc((size_t)nx * (nz - 1), sizeof(real));
  s->psi_x = (real *)calloc((size_t)(2 * h) * nz + 1, sizeof(real));
  s->psi_z = (real *)calloc((size_t)nx * (2 * h) + 1, sizeof(real));
  s->xi_x = (real *)calloc((size_t)(2 * h + 2) * nz, sizeof(real));
  s->xi_z = (real *)calloc((size_t)nx * (2 * h + 2), sizeof(real));
}
static void SFX(vd2d_state_free)(SFX(vd2d_state) *s) {
  free(s->p); free(s->vx); free(s->vz); free(s->psi_x); free(s->psi_z); free(s->xi_x); free(s->xi_z);
}
static void SFX(vd2d_state_copy)(const oracle_geom *g, SFX(vd2d_state) *dst, const SFX(vd2d_state) *src) {
  const int64_t nx = g->n[0], nz = g->n[1];
  const int h = g->halo;
  memcpy(dst->p, src->p, (size_t)nx * nz * sizeof(real));
  memcpy(dst->vx, src->vx, (size_t)(nx - 1) * nz * sizeof(real));
  memcpy(dst->vz, src->vz, (size_t)nx * (nz - 1) * sizeof(real));
  memcpy(dst->psi_x, src->psi_x, ((size_t)(2 * h) * nz + 1) * sizeof(real));
  memcpy(dst->psi_z, src->psi_z, ((size_t)nx * (2 * h) + 1) * sizeof(real));
  memcpy(dst->xi_x, src->xi_x, (size_t)(2 * h + 2) * nz * sizeof(real));
  memcpy(dst->xi_z, src->xi_z, (size_t)nx * (2 * h + 2) * sizeof(real));
}

/* swforward_1shot! for VD (acou_forward.jl:83-125) */
int SFX(oracle_vd2d_forward)(const oracle_geom *g, const real *m0, const real *m1x, const real *m1z,
                             const SFX(cpml) *cx, const SFX(cpml) *cz, int nt, int nsrc, const int64_t *src,
                             const real *srctf, int nrec, const int64_t *rec, real *traces) {
  SFX(vd2d_state) s;
  SFX(vd2d_state_alloc)(g, &s);
  for (int it = 1; it <= nt; it++) {
    SFX(vd2d_fwd_step)(g, m0, m1x, m1z, &s, cx, cz, nsrc, src, srctf, nt, it, nrec, rec, traces);
    if (snap_due(it)) {   /* "pcur", then "vcur" = [vx, vz] (acou_forward.jl:116-120) */
      const int64_t nx = g->n[0], nz = g->n[1];
      snap_put(s.p, (size_t)nx * nz * sizeof(real));
      snap_put(s.vx, (size_t)(nx - 1) * nz * sizeof(real));
      snap_put(s.vz, (size_t)nx * (nz - 1) * sizeof(real));
    }
    oracle_progress("vd2d forward", it, nt);
  }
  SFX(vd2d_state_free)(&s);
  return 0;
}

/* swgradient_1shot! for VD (acou_gradient.jl:97-184) up to the raw
 * grad_m0 and grad_m1_stag fields, with the LinearCheckpointer of
 * acou_models.jl:409-423 (pcur width 1, buffered pcur; pcur, vcur, psi, xi
 * checkpointed every cf steps). */
int SFX(oracle_vd2d_gradient)(const oracle_geom *g, const real *m0, const real *m1x, const real *m1z,
                              const SFX(cpml) *cx, const SFX(cpml) *cz, int nt, int nsrc, const int64_t *src,
                              const real *srctf, int nrec, const int64_t *rec, const real *adjsrc, int cf,
                              real *g_m0, real *g_m1x, real *g_m1z, real *traces) {
  if (cf < 1 || cf >= nt) return -1;
  const int64_t nx = g->n[0], nz = g->n[1];
  const size_t N = (size_t)nx * nz;
  SFX(vd2d_state) s, a;
  SFX(vd2d_state_alloc)(g, &s);
  SFX(vd2d_state_alloc)(g, &a);
  const int last = (nt / cf) * cf;
  int curr = last;
  const int nk = (nt + 1) / cf + 1;
  SFX(vd2d_state) *ck = (SFX(vd2d_state) *)calloc(nk, sizeof(SFX(vd2d_state)));
  for (int k = 0; k < nk; k++) SFX(vd2d_state_alloc)(g, &ck[k]);   /* ck[0] = zero state */
  real **buf = (real **)calloc(cf + 1, sizeof(real *));
  for (int k = 0; k <= cf; k++) buf[k] = (real *)calloc(N, sizeof(real));
  real *own_tr = traces ? traces : (real *)calloc((size_t)nt * nrec, sizeof(real));
  for (int it = 1; it <= nt; it++) {
    SFX(vd2d_fwd_step)(g, m0, m1x, m1z, &s, cx, cz, nsrc, src, srctf, nt, it, nrec, rec, own_tr);
    if (it % cf == 0) SFX(vd2d_state_copy)(g, &ck[it / cf], &s);
    if (it >= last) memcpy(buf[it - last], s.p, N * sizeof(real));
    oracle_progress("vd2d gradient, forward", it, nt);
  }
  real *res = (real *)malloc((size_t)nt * nrec * sizeof(real));
  for (int r = 0; r < nrec; r++) {
    real f = A2(m0, nx, rec[r], rec[r + nrec]);
    for (int it = 0; it < nt; it++) res[it + (size_t)nt * r] = adjsrc[it + (size_t)nt * r] * f;
  }
  memset(g_m0, 0, N * sizeof(real));
  memset(g_m1x, 0, (size_t)(nx - 1) * nz * sizeof(real));
  memset(g_m1z, 0, (size_t)nx * (nz - 1) * sizeof(real));
  const real _dt = (real)(1.0 / g->dt);
  const real _dx = (real)(1.0 / g->h[0]), _dz = (real)(1.0 / g->h[1]);
#define VD_SAVED(t) (((t) % cf == 0) ? ck[(t) / cf].p : buf[(t) - curr])
  for (int it = nt; it >= 1; it--) {
    SFX(vd2d_adj_step)(g, m0, m1x, m1z, &a, cx, cz, nrec, rec, res, nt, it);
    oracle_progress("vd2d gradient, adjoint", nt - it + 1, nt);
    int t = it - 1;
    int saved = (curr <= t && t <= curr + cf) || (t % cf == 0);
    if (!saved) {
      int old = curr;
      curr -= cf;
      memcpy(buf[0], ck[curr / cf].p, N * sizeof(real));
      memcpy(buf[cf], ck[old / cf].p, N * sizeof(real));
      SFX(vd2d_state_copy)(g, &s, &ck[curr / cf]);
      for (int rit = curr + 1; rit <= curr + cf - 1; rit++) {
        SFX(vd2d_fwd_step)(g, m0, m1x, m1z, &s, cx, cz, nsrc, src, srctf, nt, rit, 0, NULL, NULL);
        memcpy(buf[rit - curr], s.p, N * sizeof(real));
      }
    }
    const real *pc = VD_SAVED(it);
    const real *po = VD_SAVED(it - 1);
    /* correlate_gradient_m0! (correlate_gradient_xPU.jl:12-21) */
#pragma omp parallel for schedule(static)
    for (size_t q = 0; q < N; q++) g_m0[q] = g_m0[q] - (a.p[q] * (pc[q] - po[q]) * _dt);
    /* correlate_gradient_m1! (acoustic2D_VD_xPU.jl:180-198) */
#pragma omp parallel for schedule(static)
    for (int64_t j = 1; j <= nz; j++)
      for (int64_t i = 1; i <= nx - 1; i++)
        A2(g_m1x, nx - 1, i, j) = A2(g_m1x, nx - 1, i, j) + A2(a.vx, nx - 1, i, j) * (SFX(d4x)(pc, nx, nx, i, j) * _dx);
#pragma omp parallel for schedule(static)
    for (int64_t j = 1; j <= nz - 1; j++)
      for (int64_t i = 1; i <= nx; i++)
        A2(g_m1z, nx, i, j) = A2(g_m1z, nx, i, j) + A2(a.vz, nx, i, j) * (SFX(d4z)(pc, nx, nz, i, j) * _dz);
  }
#undef VD_SAVED
  free(res);
  if (!traces) free(own_tr);
  for (int k = 0; k <= cf; k++) free(buf[k]);
  free(buf);
  for (int k = 0; k < nk; k++) SFX(vd2d_state_free)(&ck[k]);
  free(ck);
  SFX(vd2d_state_free)(&s);
  SFX(vd2d_state_free)(&a);
  return 0;
}
