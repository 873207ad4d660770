/*
 * CPU restatement of the 2D isotropic elastic path of SeismicWaves.jl
 * (TEST INFRASTRUCTURE: the checker, never the product path).  Included by
 * swx_oracle.c once per precision.  1-based loops as in the reference.
 *
 * Reference: src/models/elastic/backends/shared/elastic2D_iso_xPU.jl:1-448
 * (kernels and steps), freesurface_derivatives_4th_mirror.jl:1-234
 * (derivatives and free-surface mirror rules), fdgenerated.jl:178-199
 * (C-PML wrapper), correlate_gradient_xPU.jl:1-82 (elastic correlation),
 * src/models/elastic/ela_gradient.jl (driver), checkpointers.jl.
 *
 * Fields: sxx, szz (nx x nz), sxz (nx-1 x nz-1), ux (nx-1 x nz), uz (nx x nz-1).
 * Material: lam, mu (nx x nz), rho_ih (nx-1 x nz), rho_jh (nx x nz-1),
 * mu_ihjh (nx-1 x nz-1).  C-PML memory variables (psi):
 *   0 dsxx/dx  (2h x nz)      1 dsxz/dx  (2h+2 x nz-1)
 *   2 dszz/dz  (nx x 2h)      3 dsxz/dz  (nx-1 x 2h+2)
 *   4 dux/dx   (2h+2 x nz)    5 duz/dx   (2h x nz-1)
 *   6 dux/dz   (nx-1 x 2h)    7 duz/dz   (nx x 2h+2)
 * Points: each source/receiver is a list of taps (1-based i, j, coefficient)
 * per staggered field, as built by spread_positions (utils.jl:168-214).
 */

#define EL_A(arr, n1, i, j) (arr)[((i) - 1) + (size_t)(n1) * ((j) - 1)]

/* ORACLE_PLANT (tests only, oracle/Makefile target `planted`) builds a copy
 * of the oracle with ONE free-surface mirror sign flipped, to show that the
 * half-space check of the free surface fails on such a mistake:
 *   1: the sigma_zz mirror of d_szz_dz at j = 1
 *   2: the sigma_xz mirror of d_sxz_dz at j = 2
 *   3: the ux mirror of d_ux_dz at j = 1
 *   4: the uz mirror of d_uz_dz at j = 2                                    */
#ifndef ORACLE_PLANT
#define ORACLE_PLANT 0
#endif
#define EL_FLIP(k) (ORACLE_PLANT == (k) ? -1 : 1)

static inline real SFX(el_inner)(real f1, real f2, real f3, real f4, real d) {
  /* ∂x4th_inner (freesurface_derivatives_4th_mirror.jl:2-7) */
  return ((real)(1.0 / 24.0) * f1 - (real)(27.0 / 24.0) * f2 + (real)(27.0 / 24.0) * f3 - (real)(1.0 / 24.0) * f4) * d;
}

typedef struct {
  int64_t nx, nz;
  int halo, freetop;
  real dx_, dz_;                  /* 1/dx, 1/dz */
} SFX(el_geom);

/* --- derivative wrappers (freesurface_derivatives_4th_mirror.jl:72-234) --- */
static real SFX(d_sxx_dx)(const SFX(el_geom) *g, const real *s, int64_t i, int64_t j) {
  const int64_t nx = g->nx;
#define S(a) EL_A(s, nx, (a), j)
  if (i == 1) return SFX(el_inner)(0, S(i), S(i + 1), S(i + 2), g->dx_);
  if (i == nx - 1) return SFX(el_inner)(S(i - 1), S(i), S(i + 1), 0, g->dx_);
  return SFX(el_inner)(S(i - 1), S(i), S(i + 1), S(i + 2), g->dx_);
#undef S
}
static real SFX(d_szz_dz)(const SFX(el_geom) *g, const real *s, int64_t i, int64_t j) {
  const int64_t nx = g->nx, nz = g->nz;
#define S(b) EL_A(s, nx, i, (b))
  if (j == 1) {
    if (g->freetop) return SFX(el_inner)(-EL_FLIP(1) * S(j + 1), S(j), S(j + 1), S(j + 2), g->dz_);
    return SFX(el_inner)(0, S(j), S(j + 1), S(j + 2), g->dz_);
  }
  if (j == nz - 1) return SFX(el_inner)(S(j - 1), S(j), S(j + 1), 0, g->dz_);
  return SFX(el_inner)(S(j - 1), S(j), S(j + 1), S(j + 2), g->dz_);
#undef S
}
static real SFX(d_sxz_dx)(const SFX(el_geom) *g, const real *s, int64_t i, int64_t j) {
  const int64_t nx = g->nx, n1 = g->nx - 1;
#define S(a) EL_A(s, n1, (a), j)
  if (i == 1) return SFX(el_inner)(0, 0, S(i), S(i + 1), g->dx_);
  if (i == 2) return SFX(el_inner)(0, S(i - 1), S(i), S(i + 1), g->dx_);
  if (i == nx - 1) return SFX(el_inner)(S(i - 2), S(i - 1), S(i), 0, g->dx_);
  if (i == nx) return SFX(el_inner)(S(i - 2), S(i - 1), 0, 0, g->dx_);
  return SFX(el_inner)(S(i - 2), S(i - 1), S(i), S(i + 1), g->dx_);
#undef S
}
static real SFX(d_sxz_dz)(const SFX(el_geom) *g, const real *s, int64_t i, int64_t j) {
  const int64_t nz = g->nz, n1 = g->nx - 1;
#define S(b) EL_A(s, n1, i, (b))
  if (j == 1) {
    if (g->freetop) return SFX(el_inner)(-S(j + 1), -S(j), S(j), S(j + 1), g->dz_);
    return SFX(el_inner)(0, 0, S(j), S(j + 1), g->dz_);
  }
  if (j == 2) {
    if (g->freetop) return SFX(el_inner)(-EL_FLIP(2) * S(j - 1), S(j - 1), S(j), S(j + 1), g->dz_);
    return SFX(el_inner)(0, S(j - 1), S(j), S(j + 1), g->dz_);
  }
  if (j == nz - 1) return SFX(el_inner)(S(j - 2), S(j - 1), S(j), 0, g->dz_);
  if (j == nz) return SFX(el_inner)(S(j - 2), S(j - 1), 0, 0, g->dz_);
  return SFX(el_inner)(S(j - 2), S(j - 1), S(j), S(j + 1), g->dz_);
#undef S
}
static real SFX(d_ux_dx)(const SFX(el_geom) *g, const real *u, int64_t i, int64_t j) {
  const int64_t nx = g->nx, n1 = g->nx - 1;
#define U(a) EL_A(u, n1, (a), j)
  if (i == 1) return SFX(el_inner)(0, 0, U(i), U(i + 1), g->dx_);
  if (i == 2) return SFX(el_inner)(0, U(i - 1), U(i), U(i + 1), g->dx_);
  if (i == nx - 1) return SFX(el_inner)(U(i - 2), U(i - 1), U(i), 0, g->dx_);
  if (i == nx) return SFX(el_inner)(U(i - 2), U(i - 1), 0, 0, g->dx_);
  return SFX(el_inner)(U(i - 2), U(i - 1), U(i), U(i + 1), g->dx_);
#undef U
}
static real SFX(d_uz_dz)(const SFX(el_geom) *g, const real *ux, const real *uz, const real *lam, const real *mu,
                         int64_t i, int64_t j) {
  const int64_t nx = g->nx, nz = g->nz;
#define U(b) EL_A(uz, nx, i, (b))
  if (j == 1) {
    if (g->freetop) {
      /* sigma_zz = 0 at the free surface: duz/dz = -lambda/(lambda+2mu) dux/dx */
      const real dux = SFX(d_ux_dx)(g, ux, i, j);
      return (-EL_A(lam, nx, i, j) / (EL_A(lam, nx, i, j) + 2 * EL_A(mu, nx, i, j))) * dux;
    }
    return SFX(el_inner)(0, 0, U(j), U(j + 1), g->dz_);
  }
  if (j == 2) {
    if (g->freetop) return SFX(el_inner)(EL_FLIP(4) * U(j - 1), U(j - 1), U(j), U(j + 1), g->dz_);
    return SFX(el_inner)(0, U(j - 1), U(j), U(j + 1), g->dz_);
  }
  if (j == nz - 1) return SFX(el_inner)(U(j - 2), U(j - 1), U(j), 0, g->dz_);
  if (j == nz) return SFX(el_inner)(U(j - 2), U(j - 1), 0, 0, g->dz_);
  return SFX(el_inner)(U(j - 2), U(j - 1), U(j), U(j + 1), g->dz_);
#undef U
}
static real SFX(d_ux_dz)(const SFX(el_geom) *g, const real *u, int64_t i, int64_t j) {
  const int64_t nz = g->nz, n1 = g->nx - 1;
#define U(b) EL_A(u, n1, i, (b))
  if (j == 1) {
    if (g->freetop) return SFX(el_inner)(EL_FLIP(3) * U(j + 1), U(j), U(j + 1), U(j + 2), g->dz_);
    return SFX(el_inner)(0, U(j), U(j + 1), U(j + 2), g->dz_);
  }
  if (j == nz - 1) return SFX(el_inner)(U(j - 1), U(j), U(j + 1), 0, g->dz_);
  return SFX(el_inner)(U(j - 1), U(j), U(j + 1), U(j + 2), g->dz_);
#undef U
}
static real SFX(d_uz_dx)(const SFX(el_geom) *g, const real *u, int64_t i, int64_t j) {
  const int64_t nx = g->nx;
#define U(a) EL_A(u, nx, (a), j)
  if (i == 1) return SFX(el_inner)(0, U(i), U(i + 1), U(i + 2), g->dx_);
  if (i == nx - 1) return SFX(el_inner)(U(i - 1), U(i), U(i + 1), 0, g->dx_);
  return SFX(el_inner)(U(i - 1), U(i), U(i + 1), U(i + 2), g->dx_);
#undef U
}

/* ∂̃4th (fdgenerated.jl:178-199): C-PML memory update along `dir` (0 x, 1 z)
 * of a field whose size along dir is `ndim`; `I` is the base index along
 * dir, `other` the index along the other axis; psi has `pn1` rows.          */
static real SFX(el_cpml)(const SFX(el_geom) *g, real d, const real *a, const real *b, real *psi, int64_t pn1,
                         int dir, int64_t I, int64_t other, int64_t ndim, int half) {
  const int h = g->halo;
  const int plusone = half ? 1 : 0;
  const int64_t idim = I + plusone;
  const int64_t iidim = I - (ndim - h) + 1 + (h + plusone);
  int64_t k;
  if (idim <= h + plusone) k = idim;
  else if (idim >= ndim - h) k = iidim;
  else return d;
  real *p = dir == 0 ? &EL_A(psi, pn1, k, other) : &EL_A(psi, pn1, other, k);
  *p = b[k - 1] * *p + a[k - 1] * d;
  return d + *p;
}

typedef struct {
  real *ux[3], *uz[3];            /* uold, ucur, unew (rotated)                */
  real *sxx, *szz, *sxz;
  real *psi[8];
} SFX(el_state);

typedef struct {
  const SFX(cpml) *cx, *cz;
  const real *lam, *mu, *rho_ih, *rho_jh, *mu_ihjh;
  real dt;
} SFX(el_model);

/* one point family: npts points, tap lists concatenated (1-based ij, coefs) */
typedef struct {
  int npts;
  const int *ptr;                 /* npts+1 */
  const int64_t *ij;              /* ntaps x 2, column-major (i then j)       */
  int64_t ntaps;
  const real *coef;
} SFX(el_taps);

static size_t SFX(el_psi_size)(const SFX(el_geom) *g, int k) {
  const int64_t nx = g->nx, nz = g->nz, h = g->halo;
  switch (k) {
    case 0: return (size_t)(2 * h) * nz;
    case 1: return (size_t)(2 * h + 2) * (nz - 1);
    case 2: return (size_t)nx * (2 * h);
    case 3: return (size_t)(nx - 1) * (2 * h + 2);
    case 4: return (size_t)(2 * h + 2) * nz;
    case 5: return (size_t)(2 * h) * (nz - 1);
    case 6: return (size_t)(nx - 1) * (2 * h);
    default: return (size_t)nx * (2 * h + 2);
  }
}

static void SFX(el_state_alloc)(const SFX(el_geom) *g, SFX(el_state) *s) {
  const int64_t nx = g->nx, nz = g->nz;
  for (int k = 0; k < 3; k++) {
    s->ux[k] = (real *)calloc((size_t)(nx - 1) * nz, sizeof(real));
    s->uz[k] = (real *)calloc((size_t)nx * (nz - 1), sizeof(real));
  }
  s->sxx = (real *)calloc((size_t)nx * nz, sizeof(real));
  s->szz = (real *)calloc((size_t)nx * nz, sizeof(real));
  s->sxz = (real *)calloc((size_t)(nx - 1) * (nz - 1), sizeof(real));
  for (int k = 0; k < 8; k++) s->psi[k] = (real *)calloc(SFX(el_psi_size)(g, k) + 1, sizeof(real));
}
static void SFX(el_state_free)(SFX(el_state) *s) {
  for (int k = 0; k < 3; k++) { free(s->ux[k]); free(s->uz[k]); }
  free(s->sxx); free(s->szz); free(s->sxz);
  for (int k = 0; k < 8; k++) free(s->psi[k]);
}
static void SFX(el_rotate)(SFX(el_state) *s) {
  real *o = s->ux[0]; s->ux[0] = s->ux[1]; s->ux[1] = s->ux[2]; s->ux[2] = o;
  o = s->uz[0]; s->uz[0] = s->uz[1]; s->uz[1] = s->uz[2]; s->uz[2] = o;
}

/* update_σxx_σzz! and update_σxz! (elastic2D_iso_xPU.jl:39-79) from ucur */
static void SFX(el_update_stress)(const SFX(el_geom) *g, const SFX(el_model) *m, SFX(el_state) *s) {
  const int64_t nx = g->nx, nz = g->nz, h = g->halo;
  const real *ux = s->ux[1], *uz = s->uz[1];
  const int64_t j0 = g->freetop ? 1 : 2;
#pragma omp parallel for schedule(static)
  for (int64_t j = j0; j <= nz - 1; j++)
    for (int64_t i = 2; i <= nx - 1; i++) {
      const real dux = SFX(d_ux_dx)(g, ux, i, j);
      const real duz = SFX(d_uz_dz)(g, ux, uz, m->lam, m->mu, i, j);
      const real duxc = SFX(el_cpml)(g, dux, m->cx->a, m->cx->b, s->psi[4], 2 * h + 2, 0, i - 1, j, nx - 1, 1);
      const real duzc = SFX(el_cpml)(g, duz, m->cz->a, m->cz->b, s->psi[7], nx, 1, j - 1, i, nz - 1, 1);
      const real l = EL_A(m->lam, nx, i, j), u = EL_A(m->mu, nx, i, j);
      EL_A(s->sxx, nx, i, j) = (l + 2 * u) * duxc + l * duzc;
      if (j == 1) EL_A(s->szz, nx, i, j) = 0;
      else EL_A(s->szz, nx, i, j) = l * duxc + (l + 2 * u) * duzc;
    }
#pragma omp parallel for schedule(static)
  for (int64_t j = 1; j <= nz - 1; j++)
    for (int64_t i = 1; i <= nx - 1; i++) {
      const real duzx = SFX(d_uz_dx)(g, uz, i, j);
      const real duxz = SFX(d_ux_dz)(g, ux, i, j);
      const real a1 = SFX(el_cpml)(g, duzx, m->cx->a_h, m->cx->b_h, s->psi[5], 2 * h, 0, i, j, nx, 0);
      const real a2 = SFX(el_cpml)(g, duxz, m->cz->a_h, m->cz->b_h, s->psi[6], nx - 1, 1, j, i, nz, 0);
      EL_A(s->sxz, nx - 1, i, j) = EL_A(m->mu_ihjh, nx - 1, i, j) * (a1 + a2);
    }
}

/* update_ux! and update_uz! (elastic2D_iso_xPU.jl:1-37) into unew */
static void SFX(el_update_disp)(const SFX(el_geom) *g, const SFX(el_model) *m, SFX(el_state) *s) {
  const int64_t nx = g->nx, nz = g->nz, h = g->halo;
  const real dt2 = m->dt * m->dt;
#pragma omp parallel for schedule(static)
  for (int64_t j = 1; j <= nz; j++)
    for (int64_t i = 1; i <= nx - 1; i++) {
      const real d1 = SFX(d_sxx_dx)(g, s->sxx, i, j);
      const real d2 = SFX(d_sxz_dz)(g, s->sxz, i, j);
      const real c1 = SFX(el_cpml)(g, d1, m->cx->a_h, m->cx->b_h, s->psi[0], 2 * h, 0, i, j, nx, 0);
      const real c2 = SFX(el_cpml)(g, d2, m->cz->a, m->cz->b, s->psi[3], nx - 1, 1, j - 1, i, nz - 1, 1);
      EL_A(s->ux[2], nx - 1, i, j) = 2 * EL_A(s->ux[1], nx - 1, i, j) - EL_A(s->ux[0], nx - 1, i, j) +
                                     dt2 / EL_A(m->rho_ih, nx - 1, i, j) * (c1 + c2);
    }
#pragma omp parallel for schedule(static)
  for (int64_t j = 1; j <= nz - 1; j++)
    for (int64_t i = 1; i <= nx; i++) {
      const real d1 = SFX(d_sxz_dx)(g, s->sxz, i, j);
      const real d2 = SFX(d_szz_dz)(g, s->szz, i, j);
      const real c1 = SFX(el_cpml)(g, d1, m->cx->a, m->cx->b, s->psi[1], 2 * h + 2, 0, i - 1, j, nx - 1, 1);
      const real c2 = SFX(el_cpml)(g, d2, m->cz->a_h, m->cz->b_h, s->psi[2], nx, 1, j, i, nz, 0);
      EL_A(s->uz[2], nx, i, j) = 2 * EL_A(s->uz[1], nx, i, j) - EL_A(s->uz[0], nx, i, j) +
                                 dt2 / EL_A(m->rho_jh, nx, i, j) * (c1 + c2);
    }
}

/* inject_external_sources2D_{ux,uz}! (elastic2D_iso_xPU.jl:96-106):
 * u += coef * tf[it, comp, s] / rho * dt^2, sources in order */
static void SFX(el_inject_force)(const SFX(el_geom) *g, const SFX(el_model) *m, SFX(el_state) *s,
                                 const SFX(el_taps) *tx, const SFX(el_taps) *tz, const real *tf, int nt, int it) {
  const int64_t nx = g->nx;
  const real dt2 = m->dt * m->dt;
  for (int p = 0; p < tx->npts; p++)
    for (int q = tx->ptr[p]; q < tx->ptr[p + 1]; q++) {
      int64_t i = tx->ij[q], j = tx->ij[q + tx->ntaps];
      EL_A(s->ux[2], nx - 1, i, j) += tx->coef[q] * tf[(it - 1) + (size_t)nt * (0 + 2 * (size_t)p)] /
                                      EL_A(m->rho_ih, nx - 1, i, j) * dt2;
    }
  for (int p = 0; p < tz->npts; p++)
    for (int q = tz->ptr[p]; q < tz->ptr[p + 1]; q++) {
      int64_t i = tz->ij[q], j = tz->ij[q + tz->ntaps];
      EL_A(s->uz[2], nx, i, j) += tz->coef[q] * tf[(it - 1) + (size_t)nt * (1 + 2 * (size_t)p)] /
                                  EL_A(m->rho_jh, nx, i, j) * dt2;
    }
}

/* record_receivers2D_{ux,uz}! + reduction (elastic2D_iso_xPU.jl:108-118, :216-231) */
static void SFX(el_record)(const SFX(el_geom) *g, const SFX(el_state) *s, const SFX(el_taps) *rx,
                           const SFX(el_taps) *rz, real *traces, int nt, int it) {
  const int64_t nx = g->nx;
  for (int r = 0; r < rx->npts; r++) {
    real acc = 0;
    for (int q = rx->ptr[r]; q < rx->ptr[r + 1]; q++)
      acc += rx->coef[q] * EL_A(s->ux[2], nx - 1, rx->ij[q], rx->ij[q + rx->ntaps]);
    traces[(it - 1) + (size_t)nt * (0 + 2 * (size_t)r)] = acc;
    acc = 0;
    for (int q = rz->ptr[r]; q < rz->ptr[r + 1]; q++)
      acc += rz->coef[q] * EL_A(s->uz[2], nx, rz->ij[q], rz->ij[q + rz->ntaps]);
    traces[(it - 1) + (size_t)nt * (1 + 2 * (size_t)r)] = acc;
  }
}

/* forward_onestep_CPML! (MT: :120-239, EF: :241-357).  kind 0 = moment
 * tensor (taps sxx/szz in s1, sxz in s2, M[3*s..] = Mxx, Mzz, Mxz, tf nt x nsrc),
 * kind 1 = external force (taps ux in s1, uz in s2, tf nt x 2 x nsrc).      */
static void SFX(el_fwd_step)(const SFX(el_geom) *g, const SFX(el_model) *m, SFX(el_state) *s, int kind,
                             const SFX(el_taps) *s1, const SFX(el_taps) *s2, const real *M, const real *tf, int nt,
                             int it, const SFX(el_taps) *rx, const SFX(el_taps) *rz, real *traces) {
  const int64_t nx = g->nx;
  SFX(el_update_stress)(g, m, s);
  if (kind == 0) {
    /* inject_momten_sources2D_* (:81-94), sources in order */
    for (int p = 0; p < s1->npts; p++) {
      const real f = tf[(it - 1) + (size_t)nt * p];
      for (int q = s1->ptr[p]; q < s1->ptr[p + 1]; q++) {
        int64_t i = s1->ij[q], j = s1->ij[q + s1->ntaps];
        EL_A(s->sxx, nx, i, j) += M[3 * p + 0] * s1->coef[q] * f;
        EL_A(s->szz, nx, i, j) += M[3 * p + 1] * s1->coef[q] * f;
      }
      for (int q = s2->ptr[p]; q < s2->ptr[p + 1]; q++) {
        int64_t i = s2->ij[q], j = s2->ij[q + s2->ntaps];
        EL_A(s->sxz, nx - 1, i, j) += M[3 * p + 2] * s2->coef[q] * f;
      }
    }
  }
  SFX(el_update_disp)(g, m, s);
  if (kind == 1) SFX(el_inject_force)(g, m, s, s1, s2, tf, nt, it);
  if (traces) SFX(el_record)(g, s, rx, rz, traces, nt, it);
  SFX(el_rotate)(s);
}

int SFX(oracle_el2d_forward)(const oracle_geom *og, const real *lam, const real *mu, const real *rho_ih,
                             const real *rho_jh, const real *mu_ihjh, const SFX(cpml) *cx, const SFX(cpml) *cz,
                             int nt, int kind, const SFX(el_taps) *s1, const SFX(el_taps) *s2, const real *M,
                             const real *tf, const SFX(el_taps) *rx, const SFX(el_taps) *rz, real *traces) {
  SFX(el_geom) g = {og->n[0], og->n[1], og->halo, og->freetop, (real)(1.0 / og->h[0]), (real)(1.0 / og->h[1])};
  SFX(el_model) m = {cx, cz, lam, mu, rho_ih, rho_jh, mu_ihjh, (real)og->dt};
  SFX(el_state) s;
  SFX(el_state_alloc)(&g, &s);
  for (int it = 1; it <= nt; it++) {
    SFX(el_fwd_step)(&g, &m, &s, kind, s1, s2, M, tf, nt, it, rx, rz, traces);
    if (snap_due(it)) {   /* "ucur" = [ux, uz], "σ" = [σxx, σzz, σxz] (ela_forward.jl:74-77) */
      const int64_t nx = g.nx, nz = g.nz;
      snap_put(s.ux[1], (size_t)(nx - 1) * nz * sizeof(real));
      snap_put(s.uz[1], (size_t)nx * (nz - 1) * sizeof(real));
      snap_put(s.sxx, (size_t)nx * nz * sizeof(real));
      snap_put(s.szz, (size_t)nx * nz * sizeof(real));
      snap_put(s.sxz, (size_t)(nx - 1) * (nz - 1) * sizeof(real));
    }
    oracle_progress("elastic forward", it, nt);
  }
  SFX(el_state_free)(&s);
  return 0;
}

/* correlate_gradients! (correlate_gradient_xPU.jl(elastic):1-82) */
static void SFX(el_correlate)(const SFX(el_geom) *g, const SFX(el_model) *m, const SFX(el_state) *a,
                              const real *uxo, const real *uzo, const real *uxc, const real *uzc, const real *uxn,
                              const real *uzn, real *g_rih, real *g_rjh, real *g_lam, real *g_mu, real *g_muhh) {
  const int64_t nx = g->nx, nz = g->nz;
  const real _dt2 = (real)(1.0 / ((double)m->dt * (double)m->dt));
  const real *aux = a->ux[1], *auz = a->uz[1];
#pragma omp parallel for schedule(static)
  for (int64_t j = 1; j <= nz; j++)
    for (int64_t i = 1; i <= nx - 1; i++) {
      const real v = EL_A(aux, nx - 1, i, j) *
                     (EL_A(uxo, nx - 1, i, j) - 2 * EL_A(uxc, nx - 1, i, j) + EL_A(uxn, nx - 1, i, j)) * _dt2;
      EL_A(g_rih, nx - 1, i, j) += (j == 1 && g->freetop) ? v / 2 : v;
    }
#pragma omp parallel for schedule(static)
  for (int64_t j = 1; j <= nz - 1; j++)
    for (int64_t i = 1; i <= nx; i++)
      EL_A(g_rjh, nx, i, j) += EL_A(auz, nx, i, j) *
                               (EL_A(uzo, nx, i, j) - 2 * EL_A(uzc, nx, i, j) + EL_A(uzn, nx, i, j)) * _dt2;
  const int64_t j0 = g->freetop ? 1 : 2;
#pragma omp parallel for schedule(static)
  for (int64_t j = j0; j <= nz - 1; j++)
    for (int64_t i = 2; i <= nx - 1; i++) {
      const real exx = SFX(d_ux_dx)(g, uxc, i, j), exxa = SFX(d_ux_dx)(g, aux, i, j);
      const real ezz = SFX(d_uz_dz)(g, uxc, uzc, m->lam, m->mu, i, j);
      const real ezza = SFX(d_uz_dz)(g, aux, auz, m->lam, m->mu, i, j);
      const real div = exx + ezz, diva = exxa + ezza;
      if (j == 1 && g->freetop) {
        EL_A(g_lam, nx, i, j) += div * diva / 2;
        EL_A(g_mu, nx, i, j) += (exx * exxa + ezz * ezza);
      } else {
        EL_A(g_lam, nx, i, j) += div * diva;
        EL_A(g_mu, nx, i, j) += 2 * (exx * exxa + ezz * ezza);
      }
    }
#pragma omp parallel for schedule(static)
  for (int64_t j = 1; j <= nz - 1; j++)
    for (int64_t i = 1; i <= nx - 1; i++) {
      const real exz = (SFX(d_uz_dx)(g, uzc, i, j) + SFX(d_ux_dz)(g, uxc, i, j)) / 2;
      const real exza = (SFX(d_uz_dx)(g, auz, i, j) + SFX(d_ux_dz)(g, aux, i, j)) / 2;
      EL_A(g_muhh, nx - 1, i, j) += 2 * (exz * exza + exz * exza);
    }
}

/* swgradient_1shot! for elastic (ela_gradient.jl:4-186 MT, :188-362 EF) up
 * to the raw grad fields; checkpointer of ela_models.jl:370-385 (ucur width
 * 2, the four psi groups every cf steps, ucur buffered).  adjsrc is
 * -dchi/du (nt x 2 x nrec), injected as external forces at the receivers. */
int SFX(oracle_el2d_gradient)(const oracle_geom *og, const real *lam, const real *mu, const real *rho_ih,
                              const real *rho_jh, const real *mu_ihjh, const SFX(cpml) *cx, const SFX(cpml) *cz,
                              int nt, int kind, const SFX(el_taps) *s1, const SFX(el_taps) *s2, const real *M,
                              const real *tf, const SFX(el_taps) *rx, const SFX(el_taps) *rz, const real *adjsrc,
                              int cf, real *g_rih, real *g_rjh, real *g_lam, real *g_mu, real *g_muhh, real *traces) {
  if (cf < 1 || cf >= nt) return -1;
  SFX(el_geom) g = {og->n[0], og->n[1], og->halo, og->freetop, (real)(1.0 / og->h[0]), (real)(1.0 / og->h[1])};
  SFX(el_model) m = {cx, cz, lam, mu, rho_ih, rho_jh, mu_ihjh, (real)og->dt};
  const int64_t nx = g.nx, nz = g.nz;
  const size_t Nx = (size_t)(nx - 1) * nz, Nz = (size_t)nx * (nz - 1);
  SFX(el_state) s, a;
  SFX(el_state_alloc)(&g, &s);
  SFX(el_state_alloc)(&g, &a);
  const int last = (nt / cf) * cf;
  int curr = last;
#define HAS_U(t) ((t) >= -1 && (t) <= nt + 1 && ((((t) >= 0) && (t) % cf == 0) || ((t) + 1) % cf == 0))
  real **ckx = (real **)calloc(nt + 3, sizeof(real *)), **ckz = (real **)calloc(nt + 3, sizeof(real *));
  for (int t = -1; t <= nt + 1; t++)
    if (HAS_U(t)) {
      ckx[t + 1] = (real *)calloc(Nx, sizeof(real));
      ckz[t + 1] = (real *)calloc(Nz, sizeof(real));
    }
  const int nk = (nt + 1) / cf + 1;
  real **ckpsi = (real **)calloc((size_t)nk * 8, sizeof(real *));
  for (int k = 0; k < nk * 8; k++) ckpsi[k] = (real *)calloc(SFX(el_psi_size)(&g, k % 8) + 1, sizeof(real));
  real **bx = (real **)calloc(cf + 1, sizeof(real *)), **bz = (real **)calloc(cf + 1, sizeof(real *));
  for (int k = 0; k <= cf; k++) {
    bx[k] = (real *)calloc(Nx, sizeof(real));
    bz[k] = (real *)calloc(Nz, sizeof(real));
  }
  real *own_tr = traces ? traces : (real *)calloc((size_t)nt * 2 * rx->npts, sizeof(real));
  for (int it = 1; it <= nt; it++) {
    SFX(el_fwd_step)(&g, &m, &s, kind, s1, s2, M, tf, nt, it, rx, rz, own_tr);
    for (int itw = it; itw <= it + 1; itw++)
      if (itw % cf == 0) {
        memcpy(ckx[it + 1], s.ux[1], Nx * sizeof(real));
        memcpy(ckz[it + 1], s.uz[1], Nz * sizeof(real));
      }
    if (it % cf == 0)
      for (int k = 0; k < 8; k++) memcpy(ckpsi[(it / cf) * 8 + k], s.psi[k], (SFX(el_psi_size)(&g, k) + 1) * sizeof(real));
    if (it >= last) {
      memcpy(bx[it - last], s.ux[1], Nx * sizeof(real));
      memcpy(bz[it - last], s.uz[1], Nz * sizeof(real));
    }
    oracle_progress("elastic gradient, forward", it, nt);
  }
  memset(g_rih, 0, Nx * sizeof(real));
  memset(g_rjh, 0, Nz * sizeof(real));
  memset(g_lam, 0, (size_t)nx * nz * sizeof(real));
  memset(g_mu, 0, (size_t)nx * nz * sizeof(real));
  memset(g_muhh, 0, (size_t)(nx - 1) * (nz - 1) * sizeof(real));
#define SX(t) (HAS_U(t) ? ckx[(t) + 1] : bx[(t) - curr])
#define SZ(t) (HAS_U(t) ? ckz[(t) + 1] : bz[(t) - curr])
  for (int it = nt; it >= 1; it--) {
    /* adjoint_onestep_CPML! (:359-448): residuals injected as external forces */
    SFX(el_update_stress)(&g, &m, &a);
    SFX(el_update_disp)(&g, &m, &a);
    SFX(el_inject_force)(&g, &m, &a, rx, rz, adjsrc, nt, it);
    SFX(el_rotate)(&a);
    oracle_progress("elastic gradient, adjoint", nt - it + 1, nt);
    int t = it - 2;
    if (!((curr <= t && t <= curr + cf) || HAS_U(t))) {
      int old = curr;
      curr -= cf;
      memcpy(bx[0], ckx[curr + 1], Nx * sizeof(real));
      memcpy(bz[0], ckz[curr + 1], Nz * sizeof(real));
      memcpy(bx[cf], ckx[old + 1], Nx * sizeof(real));
      memcpy(bz[cf], ckz[old + 1], Nz * sizeof(real));
      memcpy(s.ux[0], ckx[curr], Nx * sizeof(real));
      memcpy(s.uz[0], ckz[curr], Nz * sizeof(real));
      memcpy(s.ux[1], ckx[curr + 1], Nx * sizeof(real));
      memcpy(s.uz[1], ckz[curr + 1], Nz * sizeof(real));
      for (int k = 0; k < 8; k++) memcpy(s.psi[k], ckpsi[(curr / cf) * 8 + k], (SFX(el_psi_size)(&g, k) + 1) * sizeof(real));
      for (int rit = curr + 1; rit <= curr + cf - 1; rit++) {
        SFX(el_fwd_step)(&g, &m, &s, kind, s1, s2, M, tf, nt, rit, rx, rz, NULL);
        memcpy(bx[rit - curr], s.ux[1], Nx * sizeof(real));
        memcpy(bz[rit - curr], s.uz[1], Nz * sizeof(real));
      }
    }
    SFX(el_correlate)(&g, &m, &a, SX(it - 2), SZ(it - 2), SX(it - 1), SZ(it - 1), SX(it), SZ(it), g_rih, g_rjh,
                      g_lam, g_mu, g_muhh);
  }
#undef SX
#undef SZ
#undef HAS_U
  if (!traces) free(own_tr);
  for (int k = 0; k <= cf; k++) { free(bx[k]); free(bz[k]); }
  free(bx); free(bz);
  for (int k = 0; k < nk * 8; k++) free(ckpsi[k]);
  free(ckpsi);
  for (int t = -1; t <= nt + 1; t++) { free(ckx[t + 1]); free(ckz[t + 1]); }
  free(ckx); free(ckz);
  SFX(el_state_free)(&s);
  SFX(el_state_free)(&a);
  return 0;
}
