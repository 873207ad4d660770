/*
 * Memory-safety check of the oracle (TEST INFRASTRUCTURE): the C
 * restatement built with AddressSanitizer and UndefinedBehaviorSanitizer
 * (oracle/Makefile target `asan`) runs small forward and gradient cases of
 * every physics, in both precisions, with points on the grid edges and in
 * the C-PML, free surface on and off, and a replay interval that leaves a
 * partial last segment.  Any out-of-bounds access or undefined behaviour
 * aborts with the sanitizer's report; exit status 0 means every case ran
 * clean.
 */
#include "swx_oracle.c"

#include <stdio.h>

#define CASES(T, SUF)                                                                                         \
  static int run_##SUF(int freetop) {                                                                         \
    const int nx = 46, nz = 38, h = 5, nt = 37, cf = 7;                                                       \
    oracle_geom g = {{nx, nz, 1}, {10.0, 12.0, 1.0}, 1e-3, h, freetop};                                       \
    const size_t N = (size_t)nx * nz;                                                                         \
    T *f = calloc(N, sizeof(T)), *f2 = calloc(N, sizeof(T)), *f3 = calloc(N, sizeof(T));                      \
    T *f4 = calloc(N, sizeof(T)), *f5 = calloc(N, sizeof(T));                                                 \
    for (size_t q = 0; q < N; q++) {                                                                          \
      f[q] = (T)(0.05 + 1e-4 * (q % 13));                                                                     \
      f2[q] = (T)(0.4 + 1e-3 * (q % 7));                                                                      \
      f3[q] = (T)(0.4 + 1e-3 * (q % 5));                                                                      \
      f4[q] = (T)(2.0e9 + 1e7 * (q % 3));                                                                     \
      f5[q] = (T)(1.0e9 + 1e7 * (q % 4));                                                                     \
    }                                                                                                         \
    const int la = 2 * h + 2, lh = 2 * h;                                                                     \
    T *cc = calloc(4 * (size_t)(la + lh) * 3, sizeof(T));                                                    \
    for (size_t q = 0; q < 4 * (size_t)(la + lh) * 3; q++) cc[q] = (T)(0.5 + 0.01 * (q % 9));               \
    SFX_##SUF(cpml) cx = {cc, cc + la, cc + la + lh, cc + 2 * la + lh};                                       \
    SFX_##SUF(cpml) cz = {cc + 2 * (la + lh), cc + 3 * la + 2 * lh, cc + 3 * la + 3 * lh, cc + 4 * la + 3 * lh}; \
    SFX_##SUF(cpml) c3[3] = {cx, cz, cx};                                                                     \
    /* sources: interior, edge row 1, last column, C-PML; receivers alike */                                 \
    const int ns = 4, nr = 5;                                                                                 \
    int64_t src[8] = {20, 1, nx, 3, 15, 1, nz, 30};                                                           \
    int64_t rec[10] = {1, 10, nx, 25, 30, 1, 5, nz, 19, 2};                                                   \
    T *tf = calloc((size_t)nt * ns * 2, sizeof(T)), *tr = calloc((size_t)nt * nr * 2, sizeof(T));            \
    T *adj = calloc((size_t)nt * nr * 2, sizeof(T));                                                          \
    for (int k = 0; k < nt * ns * 2; k++) tf[k] = (T)(0.01 * ((k % 11) - 5));                                \
    for (int k = 0; k < nt * nr * 2; k++) adj[k] = (T)(0.002 * ((k % 7) - 3));                               \
    T *g1 = calloc(N, sizeof(T)), *g2 = calloc(N, sizeof(T)), *g3 = calloc(N, sizeof(T));                     \
    T *g4 = calloc(N, sizeof(T)), *g5 = calloc(N, sizeof(T));                                                 \
    int rc = 0;                                                                                               \
    rc |= oracle_cd2d_forward_##SUF(&g, f, &cx, &cz, nt, ns, src, tf, nr, rec, tr);                          \
    rc |= oracle_cd2d_gradient_##SUF(&g, f, &cx, &cz, nt, ns, src, tf, nr, rec, adj, cf, g1, tr);            \
    rc |= oracle_vd2d_forward_##SUF(&g, f, f2, f3, &cx, &cz, nt, ns, src, tf, nr, rec, tr);                  \
    rc |= oracle_vd2d_gradient_##SUF(&g, f, f2, f3, &cx, &cz, nt, ns, src, tf, nr, rec, adj, cf, g1, g2, g3, tr); \
    {                                                                                                         \
      oracle_geom g3d = {{20, 18, 16}, {10.0, 10.0, 10.0}, 1e-3, 3, 0};                                       \
      const size_t N3 = 20 * 18 * 16;                                                                         \
      T *fa = calloc(N3, sizeof(T)), *gg = calloc(N3, sizeof(T));                                             \
      for (size_t q = 0; q < N3; q++) fa[q] = (T)0.05;                                                        \
      int64_t s3[6] = {10, 1, 1, 9, 16, 1};   /* nsrc x 3 column-major: (10,1,16), (1,9,1) */                \
      int64_t r3[6] = {20, 5, 1, 1, 16, 8};   /* (20,1,16), (5,1,8) */                                      \
      rc |= oracle_cd3d_forward_##SUF(&g3d, fa, c3, 20, 2, s3, tf, 2, r3, tr);                                \
      rc |= oracle_cd3d_gradient_##SUF(&g3d, fa, c3, 20, 2, s3, tf, 2, r3, adj, 6, gg, tr);                   \
      free(fa); free(gg);                                                                                     \
    }                                                                                                         \
    /* elastic: one point per family with taps on the family's grid edges */                                 \
    int ptr[3] = {0, 2, 4};                                                                                   \
    /* taps (i, j) as ntaps x 2 column-major: all i, then all j */                                           \
    int64_t ij_xx[8] = {1, nx, 2, 3, 1, nz, 1, 4};                                                            \
    int64_t ij_xz[8] = {1, nx - 1, 1, 5, 1, nz - 1, 2, 6};                                                    \
    int64_t ij_ux[8] = {1, nx - 1, 2, 4, 1, nz, nz, 1};                                                       \
    int64_t ij_uz[8] = {1, nx, nx, 3, 1, nz - 1, 1, 2};                                                       \
    T coef[4] = {(T)0.5, (T)-0.25, (T)0.3, (T)0.1};                                                           \
    T M[6] = {(T)1, (T)0.5, (T)0.2, (T)-1, (T)0.3, (T)0.7};                                                   \
    SFX_##SUF(el_taps) t_xx = {2, ptr, ij_xx, 4, coef}, t_xz = {2, ptr, ij_xz, 4, coef};                       \
    SFX_##SUF(el_taps) t_ux = {2, ptr, ij_ux, 4, coef}, t_uz = {2, ptr, ij_uz, 4, coef};                       \
    for (int kind = 0; kind < 2; kind++) {                                                                    \
      const SFX_##SUF(el_taps) *s1 = kind == 0 ? &t_xx : &t_ux, *s2 = kind == 0 ? &t_xz : &t_uz;             \
      rc |= oracle_el2d_forward_##SUF(&g, f4, f5, f2, f3, f5, &cx, &cz, nt, kind, s1, s2, M, tf, &t_ux, &t_uz, tr); \
      rc |= oracle_el2d_gradient_##SUF(&g, f4, f5, f2, f3, f5, &cx, &cz, nt, kind, s1, s2, M, tf, &t_ux, &t_uz, \
                                       adj, cf, g1, g2, g3, g4, g5, tr);                                      \
    }                                                                                                         \
    free(f); free(f2); free(f3); free(f4); free(f5); free(cc); free(tf); free(tr); free(adj);                \
    free(g1); free(g2); free(g3); free(g4); free(g5);                                                         \
    return rc;                                                                                                \
  }

#define SFX_f64(x) x##_f64
#define SFX_f32(x) x##_f32
CASES(double, f64)
CASES(float, f32)

int main(void) {
  int rc = 0;
  for (int ft = 0; ft < 2; ft++) {
    rc |= run_f64(ft);
    rc |= run_f32(ft);
  }
  printf("oracle sanitizer check: %s\n", rc ? "a case returned an error" : "all cases ran clean");
  return rc ? 1 : 0;
}
