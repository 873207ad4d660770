/*
 * FP64 oracle (and same-precision FP32 CPU baseline) for the SeismicWaves.jl
 * finite-difference path.  TEST INFRASTRUCTURE: loaded only by tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg, never by the
 * product library.  Build: oracle/Makefile (gcc -O3 -fopenmp).
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>

/* About ten progress lines per time loop on stderr when the environment
 * sets ORACLE_PROGRESS (long oracle runs at full benchmark sizes). */
static void oracle_progress(const char *phase, int done, int total) {
  static int on = -1;
  if (on < 0) on = getenv("ORACLE_PROGRESS") != NULL;
  if (!on || total <= 0) return;
  const int every = total / 10 > 0 ? total / 10 : 1;
  if (done % every == 0 || done == total) {
    fprintf(stderr, "[oracle] %s %d/%d\n", phase, done, total);
    fflush(stderr);
  }
}

typedef struct {
  int64_t n[3];   /* grid size, x (fast axis) first                         */
  double h[3];    /* grid spacing                                           */
  double dt;
  int halo;       /* C-PML thickness in cells                               */
  int freetop;    /* free surface on the last axis (elastic only)           */
} oracle_geom;

#define real double
#define SFX(x) x##_f64
typedef struct { const real *a, *a_h, *b, *b_h; } SFX(cpml);
#include "oracle_impl.h"
#include "oracle_cd3d.h"
#include "oracle_el2d.h"
#undef real
#undef SFX
#undef A3

#define real float
#define SFX(x) x##_f32
typedef struct { const real *a, *a_h, *b, *b_h; } SFX(cpml);
#include "oracle_impl.h"
#include "oracle_cd3d.h"
#include "oracle_el2d.h"
#undef real
#undef SFX

int oracle_abi_version(void) { return 1; }
