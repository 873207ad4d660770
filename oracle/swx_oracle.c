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

/* Snapshot sink of the forward loops (LinearSnapshotter, snapshotter.jl:
 * 1-41): with every > 0, after each step it with it % every == 0 the
 * snapshotted fields are appended to buf, in the order the reference's
 * snapshotter holds them.  Set by oracle_set_snapshots; 0 turns it off. */
static int g_snap_every = 0;
static char *g_snap_buf = NULL;
static size_t g_snap_off = 0;

void oracle_set_snapshots(int every, void *buf) {
  g_snap_every = every;
  g_snap_buf = (char *)buf;
  g_snap_off = 0;
}

static int snap_due(int it) { return g_snap_every > 0 && g_snap_buf && it % g_snap_every == 0; }

static void snap_put(const void *src, size_t bytes) {
  memcpy(g_snap_buf + g_snap_off, src, bytes);
  g_snap_off += bytes;
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
#include "oracle_1d.h"
#undef real
#undef SFX
#undef A3
#undef A1

#define real float
#define SFX(x) x##_f32
typedef struct { const real *a, *a_h, *b, *b_h; } SFX(cpml);
#include "oracle_impl.h"
#include "oracle_cd3d.h"
#include "oracle_el2d.h"
#include "oracle_1d.h"
#undef real
#undef SFX

int oracle_abi_version(void) { return 1; }
