/*
 * swx.h — C ABI of the MI355X (gfx950) backend for the SeismicWaves.jl
 * finite-difference forward / adjoint-gradient path.
 *
 * The reference dispatches each time step through a backend *module*
 * returned by select_backend (src/models/backend_selection.jl:2-40, GPU
 * modules registered by ext/SeismicWaves_AMDGPUBackendExt.jl:15-26).  Its
 * drivers call, once per step, backend.forward_onestep_CPML!,
 * backend.adjoint_onestep_CPML!, backend.prescale_residuals! and
 * backend.correlate_gradient[_m0|_m1]! (acou_forward.jl:49-58,
 * acou_gradient.jl:33-82 and :145-175), and keep the field state and the
 * LinearCheckpointer (src/utils/checkpointers.jl) in the wave simulation.
 *
 * This library replaces that whole per-shot loop for a *batch* of shots:
 * one plan owns the device-resident state of all shots (fields, C-PML memory
 * variables, checkpoints, replay buffers), swx_forward runs the forward time
 * loop of every shot, swx_adjoint replays from the GPU-resident checkpoints,
 * runs the adjoint loop and correlates the gradient.  The host keeps what the
 * reference keeps on the host: checks, C-PML coefficients (cpmlcoeffs.jl),
 * source scaling (possrcrec_scaletf), the misfit and its adjoint source
 * (L2Misfit.jl) and the gradient post-processing (mute, chain rule).
 *
 * Conventions
 *   - Arrays are column-major with x (dimension 1) fastest, exactly the
 *     memory layout of the Julia arrays the reference hands to its backends.
 *   - Grid indices are 1-based Int64 (Julia Matrix{Int} from
 *     find_nearest_grid_points, src/utils/utils.jl:47-58).
 *   - Scalars of type T are float (SWX_F32) or double (SWX_F64) as the
 *     plan's dtype says; `const void*` arguments point at host memory of T.
 *   - Every function returns SWX_OK (0) or an swx_status; the message of the
 *     last failure on the calling thread is returned by swx_last_error().
 *     All validation happens before any kernel is launched, mirroring the
 *     reference's Julia-side @assert checks.
 *   - A plan is bound to one device and one HIP stream; a plan is not
 *     thread-safe, distinct plans are.
 */
#ifndef SWX_H
#define SWX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SWX_ABI_VERSION 1

typedef enum {
  SWX_OK = 0,
  SWX_EINVAL = 1,       /* invalid argument (reference: AssertionError)     */
  SWX_ENOMEM = 2,       /* device memory budget exceeded                    */
  SWX_EHIP = 3,         /* HIP runtime error                                */
  SWX_ESTATE = 4,       /* call out of order (e.g. adjoint before forward)  */
  SWX_EUNSUPPORTED = 5  /* physics/dimension not built into this library    */
} swx_status;

typedef enum {
  SWX_ACOUSTIC_CD = 1,  /* AcousticCDCPMLWaveSimulation, acou_models.jl:68   */
  SWX_ACOUSTIC_VD = 2,  /* AcousticVDStaggeredCPMLWaveSimulation, :311     */
  SWX_ELASTIC_ISO = 3   /* ElasticIsoCPMLWaveSimulation, ela_models.jl:177  */
} swx_physics;

typedef enum { SWX_F32 = 1, SWX_F64 = 2 } swx_dtype;

typedef struct swx_plan swx_plan;

/* Plan description: the static part of InputParametersAcoustic
 * (acou_params.jl:9-35), CPMLBoundaryConditionParameters (bdc_params.jl:
 * 16-27) and GradParameters.check_freq (genparameters.jl:63-78), plus the
 * batch capacity. */
typedef struct {
  int32_t physics;       /* swx_physics                                     */
  int32_t dtype;         /* swx_dtype                                       */
  int32_t ndim;          /* 1, 2 or 3 (1: acoustic only; elastic: 2)        */
  int32_t halo;          /* C-PML thickness in grid points (>= 0)           */
  int64_t n[3];          /* grid points per dimension, x first              */
  double h[3];           /* grid spacing per dimension                      */
  double dt;             /* time step                                       */
  int32_t freetop;       /* free surface on the last axis                   */
  int32_t max_shots;     /* batch capacity                                  */
  int32_t nt_max;        /* largest number of time steps of any shot        */
  int32_t check_freq;    /* checkpoint interval; 1 = keep every time step   */
  int32_t gradient;      /* 1: allocate adjoint state and checkpoints       */
  int32_t device;        /* HIP device ordinal                              */
  double mem_budget;     /* bytes of HBM the plan may use, 0 = 90% of free  */
} swx_plan_desc;

/* One shot of scalar sources / scalar receivers (ScalarShot, shot.jl:9-16).
 * `srctf` is the scaled source term of possrcrec_scaletf (acou_forward.jl:
 * 6-20 for CD, :67-81 for VD), nt x nsrc column-major.  `cpml[d]` are the
 * coefficient vectors of CPMLCoefficientsAxis for axis d (cpmlcoeffs.jl:2-16)
 * as computed by init_bdc! for this shot's dominant frequency:
 * a (2(halo+1)), a_h (2halo), b (2(halo+1)), b_h (2halo).                  */
typedef struct {
  int32_t nt;
  int32_t nsrc;
  const int64_t *src_idx;   /* nsrc x ndim, 1-based                         */
  const void *srctf;        /* nt x nsrc, T                                  */
  int32_t nrec;
  const int64_t *rec_idx;   /* nrec x ndim, 1-based                         */
  const void *cpml[3][4];   /* [axis][a, a_h, b, b_h], T                     */
} swx_shot_desc;

/* Source kinds of an elastic shot: MomentTensorSources and
 * ExternalForceSources (src/shots/sources.jl:54-91). */
typedef enum {
  SWX_SRC_MOMENT_TENSOR = 1,
  SWX_SRC_EXTERNAL_FORCE = 2
} swx_source_kind;

/* Points spread over grid taps, the srccoeij_* / srccoeval_* (and reccoeij_*
 * / reccoeval_*) lists of possrcrec_scaletf (ela_models.jl:6-90) built by
 * spread_positions (src/utils/utils.jl:168-214): point p owns taps
 * ptr[p] .. ptr[p+1]-1; tap k is the 1-based cell (ij[2k], ij[2k+1]) of the
 * family's staggered grid with weight coef[k] (T). */
typedef struct {
  const int32_t *ptr;       /* npts + 1                                      */
  const int64_t *ij;        /* 2 x ntaps, (i, j) pairs                       */
  const void *coef;         /* ntaps, T                                      */
} swx_taps;

/* One elastic shot (MomentTensorShot / ExternalForceShot, shot.jl:23-44)
 * with VectorReceivers (receivers.jl:64-100).
 *   moment tensor:  src[0] taps on the sxx/szz grid (nx x nz), src[1] on
 *                   the sxz grid (nx-1 x nz-1); moments = nsrc x (Mxx, Mzz,
 *                   Mxz) (T); srctf nt x nsrc.
 *   external force: src[0] taps on the ux grid (nx-1 x nz), src[1] on the uz
 *                   grid (nx x nz-1); srctf nt x 2 x nsrc.
 *   srctf is already divided by dx*dz (possrcrec_scaletf).  rec[0] / rec[1]
 *   are the receivers' ux / uz taps.  Seismograms are nt x 2 x nrec.       */
typedef struct {
  int32_t nt;
  int32_t source_kind;      /* swx_source_kind                               */
  int32_t nsrc;
  swx_taps src[2];
  const void *moments;      /* moment tensor sources only                    */
  const void *srctf;
  int32_t nrec;
  swx_taps rec[2];
  const void *cpml[3][4];   /* as in swx_shot_desc                           */
} swx_elastic_shot_desc;

typedef struct {
  int64_t bytes_state;       /* fields + C-PML memory variables             */
  int64_t bytes_checkpoints; /* checkpoint storage                          */
  int64_t bytes_replay;      /* replay buffer                               */
  int32_t segment;           /* replay segment length actually used         */
  int32_t nshots;            /* shots currently set                         */
  double ms_forward;         /* device time of the last swx_forward         */
  double ms_adjoint;         /* device time of the last swx_adjoint         */
  /* With kernel timing on: summed device time and number of launches of
   * the forward-step kernel family [0] (forward and replay steps) and of
   * the adjoint-step kernel family [1], over the last swx_forward or
   * swx_adjoint call. */
  double ms_kernel[2];
  int64_t n_kernel[2];
  /* Launch shapes the plan's autotuner chose (or the environment pinned):
   * VD [forward, adjoint] tile codes (8/16/32 interior + other launches,
   * 308/316/332 one launch, 102/104 z-marching strips; + 1000 k: shots in
   * groups of 2^k in the block order); elastic [forward, adjoint] variant
   * codes (0/1 stress and displacement launches, 2 one fused launch); CD 2D
   * [rows per thread, 10 x adjoint + forward levels per launch]; CD 3D
   * [launch shape 8/4/2/108/104/102/208/204/202, planes per z chunk]; 0
   * where there is one shape. */
  int32_t kernel_variant[2];
} swx_plan_info;

/* Version of this ABI (SWX_ABI_VERSION). No reference counterpart. */
int swx_abi_version(void);

/* Message of the last failing call on this thread ("" if none). */
const char *swx_last_error(void);

/* Replaces the allocation done by the WaveSimulation constructors
 * (acou_models.jl:90-201, :333-451) for `max_shots` shots at once.  The
 * time stepping of the plan replaces the backends' forward_onestep_CPML! /
 * adjoint_onestep_CPML!: acoustic1D_xPU.jl, acoustic1D_VD_xPU.jl (ndim 1),
 * acoustic2D_xPU.jl, acoustic2D_VD_xPU.jl, acoustic3D_xPU.jl and
 * elastic2D_iso_xPU.jl (src/models/{acoustic,elastic}/backends/shared/). */
int swx_plan_create(const swx_plan_desc *desc, swx_plan **plan);

int swx_plan_destroy(swx_plan *plan);

/* Replaces precompute_fact! uploads (acou_models.jl:60, :292-300) and
 * precomp_elaprop! (ela_models.jl:156-173): the material fields consumed by
 * the kernels, host arrays of T.
 *   SWX_ACOUSTIC_CD: fields[0] = fact = dt^2 vp^2            (n)
 *   SWX_ACOUSTIC_VD: fields[0] = fact_m0 = vp^2 rho dt       (n)
 *                    fields[1] = fact_m1_stag[1] (x)         (nx-1, nz)
 *                    fields[2] = fact_m1_stag[2] (z)         (nx, nz-1)
 *                    (1D: fields[0] (nx), fields[1] (nx-1))
 *   SWX_ELASTIC_ISO: fields[0] = lambda, fields[1] = mu      (n)
 *                    fields[2] = rho_ihalf                   (nx-1, nz)
 *                    fields[3] = rho_jhalf                   (nx, nz-1)
 *                    fields[4] = mu_ihalf_jhalf              (nx-1, nz-1) */
int swx_set_model(swx_plan *plan, int nfields, const void *const *fields);

/* Replaces the per-shot uploads of swforward_1shot!/swgradient_1shot!
 * (acou_forward.jl:38-41) for a batch of shots.  Validates positions.
 * Acoustic plans only.  A rejected batch leaves the plan without one:
 * swx_forward and swx_adjoint then return SWX_ESTATE until a batch is set. */
int swx_set_shots(swx_plan *plan, int nshots, const swx_shot_desc *shots);

/* Elastic counterpart of swx_set_shots: the per-shot uploads of the elastic
 * swforward_1shot!/swgradient_1shot! (ela_forward.jl:21-43,
 * ela_gradient.jl:24-46).  All shots of a batch share one source kind.
 * A rejected batch leaves the plan without one, as for swx_set_shots. */
int swx_set_shots_elastic(swx_plan *plan, int nshots, const swx_elastic_shot_desc *shots);

/* Forward time loop of every shot (acou_forward.jl:49-58 per shot).
 * traces: host T array, the shots' seismograms (nt_s x nrec_s, column-major)
 * concatenated in shot order; may be NULL.  With save_state != 0 (gradient
 * plans) the checkpoints for swx_adjoint stay resident on the device. */
int swx_forward(swx_plan *plan, void *traces, int save_state);

/* Adjoint time loop with replay from the device checkpoints and gradient
 * correlation (acou_gradient.jl:51-84 for CD, :145-175 for VD,
 * ela_gradient.jl:95-150 for elastic).
 * adjsrc: host T array, -dchi/du per shot (nt_s x nrec_s, elastic
 *   nt_s x 2 x nrec_s) concatenated; the acoustic prescale by the material
 *   factor at the receivers is done here.
 * grad: host T output.  Raw gradient fields as the reference's grid fields
 *   hold them after the adjoint loop, already summed over shots unless
 *   per_shot != 0 (then nshots consecutive copies):
 *   SWX_ACOUSTIC_CD: grad_vp (n)
 *   SWX_ACOUSTIC_VD: grad_m0 (n), grad_m1_stag x (nx-1,nz), z (nx,nz-1)
 *                    (1D: grad_m0 (nx), grad_m1_stag x (nx-1))
 *   SWX_ELASTIC_ISO: grad_rho_ihalf (nx-1,nz), grad_rho_jhalf (nx,nz-1),
 *                    grad_lambda (n), grad_mu (n),
 *                    grad_mu_ihalf_jhalf (nx-1,nz-1) */
int swx_adjoint(swx_plan *plan, const void *adjsrc, void *grad, int per_shot);

/* Bytes of the static device array that swx_forward_noalloc works in. */
#define SWX_NOALLOC_SCRATCH_BYTES (64ull << 20)

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the forward time loop of ONE 2D acoustic
 * constant-density shot, the work of swx_plan_create + swx_set_model +
 * swx_set_shots + swx_forward on a 1-shot plan, with the same kernels and
 * bitwise the same seismograms, but without any device allocation.  All
 * device memory (model, time levels, C-PML memory variables and
 * coefficients, point tables, source time function, traces) is carved out
 * of a static device array of SWX_NOALLOC_SCRATCH_BYTES that every GPU the
 * library runs on holds until the process ends; a shot that needs more
 * returns SWX_ENOMEM.  desc: physics SWX_ACOUSTIC_CD, ndim 2, nt_max >=
 * shot->nt (max_shots, check_freq, gradient and mem_budget are not used).
 * fact, shot and traces are host memory as for swx_set_model, swx_set_shots
 * and swx_forward.  Work runs on the device's null stream and the call
 * returns when the traces are copied.  All calls on one GPU share its
 * array, which keeps the last call's contents: calls that overlap on one
 * GPU (from several host threads) race. */
int swx_forward_noalloc(const swx_plan_desc *desc, const void *fact, const swx_shot_desc *shot, void *traces);

/* Snapshots of the following swx_forward calls of the current batch
 * (LinearSnapshotter, src/utils/snapshotter.jl:1-41, runparams.snapevery;
 * taken right after every snapevery-th step as at acou_forward.jl:55-57 and
 * :116-120, ela_forward.jl:74-77).  snaps is host memory of `bytes`; for
 * each shot in batch order, for k = 1 .. floor(nt_s / snapevery), the
 * fields of step k * snapevery follow each other at their Julia sizes:
 *   SWX_ACOUSTIC_CD: "pcur" (n)
 *   SWX_ACOUSTIC_VD: "pcur" (nx, nz), "vcur" vx (nx-1, nz), vz (nx, nz-1)
 *                    (1D: "pcur" (nx), "vcur" vx (nx-1))
 *   SWX_ELASTIC_ISO: "ucur" ux (nx-1, nz), uz (nx, nz-1),
 *                    "σ" σxx (nx, nz), σzz (nx, nz), σxz (nx-1, nz-1)
 * snapevery = 0 turns snapshots off; swx_set_shots* also does.  snapevery
 * must be smaller than every shot's nt (snapshotter.jl:12). */
int swx_set_snapshots(swx_plan *plan, int snapevery, void *snaps, int64_t bytes);

/* Observed data of an L2 misfit for the batch last set (L2Misfit.jl:2-20):
 * observed is a host T array in the layout of swx_forward's traces (per
 * shot nt_s x nrec_s, elastic nt_s x 2 x nrec_s, concatenated).  weights is
 * NULL (identity inverse covariance, no windows) or a host double array of
 * the same layout holding window mask x diagonal of the inverse covariance
 * for every sample.  Kept on the device until the next swx_set_shots*. */
int swx_set_observed(swx_plan *plan, const void *observed, const double *weights);

/* calcmisfit of L2Misfit (L2Misfit.jl:24-62) for every shot of the batch,
 * from the traces of the last swx_forward and the observed data on the
 * device: misfit[s] = 1/2 sum w (syn - obs)^2, summed in double. */
int swx_misfit_l2(swx_plan *plan, double *misfit);

/* swx_adjoint with the adjoint source built on the device: -dchi/du =
 * -(syn - obs) * w (L2Misfit.jl:64-95, acou_gradient.jl:45-51), formed in
 * double and rounded to T, so bitwise the source the host would pass to
 * swx_adjoint; no trace or residual crosses PCIe.  misfit (may be NULL)
 * receives the per-shot misfits as from swx_misfit_l2. */
int swx_adjoint_l2(swx_plan *plan, void *grad, int per_shot, double *misfit);

/* Sizes, segment length and device timings of the last calls. */
int swx_plan_get_info(swx_plan *plan, swx_plan_info *info);

/* Record HIP events around every stencil launch of subsequent calls
 * (0 = off).  Used by the benchmark to time the dominant kernel. */
int swx_set_kernel_timing(swx_plan *plan, int enable);

#ifdef __cplusplus
}
#endif

#endif /* SWX_H */
