// Internal declarations shared by the translation units of libswx.so.
// Not part of the public ABI (that is include/swx.h).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "swx.h"

// Debug build (build.py --debug, -DSWX_DEBUG): device-side asserts on the
// indices of the shared accessors (point-table rows, C-PML coefficient
// vectors, elastic field accessor); they compile to nothing otherwise.
#ifdef SWX_DEBUG
#include <cassert>
#define SWX_DASSERT(cond) assert(cond)
#else
#define SWX_DASSERT(cond) ((void)0)
#endif

namespace swx {

// Error carrying an swx_status; converted to a return code at the C boundary.
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define SWX_HIP(call)                                                              \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess)                                                          \
      throw ::swx::Error(SWX_EHIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define SWX_REQUIRE(cond, code, msg)                 \
  do {                                               \
    if (!(cond)) throw ::swx::Error((code), (msg));  \
  } while (0)

// Message returned by swx_last_error() on the calling thread.
void set_last_error(const char *msg);

// Runs f at the C boundary: SWX_OK, or the status of the error it threw
// with its message kept for swx_last_error().
template <typename F>
int guarded(F &&f) {
  try {
    f();
    set_last_error("");
    return SWX_OK;
  } catch (const Error &e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::exception &e) {
    set_last_error(e.what());
    return SWX_EINVAL;
  }
}

// Owning device allocation (hipMalloc / hipFree).
class DevBuf {
 public:
  DevBuf() = default;
  explicit DevBuf(size_t bytes) { alloc(bytes); }
  ~DevBuf() { release(); }
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  DevBuf(DevBuf &&o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
  DevBuf &operator=(DevBuf &&o) noexcept {
    if (this != &o) { release(); p_ = o.p_; n_ = o.n_; o.p_ = nullptr; o.n_ = 0; }
    return *this;
  }
  void alloc(size_t bytes) {
    release();
    if (bytes == 0) return;
    hipError_t e = hipMalloc(&p_, bytes);
    if (e != hipSuccess) {
      p_ = nullptr;
      throw Error(SWX_ENOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed: " + hipGetErrorString(e));
    }
    n_ = bytes;
  }
  void release() {
    if (p_) (void)hipFree(p_);
    p_ = nullptr;
    n_ = 0;
  }
  template <typename T> T *as() const { return static_cast<T *>(p_); }
  void *get() const { return p_; }
  size_t bytes() const { return n_; }

 private:
  void *p_ = nullptr;
  size_t n_ = 0;
};

// Per-batch source/receiver tables shared by the acoustic engines.
// Sources and receivers that fall on the same grid cell are grouped so that
// one thread adds them in the reference's serial order (inject_sources!,
// acoustic2D_xPU.jl:48-54, runs over sources in index order).
struct PointTables {
  int nshots = 0;
  std::vector<int> nt, nsrc, nrec;
  std::vector<int64_t> off_tf, off_tr;    // offsets into srctf / traces
  int64_t total_tf = 0, total_tr = 0;
  // device arrays
  DevBuf d_nt;        // int32 [ns]
  DevBuf d_off_tf;    // int64 [ns]
  DevBuf d_off_tr;    // int64 [ns]
  DevBuf d_usrc_off;  // int32 [ns+1]  unique source cells per shot
  DevBuf d_usrc_cell; // int64 [*]
  DevBuf d_usrc_ptr;  // int32 [*+1]  into d_usrc_list (global)
  DevBuf d_usrc_list; // int32 [*]    source column within its shot
  DevBuf d_urec_off;  // same for receivers (adjoint injection)
  DevBuf d_urec_cell;
  DevBuf d_urec_ptr;
  DevBuf d_urec_list;
  DevBuf d_rec_off;   // int32 [ns+1] receivers per shot (recording)
  DevBuf d_rec_cell;  // int64 [*]
  // Row index of the unique cells (sorted by linear index): for shot s the
  // unique source cells of grid row r (all indices but the first) are
  // [srow[s*(nrows+1)+r], srow[s*(nrows+1)+r+1]); same for receivers.
  int64_t nrows = 0;
  DevBuf d_srow;      // int32 [ns*(nrows+1)]
  DevBuf d_rrow;      // int32 [ns*(nrows+1)]
  DevBuf d_srctf;     // T [total_tf]
  DevBuf d_traces;    // T [total_tr]
  DevBuf d_res;       // T [total_tr] (adjoint sources)
  std::vector<int64_t> h_rec_cell;  // host copies (for validation/debug)
};

struct PointArgsView {
  const int *nt;
  const int64_t *off_tf, *off_tr;
  const int *usrc_off; const int64_t *usrc_cell; const int *usrc_ptr; const int *usrc_list;
  const int *urec_off; const int64_t *urec_cell; const int *urec_ptr; const int *urec_list;
  const int *rec_off; const int64_t *rec_cell;
  const int *srow; const int *rrow; int64_t nrows;
};

inline PointArgsView view(const PointTables &t) {
  return PointArgsView{t.d_nt.as<int>(), t.d_off_tf.as<int64_t>(), t.d_off_tr.as<int64_t>(),
                       t.d_usrc_off.as<int>(), t.d_usrc_cell.as<int64_t>(), t.d_usrc_ptr.as<int>(),
                       t.d_usrc_list.as<int>(), t.d_urec_off.as<int>(), t.d_urec_cell.as<int64_t>(),
                       t.d_urec_ptr.as<int>(), t.d_urec_list.as<int>(), t.d_rec_off.as<int>(),
                       t.d_rec_cell.as<int64_t>(), t.d_srow.as<int>(), t.d_rrow.as<int>(), t.nrows};
}

// Checks of swx_plan_create on a plan description (throws SWX_EINVAL).
void validate_plan_desc(const swx_plan_desc &d);

// Host half of PointTables: the same tables in host vectors.
struct HostPointTables {
  int nshots = 0;
  std::vector<int> nt, nsrc, nrec;
  std::vector<int64_t> off_tf, off_tr;
  int64_t total_tf = 0, total_tr = 0, nrows = 0;
  std::vector<int> usrc_off, usrc_ptr, usrc_list, urec_off, urec_ptr, urec_list, rec_off, srow, rrow;
  std::vector<int64_t> usrc_cell, urec_cell, rec_cell;
};

// Validates the ABI shot descriptors and builds the host tables.
void host_point_tables(HostPointTables &t, int ndim, const int64_t *grid, int nshots, const swx_shot_desc *shots);

// Builds PointTables from the ABI shot descriptors.  `grid` are the point
// counts per axis of the field sources and receivers live on; indices are
// validated against it (check_positions, src/utils/checks.jl:55-70).
void build_point_tables(PointTables &t, int ndim, const int64_t *grid, int esize, int nshots,
                        const swx_shot_desc *shots, hipStream_t stream);

// Event-pair timer for the dominant kernel.
class KernelTimer {
 public:
  ~KernelTimer();
  void begin(hipStream_t s);
  // kind 0: forward-step kernel family, 1: adjoint-step kernel family
  void end(hipStream_t s, int kind);
  // The event pair of one launch made with hipExtLaunchKernelGGL, which
  // records them at the kernel's own start and end (no event packets
  // around the launch); both null when timing is off.
  void dispatch_events(int kind, hipEvent_t *start, hipEvent_t *stop);
  // Sums elapsed times of recorded pairs per kind (stream synchronized).
  void collect(double ms[2], int64_t n[2]);
  void reset() { used_ = 0; kinds_.clear(); }
  bool enabled = false;

 private:
  std::vector<hipEvent_t> ev_;
  std::vector<int> kinds_;
  size_t used_ = 0;
};

// One field of a snapshot: `rows` rows of `width` bytes, `pitch` bytes apart
// on the device (staggered fields live in nx-pitched slabs).
struct SnapPiece {
  const void *dev;
  int64_t width, pitch, rows;
};

// Engine interface: one implementation per physics x dimension x dtype.
class Engine {
 public:
  virtual ~Engine() = default;
  virtual void set_model(int nfields, const void *const *fields) = 0;
  // Receives the shot descriptors after the point tables were built (the
  // engine uploads the per-shot C-PML coefficients).
  virtual void on_shots_set(int nshots, const swx_shot_desc *shots) = 0;
  // Elastic engines build their own tap tables (and the trace buffers of
  // the PlanCore point tables) from the elastic shot descriptors.
  virtual void on_elastic_shots(int, const swx_elastic_shot_desc *) {
    throw Error(SWX_EINVAL, "swx_set_shots_elastic needs an SWX_ELASTIC_ISO plan");
  }
  virtual void forward(bool save_state) = 0;
  // Runs the adjoint loop on the device (residuals already in pts.d_res).
  virtual void adjoint(bool per_shot) = 0;
  // Copies the gradient fields of the last adjoint() to host memory.
  virtual void copy_grad(void *grad_host, bool per_shot) = 0;
  virtual int64_t bytes_state() const = 0;
  virtual int64_t bytes_checkpoints() const = 0;
  virtual int64_t bytes_replay() const = 0;
  virtual int segment() const = 0;
  // Kernel variants chosen by the autotuner (0 if a family has one).
  virtual void variants(int32_t v[2]) const { v[0] = v[1] = 0; }
  // The snapshotted fields of shot s at forward level it (valid right after
  // that forward step), in the reference snapshotter's order; returns their
  // number (0: snapshots not provided).
  virtual int snap_fields(int it, int s, SnapPiece *out) const {
    (void)it; (void)s; (void)out;
    return 0;
  }
};

struct PlanCore {
  swx_plan_desc d{};
  hipStream_t stream = nullptr;
  int esize = 4;
  PointTables pts;
  KernelTimer ktimer;
  bool forward_done = false;
  double ms_forward = 0, ms_adjoint = 0;
  double ms_kernel[2] = {0, 0};
  int64_t n_kernel[2] = {0, 0};
  size_t budget = 0;  // bytes
  // observed data of an L2 misfit for the current batch (swx_set_observed):
  // T [total_tr], optional double weights [total_tr], per-shot sample
  // offsets [ns+1] and the partial sums of the misfit reduction
  bool obs_set = false, obs_weighted = false;
  bool traces_valid = false;   // d_traces holds the last forward's traces
  // snapshots of the forward runs (swx_set_snapshots): every snap_every
  // steps the engine's snap_fields go to snap_host + snap_off[s] + k * snap_one
  int snap_every = 0;
  char *snap_host = nullptr;
  int64_t snap_one = 0;
  std::vector<int64_t> snap_off;
  DevBuf d_obs, d_obs_w, d_obs_off, d_obs_part;
};

// Copies the snapshot of forward level it of every shot still running (a
// no-op unless snapshots are set and it is a multiple of snap_every).
void take_snapshots(PlanCore &c, const Engine &e, int it);

// Uploads the observed data (and weights) of the current batch.
void set_observed(PlanCore &c, const void *observed, const double *weights);
// d_res = -(traces - observed) * weights on the device; with misfit_host,
// the per-shot misfits 1/2 sum w r^2 (copied asynchronously on c.stream).
void l2_adjoint_source(PlanCore &c, double *misfit_host);

std::unique_ptr<Engine> make_acoustic_cd2d(PlanCore &core);
std::unique_ptr<Engine> make_acoustic_vd2d(PlanCore &core);
std::unique_ptr<Engine> make_acoustic_cd3d(PlanCore &core);
std::unique_ptr<Engine> make_elastic2d(PlanCore &core);
std::unique_ptr<Engine> make_acoustic_1d(PlanCore &core);

// Helpers
inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }
inline int mod3(int k) { return ((k % 3) + 3) % 3; }

// Progress lines on stderr when the environment sets SWX_PROGRESS: about
// ten per time loop (forward steps, adjoint segments).
void progress(const char *phase, int done, int total);

// Chooses the replay segment length S for linear checkpointing given the
// bytes per replay slot (all shots) and per checkpoint (all shots), the
// user check_freq (1 = keep every step) and the memory still available.
int choose_segment(int nt_max, int check_freq, int64_t slot_bytes, int64_t ckpt_bytes, int64_t avail);

// Launch-shape tuner: times launch(k) for the candidates k = 0 .. n-1 on
// `stream`, after one warm-up launch each, in three interleaved rounds (a
// clock ramp or a busy moment then cannot decide) of at least five launches
// and at least 2 ms each (short launches are otherwise lost in the timing
// noise), keeps each candidate's best round (per launch) and returns the
// index of the fastest.  The launches must be repeatable (the engines time
// a step that reads only state it does not write, or the zero initial
// state).
template <typename F>
int fastest_candidate(hipStream_t stream, int n, F launch) {
  if (n <= 1) return 0;
  hipEvent_t e0, e1;
  SWX_HIP(hipEventCreate(&e0));
  SWX_HIP(hipEventCreate(&e1));
  auto timed = [&](int k, int reps) {
    SWX_HIP(hipEventRecord(e0, stream));
    for (int r = 0; r < reps; r++) launch(k);
    SWX_HIP(hipEventRecord(e1, stream));
    SWX_HIP(hipEventSynchronize(e1));
    float ms = 0;
    SWX_HIP(hipEventElapsedTime(&ms, e0, e1));
    return ms;
  };
  std::vector<int> reps(n, 5);
  for (int k = 0; k < n; k++) {
    const float one = timed(k, 1);   // warm-up, and the size of a round
    reps[k] = std::max(5, std::min(2000, (int)(2.0f / std::max(one, 1e-4f)) + 1));
  }
  std::vector<float> best(n, 1e30f);
  for (int round = 0; round < 3; round++)
    for (int k = 0; k < n; k++) best[k] = std::min(best[k], timed(k, reps[k]) / reps[k]);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  int k_best = 0;
  for (int k = 1; k < n; k++)
    if (best[k] < best[k_best]) k_best = k;
  return k_best;
}

}  // namespace swx
