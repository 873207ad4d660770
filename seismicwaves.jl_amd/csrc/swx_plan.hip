// C ABI entry points of libswx.so (declared in include/swx.h): plan life
// cycle, validation, shot tables and dispatch to the physics engines.
#include <algorithm>
#include <cstdlib>
#include <cstdio>
#include <map>
#include <numeric>

#include "swx_internal.hpp"

struct swx_plan {
  swx::PlanCore core;
  std::unique_ptr<swx::Engine> engine;
};

namespace {
thread_local std::string g_last_error;
}  // namespace

namespace swx {

void set_last_error(const char *msg) { g_last_error = msg; }

void take_snapshots(PlanCore &c, const Engine &e, int it) {
  if (c.snap_every <= 0 || it % c.snap_every != 0) return;
  SnapPiece pc[8];
  for (int s = 0; s < c.pts.nshots; s++) {
    if (it > c.pts.nt[s]) continue;
    char *dst = c.snap_host + c.snap_off[s] + (int64_t)(it / c.snap_every - 1) * c.snap_one;
    const int n = e.snap_fields(it, s, pc);
    for (int k = 0; k < n; k++) {
      SWX_HIP(hipMemcpy2DAsync(dst, pc[k].width, pc[k].dev, pc[k].pitch, pc[k].width, pc[k].rows,
                               hipMemcpyDeviceToHost, c.stream));
      dst += pc[k].width * pc[k].rows;
    }
  }
}

KernelTimer::~KernelTimer() {
  for (auto e : ev_) (void)hipEventDestroy(e);
}

void KernelTimer::begin(hipStream_t s) {
  if (!enabled) return;
  if (used_ + 2 > ev_.size()) {
    for (int k = 0; k < 2; k++) {
      hipEvent_t e;
      SWX_HIP(hipEventCreate(&e));
      ev_.push_back(e);
    }
  }
  SWX_HIP(hipEventRecord(ev_[used_], s));
}

void KernelTimer::end(hipStream_t s, int kind) {
  if (!enabled) return;
  SWX_HIP(hipEventRecord(ev_[used_ + 1], s));
  kinds_.push_back(kind);
  used_ += 2;
}

void KernelTimer::dispatch_events(int kind, hipEvent_t *start, hipEvent_t *stop) {
  *start = *stop = nullptr;
  if (!enabled) return;
  while (used_ + 2 > ev_.size()) {
    hipEvent_t e;
    SWX_HIP(hipEventCreate(&e));
    ev_.push_back(e);
  }
  *start = ev_[used_];
  *stop = ev_[used_ + 1];
  kinds_.push_back(kind);
  used_ += 2;
}

void KernelTimer::collect(double ms[2], int64_t n[2]) {
  ms[0] = ms[1] = 0;
  n[0] = n[1] = 0;
  for (size_t k = 0; k + 1 < used_; k += 2) {
    float t = 0;
    SWX_HIP(hipEventElapsedTime(&t, ev_[k], ev_[k + 1]));
    const int kind = kinds_[k / 2] ? 1 : 0;
    ms[kind] += t;
    n[kind] += 1;
  }
  reset();
}

void progress(const char *phase, int done, int total) {
  static const bool on = std::getenv("SWX_PROGRESS") != nullptr;
  if (!on || total <= 0) return;
  const int every = std::max(1, total / 10);
  if (done % every == 0 || done == total) {
    std::fprintf(stderr, "[swx] %s %d/%d\n", phase, done, total);
    std::fflush(stderr);
  }
}

int choose_segment(int nt_max, int check_freq, int64_t slot_bytes, int64_t ckpt_bytes, int64_t avail) {
  auto mem = [&](int64_t S) {
    int64_t K = ceil_div(nt_max, S);
    return S * slot_bytes + std::max<int64_t>(K - 2, 0) * ckpt_bytes;
  };
  int64_t S = check_freq <= 1 ? nt_max : std::min<int64_t>(check_freq, nt_max);
  if (mem(S) <= avail) return (int)S;
  // Not enough memory for the requested interval: take the interval with the
  // smallest footprint (replay is deterministic, so results do not change).
  int64_t best = 1, bestm = mem(1);
  for (int64_t s = 1; s <= nt_max; s++) {
    int64_t m = mem(s);
    if (m < bestm) { bestm = m; best = s; }
  }
  if (bestm > avail)
    throw Error(SWX_ENOMEM, "checkpoint storage does not fit the device memory budget (need " +
                                std::to_string(bestm) + " bytes, have " + std::to_string(avail) + ")");
  return (int)best;
}

void validate_plan_desc(const swx_plan_desc &d) {
  SWX_REQUIRE(d.dtype == SWX_F32 || d.dtype == SWX_F64, SWX_EINVAL, "dtype must be SWX_F32 or SWX_F64");
  SWX_REQUIRE(d.ndim >= 1 && d.ndim <= 3, SWX_EINVAL, "Dimensionality must be less than or equal to 3!");
  for (int k = 0; k < d.ndim; k++) {
    SWX_REQUIRE(d.n[k] > 0, SWX_EINVAL, "All numbers of grid points must be positive!");
    SWX_REQUIRE(d.h[k] > 0, SWX_EINVAL, "All grid spacings must be positive!");
  }
  SWX_REQUIRE(d.dt > 0, SWX_EINVAL, "Timestep size must be positive!");
  SWX_REQUIRE(d.nt_max > 0, SWX_EINVAL, "Number of timesteps must be positive!");
  SWX_REQUIRE(d.halo >= 0, SWX_EINVAL, "CPML halo size must be non-negative!");
  SWX_REQUIRE(d.max_shots > 0, SWX_EINVAL, "max_shots must be positive");
  SWX_REQUIRE(d.check_freq >= 1, SWX_EINVAL, "check_freq must be positive");
  for (int k = 0; k < d.ndim; k++)
    SWX_REQUIRE(d.n[k] >= 2 * d.halo + 3, SWX_EINVAL,
                "Number grid points in the dimensions with C-PML boundaries must be at least 2*halo+3 = " +
                    std::to_string(2 * d.halo + 3) + "!");
  int ndev = 0;
  SWX_HIP(hipGetDeviceCount(&ndev));
  SWX_REQUIRE(d.device >= 0 && d.device < ndev, SWX_EINVAL, "invalid device ordinal");
}

void host_point_tables(HostPointTables &t, int ndim, const int64_t *grid, int nshots, const swx_shot_desc *shots) {
  t.nshots = nshots;
  t.nt.assign(nshots, 0);
  t.nsrc.assign(nshots, 0);
  t.nrec.assign(nshots, 0);
  t.off_tf.assign(nshots, 0);
  t.off_tr.assign(nshots, 0);
  t.usrc_off.assign(1, 0);
  t.usrc_ptr.assign(1, 0);
  t.urec_off.assign(1, 0);
  t.urec_ptr.assign(1, 0);
  t.rec_off.assign(1, 0);
  t.usrc_list.clear(); t.urec_list.clear(); t.srow.clear(); t.rrow.clear();
  t.usrc_cell.clear(); t.urec_cell.clear(); t.rec_cell.clear();
  int64_t rows = 1;
  for (int d = 1; d < ndim; d++) rows *= grid[d];
  t.nrows = rows;
  int64_t otf = 0, otr = 0;
  auto linear = [&](const int64_t *idx, int n, int k, const char *what, int s) {
    int64_t lin = 0, stride = 1;
    for (int d = 0; d < ndim; d++) {
      int64_t v = idx[k + (int64_t)n * d];
      if (v < 1 || v > grid[d])
        throw Error(SWX_EINVAL, std::string(what) + " position of shot " + std::to_string(s + 1) +
                                    " is not inside the grid!");
      lin += (v - 1) * stride;
      stride *= grid[d];
    }
    return lin;
  };
  for (int s = 0; s < nshots; s++) {
    const swx_shot_desc &sh = shots[s];
    SWX_REQUIRE(sh.nt > 0, SWX_EINVAL, "Number of timesteps must be positive!");
    SWX_REQUIRE(sh.nsrc > 0, SWX_EINVAL, "There must be at least one source!");
    SWX_REQUIRE(sh.nrec > 0, SWX_EINVAL, "There must be at least one receiver!");
    SWX_REQUIRE(sh.src_idx && sh.rec_idx && sh.srctf, SWX_EINVAL, "null source/receiver array");
    t.nt[s] = sh.nt;
    t.nsrc[s] = sh.nsrc;
    t.nrec[s] = sh.nrec;
    t.off_tf[s] = otf;
    t.off_tr[s] = otr;
    otf += (int64_t)sh.nt * sh.nsrc;
    otr += (int64_t)sh.nt * sh.nrec;
    // group sources by cell (cells in ascending linear order, source
    // indices ascending inside a group) and index the groups by grid row
    const int64_t nrow = rows;
    std::map<int64_t, std::vector<int>> g;
    for (int k = 0; k < sh.nsrc; k++) g[linear(sh.src_idx, sh.nsrc, k, "Source", s)].push_back(k);
    for (auto &kv : g) {
      t.usrc_cell.push_back(kv.first);
      for (int k : kv.second) t.usrc_list.push_back(k);
      t.usrc_ptr.push_back((int)t.usrc_list.size());
    }
    std::map<int64_t, std::vector<int>> gr;
    for (int k = 0; k < sh.nrec; k++) {
      int64_t c = linear(sh.rec_idx, sh.nrec, k, "Receiver", s);
      t.rec_cell.push_back(c);
      gr[c].push_back(k);
    }
    for (auto &kv : gr) {
      t.urec_cell.push_back(kv.first);
      for (int k : kv.second) t.urec_list.push_back(k);
      t.urec_ptr.push_back((int)t.urec_list.size());
    }
    auto rowptr = [&](std::vector<int> &dst, const std::vector<int64_t> &cells, int c0) {
      // dst[r] = first unique cell of this shot whose row is >= r
      int c = c0;
      for (int64_t r = 0; r <= nrow; r++) {
        while (c < (int)cells.size() && cells[c] / grid[0] < r) c++;
        dst.push_back(c);
      }
    };
    rowptr(t.srow, t.usrc_cell, t.usrc_off.back());
    rowptr(t.rrow, t.urec_cell, t.urec_off.back());
    t.usrc_off.push_back((int)t.usrc_cell.size());
    t.urec_off.push_back((int)t.urec_cell.size());
    t.rec_off.push_back((int)t.rec_cell.size());
  }
  t.total_tf = otf;
  t.total_tr = otr;
}

void build_point_tables(PointTables &t, int ndim, const int64_t *grid, int esize, int nshots,
                        const swx_shot_desc *shots, hipStream_t stream) {
  HostPointTables h;
  host_point_tables(h, ndim, grid, nshots, shots);
  t.nshots = nshots;
  t.nt = h.nt;
  t.nsrc = h.nsrc;
  t.nrec = h.nrec;
  t.off_tf = h.off_tf;
  t.off_tr = h.off_tr;
  t.nrows = h.nrows;
  t.total_tf = h.total_tf;
  t.total_tr = h.total_tr;
  t.h_rec_cell = h.rec_cell;
  auto up = [&](DevBuf &b, const void *src, size_t bytes) {
    b.alloc(std::max<size_t>(bytes, 16));
    if (bytes) SWX_HIP(hipMemcpyAsync(b.get(), src, bytes, hipMemcpyHostToDevice, stream));
  };
  up(t.d_nt, h.nt.data(), sizeof(int) * nshots);
  up(t.d_off_tf, h.off_tf.data(), sizeof(int64_t) * nshots);
  up(t.d_off_tr, h.off_tr.data(), sizeof(int64_t) * nshots);
  up(t.d_usrc_off, h.usrc_off.data(), sizeof(int) * h.usrc_off.size());
  up(t.d_usrc_cell, h.usrc_cell.data(), sizeof(int64_t) * h.usrc_cell.size());
  up(t.d_usrc_ptr, h.usrc_ptr.data(), sizeof(int) * h.usrc_ptr.size());
  up(t.d_usrc_list, h.usrc_list.data(), sizeof(int) * h.usrc_list.size());
  up(t.d_urec_off, h.urec_off.data(), sizeof(int) * h.urec_off.size());
  up(t.d_urec_cell, h.urec_cell.data(), sizeof(int64_t) * h.urec_cell.size());
  up(t.d_urec_ptr, h.urec_ptr.data(), sizeof(int) * h.urec_ptr.size());
  up(t.d_urec_list, h.urec_list.data(), sizeof(int) * h.urec_list.size());
  up(t.d_rec_off, h.rec_off.data(), sizeof(int) * h.rec_off.size());
  up(t.d_rec_cell, h.rec_cell.data(), sizeof(int64_t) * h.rec_cell.size());
  up(t.d_srow, h.srow.data(), sizeof(int) * h.srow.size());
  up(t.d_rrow, h.rrow.data(), sizeof(int) * h.rrow.size());
  // source time functions: concatenated per shot
  std::vector<char> tf((size_t)h.total_tf * esize);
  for (int s = 0; s < nshots; s++)
    std::memcpy(tf.data() + (size_t)h.off_tf[s] * esize, shots[s].srctf,
                (size_t)shots[s].nt * shots[s].nsrc * esize);
  up(t.d_srctf, tf.data(), tf.size());
  t.d_traces.alloc(std::max<size_t>((size_t)h.total_tr * esize, 16));
  t.d_res.alloc(std::max<size_t>((size_t)h.total_tr * esize, 16));
  SWX_HIP(hipMemsetAsync(t.d_traces.get(), 0, (size_t)h.total_tr * esize, stream));
  SWX_HIP(hipStreamSynchronize(stream));
}

}  // namespace swx

using namespace swx;

namespace {
// After a rejected swx_set_shots / swx_set_shots_elastic the plan holds no
// batch: the engine's tables may be half rebuilt, so swx_forward and
// swx_adjoint must refuse to run (SWX_ESTATE) until a batch is set again.
void forget_shots(PlanCore &c) {
  c.pts.nshots = 0;
  c.forward_done = false;
  c.traces_valid = false;
  c.obs_set = false;
  c.snap_every = 0;
  c.snap_host = nullptr;
}

// Adjoint loop of the engine (residuals already in pts.d_res), timed, then
// the gradient copied to the host.
void run_adjoint(swx_plan *plan, void *grad, int per_shot) {
  PlanCore &c = plan->core;
  hipEvent_t e0, e1;
  SWX_HIP(hipEventCreate(&e0));
  SWX_HIP(hipEventCreate(&e1));
  c.ktimer.reset();
  SWX_HIP(hipEventRecord(e0, c.stream));
  plan->engine->adjoint(per_shot != 0);
  SWX_HIP(hipEventRecord(e1, c.stream));
  SWX_HIP(hipStreamSynchronize(c.stream));
  float ms = 0;
  SWX_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  c.ms_adjoint = ms;
  if (c.ktimer.enabled) c.ktimer.collect(c.ms_kernel, c.n_kernel);
  plan->engine->copy_grad(grad, per_shot != 0);
}
}  // namespace

extern "C" {

int swx_abi_version(void) { return SWX_ABI_VERSION; }

const char *swx_last_error(void) { return g_last_error.c_str(); }

int swx_plan_create(const swx_plan_desc *desc, swx_plan **out) {
  return guarded([&] {
    SWX_REQUIRE(desc && out, SWX_EINVAL, "null argument");
    *out = nullptr;
    const swx_plan_desc &d = *desc;
    validate_plan_desc(d);
    SWX_HIP(hipSetDevice(d.device));
    auto p = std::make_unique<swx_plan>();
    p->core.d = d;
    p->core.esize = d.dtype == SWX_F32 ? 4 : 8;
    SWX_HIP(hipStreamCreateWithFlags(&p->core.stream, hipStreamNonBlocking));
    size_t fr = 0, tot = 0;
    SWX_HIP(hipMemGetInfo(&fr, &tot));
    p->core.budget = d.mem_budget > 0 ? (size_t)d.mem_budget : (size_t)(0.9 * (double)fr);
    if ((d.physics == SWX_ACOUSTIC_CD || d.physics == SWX_ACOUSTIC_VD) && d.ndim == 1) {
      p->engine = make_acoustic_1d(p->core);
    } else if (d.physics == SWX_ACOUSTIC_CD && d.ndim == 2) {
      p->engine = make_acoustic_cd2d(p->core);
    } else if (d.physics == SWX_ACOUSTIC_CD && d.ndim == 3) {
      p->engine = make_acoustic_cd3d(p->core);
    } else if (d.physics == SWX_ACOUSTIC_VD && d.ndim == 2) {
      p->engine = make_acoustic_vd2d(p->core);
    } else if (d.physics == SWX_ELASTIC_ISO && d.ndim == 2) {
      p->engine = make_elastic2d(p->core);
    } else {
      throw Error(SWX_EUNSUPPORTED, "physics/dimension combination not supported by this build");
    }
    *out = p.release();
  });
}

int swx_plan_destroy(swx_plan *plan) {
  return guarded([&] {
    if (!plan) return;
    (void)hipSetDevice(plan->core.d.device);
    (void)hipStreamSynchronize(plan->core.stream);
    plan->engine.reset();
    hipStream_t s = plan->core.stream;
    delete plan;
    if (s) (void)hipStreamDestroy(s);
  });
}

int swx_set_model(swx_plan *plan, int nfields, const void *const *fields) {
  return guarded([&] {
    SWX_REQUIRE(plan && fields, SWX_EINVAL, "null argument");
    SWX_HIP(hipSetDevice(plan->core.d.device));
    plan->engine->set_model(nfields, fields);
  });
}

int swx_set_shots(swx_plan *plan, int nshots, const swx_shot_desc *shots) {
  return guarded([&] {
    SWX_REQUIRE(plan && shots, SWX_EINVAL, "null argument");
    SWX_REQUIRE(nshots >= 1 && nshots <= plan->core.d.max_shots, SWX_EINVAL,
                "number of shots must be in 1..max_shots");
    const swx_plan_desc &d = plan->core.d;
    for (int s = 0; s < nshots; s++)
      SWX_REQUIRE(shots[s].nt >= 1 && shots[s].nt <= d.nt_max, SWX_EINVAL,
                  "shot nt must be in 1..nt_max");
    SWX_REQUIRE(d.physics != SWX_ELASTIC_ISO, SWX_EINVAL, "elastic plans take swx_set_shots_elastic");
    SWX_HIP(hipSetDevice(d.device));
    forget_shots(plan->core);
    try {
      PointTables tables;
      build_point_tables(tables, d.ndim, d.n, plan->core.esize, nshots, shots, plan->core.stream);
      plan->core.pts = std::move(tables);
      plan->engine->on_shots_set(nshots, shots);
    } catch (...) {
      forget_shots(plan->core);
      throw;
    }
  });
}

int swx_set_shots_elastic(swx_plan *plan, int nshots, const swx_elastic_shot_desc *shots) {
  return guarded([&] {
    SWX_REQUIRE(plan && shots, SWX_EINVAL, "null argument");
    SWX_REQUIRE(nshots >= 1 && nshots <= plan->core.d.max_shots, SWX_EINVAL,
                "number of shots must be in 1..max_shots");
    const swx_plan_desc &d = plan->core.d;
    for (int s = 0; s < nshots; s++)
      SWX_REQUIRE(shots[s].nt >= 1 && shots[s].nt <= d.nt_max, SWX_EINVAL, "shot nt must be in 1..nt_max");
    SWX_HIP(hipSetDevice(d.device));
    forget_shots(plan->core);
    try {
      plan->engine->on_elastic_shots(nshots, shots);
    } catch (...) {
      forget_shots(plan->core);
      throw;
    }
  });
}

int swx_forward(swx_plan *plan, void *traces, int save_state) {
  return guarded([&] {
    SWX_REQUIRE(plan, SWX_EINVAL, "null plan");
    PlanCore &c = plan->core;
    SWX_REQUIRE(c.pts.nshots > 0, SWX_ESTATE, "swx_set_shots must be called before swx_forward");
    SWX_REQUIRE(!save_state || c.d.gradient, SWX_ESTATE, "plan was not created for gradients");
    SWX_HIP(hipSetDevice(c.d.device));
    hipEvent_t e0, e1;
    SWX_HIP(hipEventCreate(&e0));
    SWX_HIP(hipEventCreate(&e1));
    c.ktimer.reset();
    SWX_HIP(hipEventRecord(e0, c.stream));
    plan->engine->forward(save_state != 0);
    SWX_HIP(hipEventRecord(e1, c.stream));
    SWX_HIP(hipStreamSynchronize(c.stream));
    float ms = 0;
    SWX_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    c.ms_forward = ms;
    if (c.ktimer.enabled) c.ktimer.collect(c.ms_kernel, c.n_kernel);
    if (traces)
      SWX_HIP(hipMemcpy(traces, c.pts.d_traces.get(), (size_t)c.pts.total_tr * c.esize, hipMemcpyDeviceToHost));
    c.forward_done = save_state != 0;
    c.traces_valid = true;
  });
}

int swx_adjoint(swx_plan *plan, const void *adjsrc, void *grad, int per_shot) {
  return guarded([&] {
    SWX_REQUIRE(plan && adjsrc && grad, SWX_EINVAL, "null argument");
    PlanCore &c = plan->core;
    SWX_REQUIRE(c.d.gradient, SWX_ESTATE, "plan was not created for gradients");
    SWX_REQUIRE(c.forward_done, SWX_ESTATE, "swx_forward(save_state=1) must precede swx_adjoint");
    SWX_HIP(hipSetDevice(c.d.device));
    SWX_HIP(hipMemcpyAsync(c.pts.d_res.get(), adjsrc, (size_t)c.pts.total_tr * c.esize, hipMemcpyHostToDevice,
                           c.stream));
    run_adjoint(plan, grad, per_shot);
  });
}

int swx_set_snapshots(swx_plan *plan, int snapevery, void *snaps, int64_t bytes) {
  return guarded([&] {
    SWX_REQUIRE(plan, SWX_EINVAL, "null plan");
    PlanCore &c = plan->core;
    SWX_REQUIRE(snapevery >= 0, SWX_EINVAL, "snapevery must be non-negative");
    c.snap_every = 0;
    c.snap_host = nullptr;
    if (snapevery == 0) return;
    SWX_REQUIRE(c.pts.nshots > 0, SWX_ESTATE, "swx_set_shots must be called before swx_set_snapshots");
    SWX_REQUIRE(snaps, SWX_EINVAL, "null snapshot buffer");
    SnapPiece pc[8];
    const int n = plan->engine->snap_fields(1, 0, pc);
    SWX_REQUIRE(n > 0, SWX_EUNSUPPORTED, "snapshots are not provided for this physics");
    int64_t one = 0;
    for (int k = 0; k < n; k++) one += pc[k].width * pc[k].rows;
    std::vector<int64_t> off(c.pts.nshots);
    int64_t tot = 0;
    for (int s = 0; s < c.pts.nshots; s++) {
      SWX_REQUIRE(snapevery < c.pts.nt[s], SWX_EINVAL,
                  "Checkpointing frequency must be smaller than the number of timesteps!");
      off[s] = tot;
      tot += (int64_t)(c.pts.nt[s] / snapevery) * one;
    }
    SWX_REQUIRE(bytes >= tot, SWX_EINVAL, "snapshot buffer too small: need " + std::to_string(tot) + " bytes");
    c.snap_one = one;
    c.snap_off = off;
    c.snap_host = static_cast<char *>(snaps);
    c.snap_every = snapevery;
  });
}

int swx_set_observed(swx_plan *plan, const void *observed, const double *weights) {
  return guarded([&] {
    SWX_REQUIRE(plan && observed, SWX_EINVAL, "null argument");
    PlanCore &c = plan->core;
    SWX_REQUIRE(c.pts.nshots > 0, SWX_ESTATE, "swx_set_shots must be called before swx_set_observed");
    SWX_HIP(hipSetDevice(c.d.device));
    set_observed(c, observed, weights);
  });
}

int swx_misfit_l2(swx_plan *plan, double *misfit) {
  return guarded([&] {
    SWX_REQUIRE(plan && misfit, SWX_EINVAL, "null argument");
    PlanCore &c = plan->core;
    SWX_REQUIRE(c.pts.nshots > 0 && c.traces_valid, SWX_ESTATE, "swx_forward must precede swx_misfit_l2");
    SWX_REQUIRE(c.obs_set, SWX_ESTATE, "swx_set_observed must precede swx_misfit_l2");
    SWX_HIP(hipSetDevice(c.d.device));
    l2_adjoint_source(c, misfit);
    SWX_HIP(hipStreamSynchronize(c.stream));
  });
}

int swx_adjoint_l2(swx_plan *plan, void *grad, int per_shot, double *misfit) {
  return guarded([&] {
    SWX_REQUIRE(plan && grad, SWX_EINVAL, "null argument");
    PlanCore &c = plan->core;
    SWX_REQUIRE(c.d.gradient, SWX_ESTATE, "plan was not created for gradients");
    SWX_REQUIRE(c.forward_done, SWX_ESTATE, "swx_forward(save_state=1) must precede swx_adjoint_l2");
    SWX_REQUIRE(c.obs_set, SWX_ESTATE, "swx_set_observed must precede swx_adjoint_l2");
    SWX_HIP(hipSetDevice(c.d.device));
    l2_adjoint_source(c, misfit);
    run_adjoint(plan, grad, per_shot);
  });
}

int swx_plan_get_info(swx_plan *plan, swx_plan_info *info) {
  return guarded([&] {
    SWX_REQUIRE(plan && info, SWX_EINVAL, "null argument");
    std::memset(info, 0, sizeof(*info));
    info->bytes_state = plan->engine->bytes_state();
    info->bytes_checkpoints = plan->engine->bytes_checkpoints();
    info->bytes_replay = plan->engine->bytes_replay();
    info->segment = plan->engine->segment();
    info->nshots = plan->core.pts.nshots;
    info->ms_forward = plan->core.ms_forward;
    info->ms_adjoint = plan->core.ms_adjoint;
    for (int k = 0; k < 2; k++) {
      info->ms_kernel[k] = plan->core.ms_kernel[k];
      info->n_kernel[k] = plan->core.n_kernel[k];
    }
    plan->engine->variants(info->kernel_variant);
  });
}

int swx_set_kernel_timing(swx_plan *plan, int enable) {
  return guarded([&] {
    SWX_REQUIRE(plan, SWX_EINVAL, "null plan");
    plan->core.ktimer.enabled = enable != 0;
  });
}

}  // extern "C"
