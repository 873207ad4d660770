"""Benchmark: FP32 forward + adjoint gradient throughput on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

A step is one full gradient of the workload (forward with device-resident
checkpoints, host adjoint source, adjoint with replay and correlation),
all inputs uploaded to HBM before the timed region.  The printed JSON line
carries the throughput in GCell-updates/s (2 * cells * sum of shot nt per
step: one forward plus one adjoint update per cell and time step; replay is
not counted as work), the roofline fraction of the dominant stencil kernel
timed with HIP events on its own stream, and a CPU baseline (the oracle's
OpenMP restatement of the reference's threads backend on a bounded sample).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
PKG_ROOT = os.path.join(REPO, "seismicwaves.jl_amd")
for _p in (REPO, PKG_ROOT):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np  # noqa: E402

METRIC = "GCell-updates/s fwd+adjoint (batched shots) & % HBM roofline vs plain 1-shot path"
HBM_PEAK_GBS = 8000.0


# ---------------------------------------------------------------------------
# Workloads
# ---------------------------------------------------------------------------
def workload_cd2d_grad_1000(dtype=np.float32):
    """2D acoustic constant-density gradient: 1000 x 1000, dh = 5 m,
    dt = 0.5 ms, nt = 4000, seeded 8-layer model, start model smoothed with
    sigma = 100 m, one 12 Hz Ricker shot at (2500 m, 10 m), 400 receivers at
    10 m depth, checkpoint every 50 steps with recompute."""
    import seismicwaves as sw
    from seismicwaves import synthetic as syn
    nx = nz = 1000
    h = (5.0, 5.0)
    dt, nt, f0, halo = 5e-4, 4000, 12.0, 20
    rng = np.random.default_rng(1)
    vtrue = syn.layered(nx, nz, h, rng)
    v0 = syn.gaussian_smooth(vtrue, 100.0 / h[0])
    tf = syn.ricker_tf(nt, dt, f0, 1.2 / f0)[:, None]
    rec = syn.surface_receivers(400, nx, h, halo, 10.0)
    params = sw.InputParametersAcoustic(nt, dt, (nx, nz), h, sw.CPMLBoundaryConditionParameters(halo=halo))
    shots = [sw.ScalarShot(sw.ScalarSources(np.array([[2500.0, 10.0]]), tf.astype(dtype), f0),
                           sw.ScalarReceivers(rec, nt, dtype=dtype))]
    return dict(name="acoustic_cd_2d_gradient_1000x1000_nt4000", params=params,
                mp_true=sw.VpAcousticCDMaterialProperties(vtrue.astype(dtype)),
                mp0=sw.VpAcousticCDMaterialProperties(v0.astype(dtype)), shots=shots, check_freq=50,
                physics="cd2d", config={"nx": nx, "nz": nz, "dh_m": 5.0, "dt_s": dt, "nt": nt, "shots": 1,
                                        "receivers": 400, "check_freq": 50, "halo": halo})


WORKLOADS = {"cd2d_grad_1000": workload_cd2d_grad_1000}


# ---------------------------------------------------------------------------
# Algorithmic bytes of the dominant (stencil) kernel
# ---------------------------------------------------------------------------
def stencil_alg_bytes_cd2d(nx, nz, halo, e, nts, S):
    """Bytes the 2D CD stencil launches of one gradient must move, each
    array counted once per launch: forward r p^n, r p^{n-1}, r fact,
    w p^{n+1} (4e per cell); + w D when the replay slot is stored (5e);
    adjoint with fused correlation adds r D and a grad read-modify-write
    (7e).  C-PML strips: psi read+write and xi read+write per strip entry
    (4e per strip cell and axis)."""
    N = nx * nz
    strip = (2 * halo * nz + 2 * halo * nx) * 4 * e
    total = 0
    for nt in nts:
        kl = (nt - 1) // S
        last = nt - kl * S                     # steps of the last segment (stored in the forward)
        fwd = (nt - last) * 4 * e * N + last * 5 * e * N
        rep = (nt - last) * 5 * e * N          # replay of every other segment
        nseg = kl + 1
        adj = (nt - nseg) * 7 * e * N + nseg * 4 * e * N
        total += fwd + rep + adj + (nt + (nt - last) + nt) * strip
    return total


# ---------------------------------------------------------------------------
# CPU baseline: oracle OpenMP restatement of the reference's threads backend
# ---------------------------------------------------------------------------
def cpu_baseline(wl, nt_sample=400):
    import ctypes
    import oracle
    oracle.build()
    cores = int(os.environ.get("OMP_NUM_THREADS", len(os.sched_getaffinity(0))))
    os.environ["OMP_NUM_THREADS"] = str(cores)
    p = wl["params"]
    nx, nz = p.gridsize
    h = p.gridspacing
    dt = float(p.dt)
    vp = np.asarray(wl["mp0"].vp, dtype=np.float64)
    sh = wl["shots"][0]
    f0 = float(sh.srcs.domfreq)
    T = np.float32
    cp = oracle.init_bdc(vp, dt, h, p.boundcond.halo, p.boundcond.rcoef, p.boundcond.freeboundtop, f0, T)
    si = oracle.find_nearest_grid_points(sh.srcs.positions, h)
    ri = oracle.find_nearest_grid_points(sh.recs.positions, h)
    stf = oracle.scale_srctf_cd(sh.srcs.tf[:nt_sample], h, dt, vp, si, T)
    fact = ((T(dt) ** 2) * vp.astype(T) ** 2).astype(T)
    adj = np.random.default_rng(0).standard_normal((nt_sample, ri.shape[0])).astype(T) * T(1e-3)
    t0 = time.perf_counter()
    oracle.cd2d_gradient(fact, h, dt, p.boundcond.halo, cp, nt_sample, si, stf, ri, adj, wl["check_freq"], T)
    el = time.perf_counter() - t0
    return {"value": 2.0 * nx * nz * nt_sample / el / 1e9, "unit": "GCell-updates/s", "cores": cores,
            "kind": "port",
            "sample": f"FP32 gradient of the same model, 1 shot, nt={nt_sample} (of {p.ntimesteps}), "
                      f"check_freq={wl['check_freq']}, OpenMP threads={cores}; restatement of the reference's "
                      f"threads backend, not the Julia binary; {el:.2f} s"}


# ---------------------------------------------------------------------------
def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="cd2d_grad_1000", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU shot sharding is a design note only "
                                   "(see the sharding section of the design document)", "n_gpus": args.gpus}))
        sys.exit(2)

    import seismicwaves as sw
    from seismicwaves import _native

    wl = WORKLOADS[args.workload]()
    p = wl["params"]
    dtype = np.float32
    # observed data from the true model (outside the timed region)
    sw.swforward(p, wl["mp_true"], wl["shots"], runparams=sw.RunParameters())
    misfits = [sw.L2Misfit(sh.recs.seismograms.copy()) for sh in wl["shots"]]
    ws = sw.build_wavesim(p, wl["mp0"], runparams=sw.RunParameters(),
                          gradparams=sw.GradParameters(check_freq=wl["check_freq"]), gradient=True)
    ws.set_matprop(wl["mp0"])
    descs = [ws.shot_desc(sh) for sh in wl["shots"]]
    plan = ws._ensure_plan(len(descs))
    plan.set_shots(descs)
    nts = [d["nt"] for d in descs]
    ntr = sum(d["nt"] * d["rec_idx"].shape[0] for d in descs)
    traces = np.zeros(ntr, dtype=dtype)
    grad = np.zeros(int(np.prod(p.gridsize)), dtype=dtype)

    def one_step():
        plan.forward(traces, True)
        off, adj = 0, []
        for d, m in zip(descs, misfits):
            nt, nr = d["nt"], d["rec_idx"].shape[0]
            syn_ = traces[off:off + nt * nr].reshape((nt, nr), order="F")
            r = syn_.astype(np.float64) - m.observed
            adj.append(np.asfortranarray(-r, dtype=dtype).ravel(order="F"))
            off += nt * nr
        plan.adjoint(np.concatenate(adj), grad, False)
        inf = plan.info()
        return inf.ms_forward + inf.ms_adjoint

    for _ in range(args.warmup):
        one_step()
    # clean timing (no per-kernel events)
    t0 = time.perf_counter()
    dev_ms = 0.0
    for _ in range(args.steps):
        dev_ms += one_step()
    wall = time.perf_counter() - t0
    ms_step = wall / args.steps * 1e3
    # dominant-kernel timing: HIP events around every stencil launch on the
    # plan's stream, over the same number of steps
    plan.set_kernel_timing(True)
    st_ms, st_n = 0.0, 0
    for _ in range(args.steps):
        plan.forward(traces, True)
        i1 = plan.info()
        st_ms += i1.ms_stencil
        st_n += i1.n_stencil
        plan.adjoint(np.zeros(ntr, dtype=dtype), grad, False)
        i2 = plan.info()
        st_ms += i2.ms_stencil
        st_n += i2.n_stencil
    plan.set_kernel_timing(False)
    info = plan.info()
    cells = int(np.prod(p.gridsize))
    work = 2.0 * cells * sum(nts)
    value = work / (ms_step / 1e3) / 1e9
    nx, nz = p.gridsize
    alg = stencil_alg_bytes_cd2d(nx, nz, p.boundcond.halo, 4, nts, info.segment)
    launches_per_step = st_n / args.steps
    avg_launch_ms = st_ms / st_n
    alg_per_launch = alg / launches_per_step
    achieved = alg_per_launch / (avg_launch_ms / 1e3) / 1e9
    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "GCell-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded layered model, Ricker sources)",
        "config": {"workload": wl["name"], **wl["config"], "segment_used": info.segment},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "cd2d_step_kernel<float>", "avg_launch_us": round(avg_launch_ms * 1e3, 3),
                     "alg_bytes_per_launch": int(alg_per_launch), "launches_per_step": launches_per_step},
        "device_ms_per_step": round(dev_ms / args.steps, 3),
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(wl)
        out["speedup_vs_cpu"] = round(value / out["cpu_baseline"]["value"], 2)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
