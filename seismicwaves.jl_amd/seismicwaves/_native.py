"""ctypes binding of libswx.so (C ABI declared in include/swx.h).

The product path always runs through this library: there is no CPU
fallback.  If the shared library is missing, importing this module raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# SWX_LIBRARY selects another build of the library (e.g. libswx_debug.so)
LIBPATH = os.environ.get("SWX_LIBRARY") or os.path.join(_HERE, "libswx.so")

SWX_OK, SWX_EINVAL, SWX_ENOMEM, SWX_EHIP, SWX_ESTATE, SWX_EUNSUPPORTED = range(6)
SWX_ACOUSTIC_CD, SWX_ACOUSTIC_VD, SWX_ELASTIC_ISO = 1, 2, 3
SWX_F32, SWX_F64 = 1, 2
SWX_SRC_MOMENT_TENSOR, SWX_SRC_EXTERNAL_FORCE = 1, 2

# Every symbol the header declares (checked by the CPU test suite).
EXPORTS = (
    "swx_abi_version", "swx_last_error", "swx_plan_create", "swx_plan_destroy",
    "swx_set_model", "swx_set_shots", "swx_set_shots_elastic", "swx_forward", "swx_adjoint",
    "swx_plan_get_info", "swx_set_kernel_timing", "swx_forward_noalloc",
    "swx_set_observed", "swx_misfit_l2", "swx_adjoint_l2", "swx_set_snapshots",
)
# Bytes of the static device array swx_forward_noalloc works in.
NOALLOC_SCRATCH_BYTES = 64 << 20


class PlanDesc(ctypes.Structure):
    _fields_ = [
        ("physics", ctypes.c_int32), ("dtype", ctypes.c_int32), ("ndim", ctypes.c_int32),
        ("halo", ctypes.c_int32), ("n", ctypes.c_int64 * 3), ("h", ctypes.c_double * 3),
        ("dt", ctypes.c_double), ("freetop", ctypes.c_int32), ("max_shots", ctypes.c_int32),
        ("nt_max", ctypes.c_int32), ("check_freq", ctypes.c_int32), ("gradient", ctypes.c_int32),
        ("device", ctypes.c_int32), ("mem_budget", ctypes.c_double),
    ]


class ShotDesc(ctypes.Structure):
    _fields_ = [
        ("nt", ctypes.c_int32), ("nsrc", ctypes.c_int32), ("src_idx", ctypes.c_void_p),
        ("srctf", ctypes.c_void_p), ("nrec", ctypes.c_int32), ("rec_idx", ctypes.c_void_p),
        ("cpml", (ctypes.c_void_p * 4) * 3),
    ]


class Taps(ctypes.Structure):
    _fields_ = [("ptr", ctypes.c_void_p), ("ij", ctypes.c_void_p), ("coef", ctypes.c_void_p)]


class ElasticShotDesc(ctypes.Structure):
    _fields_ = [
        ("nt", ctypes.c_int32), ("source_kind", ctypes.c_int32), ("nsrc", ctypes.c_int32),
        ("src", Taps * 2), ("moments", ctypes.c_void_p), ("srctf", ctypes.c_void_p),
        ("nrec", ctypes.c_int32), ("rec", Taps * 2), ("cpml", (ctypes.c_void_p * 4) * 3),
    ]


class PlanInfo(ctypes.Structure):
    _fields_ = [
        ("bytes_state", ctypes.c_int64), ("bytes_checkpoints", ctypes.c_int64),
        ("bytes_replay", ctypes.c_int64), ("segment", ctypes.c_int32), ("nshots", ctypes.c_int32),
        ("ms_forward", ctypes.c_double), ("ms_adjoint", ctypes.c_double),
        ("ms_kernel", ctypes.c_double * 2), ("n_kernel", ctypes.c_int64 * 2),
        ("kernel_variant", ctypes.c_int32 * 2),
    ]


class SwxError(RuntimeError):
    """Failure reported by libswx (HIP error or out-of-order call)."""


_lib = None


def load(path: str | None = None):
    """Load libswx.so once; raise ImportError if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    p = path or LIBPATH
    if not os.path.exists(p):
        raise ImportError(f"libswx.so not found at {p}: build it with `python seismicwaves.jl_amd/build.py`")
    lib = ctypes.CDLL(p)
    vp = ctypes.c_void_p
    lib.swx_abi_version.restype = ctypes.c_int
    lib.swx_last_error.restype = ctypes.c_char_p
    lib.swx_plan_create.argtypes = [ctypes.POINTER(PlanDesc), ctypes.POINTER(vp)]
    lib.swx_plan_destroy.argtypes = [vp]
    lib.swx_set_model.argtypes = [vp, ctypes.c_int, ctypes.POINTER(vp)]
    lib.swx_set_shots.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ShotDesc)]
    lib.swx_set_shots_elastic.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ElasticShotDesc)]
    lib.swx_forward.argtypes = [vp, vp, ctypes.c_int]
    lib.swx_adjoint.argtypes = [vp, vp, vp, ctypes.c_int]
    lib.swx_plan_get_info.argtypes = [vp, ctypes.POINTER(PlanInfo)]
    lib.swx_set_kernel_timing.argtypes = [vp, ctypes.c_int]
    lib.swx_forward_noalloc.argtypes = [ctypes.POINTER(PlanDesc), vp, ctypes.POINTER(ShotDesc), vp]
    lib.swx_set_observed.argtypes = [vp, vp, vp]
    lib.swx_misfit_l2.argtypes = [vp, vp]
    lib.swx_adjoint_l2.argtypes = [vp, vp, ctypes.c_int, vp]
    lib.swx_set_snapshots.argtypes = [vp, ctypes.c_int, vp, ctypes.c_int64]
    for name in EXPORTS:
        getattr(lib, name).restype = ctypes.c_int if name != "swx_last_error" else ctypes.c_char_p
    _lib = lib
    return lib


def check(rc: int):
    if rc == SWX_OK:
        return
    msg = load().swx_last_error().decode(errors="replace")
    if rc == SWX_EINVAL:
        raise AssertionError(msg)           # the reference's @assert failures
    if rc == SWX_ENOMEM:
        raise MemoryError(msg)
    if rc == SWX_EUNSUPPORTED:
        raise NotImplementedError(msg)
    raise SwxError(f"libswx error {rc}: {msg}")


def ptr(a: np.ndarray | None):
    return ctypes.c_void_p(0) if a is None else ctypes.c_void_p(a.ctypes.data)


def plan_desc(*, physics, dtype, n, h, dt, halo, freetop, max_shots, nt_max, check_freq=1, gradient=False,
              device=0, mem_budget=0.0) -> PlanDesc:
    d = PlanDesc()
    d.physics = physics
    d.dtype = SWX_F32 if np.dtype(dtype) == np.float32 else SWX_F64
    d.ndim = len(n)
    d.halo = int(halo)
    for k in range(3):
        d.n[k] = int(n[k]) if k < len(n) else 1
        d.h[k] = float(h[k]) if k < len(h) else 1.0
    d.dt = float(dt)
    d.freetop = int(bool(freetop))
    d.max_shots = int(max_shots)
    d.nt_max = int(nt_max)
    d.check_freq = int(check_freq)
    d.gradient = int(bool(gradient))
    d.device = int(device)
    d.mem_budget = float(mem_budget)
    return d


def shot_descs(shots, dtype):
    """ShotDesc array for dicts with nt, src_idx (nsrc, ndim) 1-based, srctf
    (nt, nsrc), rec_idx (nrec, ndim), cpml [axis] -> dict a, a_h, b, b_h;
    returns (descs, arrays that must outlive the call)."""
    keep = []
    descs = (ShotDesc * len(shots))()
    for k, sh in enumerate(shots):
        si = np.asfortranarray(sh["src_idx"], dtype=np.int64)
        ri = np.asfortranarray(sh["rec_idx"], dtype=np.int64)
        tf = np.asfortranarray(sh["srctf"], dtype=dtype)
        keep += [si, ri, tf]
        dd = descs[k]
        dd.nt = int(sh["nt"])
        dd.nsrc = si.shape[0]
        dd.src_idx = si.ctypes.data
        dd.srctf = tf.ctypes.data
        dd.nrec = ri.shape[0]
        dd.rec_idx = ri.ctypes.data
        for ax, co in enumerate(sh["cpml"]):
            for j, key in enumerate(("a", "a_h", "b", "b_h")):
                arr = np.ascontiguousarray(co[key], dtype=dtype)
                keep.append(arr)
                dd.cpml[ax][j] = arr.ctypes.data if arr.size else None
    return descs, keep


def forward_noalloc(*, dtype, n, h, dt, halo, fact, shot, device=0):
    """swx_forward_noalloc: the forward of one 2D acoustic CD shot (a dict as
    for Plan.set_shots) on the model factor `fact` (dt^2 vp^2), run in the
    library's static scratch array; returns the traces (nt, nrec)."""
    d = plan_desc(physics=SWX_ACOUSTIC_CD, dtype=dtype, n=n, h=h, dt=dt, halo=halo, freetop=False, max_shots=1,
                  nt_max=int(shot["nt"]), device=device)
    descs, keep = shot_descs([shot], dtype)
    f = np.asfortranarray(fact, dtype=dtype)
    nt, nrec = int(shot["nt"]), np.atleast_2d(shot["rec_idx"]).shape[0]
    traces = np.zeros((nt, nrec), dtype=dtype, order="F")
    check(load().swx_forward_noalloc(ctypes.byref(d), ptr(f), descs, ptr(traces)))
    del keep
    return np.ascontiguousarray(traces)


class Plan:
    """Owner of one swx_plan (device state for a batch of shots)."""

    def __init__(self, *, physics, dtype, n, h, dt, halo, freetop, max_shots, nt_max,
                 check_freq=1, gradient=False, device=0, mem_budget=0.0):
        lib = load()
        d = plan_desc(physics=physics, dtype=dtype, n=n, h=h, dt=dt, halo=halo, freetop=freetop,
                      max_shots=max_shots, nt_max=nt_max, check_freq=check_freq, gradient=gradient, device=device,
                      mem_budget=mem_budget)
        self.dtype = np.dtype(dtype)
        self.ndim = len(n)
        self._h = ctypes.c_void_p()
        check(lib.swx_plan_create(ctypes.byref(d), ctypes.byref(self._h)))
        self._keep = []

    def close(self):
        if self._h:
            load().swx_plan_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_model(self, fields):
        arrs = [np.asfortranarray(f, dtype=self.dtype) for f in fields]
        ptrs = (ctypes.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])
        check(load().swx_set_model(self._h, len(arrs), ptrs))

    def set_shots(self, shots):
        """shots: list of dicts with nt, src_idx (nsrc, ndim) 1-based, srctf
        (nt, nsrc), rec_idx (nrec, ndim), cpml [axis] -> dict a, a_h, b, b_h."""
        descs, keep = shot_descs(shots, self.dtype)
        check(load().swx_set_shots(self._h, len(shots), descs))
        self._keep = keep

    def _cpml(self, dd, cpml, keep):
        for ax, co in enumerate(cpml):
            for j, key in enumerate(("a", "a_h", "b", "b_h")):
                arr = np.ascontiguousarray(co[key], dtype=self.dtype)
                keep.append(arr)
                dd.cpml[ax][j] = arr.ctypes.data if arr.size else None

    def _taps(self, t, fam, keep):
        """fam: list over points of (ij (ntaps, 2) 1-based, coef (ntaps,))."""
        ptr_ = np.zeros(len(fam) + 1, dtype=np.int32)
        for k, (ij, c) in enumerate(fam):
            ptr_[k + 1] = ptr_[k] + len(c)
        ij = np.ascontiguousarray(np.concatenate([np.asarray(f[0], dtype=np.int64).reshape(-1, 2) for f in fam]))
        coef = np.ascontiguousarray(np.concatenate([np.asarray(f[1]) for f in fam]), dtype=self.dtype)
        keep += [ptr_, ij, coef]
        t.ptr, t.ij, t.coef = ptr_.ctypes.data, ij.ctypes.data, coef.ctypes.data

    def set_shots_elastic(self, shots):
        """shots: list of dicts with nt, kind (SWX_SRC_*), src (two tap
        families), moments (nsrc, 3) for moment tensors, srctf ((nt, nsrc) or
        (nt, 2, nsrc), divided by dx dz), rec (ux and uz tap families), cpml."""
        keep = []
        descs = (ElasticShotDesc * len(shots))()
        for k, sh in enumerate(shots):
            dd = descs[k]
            dd.nt = int(sh["nt"])
            dd.source_kind = int(sh["kind"])
            dd.nsrc = len(sh["src"][0])
            for f in range(2):
                self._taps(dd.src[f], sh["src"][f], keep)
                self._taps(dd.rec[f], sh["rec"][f], keep)
            if sh.get("moments") is not None:
                mom = np.ascontiguousarray(sh["moments"], dtype=self.dtype)
                keep.append(mom)
                dd.moments = mom.ctypes.data
            tf = np.asfortranarray(sh["srctf"], dtype=self.dtype)
            keep.append(tf)
            dd.srctf = tf.ctypes.data
            dd.nrec = len(sh["rec"][0])
            self._cpml(dd, sh["cpml"], keep)
        check(load().swx_set_shots_elastic(self._h, len(shots), descs))
        self._keep = keep

    def forward(self, traces: np.ndarray | None, save_state: bool):
        check(load().swx_forward(self._h, ptr(traces), int(save_state)))

    def adjoint(self, adjsrc: np.ndarray, grad: np.ndarray, per_shot: bool):
        check(load().swx_adjoint(self._h, ptr(adjsrc), ptr(grad), int(per_shot)))

    def set_snapshots(self, snapevery: int, buf: np.ndarray | None):
        """Snapshots of the next forward runs into buf (host, the layout of
        swx_set_snapshots); snapevery = 0 turns them off."""
        if snapevery == 0:
            check(load().swx_set_snapshots(self._h, 0, ctypes.c_void_p(0), 0))
            return
        self._snapbuf = buf
        check(load().swx_set_snapshots(self._h, int(snapevery), ptr(buf), int(buf.nbytes)))

    def set_observed(self, observed: np.ndarray, weights: np.ndarray | None = None):
        """Observed data of an L2 misfit (traces layout), optional float64
        weights of the same layout (window mask x diagonal inverse
        covariance); stays on the device for this batch."""
        obs = np.ascontiguousarray(observed, dtype=self.dtype)
        w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
        if w is not None:
            assert w.size == obs.size
        check(load().swx_set_observed(self._h, ptr(obs), ptr(w)))

    def misfit_l2(self, nshots: int) -> np.ndarray:
        out = np.zeros(nshots, dtype=np.float64)
        check(load().swx_misfit_l2(self._h, ptr(out)))
        return out

    def adjoint_l2(self, grad: np.ndarray, per_shot: bool, nshots: int) -> np.ndarray:
        """Adjoint with the L2 adjoint source formed on the device; returns
        the per-shot misfits."""
        mis = np.zeros(nshots, dtype=np.float64)
        check(load().swx_adjoint_l2(self._h, ptr(grad), int(per_shot), ptr(mis)))
        return mis

    def info(self) -> PlanInfo:
        inf = PlanInfo()
        check(load().swx_plan_get_info(self._h, ctypes.byref(inf)))
        return inf

    def set_kernel_timing(self, on: bool):
        check(load().swx_set_kernel_timing(self._h, int(on)))
