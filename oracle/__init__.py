"""CPU oracle for the SeismicWaves.jl finite-difference forward/adjoint path.

TEST INFRASTRUCTURE ONLY.  This package is the checker: tests/, the smoke()
entry of __graft_entry__ and the cpu_baseline leg of bench.py may import it;
the product library (seismicwaves.jl_amd) never does.

It has two halves, both restating the reference from source:

* host-side preparation in numpy (C-PML coefficients, nearest-grid-point
  lookup, source scaling, L2 misfit and its adjoint source, gradient
  post-processing), each function citing the reference file:line it follows;
* the time loops in C (swx_oracle.c / oracle_impl.h, built by oracle/Makefile
  into liboracle.so), called through ctypes.

Parity pinning of every function is listed in `oracle/PINNING.md`.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile (gcc, no GPU)."""
    so = os.path.join(_HERE, "liboracle.so")
    if force or not os.path.exists(so):
        subprocess.run(["make", "-B" if force else "-s", "-C", _HERE, "liboracle.so"],
                       check=True, stdout=subprocess.DEVNULL)
    return so


def lib():
    global _LIB
    if _LIB is None:
        so = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(so):
            build()
        _LIB = ctypes.CDLL(so)
    return _LIB


def with_snapshots(every, nt, shapes, dtype, run):
    """Runs run() -- one oracle forward loop of nt steps -- with the snapshot
    sink of swx_oracle.c on (LinearSnapshotter, snapshotter.jl:1-41): after
    every step it with it % every == 0 the fields of `shapes` (the
    reference's snapshotted fields in its order) are kept.  Returns
    {it: [arrays]}."""
    count = nt // every
    sizes = [int(np.prod(sh)) for sh in shapes]
    buf = np.zeros(max(count * sum(sizes), 1), dtype=dtype)
    lib().oracle_set_snapshots(ctypes.c_int(every), ctypes.c_void_p(buf.ctypes.data))
    try:
        run()
    finally:
        lib().oracle_set_snapshots(ctypes.c_int(0), ctypes.c_void_p(0))
    out, off = {}, 0
    for k in range(1, count + 1):
        arrs = []
        for sh, n in zip(shapes, sizes):
            arrs.append(buf[off:off + n].reshape(sh, order="F").copy())
            off += n
        out[k * every] = arrs
    return out


# ---------------------------------------------------------------------------
# ctypes structures mirroring swx_oracle.c
# ---------------------------------------------------------------------------
class Geom(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64 * 3), ("h", ctypes.c_double * 3),
                ("dt", ctypes.c_double), ("halo", ctypes.c_int), ("freetop", ctypes.c_int)]


class _Cpml(ctypes.Structure):
    _fields_ = [("a", ctypes.c_void_p), ("a_h", ctypes.c_void_p),
                ("b", ctypes.c_void_p), ("b_h", ctypes.c_void_p)]


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else ctypes.c_void_p(0)


def _geom(n, h, dt, halo, freetop=0):
    g = Geom()
    for d in range(3):
        g.n[d] = int(n[d]) if d < len(n) else 1
        g.h[d] = float(h[d]) if d < len(h) else 1.0
    g.dt = float(dt)
    g.halo = int(halo)
    g.freetop = int(freetop)
    return g


def _cpml_struct(coeffs, dtype):
    arrs = [np.ascontiguousarray(coeffs[k], dtype=dtype) for k in ("a", "a_h", "b", "b_h")]
    st = _Cpml(*[_ptr(a) for a in arrs])
    return st, arrs


# ---------------------------------------------------------------------------
# Host-side restatements (numpy)
# ---------------------------------------------------------------------------
def rickerstf(t, t0, f0):
    """src/utils/utils.jl:6 (Ricker wavelet)."""
    arg = (np.pi * f0 * (t - t0)) ** 2
    return (1.0 - 2.0 * arg) * np.exp(-arg)


def gaussderivstf(t, t0, f0):
    """src/utils/utils.jl:13."""
    return (t - t0) * np.exp(-(np.pi * f0 * (t - t0)) ** 2)


def calc_kab_cpml_axis(halo, dt, npower, d0, alpha_max, onwhere):
    """calc_Kab_CPML_axis, src/models/cpmlcoeffs.jl:43-100 (K_max = nothing)."""
    size = halo
    if onwhere == "halfgrd":
        shift = 0.5
    elif onwhere == "ongrd":
        shift = 0.0
        size += 1
    else:
        raise ValueError("Wrong onwhere parameter!")
    dist = np.linspace(0 + shift, size + shift - 1, size) if size > 1 else np.array([0 + shift] * size)
    if halo != 0:
        nleft = dist[::-1] / halo
        nright = dist / halo
    else:
        nleft = dist[::-1]
        nright = dist

    def side(nd):
        d = d0 * nd ** npower
        alpha = alpha_max * (1.0 - nd)
        b = np.exp(-(d / 1.0 + alpha) * dt)
        with np.errstate(invalid="ignore", divide="ignore"):
            a = d * (b - 1.0) / (1.0 * (d + 1.0 * alpha))
        return a, b

    al, bl = side(nleft)
    ar, br = side(nright)
    return al, ar, bl, br


def cpml_coefficients_axis(vel_max, dt, halo, rcoef, thickness, f0, dtype=np.float64):
    """compute_CPML_coefficientsAxis!, src/models/cpmlcoeffs.jl:18-41.

    Scalar inputs are first rounded to `dtype` as the reference converts them
    to T; the profile itself is evaluated in Float64 (LinRange) and rounded.
    """
    T = dtype
    alpha_max = T(np.pi * T(f0))
    npower = T(2.0)
    d0 = T(-(npower + 1) * T(vel_max) * np.log(T(rcoef)) / (2.0 * T(thickness))) if halo != 0 else T(0.0)
    a_l, a_r, b_l, b_r = calc_kab_cpml_axis(halo, float(T(dt)), float(npower), float(d0), float(alpha_max), "ongrd")
    ah_l, ah_r, bh_l, bh_r = calc_kab_cpml_axis(halo, float(T(dt)), float(npower), float(d0), float(alpha_max), "halfgrd")
    return {
        "a": np.concatenate([a_l, a_r]).astype(T),
        "a_h": np.concatenate([ah_l, ah_r]).astype(T),
        "b": np.concatenate([b_l, b_r]).astype(T),
        "b_h": np.concatenate([bh_l, bh_r]).astype(T),
    }


def init_bdc(vp, dt, spacing, halo, rcoef, freeboundtop, f0, dtype=np.float64):
    """init_bdc! for CPML, src/models/acoustic/acou_init_bc.jl:7-40."""
    vel_max = float(np.max(vp))
    out = []
    for h in spacing:
        out.append(cpml_coefficients_axis(vel_max, dt, halo, rcoef, h * halo, f0, dtype))
    if freeboundtop and len(out) >= 1:
        last = out[-1]
        for k, v in (("a", 0.0), ("a_h", 0.0), ("b", 1.0), ("b_h", 1.0)):
            last[k][: len(last[k]) // 2] = v
    return out


def find_nearest_grid_points(positions, spacing):
    """src/utils/utils.jl:47-58: round(pos/h + 1, RoundNearestTiesUp), 1-based."""
    positions = np.atleast_2d(np.asarray(positions, dtype=np.float64))
    idx = np.floor(positions / np.asarray(spacing, dtype=np.float64)[None, :] + 1 + 0.5)
    return idx.astype(np.int64)


def scale_srctf_cd(tf, spacing, dt, vp, src_idx, dtype=np.float64):
    """possrcrec_scaletf for CD, src/models/acoustic/acou_forward.jl:6-20."""
    T = dtype
    tf = np.asarray(tf, dtype=T)
    out = tf / T(np.prod(np.asarray(spacing, dtype=T))) * (T(dt) ** 2)
    vp = np.asarray(vp, dtype=T)
    for s in range(out.shape[1]):
        out[:, s] *= vp[tuple(src_idx[s] - 1)] ** 2
    return out.astype(T)


def l2_misfit(syn, obs, invcov=None, windows=()):
    """calcmisfit, src/inversion/misfits/L2Misfit.jl:24-42."""
    res = np.asarray(syn, dtype=np.float64) - np.asarray(obs, dtype=np.float64)
    if windows:
        mask = np.zeros(res.shape[0])
        for a, b in windows:
            mask[a - 1:b] = 1.0
        res = mask.reshape((-1,) + (1,) * (res.ndim - 1)) * res
    if invcov is None:
        return 0.5 * float(np.sum(res * res))
    # vector receivers: sum over components (L2Misfit.jl:44-62)
    return 0.5 * float(np.sum(res * np.tensordot(np.asarray(invcov), res, axes=(1, 0))))


def l2_dchi_du(syn, obs, invcov=None, windows=()):
    """∂χ_∂u, src/inversion/misfits/L2Misfit.jl:64-95 (per component for 3D data)."""
    res = np.asarray(syn, dtype=np.float64) - np.asarray(obs, dtype=np.float64)
    if windows:
        mask = np.zeros(res.shape[0])
        for a, b in windows:
            mask[a - 1:b] = 1.0
        res = mask.reshape((-1,) + (1,) * (res.ndim - 1)) * res
    if invcov is None:
        return res
    return np.tensordot(np.asarray(invcov), res, axes=(1, 0))


def mute_around_points(grad, points, spacing, radius):
    """mutearoundmultiplepoints!, src/utils/mute_grad.jl:3-80."""
    if radius == 0:
        return grad
    grad = grad.copy()
    n = grad.shape
    N = len(n)
    extent = [(n[d] - 1) * spacing[d] for d in range(N)]
    rmax = radius * max(spacing)
    for pt in np.atleast_2d(points):
        lo, hi = [], []
        for d in range(N):
            if not (0.0 <= pt[d] <= extent[d]):
                raise ValueError("mutearoundpoint!(): The point lies outside the grid")
            ijk = int(np.floor(np.floor(pt[d] / spacing[d]))) + 1
            lo.append(ijk - radius)
            hi.append(ijk + radius)
        for idx in np.ndindex(*[hi[d] - lo[d] + 1 for d in range(N)]):
            cell = [lo[d] + idx[d] for d in range(N)]
            if not all(1 <= cell[d] <= n[d] for d in range(N)):
                continue
            xyz = [(cell[d] - 1) * spacing[d] for d in range(N)]
            r = np.sqrt(sum((pt[d] - xyz[d]) ** 2 for d in range(N)))
            if r <= rmax:
                grad[tuple(c - 1 for c in cell)] *= r / rmax
    return grad


# ---------------------------------------------------------------------------
# 2D acoustic constant density: time loops in C
# ---------------------------------------------------------------------------
def _fn(name, dtype):
    return getattr(lib(), name + ("_f64" if dtype == np.float64 else "_f32"))


def _f_order(a, dtype):
    return np.asfortranarray(np.asarray(a, dtype=dtype))


def cd2d_forward(fact, spacing, dt, halo, cpml, nt, src_idx, srctf, rec_idx, dtype=np.float64):
    """Run swforward_1shot! (acou_forward.jl:22-62) on prepared inputs.

    fact: (nx, nz) = dt^2 vp^2; cpml: list of two coefficient dicts (x, z);
    src_idx/rec_idx: 1-based (n, 2) integer arrays; srctf: scaled (nt, nsrc).
    Returns traces (nt, nrec).
    """
    fact = _f_order(fact, dtype)
    nx, nz = fact.shape
    g = _geom((nx, nz), spacing, dt, halo)
    cx, kx = _cpml_struct(cpml[0], dtype)
    cz, kz = _cpml_struct(cpml[1], dtype)
    src = np.asfortranarray(src_idx, dtype=np.int64)
    rec = np.asfortranarray(rec_idx, dtype=np.int64)
    stf = _f_order(srctf, dtype)
    traces = np.zeros((nt, rec.shape[0]), dtype=dtype, order="F")
    rc = _fn("oracle_cd2d_forward", dtype)(
        ctypes.byref(g), _ptr(fact), ctypes.byref(cx), ctypes.byref(cz), ctypes.c_int(nt),
        ctypes.c_int(src.shape[0]), _ptr(src), _ptr(stf), ctypes.c_int(rec.shape[0]), _ptr(rec), _ptr(traces))
    assert rc == 0
    return np.ascontiguousarray(traces)


def cd2d_gradient(fact, spacing, dt, halo, cpml, nt, src_idx, srctf, rec_idx, adjsrc,
                  check_freq, dtype=np.float64):
    """Run swgradient_1shot! (acou_gradient.jl:4-95) up to the raw grad_vp
    field (before muting and the 2/vp^3 rescale).  adjsrc = -dchi/du.
    Returns (raw_grad (nx, nz), traces (nt, nrec))."""
    fact = _f_order(fact, dtype)
    nx, nz = fact.shape
    g = _geom((nx, nz), spacing, dt, halo)
    cx, kx = _cpml_struct(cpml[0], dtype)
    cz, kz = _cpml_struct(cpml[1], dtype)
    src = np.asfortranarray(src_idx, dtype=np.int64)
    rec = np.asfortranarray(rec_idx, dtype=np.int64)
    stf = _f_order(srctf, dtype)
    adj = _f_order(adjsrc, dtype)
    traces = np.zeros((nt, rec.shape[0]), dtype=dtype, order="F")
    grad = np.zeros((nx, nz), dtype=dtype, order="F")
    rc = _fn("oracle_cd2d_gradient", dtype)(
        ctypes.byref(g), _ptr(fact), ctypes.byref(cx), ctypes.byref(cz), ctypes.c_int(nt),
        ctypes.c_int(src.shape[0]), _ptr(src), _ptr(stf), ctypes.c_int(rec.shape[0]), _ptr(rec),
        _ptr(adj), ctypes.c_int(check_freq), _ptr(grad), _ptr(traces))
    if rc != 0:
        raise ValueError("Checkpointing frequency must be smaller than the number of timesteps!")
    return np.array(grad), np.ascontiguousarray(traces)


# ---------------------------------------------------------------------------
# Whole-API restatements (numpy inputs) used by the parity tests
# ---------------------------------------------------------------------------
def acoustic_cd_forward(vp, spacing, dt, nt, halo, rcoef, freetop, shots, dtype=np.float64):
    """run_swforward! (src/apis/forward.jl:71-115) for CD 2D.

    shots: list of dicts {src_pos (nsrc,2), srctf (nt,nsrc), f0, rec_pos (nrec,2)}.
    Returns a list of seismograms (nt, nrec)."""
    T = dtype
    vp = np.asarray(vp, dtype=T)
    fact = (T(dt) ** 2) * vp ** 2                                  # acou_models.jl:60
    out = []
    for sh in shots:
        cpml = init_bdc(vp, dt, spacing, halo, rcoef, freetop, sh["f0"], T)
        si = find_nearest_grid_points(sh["src_pos"], spacing)
        ri = find_nearest_grid_points(sh["rec_pos"], spacing)
        stf = scale_srctf_cd(sh["srctf"], spacing, dt, vp, si, T)
        out.append(cd2d_forward(fact, spacing, dt, halo, cpml, nt, si, stf, ri, T))
    return out


def acoustic_cd_gradient(vp, spacing, dt, nt, halo, rcoef, freetop, shots, check_freq=1,
                         mute_radius_src=0, mute_radius_rec=0, dtype=np.float64):
    """run_swgradient! (src/apis/gradient.jl:93-136) for CD 2D with L2 misfits.

    Each shot dict also carries `observed` (nt, nrec) and optionally `invcov`.
    Returns (gradient dict {"vp": ...}, total misfit, list of seismograms)."""
    T = dtype
    vp = np.asarray(vp, dtype=T)
    fact = (T(dt) ** 2) * vp ** 2
    tot = np.zeros(vp.shape, dtype=T)
    misfit = 0.0
    seis = []
    for sh in shots:
        cpml = init_bdc(vp, dt, spacing, halo, rcoef, freetop, sh["f0"], T)
        si = find_nearest_grid_points(sh["src_pos"], spacing)
        ri = find_nearest_grid_points(sh["rec_pos"], spacing)
        stf = scale_srctf_cd(sh["srctf"], spacing, dt, vp, si, T)
        syn = cd2d_forward(fact, spacing, dt, halo, cpml, nt, si, stf, ri, T)
        adj = -l2_dchi_du(syn, sh["observed"], sh.get("invcov"))          # acou_gradient.jl:48
        g, tr = cd2d_gradient(fact, spacing, dt, halo, cpml, nt, si, stf, ri, adj.astype(T), check_freq, T)
        g = mute_around_points(g, sh["src_pos"], spacing, mute_radius_src)
        g = mute_around_points(g, sh["rec_pos"], spacing, mute_radius_rec)
        g = (T(2.0) / vp ** 3) * g                                        # acou_gradient.jl:93
        tot += g
        misfit += l2_misfit(tr, sh["observed"], sh.get("invcov"))
        seis.append(tr)
    return {"vp": tot}, misfit, seis


# ---------------------------------------------------------------------------
# 2D acoustic variable density (staggered)
# ---------------------------------------------------------------------------
def interp_stag(m, dim, method="arithmetic"):
    """ArithmeticAverageInterpolation / HarmonicAverageInterpolation at the
    half points along `dim` (1-based), src/utils/interpolations.jl:35-43."""
    ax = dim - 1
    n = m.shape[ax]
    lo = np.take(m, np.arange(n - 1), axis=ax)
    hi = np.take(m, np.arange(1, n), axis=ax)
    if method == "arithmetic":
        return 0.5 * lo + 0.5 * hi
    return 1.0 / (0.5 / lo + 0.5 / hi)


def back_interp(m, g, dim, method="arithmetic"):
    """back_interp, interpolations.jl:15-28 with ∂f∂m :45-47."""
    ax = dim - 1
    n = m.shape[ax]
    out = np.zeros(m.shape)
    lo = np.take(m, np.arange(n - 1), axis=ax)
    hi = np.take(m, np.arange(1, n), axis=ax)
    if method == "arithmetic":
        wl = wh = 0.5
    else:
        itp = interp_stag(m, dim, method)
        wl, wh = itp ** 2 / lo ** 2 / 2, itp ** 2 / hi ** 2 / 2
    idx_lo = [slice(None)] * m.ndim
    idx_hi = [slice(None)] * m.ndim
    idx_lo[ax] = slice(0, n - 1)
    idx_hi[ax] = slice(1, n)
    out[tuple(idx_lo)] += g * wl
    out[tuple(idx_hi)] += g * wh
    return out


def vd_model_fields(vp, rho, dt, method="arithmetic", dtype=np.float64):
    """precompute_fact! for VD, acou_models.jl:292-300."""
    T = dtype
    vp = np.asarray(vp, dtype=T)
    rho = np.asarray(rho, dtype=T)
    m0 = vp ** 2 * rho * T(dt)
    inv = T(1) / rho
    return m0, (interp_stag(inv, 1, method) * T(dt)).astype(T), (interp_stag(inv, 2, method) * T(dt)).astype(T)


def scale_srctf_vd(tf, spacing, dt, vp, rho, src_idx, dtype=np.float64):
    """possrcrec_scaletf for VD, acou_forward.jl:67-81."""
    T = dtype
    out = np.asarray(tf, dtype=T) / T(np.prod(np.asarray(spacing, dtype=T))) * T(dt)
    for s in range(out.shape[1]):
        ix = tuple(src_idx[s] - 1)
        out[:, s] *= T(vp[ix]) ** 2 * T(rho[ix])
    return out.astype(T)


def vd2d_forward(m0, m1x, m1z, spacing, dt, halo, cpml, nt, src_idx, srctf, rec_idx, dtype=np.float64):
    m0 = _f_order(m0, dtype)
    m1x = _f_order(m1x, dtype)
    m1z = _f_order(m1z, dtype)
    nx, nz = m0.shape
    g = _geom((nx, nz), spacing, dt, halo)
    cx, kx = _cpml_struct(cpml[0], dtype)
    cz, kz = _cpml_struct(cpml[1], dtype)
    src = np.asfortranarray(src_idx, dtype=np.int64)
    rec = np.asfortranarray(rec_idx, dtype=np.int64)
    stf = _f_order(srctf, dtype)
    traces = np.zeros((nt, rec.shape[0]), dtype=dtype, order="F")
    rc = _fn("oracle_vd2d_forward", dtype)(
        ctypes.byref(g), _ptr(m0), _ptr(m1x), _ptr(m1z), ctypes.byref(cx), ctypes.byref(cz), ctypes.c_int(nt),
        ctypes.c_int(src.shape[0]), _ptr(src), _ptr(stf), ctypes.c_int(rec.shape[0]), _ptr(rec), _ptr(traces))
    assert rc == 0
    return np.ascontiguousarray(traces)


def vd2d_gradient(m0, m1x, m1z, spacing, dt, halo, cpml, nt, src_idx, srctf, rec_idx, adjsrc, check_freq,
                  dtype=np.float64):
    """Raw grad_m0, grad_m1_stag (x, z) and the traces of swgradient_1shot! (VD)."""
    m0 = _f_order(m0, dtype)
    m1x = _f_order(m1x, dtype)
    m1z = _f_order(m1z, dtype)
    nx, nz = m0.shape
    g = _geom((nx, nz), spacing, dt, halo)
    cx, kx = _cpml_struct(cpml[0], dtype)
    cz, kz = _cpml_struct(cpml[1], dtype)
    src = np.asfortranarray(src_idx, dtype=np.int64)
    rec = np.asfortranarray(rec_idx, dtype=np.int64)
    stf = _f_order(srctf, dtype)
    adj = _f_order(adjsrc, dtype)
    traces = np.zeros((nt, rec.shape[0]), dtype=dtype, order="F")
    g0 = np.zeros((nx, nz), dtype=dtype, order="F")
    g1x = np.zeros((nx - 1, nz), dtype=dtype, order="F")
    g1z = np.zeros((nx, nz - 1), dtype=dtype, order="F")
    rc = _fn("oracle_vd2d_gradient", dtype)(
        ctypes.byref(g), _ptr(m0), _ptr(m1x), _ptr(m1z), ctypes.byref(cx), ctypes.byref(cz), ctypes.c_int(nt),
        ctypes.c_int(src.shape[0]), _ptr(src), _ptr(stf), ctypes.c_int(rec.shape[0]), _ptr(rec), _ptr(adj),
        ctypes.c_int(check_freq), _ptr(g0), _ptr(g1x), _ptr(g1z), _ptr(traces))
    if rc != 0:
        raise ValueError("Checkpointing frequency must be smaller than the number of timesteps!")
    return (np.array(g0), np.array(g1x), np.array(g1z)), np.ascontiguousarray(traces)


def acoustic_vd_forward(vp, rho, spacing, dt, nt, halo, rcoef, freetop, shots, method="arithmetic",
                        dtype=np.float64):
    """run_swforward! for VD 2D; returns the seismograms of each shot."""
    T = dtype
    m0, m1x, m1z = vd_model_fields(vp, rho, dt, method, T)
    out = []
    for sh in shots:
        nt_s = sh.get("nt", nt)
        cpml = init_bdc(vp, dt, spacing, halo, rcoef, freetop, sh["f0"], T)
        si = find_nearest_grid_points(sh["src_pos"], spacing)
        ri = find_nearest_grid_points(sh["rec_pos"], spacing)
        stf = scale_srctf_vd(sh["srctf"][:nt_s], spacing, dt, vp, rho, si, T)
        out.append(vd2d_forward(m0, m1x, m1z, spacing, dt, halo, cpml, nt_s, si, stf, ri, T))
    return out


def acoustic_vd_gradient(vp, rho, spacing, dt, nt, halo, rcoef, freetop, shots, check_freq=1,
                         mute_radius_src=0, mute_radius_rec=0, method="arithmetic", dtype=np.float64):
    """run_swgradient! for VD 2D (gradient.jl:93-136, acou_gradient.jl:97-203).
    Returns ({"vp", "rho"}, misfit, seismograms)."""
    T = dtype
    vp = np.asarray(vp, dtype=T)
    rho = np.asarray(rho, dtype=T)
    m0, m1x, m1z = vd_model_fields(vp, rho, dt, method, T)
    gvp = np.zeros(vp.shape, dtype=T)
    grho = np.zeros(vp.shape, dtype=T)
    misfit, seis = 0.0, []
    for sh in shots:
        nt_s = sh.get("nt", nt)
        cpml = init_bdc(vp, dt, spacing, halo, rcoef, freetop, sh["f0"], T)
        si = find_nearest_grid_points(sh["src_pos"], spacing)
        ri = find_nearest_grid_points(sh["rec_pos"], spacing)
        stf = scale_srctf_vd(sh["srctf"][:nt_s], spacing, dt, vp, rho, si, T)
        syn = vd2d_forward(m0, m1x, m1z, spacing, dt, halo, cpml, nt_s, si, stf, ri, T)
        adj = -l2_dchi_du(syn, sh["observed"], sh.get("invcov"))
        (g0, g1x, g1z), tr = vd2d_gradient(m0, m1x, m1z, spacing, dt, halo, cpml, nt_s, si, stf, ri,
                                           adj.astype(T), min(check_freq, nt_s - 1), T)
        inv = 1.0 / rho.astype(np.float64)
        g1 = back_interp(inv, g1x, 1, method) + back_interp(inv, g1z, 2, method)
        g0 = g0.astype(np.float64)
        g0 = mute_around_points(mute_around_points(g0, sh["src_pos"], spacing, mute_radius_src),
                                sh["rec_pos"], spacing, mute_radius_rec)
        g1 = mute_around_points(mute_around_points(g1, sh["src_pos"], spacing, mute_radius_src),
                                sh["rec_pos"], spacing, mute_radius_rec)
        gvp += (-2.0 * g0 / (vp ** 3 * rho)).astype(T)
        grho += (-g0 / (vp ** 2 * rho ** 2) - g1 / rho).astype(T)
        misfit += l2_misfit(tr, sh["observed"], sh.get("invcov"))
        seis.append(tr)
    return {"vp": gvp, "rho": grho}, misfit, seis


# ---------------------------------------------------------------------------
# 3D acoustic constant density
# ---------------------------------------------------------------------------
class _Cpml3(ctypes.Structure):
    _fields_ = [("ax", _Cpml * 3)]


def _cpml3(cpml, dtype):
    keep = []
    st = (_Cpml * 3)()
    for d in range(3):
        s, arrs = _cpml_struct(cpml[d], dtype)
        st[d] = s
        keep += arrs
    return st, keep


def cd3d_forward(fact, spacing, dt, halo, cpml, nt, src_idx, srctf, rec_idx, dtype=np.float64):
    """swforward_1shot! for 3D CD (acoustic3D_xPU.jl) on prepared inputs."""
    fact = _f_order(fact, dtype)
    g = _geom(fact.shape, spacing, dt, halo)
    c, keep = _cpml3(cpml, dtype)
    src = np.asfortranarray(src_idx, dtype=np.int64)
    rec = np.asfortranarray(rec_idx, dtype=np.int64)
    stf = _f_order(srctf, dtype)
    traces = np.zeros((nt, rec.shape[0]), dtype=dtype, order="F")
    rc = _fn("oracle_cd3d_forward", dtype)(
        ctypes.byref(g), _ptr(fact), c, ctypes.c_int(nt), ctypes.c_int(src.shape[0]), _ptr(src), _ptr(stf),
        ctypes.c_int(rec.shape[0]), _ptr(rec), _ptr(traces))
    assert rc == 0
    return np.ascontiguousarray(traces)


def cd3d_gradient(fact, spacing, dt, halo, cpml, nt, src_idx, srctf, rec_idx, adjsrc, check_freq, dtype=np.float64):
    fact = _f_order(fact, dtype)
    g = _geom(fact.shape, spacing, dt, halo)
    c, keep = _cpml3(cpml, dtype)
    src = np.asfortranarray(src_idx, dtype=np.int64)
    rec = np.asfortranarray(rec_idx, dtype=np.int64)
    stf = _f_order(srctf, dtype)
    adj = _f_order(adjsrc, dtype)
    traces = np.zeros((nt, rec.shape[0]), dtype=dtype, order="F")
    grad = np.zeros(fact.shape, dtype=dtype, order="F")
    rc = _fn("oracle_cd3d_gradient", dtype)(
        ctypes.byref(g), _ptr(fact), c, ctypes.c_int(nt), ctypes.c_int(src.shape[0]), _ptr(src), _ptr(stf),
        ctypes.c_int(rec.shape[0]), _ptr(rec), _ptr(adj), ctypes.c_int(check_freq), _ptr(grad), _ptr(traces))
    if rc != 0:
        raise ValueError("Checkpointing frequency must be smaller than the number of timesteps!")
    return np.array(grad), np.ascontiguousarray(traces)


def acoustic_cd3d_forward(vp, spacing, dt, nt, halo, rcoef, freetop, shots, dtype=np.float64):
    """run_swforward! for CD 3D."""
    T = dtype
    vp = np.asarray(vp, dtype=T)
    fact = (T(dt) ** 2) * vp ** 2
    out = []
    for sh in shots:
        nt_s = sh.get("nt", nt)
        cpml = init_bdc(vp, dt, spacing, halo, rcoef, freetop, sh["f0"], T)
        si = find_nearest_grid_points(sh["src_pos"], spacing)
        ri = find_nearest_grid_points(sh["rec_pos"], spacing)
        stf = scale_srctf_cd(sh["srctf"][:nt_s], spacing, dt, vp, si, T)
        out.append(cd3d_forward(fact, spacing, dt, halo, cpml, nt_s, si, stf, ri, T))
    return out


def acoustic_cd3d_gradient(vp, spacing, dt, nt, halo, rcoef, freetop, shots, check_freq=1, dtype=np.float64):
    """run_swgradient! for CD 3D; returns ({"vp"}, misfit, seismograms)."""
    T = dtype
    vp = np.asarray(vp, dtype=T)
    fact = (T(dt) ** 2) * vp ** 2
    tot = np.zeros(vp.shape, dtype=T)
    misfit, seis = 0.0, []
    for sh in shots:
        nt_s = sh.get("nt", nt)
        cpml = init_bdc(vp, dt, spacing, halo, rcoef, freetop, sh["f0"], T)
        si = find_nearest_grid_points(sh["src_pos"], spacing)
        ri = find_nearest_grid_points(sh["rec_pos"], spacing)
        stf = scale_srctf_cd(sh["srctf"][:nt_s], spacing, dt, vp, si, T)
        syn = cd3d_forward(fact, spacing, dt, halo, cpml, nt_s, si, stf, ri, T)
        adj = -l2_dchi_du(syn, sh["observed"], sh.get("invcov"))
        g, tr = cd3d_gradient(fact, spacing, dt, halo, cpml, nt_s, si, stf, ri, adj.astype(T),
                              min(check_freq, nt_s - 1), T)
        tot += (T(2.0) / vp ** 3) * g
        misfit += l2_misfit(tr, sh["observed"], sh.get("invcov"))
        seis.append(tr)
    return {"vp": tot}, misfit, seis


# ---------------------------------------------------------------------------
# 2D isotropic elastic
# ---------------------------------------------------------------------------
def kaiser(x, r, beta):
    """Kaiser window, src/utils/utils.jl:71 (besseli(0, .) = modified Bessel I0)."""
    if -r <= x <= r:
        return float(np.i0(beta * np.sqrt(1.0 - (x / r) ** 2)) / np.i0(beta))
    return 0.0


def modd(x, x0, r, beta, dx):
    """Band-limited delta, utils.jl:84-87: kaiser(x-x0, r dx, beta) sinc((x-x0)/dx)."""
    return kaiser(x - x0, r * dx, beta) * float(np.sinc((x - x0) / dx))


def coeffsinc1d(x0, dx, nx, r, beta, xstart, mirror, xbl, xbr):
    """coeffsinc1D, utils.jl:105-154, literally: taps i0-r-1..i0+r+1 around
    the nearest grid point, |coef| <= 1e-15 dropped, taps beyond [xbl, xbr]
    folded back (sign flipped when `mirror`), coincident taps summed.
    Returns (1-based indices, coefficients) sorted by index."""
    xs = xstart + dx * np.arange(nx)

    def findnearest(x):
        return int(np.argmin(np.abs(x - xs))) + 1

    i0 = findnearest(x0)
    idxs, coeffs = [], []
    for idx in range(i0 - r - 1, i0 + r + 2):
        xcurr = (idx - 1) * dx + xstart
        coe = modd(xcurr, x0, r, beta, dx)
        if not abs(coe) <= 1e-15:
            idxs.append(idx)
            coeffs.append(coe)
    acc = {}
    for idx, coe in zip(idxs, coeffs):
        xcurr = (idx - 1) * dx + xstart
        if xcurr < xbl:
            idx = findnearest(xbl + (xbl - xcurr))
            coe = -coe if mirror else coe
        elif xcurr > xbr:
            idx = findnearest(xbr - (xcurr - xbr))
            coe = -coe if mirror else coe
        acc[idx] = acc.get(idx, 0.0) + coe
    keys = sorted(acc)
    return np.array(keys, dtype=np.int64), np.array([acc[k] for k in keys])


def spread_positions(positions, spacing, gridsize, shift, mirror, r=4, beta=6.31):
    """spread_positions with freesurfposition=:ongridbound, utils.jl:168-214.
    Returns one (ij (ntaps, 2) 1-based, coef (ntaps,)) pair per position,
    tensor product with the first dimension fastest."""
    positions = np.atleast_2d(np.asarray(positions, dtype=np.float64))
    N = positions.shape[1]
    extent = [(gridsize[d] - 1) * spacing[d] for d in range(N)]
    out = []
    for p in positions:
        per = [coeffsinc1d(p[d], spacing[d], gridsize[d], r, beta, shift[d], mirror, 0.0, extent[d])
               for d in range(N)]
        ii, jj = np.meshgrid(per[0][0], per[1][0], indexing="ij")
        ci, cj = np.meshgrid(per[0][1], per[1][1], indexing="ij")
        ij = np.stack([ii.ravel(order="F"), jj.ravel(order="F")], 1)
        out.append((ij, (ci * cj).ravel(order="F")))
    return out


def el_possrcrec(kind, src_pos, rec_pos, spacing, gridsize, sincinterp=True):
    """possrcrec_scaletf point part, src/models/elastic/ela_models.jl:6-90.
    kind "mt": sources on sxx/szz (shift 0, mirror) and sxz (shift h/2, mirror);
    kind "ef": sources on ux (shift dx/2) and uz (shift dz/2).  Receivers on
    ux and uz.  Returns (src_fam0, src_fam1, rec_ux, rec_uz)."""
    dx, dz = spacing
    if not sincinterp:
        si = find_nearest_grid_points(src_pos, spacing)
        ri = find_nearest_grid_points(rec_pos, spacing)
        one = lambda idx: [(idx[k:k + 1].copy(), np.ones(1)) for k in range(idx.shape[0])]
        return one(si), one(si), one(ri), one(ri)
    if kind == "mt":
        s0 = spread_positions(src_pos, spacing, gridsize, (0.0, 0.0), True)
        s1 = spread_positions(src_pos, spacing, gridsize, (dx / 2, dz / 2), True)
    else:
        s0 = spread_positions(src_pos, spacing, gridsize, (dx / 2, 0.0), False)
        s1 = spread_positions(src_pos, spacing, gridsize, (0.0, dz / 2), False)
    r0 = spread_positions(rec_pos, spacing, gridsize, (dx / 2, 0.0), False)
    r1 = spread_positions(rec_pos, spacing, gridsize, (0.0, dz / 2), False)
    return s0, s1, r0, r1


def interp_dims(m, dims, method="arithmetic"):
    """interp with a list of dims, interpolations.jl:4-5 and :34-42: linear
    B-spline at the half points (mean of the 2^len(dims) corners)."""
    def avg(a):
        for d in dims:
            ax = d - 1
            n = a.shape[ax]
            a = (np.take(a, range(0, n - 1), axis=ax) + np.take(a, range(1, n), axis=ax)) / 2
        return a
    if method == "arithmetic":
        return avg(np.asarray(m, dtype=np.float64))
    return 1.0 / avg(1.0 / np.asarray(m, dtype=np.float64))


def back_interp_dims(m, g, dims, method="arithmetic"):
    """back_interp with a list of dims, interpolations.jl:15-28 and :45-47."""
    m = np.asarray(m, dtype=np.float64)
    res = np.zeros(m.shape)
    k = len(dims)
    if method == "arithmetic":
        w = np.full(g.shape, 1.0 / 2 ** k)
    else:
        itp = interp_dims(m, dims, method)
    for perm in range(2 ** k):
        sl = [slice(None)] * m.ndim
        for b, d in enumerate(dims):
            off = (perm >> (k - 1 - b)) & 1
            sl[d - 1] = slice(off, m.shape[d - 1] - 1 + off)
        if method != "arithmetic":
            w = itp ** 2 / m[tuple(sl)] ** 2 / 2 ** k
        res[tuple(sl)] += g * w
    return res


def el_model_fields(lam, mu, rho, interp_rho="arithmetic", interp_mu="arithmetic", dtype=np.float64):
    """precomp_elaprop!, ela_models.jl:156-173: lam, mu, rho at (i+1/2, j),
    rho at (i, j+1/2), mu at (i+1/2, j+1/2)."""
    T = dtype
    return (np.asarray(lam, dtype=T), np.asarray(mu, dtype=T),
            interp_dims(rho, [1], interp_rho).astype(T), interp_dims(rho, [2], interp_rho).astype(T),
            interp_dims(mu, [1, 2], interp_mu).astype(T))


class _Taps(ctypes.Structure):
    _fields_ = [("npts", ctypes.c_int), ("ptr", ctypes.c_void_p), ("ij", ctypes.c_void_p),
                ("ntaps", ctypes.c_int64), ("coef", ctypes.c_void_p)]


def _taps(fam, dtype):
    ptr = np.zeros(len(fam) + 1, dtype=np.int32)
    for k, (ij, c) in enumerate(fam):
        ptr[k + 1] = ptr[k] + len(c)
    ij = np.concatenate([f[0] for f in fam], 0).astype(np.int64)
    ij = np.asfortranarray(ij)
    coef = np.ascontiguousarray(np.concatenate([f[1] for f in fam]), dtype=dtype)
    st = _Taps(len(fam), _ptr(ptr), _ptr(ij), ij.shape[0], _ptr(coef))
    return st, (ptr, ij, coef)


def _check_taps(kind, fams, nx, nz):
    """Every tap lies on its family's staggered grid (the C loops do not check)."""
    grids = ([(nx, nz), (nx - 1, nz - 1)] if kind == "mt" else [(nx - 1, nz), (nx, nz - 1)]) + \
        [(nx - 1, nz), (nx, nz - 1)]
    for fam, (gi, gj) in zip(fams, grids):
        for ij, _ in fam:
            if ij.size and not (ij[:, 0].min() >= 1 and ij[:, 0].max() <= gi and ij[:, 1].min() >= 1
                                and ij[:, 1].max() <= gj):
                raise ValueError("point taps outside the staggered grid")


def _el_args(fields, spacing, dt, halo, freetop, cpml, dtype):
    lam, mu, rih, rjh, mhh = [_f_order(f, dtype) for f in fields]
    nx, nz = lam.shape
    g = _geom((nx, nz), spacing, dt, halo, freetop)
    cx, kx = _cpml_struct(cpml[0], dtype)
    cz, kz = _cpml_struct(cpml[1], dtype)
    keep = (lam, mu, rih, rjh, mhh, kx, kz, cx, cz)
    return g, keep


def el2d_forward(fields, spacing, dt, halo, freetop, cpml, nt, kind, fams, moments, srctf, dtype=np.float64):
    """Forward loop of swforward_1shot! (ela_forward.jl:4-159) on prepared
    inputs.  fields = el_model_fields(...); fams = el_possrcrec(...);
    srctf already divided by dx dz (MT (nt, nsrc), EF (nt, 2, nsrc)).
    Returns traces (nt, 2, nrec)."""
    g, keep = _el_args(fields, spacing, dt, halo, freetop, cpml, dtype)
    lam, mu, rih, rjh, mhh, _, _, cx, cz = keep
    _check_taps(kind, fams, *lam.shape)
    t = [_taps(f, dtype) for f in fams]
    M = np.ascontiguousarray(moments if moments is not None else np.zeros((1, 3)), dtype=dtype)
    stf = _f_order(srctf, dtype)
    nrec = len(fams[2])
    traces = np.zeros((nt, 2, nrec), dtype=dtype, order="F")
    rc = _fn("oracle_el2d_forward", dtype)(
        ctypes.byref(g), _ptr(lam), _ptr(mu), _ptr(rih), _ptr(rjh), _ptr(mhh), ctypes.byref(cx), ctypes.byref(cz),
        ctypes.c_int(nt), ctypes.c_int(0 if kind == "mt" else 1), ctypes.byref(t[0][0]), ctypes.byref(t[1][0]),
        _ptr(M), _ptr(stf), ctypes.byref(t[2][0]), ctypes.byref(t[3][0]), _ptr(traces))
    assert rc == 0
    return np.array(traces)


def el2d_gradient(fields, spacing, dt, halo, freetop, cpml, nt, kind, fams, moments, srctf, adjsrc, check_freq,
                  dtype=np.float64):
    """swgradient_1shot! (ela_gradient.jl:4-362) up to the raw grad fields.
    Returns ((g_rho_ih, g_rho_jh, g_lam, g_mu, g_mu_ihjh), traces)."""
    g, keep = _el_args(fields, spacing, dt, halo, freetop, cpml, dtype)
    lam, mu, rih, rjh, mhh, _, _, cx, cz = keep
    nx, nz = lam.shape
    _check_taps(kind, fams, nx, nz)
    t = [_taps(f, dtype) for f in fams]
    M = np.ascontiguousarray(moments if moments is not None else np.zeros((1, 3)), dtype=dtype)
    stf = _f_order(srctf, dtype)
    adj = _f_order(adjsrc, dtype)
    nrec = len(fams[2])
    traces = np.zeros((nt, 2, nrec), dtype=dtype, order="F")
    gr = [np.zeros(s, dtype=dtype, order="F") for s in ((nx - 1, nz), (nx, nz - 1), (nx, nz), (nx, nz),
                                                         (nx - 1, nz - 1))]
    rc = _fn("oracle_el2d_gradient", dtype)(
        ctypes.byref(g), _ptr(lam), _ptr(mu), _ptr(rih), _ptr(rjh), _ptr(mhh), ctypes.byref(cx), ctypes.byref(cz),
        ctypes.c_int(nt), ctypes.c_int(0 if kind == "mt" else 1), ctypes.byref(t[0][0]), ctypes.byref(t[1][0]),
        _ptr(M), _ptr(stf), ctypes.byref(t[2][0]), ctypes.byref(t[3][0]), _ptr(adj), ctypes.c_int(check_freq),
        *[_ptr(x) for x in gr], _ptr(traces))
    if rc != 0:
        raise ValueError("Checkpointing frequency must be smaller than the number of timesteps!")
    return tuple(np.array(x) for x in gr), np.array(traces)


def _el_prepare(lam, mu, rho, spacing, dt, halo, rcoef, freetop, sh, sincinterp, dtype):
    T = dtype
    fields = el_model_fields(lam, mu, rho, dtype=T)
    vp = np.sqrt((np.asarray(lam, np.float64) + 2 * np.asarray(mu, np.float64)) / np.asarray(rho, np.float64))
    cpml = init_bdc(vp.astype(T), dt, spacing, halo, rcoef, freetop, sh["f0"], T)      # ela_init_bc.jl:7-40
    kind = "mt" if "moments" in sh else "ef"
    n = np.asarray(lam).shape
    fams = el_possrcrec(kind, sh["src_pos"], sh["rec_pos"], spacing, n, sincinterp)
    nt_s = sh.get("nt", None)
    stf = np.asarray(sh["srctf"], dtype=T)[:nt_s] / T(np.prod(np.asarray(spacing, dtype=T)))  # ela_models.jl:46,89
    mom = None if kind == "ef" else np.asarray(sh["moments"], dtype=T)
    return fields, cpml, kind, fams, stf, mom


def elastic_forward(lam, mu, rho, spacing, dt, nt, halo, rcoef, freetop, shots, sincinterp=True, dtype=np.float64):
    """run_swforward! for ElasticIsoCPMLWaveSimulation 2D.  Shot dicts:
    src_pos (nsrc, 2), srctf ((nt, nsrc) with `moments` (nsrc, 3) = Mxx, Mzz,
    Mxz for moment-tensor sources, or (nt, 2, nsrc) for external forces),
    f0, rec_pos (nrec, 2), optionally their own nt.  Returns seismograms
    (nt, 2, nrec) per shot."""
    out = []
    for sh in shots:
        fields, cpml, kind, fams, stf, mom = _el_prepare(lam, mu, rho, spacing, dt, halo, rcoef, freetop, sh,
                                                         sincinterp, dtype)
        out.append(el2d_forward(fields, spacing, dt, halo, freetop, cpml, sh.get("nt", nt), kind, fams, mom, stf,
                                dtype))
    return out


def elastic_gradient(lam, mu, rho, spacing, dt, nt, halo, rcoef, freetop, shots, check_freq=1, sincinterp=True,
                     mute_radius_src=0, mute_radius_rec=0, dtype=np.float64):
    """run_swgradient! for elastic 2D with L2 misfits (ela_gradient.jl:4-186,
    :188-362): raw correlations, chain rule through the staggered
    interpolations (:161-165), muting (:167-178).  Returns (gradient dict
    {"rho", "lambda", "mu"}, total misfit, seismograms per shot)."""
    T = dtype
    n = np.asarray(lam).shape
    tot = {k: np.zeros(n, dtype=T) for k in ("rho", "lambda", "mu")}
    misfit = 0.0
    seis = []
    for sh in shots:
        fields, cpml, kind, fams, stf, mom = _el_prepare(lam, mu, rho, spacing, dt, halo, rcoef, freetop, sh,
                                                         sincinterp, T)
        nt_s = sh.get("nt", nt)
        syn = el2d_forward(fields, spacing, dt, halo, freetop, cpml, nt_s, kind, fams, mom, stf, T)
        adj = -l2_dchi_du(syn, sh["observed"], sh.get("invcov"))
        (grih, grjh, glam, gmu, gmhh), tr = el2d_gradient(fields, spacing, dt, halo, freetop, cpml, nt_s, kind,
                                                          fams, mom, stf, adj.astype(T), check_freq, T)
        g_rho = (back_interp_dims(rho, grih, [1]) + back_interp_dims(rho, grjh, [2])).astype(T)
        g_mu = (gmu + back_interp_dims(mu, gmhh, [1, 2])).astype(T)
        g_lam = glam.astype(T)
        g_rho, g_lam, g_mu = [mute_around_points(mute_around_points(g, sh["src_pos"], spacing, mute_radius_src),
                                                 sh["rec_pos"], spacing, mute_radius_rec) for g in (g_rho, g_lam, g_mu)]
        tot["rho"] += g_rho
        tot["lambda"] += g_lam
        tot["mu"] += g_mu
        misfit += l2_misfit(tr, sh["observed"], sh.get("invcov"))
        seis.append(tr)
    return tot, misfit, seis


# ---------------------------------------------------------------------------
# 1D acoustic, constant and variable density
# ---------------------------------------------------------------------------
def cd1d_forward(fact, spacing, dt, halo, cpml, nt, src_idx, srctf, rec_idx, dtype=np.float64):
    """swforward_1shot! for CD 1D (acoustic1D_xPU.jl) on prepared inputs;
    src_idx/rec_idx 1-based (n, 1).  Returns traces (nt, nrec)."""
    fact = _f_order(fact, dtype)
    g = _geom(fact.shape, spacing, dt, halo)
    cx, keep = _cpml_struct(cpml[0], dtype)
    src = np.asfortranarray(src_idx, dtype=np.int64)
    rec = np.asfortranarray(rec_idx, dtype=np.int64)
    stf = _f_order(srctf, dtype)
    traces = np.zeros((nt, rec.shape[0]), dtype=dtype, order="F")
    rc = _fn("oracle_cd1d_forward", dtype)(
        ctypes.byref(g), _ptr(fact), ctypes.byref(cx), ctypes.c_int(nt), ctypes.c_int(src.shape[0]), _ptr(src),
        _ptr(stf), ctypes.c_int(rec.shape[0]), _ptr(rec), _ptr(traces))
    assert rc == 0
    return np.ascontiguousarray(traces)


def cd1d_gradient(fact, spacing, dt, halo, cpml, nt, src_idx, srctf, rec_idx, adjsrc, check_freq, dtype=np.float64):
    """Raw grad_vp (before muting and 2/vp^3) and traces of swgradient_1shot!
    for CD 1D."""
    fact = _f_order(fact, dtype)
    g = _geom(fact.shape, spacing, dt, halo)
    cx, keep = _cpml_struct(cpml[0], dtype)
    src = np.asfortranarray(src_idx, dtype=np.int64)
    rec = np.asfortranarray(rec_idx, dtype=np.int64)
    stf = _f_order(srctf, dtype)
    adj = _f_order(adjsrc, dtype)
    traces = np.zeros((nt, rec.shape[0]), dtype=dtype, order="F")
    grad = np.zeros(fact.shape, dtype=dtype)
    rc = _fn("oracle_cd1d_gradient", dtype)(
        ctypes.byref(g), _ptr(fact), ctypes.byref(cx), ctypes.c_int(nt), ctypes.c_int(src.shape[0]), _ptr(src),
        _ptr(stf), ctypes.c_int(rec.shape[0]), _ptr(rec), _ptr(adj), ctypes.c_int(check_freq), _ptr(grad),
        _ptr(traces))
    if rc != 0:
        raise ValueError("Checkpointing frequency must be smaller than the number of timesteps!")
    return grad, np.ascontiguousarray(traces)


def vd1d_forward(m0, m1, spacing, dt, halo, cpml, nt, src_idx, srctf, rec_idx, dtype=np.float64):
    m0 = _f_order(m0, dtype)
    m1 = _f_order(m1, dtype)
    g = _geom(m0.shape, spacing, dt, halo)
    cx, keep = _cpml_struct(cpml[0], dtype)
    src = np.asfortranarray(src_idx, dtype=np.int64)
    rec = np.asfortranarray(rec_idx, dtype=np.int64)
    stf = _f_order(srctf, dtype)
    traces = np.zeros((nt, rec.shape[0]), dtype=dtype, order="F")
    rc = _fn("oracle_vd1d_forward", dtype)(
        ctypes.byref(g), _ptr(m0), _ptr(m1), ctypes.byref(cx), ctypes.c_int(nt), ctypes.c_int(src.shape[0]),
        _ptr(src), _ptr(stf), ctypes.c_int(rec.shape[0]), _ptr(rec), _ptr(traces))
    assert rc == 0
    return np.ascontiguousarray(traces)


def vd1d_gradient(m0, m1, spacing, dt, halo, cpml, nt, src_idx, srctf, rec_idx, adjsrc, check_freq,
                  dtype=np.float64):
    """Raw grad_m0, grad_m1_stag and traces of swgradient_1shot! (VD 1D)."""
    m0 = _f_order(m0, dtype)
    m1 = _f_order(m1, dtype)
    nx = m0.shape[0]
    g = _geom(m0.shape, spacing, dt, halo)
    cx, keep = _cpml_struct(cpml[0], dtype)
    src = np.asfortranarray(src_idx, dtype=np.int64)
    rec = np.asfortranarray(rec_idx, dtype=np.int64)
    stf = _f_order(srctf, dtype)
    adj = _f_order(adjsrc, dtype)
    traces = np.zeros((nt, rec.shape[0]), dtype=dtype, order="F")
    g0 = np.zeros(nx, dtype=dtype)
    g1 = np.zeros(nx - 1, dtype=dtype)
    rc = _fn("oracle_vd1d_gradient", dtype)(
        ctypes.byref(g), _ptr(m0), _ptr(m1), ctypes.byref(cx), ctypes.c_int(nt), ctypes.c_int(src.shape[0]),
        _ptr(src), _ptr(stf), ctypes.c_int(rec.shape[0]), _ptr(rec), _ptr(adj), ctypes.c_int(check_freq),
        _ptr(g0), _ptr(g1), _ptr(traces))
    if rc != 0:
        raise ValueError("Checkpointing frequency must be smaller than the number of timesteps!")
    return (g0, g1), np.ascontiguousarray(traces)


def _shot_1d(vp, spacing, dt, halo, rcoef, freetop, sh, T):
    cpml = init_bdc(vp, dt, spacing, halo, rcoef, freetop, sh["f0"], T)
    si = find_nearest_grid_points(sh["src_pos"], spacing)
    ri = find_nearest_grid_points(sh["rec_pos"], spacing)
    return cpml, si, ri


def acoustic_cd1d_forward(vp, spacing, dt, nt, halo, rcoef, freetop, shots, dtype=np.float64):
    """run_swforward! for CD 1D; shots as acoustic_cd_forward (positions (n, 1))."""
    T = dtype
    vp = np.asarray(vp, dtype=T)
    fact = (T(dt) ** 2) * vp ** 2
    out = []
    for sh in shots:
        nt_s = sh.get("nt", nt)
        cpml, si, ri = _shot_1d(vp, spacing, dt, halo, rcoef, freetop, sh, T)
        stf = scale_srctf_cd(sh["srctf"][:nt_s], spacing, dt, vp, si, T)
        out.append(cd1d_forward(fact, spacing, dt, halo, cpml, nt_s, si, stf, ri, T))
    return out


def acoustic_cd1d_gradient(vp, spacing, dt, nt, halo, rcoef, freetop, shots, check_freq=1,
                           mute_radius_src=0, mute_radius_rec=0, dtype=np.float64):
    """run_swgradient! for CD 1D with L2 misfits; returns ({"vp"}, misfit, seismograms)."""
    T = dtype
    vp = np.asarray(vp, dtype=T)
    fact = (T(dt) ** 2) * vp ** 2
    tot = np.zeros(vp.shape, dtype=T)
    misfit, seis = 0.0, []
    for sh in shots:
        nt_s = sh.get("nt", nt)
        cpml, si, ri = _shot_1d(vp, spacing, dt, halo, rcoef, freetop, sh, T)
        stf = scale_srctf_cd(sh["srctf"][:nt_s], spacing, dt, vp, si, T)
        syn = cd1d_forward(fact, spacing, dt, halo, cpml, nt_s, si, stf, ri, T)
        adj = -l2_dchi_du(syn, sh["observed"], sh.get("invcov"))
        g, tr = cd1d_gradient(fact, spacing, dt, halo, cpml, nt_s, si, stf, ri, adj.astype(T),
                              min(check_freq, nt_s - 1), T)
        g = mute_around_points(g, sh["src_pos"], spacing, mute_radius_src)
        g = mute_around_points(g, sh["rec_pos"], spacing, mute_radius_rec)
        tot += (T(2.0) / vp ** 3) * g
        misfit += l2_misfit(tr, sh["observed"], sh.get("invcov"))
        seis.append(tr)
    return {"vp": tot}, misfit, seis


def acoustic_vd1d_forward(vp, rho, spacing, dt, nt, halo, rcoef, freetop, shots, method="arithmetic",
                          dtype=np.float64):
    """run_swforward! for VD 1D."""
    T = dtype
    vp = np.asarray(vp, dtype=T)
    rho = np.asarray(rho, dtype=T)
    m0 = vp ** 2 * rho * T(dt)
    m1 = (interp_stag(T(1) / rho, 1, method) * T(dt)).astype(T)
    out = []
    for sh in shots:
        nt_s = sh.get("nt", nt)
        cpml, si, ri = _shot_1d(vp, spacing, dt, halo, rcoef, freetop, sh, T)
        stf = scale_srctf_vd(sh["srctf"][:nt_s], spacing, dt, vp, rho, si, T)
        out.append(vd1d_forward(m0, m1, spacing, dt, halo, cpml, nt_s, si, stf, ri, T))
    return out


def acoustic_vd1d_gradient(vp, rho, spacing, dt, nt, halo, rcoef, freetop, shots, check_freq=1,
                           mute_radius_src=0, mute_radius_rec=0, method="arithmetic", dtype=np.float64):
    """run_swgradient! for VD 1D (acou_gradient.jl:97-203); returns
    ({"vp", "rho"}, misfit, seismograms)."""
    T = dtype
    vp = np.asarray(vp, dtype=T)
    rho = np.asarray(rho, dtype=T)
    m0 = vp ** 2 * rho * T(dt)
    m1 = (interp_stag(T(1) / rho, 1, method) * T(dt)).astype(T)
    gvp = np.zeros(vp.shape, dtype=T)
    grho = np.zeros(vp.shape, dtype=T)
    misfit, seis = 0.0, []
    for sh in shots:
        nt_s = sh.get("nt", nt)
        cpml, si, ri = _shot_1d(vp, spacing, dt, halo, rcoef, freetop, sh, T)
        stf = scale_srctf_vd(sh["srctf"][:nt_s], spacing, dt, vp, rho, si, T)
        syn = vd1d_forward(m0, m1, spacing, dt, halo, cpml, nt_s, si, stf, ri, T)
        adj = -l2_dchi_du(syn, sh["observed"], sh.get("invcov"))
        (g0, g1), tr = vd1d_gradient(m0, m1, spacing, dt, halo, cpml, nt_s, si, stf, ri, adj.astype(T),
                                     min(check_freq, nt_s - 1), T)
        g1 = back_interp(1.0 / rho.astype(np.float64), g1, 1, method)
        g0 = g0.astype(np.float64)
        g0 = mute_around_points(mute_around_points(g0, sh["src_pos"], spacing, mute_radius_src),
                                sh["rec_pos"], spacing, mute_radius_rec)
        g1 = mute_around_points(mute_around_points(g1, sh["src_pos"], spacing, mute_radius_src),
                                sh["rec_pos"], spacing, mute_radius_rec)
        gvp += (-2.0 * g0 / (vp ** 3 * rho)).astype(T)
        grho += (-g0 / (vp ** 2 * rho ** 2) - g1 / rho).astype(T)
        misfit += l2_misfit(tr, sh["observed"], sh.get("invcov"))
        seis.append(tr)
    return {"vp": gvp, "rho": grho}, misfit, seis
