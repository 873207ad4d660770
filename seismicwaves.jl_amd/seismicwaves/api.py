"""Public API: build_wavesim, swforward, swmisfit, swgradient.

Mirrors src/apis/{build,forward,misfit,gradient}.jl of the reference with
the MI355X backend below: instead of the reference's per-shot loop
(gradient.jl:115-131) the shots of a call are run as batches on the device
(RunParameters.max_batch shots per launch, default all of them).

Like the reference, swforward/swgradient fill `shot.recs.seismograms` in
place.  A shot's number of time steps is the number of rows of its
seismogram array (at most params.ntimesteps), which lets one call batch
shots of uneven length; the reference requires them all to equal nt.
"""
from __future__ import annotations

import math
import warnings

import numpy as np

from . import _native
from .cpml import init_bdc
from .misfits import L2Misfit
from .params import (ARITHMETIC, CPMLBoundaryConditionParameters, GradParameters, InputParametersAcoustic,
                     RunParameters, VpAcousticCDMaterialProperties, VpRhoAcousticVDMaterialProperties)
from .shots import ScalarShot

BACKENDS = ("mi355x",)


def _dtype_of(matprop):
    dt = np.asarray(matprop.vp).dtype
    return np.float32 if dt == np.float32 else np.float64


def find_nearest_grid_points(positions, spacing):
    """utils.jl:47-58: round(pos/h + 1, RoundNearestTiesUp) -> 1-based indices."""
    positions = np.atleast_2d(np.asarray(positions, dtype=np.float64))
    return np.floor(positions / np.asarray(spacing, dtype=np.float64)[None, :] + 1.5).astype(np.int64)


# ---------------------------------------------------------------------------
# Material interpolation for the staggered VD grid (interpolations.jl:1-47)
# ---------------------------------------------------------------------------
def interp_stag(m, dim, method):
    """Value of m at the half points along `dim` (1-based dim as in Julia)."""
    ax = dim - 1
    lo = np.take(m, range(0, m.shape[ax] - 1), axis=ax)
    hi = np.take(m, range(1, m.shape[ax]), axis=ax)
    if method == ARITHMETIC:
        return (lo + hi) / 2
    return 1 / ((1 / lo + 1 / hi) / 2)


def back_interp(method, m, g_stag, dim):
    """Adjoint of interp_stag (interpolations.jl:15-28, ∂f∂m :45-47)."""
    ax = dim - 1
    n = m.shape[ax]
    res = np.zeros(m.shape, dtype=np.result_type(m, g_stag))
    lo = np.take(m, range(0, n - 1), axis=ax)
    hi = np.take(m, range(1, n), axis=ax)
    if method == ARITHMETIC:
        dlo = dhi = np.full(g_stag.shape, 0.5, dtype=res.dtype)
    else:
        itp = interp_stag(m, dim, method)
        dlo = itp ** 2 / lo ** 2 / 2
        dhi = itp ** 2 / hi ** 2 / 2
    sl_lo = [slice(None)] * m.ndim
    sl_hi = [slice(None)] * m.ndim
    sl_lo[ax] = slice(0, n - 1)
    sl_hi[ax] = slice(1, n)
    res[tuple(sl_lo)] += g_stag * dlo
    res[tuple(sl_hi)] += g_stag * dhi
    return res


def mute_around_points(grad, points, spacing, radius):
    """mutearoundmultiplepoints! (mute_grad.jl:49-126): scale by r/rmax
    inside a ball of `radius` grid points around each point."""
    if radius == 0:
        return grad
    assert radius > 0, "mutearoundpoint!(): The smoothing radius must be positive."
    n = grad.shape
    N = len(n)
    rmax = radius * max(spacing)
    for pt in np.atleast_2d(points):
        ranges = []
        for d in range(N):
            extent = (n[d] - 1) * spacing[d]
            if not (0.0 <= pt[d] <= extent):
                raise ValueError(f"mutearoundpoint!(): The point lies outside the grid on dimension {d + 1} at position {pt[d]}.")
            ijk = int(np.floor(pt[d] / spacing[d])) + 1
            ranges.append(np.arange(max(ijk - radius, 1), min(ijk + radius, n[d]) + 1))
        mesh = np.meshgrid(*ranges, indexing="ij")
        r2 = sum((pt[d] - (mesh[d] - 1) * spacing[d]) ** 2 for d in range(N))
        r = np.sqrt(r2)
        sel = r <= rmax
        idx = tuple(m[sel] - 1 for m in mesh)
        grad[idx] *= (r[sel] / rmax).astype(grad.dtype)
    return grad


# ---------------------------------------------------------------------------
# Wave simulations
# ---------------------------------------------------------------------------
class _AcousticCPMLWaveSimulation:
    physics = None
    cfl_factor = 1.0

    def __init__(self, params: InputParametersAcoustic, matprop, cpmlparams: CPMLBoundaryConditionParameters, *,
                 runparams: RunParameters, gradparams: GradParameters | None = None, gradient: bool = False):
        # acou_models.jl:98-108
        assert cpmlparams.halo >= 0, "CPML halo size must be non-negative!"
        ns_cpml = params.gridsize[:-1] if cpmlparams.freeboundtop else params.gridsize
        h2 = 2 * cpmlparams.halo + 3
        assert all(n >= h2 for n in ns_cpml), \
            f"Number grid points in the dimensions with C-PML boundaries must be at least 2*halo+3 = {h2}!"
        if runparams.parall not in BACKENDS:
            raise ValueError(f"No backend found for model of type {type(self).__name__} and `parall =` "
                             f"{runparams.parall}. Argument `parall` must be one of the following: {BACKENDS}.")
        if runparams.snapevery is not None:
            raise NotImplementedError("snapshots (snapevery) are not provided by the mi355x backend")
        if runparams.infoevery is not None:
            assert 1 <= runparams.infoevery <= params.ntimesteps, \
                "Infoevery parameter must be positive and less then nt!"
        self.params = params
        self.cpmlparams = cpmlparams
        self.runparams = runparams
        self.gradparams = gradparams
        self.gradient = gradient
        self.nt = params.ntimesteps
        self.dtype = _dtype_of(matprop)
        self.dt = self.dtype(params.dt)
        self.spacing = tuple(float(self.dtype(h)) for h in params.gridspacing)
        self.gridsize = params.gridsize
        self.matprop = None
        self._plan = None
        self._plan_shots = 0

    # -- checks (acou_models.jl:5-38) --
    def check_courant_condition(self, vp):
        courant = float(np.max(vp)) * float(self.dt) * math.sqrt(sum(1 / h ** 2 for h in self.spacing)) * self.cfl_factor
        self.runparams.logger.info("Courant number: %s", courant)
        if self.runparams.erroronCFL:
            assert courant < 1, f"Courant condition not satisfied! [{courant}]"
        elif courant > 1:
            warnings.warn(f"CFL condition not satisfied! [{courant}]")

    def check_numerics(self, shot):
        vel_min = float(np.min(self.matprop.vp))
        fmax = shot.srcs.domfreq * 2.0
        ppw = vel_min / (fmax * max(self.spacing))
        minppw = self.runparams.minPPW
        msg = (f"Not enough points per wavelength (assuming fmax = 2*domfreq)! \n [{round(ppw, 1)} instead of "
               f">= {minppw}]\n  Grid spacing should be <= {round(vel_min / (minppw * fmax), 2)}")
        if self.runparams.erroronPPW:
            assert ppw >= minppw, msg
        elif ppw < minppw:
            warnings.warn(msg)

    def check_positions(self, positions):
        """checks.jl:55-94."""
        positions = np.atleast_2d(positions)
        N = len(self.gridsize)
        assert positions.shape[1] == N, "Positions matrix do not match the dimension of the model!"
        extent = [(n - 1) * h for n, h in zip(self.gridsize, self.spacing)]
        halo = self.cpmlparams.halo
        for p in positions:
            for c in range(N):
                assert 0 <= p[c] <= extent[c], f"Position {p} is not inside the grid!"
                if not (c == N - 1 and self.cpmlparams.freeboundtop):
                    if not (self.spacing[c] * halo <= p[c] <= extent[c] - self.spacing[c] * halo):
                        self.runparams.logger.warning("Position %s is inside the CPML region!", p)

    def check_shot(self, shot):
        self.check_numerics(shot)
        self.check_positions(shot.srcs.positions)
        self.check_positions(shot.recs.positions)

    # -- plan management --
    def _ensure_plan(self, nshots):
        if self._plan is not None and self._plan_shots >= nshots:
            return self._plan
        if self._plan is not None:
            self._plan.close()
        gp = self.gradparams
        self._plan = _native.Plan(
            physics=self.physics, dtype=self.dtype, n=self.gridsize, h=self.spacing, dt=float(self.dt),
            halo=self.cpmlparams.halo, freetop=self.cpmlparams.freeboundtop, max_shots=nshots,
            nt_max=self.nt, check_freq=(gp.check_freq if gp else 1), gradient=self.gradient,
            device=self.runparams.device, mem_budget=self.runparams.mem_budget)
        self._plan_shots = nshots
        if self.matprop is not None:
            self._plan.set_model(self.model_fields())
        return self._plan

    def set_matprop(self, matprop):
        """set_wavesim_matprop! (apis/utils.jl:3-8)."""
        self.check_matprop(matprop)
        self.matprop = matprop
        if self._plan is not None:
            self._plan.set_model(self.model_fields())

    def shot_nt(self, shot):
        nt_s = shot.recs.seismograms.shape[0]
        assert 1 <= nt_s <= self.nt, "number of seismogram rows must be in 1..ntimesteps"
        assert shot.srcs.tf.shape[0] >= nt_s, "source time functions shorter than the seismograms"
        return nt_s

    def shot_desc(self, shot):
        """init_shot! + possrcrec_scaletf (shot.jl:248-253, acou_forward.jl:6-20)."""
        self.check_shot(shot)
        T = self.dtype
        vel_max = self.cpmlparams.vel_max if self.cpmlparams.vel_max is not None else float(np.max(self.matprop.vp))
        cp = init_bdc(vel_max, self.dt, self.spacing, self.cpmlparams.halo, self.cpmlparams.rcoef,
                      self.cpmlparams.freeboundtop, shot.srcs.domfreq, T)
        si = find_nearest_grid_points(shot.srcs.positions, self.spacing)
        ri = find_nearest_grid_points(shot.recs.positions, self.spacing)
        nt_s = self.shot_nt(shot)
        stf = self.scale_srctf(np.asarray(shot.srcs.tf[:nt_s], dtype=T), si)
        return {"nt": nt_s, "src_idx": si, "rec_idx": ri, "srctf": stf, "cpml": cp}

    def batches(self, nshots):
        b = self.runparams.max_batch or nshots
        return [list(range(k, min(k + b, nshots))) for k in range(0, nshots, b)]


class AcousticCDCPMLWaveSimulation(_AcousticCPMLWaveSimulation):
    """AcousticCDCPMLWaveSimulation, acou_models.jl:68-202."""
    physics = _native.SWX_ACOUSTIC_CD
    cfl_factor = 1.0

    def check_matprop(self, matprop):
        vp = np.asarray(matprop.vp)
        assert vp.ndim == len(self.gridsize), "Material property dimensionality must be the same as the wavesim!"
        assert vp.shape == self.gridsize, "Material property number of grid points must be the same as the wavesim!"
        assert np.all(vp > 0), "Pressure velocity material property must be positive!"
        self.check_courant_condition(vp)

    def model_fields(self):
        T = self.dtype
        vp = np.asarray(self.matprop.vp, dtype=T)
        return [(T(self.dt) ** 2) * vp ** 2]            # precompute_fact!, acou_models.jl:60

    def scale_srctf(self, tf, si):
        T = self.dtype
        out = tf / T(np.prod(np.asarray(self.spacing, dtype=T))) * (T(self.dt) ** 2)
        vp = np.asarray(self.matprop.vp, dtype=T)
        for s in range(out.shape[1]):
            out[:, s] *= vp[tuple(si[s] - 1)] ** 2
        return out

    def grad_fields_shapes(self):
        return [self.gridsize]

    def finish_gradient(self, raw, shots_in_batch, per_shot):
        """Muting (acou_gradient.jl:87-91) and chain rule 2/vp^3 (:93)."""
        T = self.dtype
        vp = np.asarray(self.matprop.vp, dtype=T)
        gp = self.gradparams
        if per_shot:
            tot = np.zeros(self.gridsize, dtype=T)
            for k, shot in enumerate(shots_in_batch):
                g = raw[k][0].copy()
                g = mute_around_points(g, shot.srcs.positions, self.spacing, gp.mute_radius_src)
                g = mute_around_points(g, shot.recs.positions, self.spacing, gp.mute_radius_rec)
                tot += (T(2.0) / vp ** 3) * g
            return {"vp": tot}
        return {"vp": (T(2.0) / vp ** 3) * raw[0]}


class AcousticVDStaggeredCPMLWaveSimulation(_AcousticCPMLWaveSimulation):
    """AcousticVDStaggeredCPMLWaveSimulation, acou_models.jl:311-452."""
    physics = _native.SWX_ACOUSTIC_VD
    cfl_factor = 7.0 / 6.0

    def check_matprop(self, matprop):
        vp, rho = np.asarray(matprop.vp), np.asarray(matprop.rho)
        assert vp.ndim == rho.ndim == len(self.gridsize), "Material property dimensionality must be the same as the wavesim!"
        assert vp.shape == rho.shape == self.gridsize, "Material property number of grid points must be the same as the wavesim!"
        assert np.all(vp > 0), "Pressure velocity material property must be positive!"
        assert np.all(rho > 0), "Density material property must be positive!"
        self.check_courant_condition(vp)

    def model_fields(self):
        """precompute_fact! (acou_models.jl:292-300)."""
        T = self.dtype
        vp = np.asarray(self.matprop.vp, dtype=T)
        rho = np.asarray(self.matprop.rho, dtype=T)
        m0 = vp ** 2 * rho * T(self.dt)
        inv_rho = T(1) / rho
        out = [m0]
        for d in range(1, len(self.gridsize) + 1):
            out.append(interp_stag(inv_rho, d, self.matprop.interp_method).astype(T) * T(self.dt))
        return out

    def scale_srctf(self, tf, si):
        """possrcrec_scaletf for VD (acou_forward.jl:67-81)."""
        T = self.dtype
        out = tf / T(np.prod(np.asarray(self.spacing, dtype=T))) * T(self.dt)
        vp = np.asarray(self.matprop.vp, dtype=T)
        rho = np.asarray(self.matprop.rho, dtype=T)
        for s in range(out.shape[1]):
            ix = tuple(si[s] - 1)
            out[:, s] *= vp[ix] ** 2 * rho[ix]
        return out

    def grad_fields_shapes(self):
        n = self.gridsize
        shapes = [n]
        for d in range(len(n)):
            shapes.append(tuple(n[k] - (1 if k == d else 0) for k in range(len(n))))
        return shapes

    def finish_gradient(self, raw, shots_in_batch, per_shot):
        """acou_gradient.jl:302-327: back-interpolate the staggered m1
        gradients, mute, and apply the chain rule to vp and rho."""
        T = self.dtype
        vp = np.asarray(self.matprop.vp, dtype=T)
        rho = np.asarray(self.matprop.rho, dtype=T)
        gp = self.gradparams
        inv_rho = T(1) / rho
        groups = raw if per_shot else [raw]
        tot_vp = np.zeros(self.gridsize, dtype=T)
        tot_rho = np.zeros(self.gridsize, dtype=T)
        for k, fields in enumerate(groups):
            g_m0 = fields[0].copy()
            g_m1 = np.zeros(self.gridsize, dtype=T)
            for d in range(1, len(self.gridsize) + 1):
                g_m1 += back_interp(self.matprop.interp_method, inv_rho, fields[d], d).astype(T)
            if per_shot:
                shot = shots_in_batch[k]
                for arr in (g_m0, g_m1):
                    mute_around_points(arr, shot.srcs.positions, self.spacing, gp.mute_radius_src)
                    mute_around_points(arr, shot.recs.positions, self.spacing, gp.mute_radius_rec)
            tot_vp += -T(2.0) * g_m0 / (vp ** 3 * rho)
            tot_rho += -g_m0 / (vp ** 2 * rho ** 2) - g_m1 / rho
        return {"vp": tot_vp, "rho": tot_rho}


# ---------------------------------------------------------------------------
# build / run (apis/build.jl, forward.jl, misfit.jl, gradient.jl)
# ---------------------------------------------------------------------------
def build_wavesim(params, matprop, *, runparams: RunParameters, gradparams: GradParameters | None = None,
                  gradient: bool = False):
    """build_wavesim (apis/build.jl:18-40)."""
    if gradparams is None and gradient:
        gradparams = GradParameters()
    elif gradparams is not None:
        assert gradient, "build_wavesim(...) specifies GradParameters, however, gradient keyword argument is set to false."
    if isinstance(matprop, VpAcousticCDMaterialProperties):
        cls = AcousticCDCPMLWaveSimulation
    elif isinstance(matprop, VpRhoAcousticVDMaterialProperties):
        cls = AcousticVDStaggeredCPMLWaveSimulation
    else:
        raise TypeError(f"no wave simulation for material properties {type(matprop).__name__}")
    if not isinstance(params.boundcond, CPMLBoundaryConditionParameters):
        raise NotImplementedError("the mi355x backend implements the C-PML boundary models")
    return cls(params, matprop, params.boundcond, runparams=runparams, gradparams=gradparams, gradient=gradient)


def check_sim_consistency(wavesim, matprop, shots):
    """checks.jl:2-41."""
    t0 = (type(shots[0].srcs), type(shots[0].recs))
    for sh in shots:
        if (type(sh.srcs), type(sh.recs)) != t0:
            raise ValueError("Types of shots are inconsistent.")
    ok = ((isinstance(wavesim, AcousticCDCPMLWaveSimulation) and isinstance(matprop, VpAcousticCDMaterialProperties))
          or (isinstance(wavesim, AcousticVDStaggeredCPMLWaveSimulation)
              and isinstance(matprop, VpRhoAcousticVDMaterialProperties)))
    if not ok or not all(isinstance(sh, ScalarShot) for sh in shots):
        raise ValueError("Types of WaveSimulation, MaterialProperties and Sources/Receivers are inconsistent")


def run_swforward(wavesim, matprop, shots):
    """run_swforward! (forward.jl:161-199), batched."""
    check_sim_consistency(wavesim, matprop, shots)
    wavesim.set_matprop(matprop)
    for batch in wavesim.batches(len(shots)):
        descs = [wavesim.shot_desc(shots[s]) for s in batch]
        plan = wavesim._ensure_plan(len(batch))
        plan.set_shots(descs)
        traces = np.zeros(sum(d["nt"] * d["rec_idx"].shape[0] for d in descs), dtype=wavesim.dtype)
        plan.forward(traces, save_state=False)
        _scatter_traces(traces, descs, [shots[s] for s in batch])
    return None


def _scatter_traces(traces, descs, shots):
    off = 0
    for d, sh in zip(descs, shots):
        nt, nr = d["nt"], d["rec_idx"].shape[0]
        blk = traces[off:off + nt * nr].reshape((nt, nr), order="F")
        sh.recs.seismograms[:, :] = blk.astype(sh.recs.seismograms.dtype)
        off += nt * nr


def swforward(params_or_wavesim, matprop, shots, *, runparams: RunParameters | None = None):
    """swforward! (forward.jl:114-156): fills shot.recs.seismograms."""
    if isinstance(params_or_wavesim, InputParametersAcoustic):
        assert runparams is not None
        ws = build_wavesim(params_or_wavesim, matprop, runparams=runparams, gradient=False)
    else:
        ws = params_or_wavesim
    return run_swforward(ws, matprop, shots)


def swmisfit(params_or_wavesim, matprop, shots, misfit, *, runparams: RunParameters | None = None):
    """swmisfit! (misfit.jl:26-104).  The total is summed over all shots
    (the reference's loop `for s in length(shots)` counts only shot 1,
    misfit.jl:97-99; the gradient path sums them, gradient.jl:129-131)."""
    swforward(params_or_wavesim, matprop, shots, runparams=runparams)
    return sum(m.calcmisfit(sh.recs) for sh, m in zip(shots, misfit))


def run_swgradient(wavesim, matprop, shots, misfit):
    """run_swgradient! (gradient.jl:93-136), batched: forward with device
    checkpoints, host adjoint source, adjoint with replay, correlation."""
    check_sim_consistency(wavesim, matprop, shots)
    assert len(misfit) == len(shots)
    wavesim.set_matprop(matprop)
    gp = wavesim.gradparams
    per_shot = gp.mute_radius_src > 0 or gp.mute_radius_rec > 0
    totgrad = None
    totmisfit = 0.0
    T = wavesim.dtype
    shapes = wavesim.grad_fields_shapes()
    for batch in wavesim.batches(len(shots)):
        bshots = [shots[s] for s in batch]
        descs = [wavesim.shot_desc(sh) for sh in bshots]
        plan = wavesim._ensure_plan(len(batch))
        plan.set_shots(descs)
        traces = np.zeros(sum(d["nt"] * d["rec_idx"].shape[0] for d in descs), dtype=T)
        plan.forward(traces, save_state=True)
        _scatter_traces(traces, descs, bshots)
        adj = []
        for sh, d in zip(bshots, [misfit[s] for s in batch]):
            adj.append(np.asfortranarray(-d.dchi_du(sh.recs), dtype=T).ravel(order="F"))  # acou_gradient.jl:48
        adjsrc = np.concatenate(adj).astype(T)
        ncopy = len(batch) if per_shot else 1
        size1 = sum(int(np.prod(s)) for s in shapes)
        rawflat = np.zeros(ncopy * size1, dtype=T)
        plan.adjoint(adjsrc, rawflat, per_shot)
        raw = []
        for k in range(ncopy):
            fields, off = [], k * size1
            for s in shapes:
                n = int(np.prod(s))
                fields.append(rawflat[off:off + n].reshape(s, order="F").copy())
                off += n
            raw.append(fields)
        g = wavesim.finish_gradient(raw if per_shot else raw[0], bshots, per_shot)
        if totgrad is None:
            totgrad = g
        else:
            for key in totgrad:
                totgrad[key] += g[key]
        if gp.compute_misfit:
            totmisfit += sum(misfit[s].calcmisfit(shots[s].recs) for s in batch)
    return (totgrad, totmisfit) if gp.compute_misfit else totgrad


def swgradient(params_or_wavesim, matprop, shots, misfit, *, runparams: RunParameters | None = None,
               gradparams: GradParameters | None = None):
    """swgradient! (gradient.jl:31-88)."""
    if isinstance(params_or_wavesim, InputParametersAcoustic):
        ws = build_wavesim(params_or_wavesim, matprop, runparams=runparams or RunParameters(),
                           gradparams=gradparams or GradParameters(), gradient=True)
    else:
        ws = params_or_wavesim
    return run_swgradient(ws, matprop, shots, misfit)


__all__ = [
    "build_wavesim", "swforward", "swmisfit", "swgradient", "run_swforward", "run_swgradient",
    "AcousticCDCPMLWaveSimulation", "AcousticVDStaggeredCPMLWaveSimulation", "L2Misfit",
    "find_nearest_grid_points", "interp_stag", "back_interp", "mute_around_points",
]
