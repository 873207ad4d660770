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
from .params import (ARITHMETIC, CPMLBoundaryConditionParameters, ElasticIsoMaterialProperties, GradParameters,
                     InputParametersAcoustic, InputParametersElastic, RunParameters, VpAcousticCDMaterialProperties,
                     VpRhoAcousticVDMaterialProperties)
from .points import nearest_taps, spread_positions
from .shots import ExternalForceShot, MomentTensorShot, ScalarShot

BACKENDS = ("mi355x",)


def _dtype_of(matprop):
    dt = np.asarray(matprop.lam if isinstance(matprop, ElasticIsoMaterialProperties) else matprop.vp).dtype
    return np.float32 if dt == np.float32 else np.float64


def find_nearest_grid_points(positions, spacing):
    """utils.jl:47-58: round(pos/h + 1, RoundNearestTiesUp) -> 1-based indices."""
    positions = np.atleast_2d(np.asarray(positions, dtype=np.float64))
    return np.floor(positions / np.asarray(spacing, dtype=np.float64)[None, :] + 1.5).astype(np.int64)


# ---------------------------------------------------------------------------
# Material interpolation for the staggered VD grid (interpolations.jl:1-47)
# ---------------------------------------------------------------------------
def _dims(dim):
    return [dim] if np.isscalar(dim) else list(dim)


def _corners(shape, dims):
    """Slices selecting, for every corner of the averaging cell along
    `dims`, the lower (0) or upper (1) neighbour of each half point."""
    for bits in range(2 ** len(dims)):
        sl = [slice(None)] * len(shape)
        for b, d in enumerate(dims):
            off = (bits >> (len(dims) - 1 - b)) & 1
            sl[d - 1] = slice(off, shape[d - 1] - 1 + off)
        yield tuple(sl)


def interp_stag(m, dim, method):
    """Value of m at the half points along `dim` (an axis or a list of axes,
    1-based as in Julia): mean of the 2^k cell corners (linear B-spline at
    the half points), or its harmonic counterpart (interpolations.jl:4-43)."""
    dims = _dims(dim)
    m = np.asarray(m)
    if method == ARITHMETIC:
        return sum(m[sl] for sl in _corners(m.shape, dims)) / 2 ** len(dims)
    return 1 / (sum(1 / m[sl] for sl in _corners(m.shape, dims)) / 2 ** len(dims))


def back_interp(method, m, g_stag, dim):
    """Adjoint of interp_stag (interpolations.jl:15-28, ∂f∂m :45-47)."""
    dims = _dims(dim)
    m = np.asarray(m)
    res = np.zeros(m.shape, dtype=np.result_type(m, g_stag))
    k = 2 ** len(dims)
    itp = None if method == ARITHMETIC else interp_stag(m, dims, method)
    for sl in _corners(m.shape, dims):
        w = (1.0 / k) if itp is None else itp ** 2 / m[sl] ** 2 / k
        res[sl] += g_stag * w
    return res


def mute_around_points(grad, points, spacing, radius):
    """mutearoundmultiplepoints! (mute_grad.jl:3-80): scale by r/rmax
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
            # LinearSnapshotter (snapshotter.jl:12)
            assert runparams.snapevery < params.ntimesteps, \
                "Checkpointing frequency must be smaller than the number of timesteps!"
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
        self._fit_batch = 0
        # L2 misfits (identity or diagonal inverse covariance) are evaluated
        # on the device; False takes the host path (calcmisfit / ∂χ_∂u)
        self.device_misfit = True

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
        """init_shot! + possrcrec_scaletf (shot.jl:46-51, acou_forward.jl:6-20)."""
        self.check_shot(shot)
        T = self.dtype
        vel_max = self.cpmlparams.vel_max if self.cpmlparams.vel_max is not None else float(np.max(self.matprop.vp))
        cp = init_bdc(vel_max, self.dt, self.spacing, self.cpmlparams.halo, self.cpmlparams.rcoef,
                      self.cpmlparams.freeboundtop, shot.srcs.domfreq, T)
        si = find_nearest_grid_points(shot.srcs.positions, self.spacing)
        ri = find_nearest_grid_points(shot.recs.positions, self.spacing)
        nt_s = self.shot_nt(shot)
        stf = self.scale_srctf(np.asarray(shot.srcs.tf[:nt_s], dtype=T), si)
        return {"nt": nt_s, "src_idx": si, "rec_idx": ri, "srctf": stf, "cpml": cp, "tshape": (nt_s, ri.shape[0])}

    def upload_shots(self, plan, descs):
        plan.set_shots(descs)

    def snapshot_layout(self):
        """The reference snapshotter's fields (acou_models.jl:177, :427):
        [(name, [shapes], several)]."""
        if isinstance(self, AcousticVDStaggeredCPMLWaveSimulation):
            n = self.gridsize
            stag = [tuple(n[k] - (1 if k == d else 0) for k in range(len(n))) for d in range(len(n))]
            return [("pcur", [n], False), ("vcur", stag, True)]
        return [("pcur", [self.gridsize], False)]

    def batches(self, nshots):
        """Yields (shot indices, plan) for consecutive batches of the call's
        shots.  A batch holds RunParameters.max_batch shots (all of them if
        unset).  When a plan for that many shots does not fit the HBM budget
        (the library reports SWX_ENOMEM at plan creation, state and
        checkpoints included), the batch is halved until it fits, and the
        size that fitted is kept for the next calls on this wave simulation.
        Shots are independent, so the results do not depend on the split."""
        b = self.runparams.max_batch or nshots
        if self._fit_batch:
            b = min(b, self._fit_batch)
        k = 0
        while k < nshots:
            size = min(b, nshots - k)
            try:
                plan = self._ensure_plan(size)
            except MemoryError:
                if size == 1:
                    raise
                b = (size + 1) // 2
                self._fit_batch = b
                self.runparams.logger.info("batch of %d shots does not fit the device memory budget; "
                                           "trying %d", size, b)
                continue
            yield list(range(k, k + size)), plan
            k += size


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
        """acou_gradient.jl:176-202: back-interpolate the staggered m1
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


class ElasticIsoCPMLWaveSimulation(_AcousticCPMLWaveSimulation):
    """ElasticIsoCPMLWaveSimulation 2D, src/models/elastic/ela_models.jl:177-436
    (sincinterp = true as the reference's default, :208)."""
    physics = _native.SWX_ELASTIC_ISO
    cfl_factor = 7.0 / 6.0

    def __init__(self, params, matprop, cpmlparams, *, runparams, gradparams=None, gradient=False,
                 sincinterp: bool = True):
        assert len(params.gridsize) == 2, "Only elastic 2D is currently implemented."
        super().__init__(params, matprop, cpmlparams, runparams=runparams, gradparams=gradparams, gradient=gradient)
        self.sincinterp = sincinterp

    @staticmethod
    def _vp(matprop):
        return np.sqrt((np.asarray(matprop.lam, np.float64) + 2.0 * np.asarray(matprop.mu, np.float64))
                       / np.asarray(matprop.rho, np.float64))

    def check_matprop(self, matprop):
        """check_matprop, ela_models.jl:93-114."""
        lam, mu, rho = (np.asarray(a) for a in (matprop.lam, matprop.mu, matprop.rho))
        assert np.all(lam >= 0), "Lamè coefficient λ must be positive!"
        assert np.all(mu >= 0), "Lamè coefficient μ must be positive!"
        assert np.all(rho > 0), "Density must be positive!"
        vp = self._vp(matprop)
        assert vp.ndim == len(self.gridsize), "Material property dimensionality must be the same as the wavesim!"
        assert lam.shape == mu.shape == rho.shape == self.gridsize, \
            "Material property number of grid points must be the same as the wavesim!"
        self.check_courant_condition(vp)

    def check_numerics(self, shot):
        """check_numerics, ela_models.jl:116-143: minimum S velocity (P if
        the medium has no shear modulus)."""
        vs = np.sqrt(np.asarray(self.matprop.mu, np.float64) / np.asarray(self.matprop.rho, np.float64))
        vel_min = float(np.min(vs))
        if vel_min == 0:
            vel_min = float(np.min(self._vp(self.matprop)))
        fmax = shot.srcs.domfreq * 2.0
        ppw = vel_min / (fmax * max(self.spacing))
        minppw = self.runparams.minPPW
        msg = (f"Not enough points per wavelength (assuming fmax = 2*domfreq)! \n [{round(ppw, 1)} instead of "
               f">= {minppw}]\n  Grid spacing should be <= {round(vel_min / (minppw * fmax), 2)}")
        if self.runparams.erroronPPW:
            assert ppw >= minppw, msg
        elif ppw < minppw:
            warnings.warn(msg)

    def model_fields(self):
        """precomp_elaprop!, ela_models.jl:156-173."""
        T = self.dtype
        mp = self.matprop
        rho = np.asarray(mp.rho, dtype=T)
        mu = np.asarray(mp.mu, dtype=T)
        return [np.asarray(mp.lam, dtype=T), mu,
                interp_stag(rho, 1, mp.interp_method_rho).astype(T),
                interp_stag(rho, 2, mp.interp_method_rho).astype(T),
                interp_stag(mu, [1, 2], mp.interp_method_mu).astype(T)]

    def shot_nt(self, shot):
        nt_s = shot.recs.seismograms.shape[0]
        assert 1 <= nt_s <= self.nt, "number of seismogram rows must be in 1..ntimesteps"
        assert shot.srcs.tf.shape[0] >= nt_s, "source time functions shorter than the seismograms"
        return nt_s

    def shot_desc(self, shot):
        """init_shot! + init_bdc! (ela_init_bc.jl:7-40) + possrcrec_scaletf
        (ela_models.jl:6-90)."""
        self.check_shot(shot)
        T = self.dtype
        vel_max = self.cpmlparams.vel_max if self.cpmlparams.vel_max is not None else float(np.max(self._vp(self.matprop)))
        cp = init_bdc(vel_max, self.dt, self.spacing, self.cpmlparams.halo, self.cpmlparams.rcoef,
                      self.cpmlparams.freeboundtop, shot.srcs.domfreq, T)
        nt_s = self.shot_nt(shot)
        sp, n = self.spacing, self.gridsize
        mt = isinstance(shot, MomentTensorShot)
        if self.sincinterp:
            if mt:
                src = [spread_positions(shot.srcs.positions, sp, n, (0.0, 0.0), True),
                       spread_positions(shot.srcs.positions, sp, n, (sp[0] / 2, sp[1] / 2), True)]
            else:
                src = [spread_positions(shot.srcs.positions, sp, n, (sp[0] / 2, 0.0), False),
                       spread_positions(shot.srcs.positions, sp, n, (0.0, sp[1] / 2), False)]
            rec = [spread_positions(shot.recs.positions, sp, n, (sp[0] / 2, 0.0), False),
                   spread_positions(shot.recs.positions, sp, n, (0.0, sp[1] / 2), False)]
        else:
            si = find_nearest_grid_points(shot.srcs.positions, sp)
            ri = find_nearest_grid_points(shot.recs.positions, sp)
            src = [nearest_taps(si), nearest_taps(si)]
            rec = [nearest_taps(ri), nearest_taps(ri)]
        stf = np.asarray(shot.srcs.tf[:nt_s], dtype=T) / T(np.prod(np.asarray(sp, dtype=T)))
        mom = None
        if mt:
            mom = np.array([[m.Mxx, m.Mzz, m.Mxz] for m in shot.srcs.momtens], dtype=T)
        nrec = shot.recs.positions.shape[0]
        return {"nt": nt_s, "kind": _native.SWX_SRC_MOMENT_TENSOR if mt else _native.SWX_SRC_EXTERNAL_FORCE,
                "src": src, "rec": rec, "moments": mom, "srctf": stf, "cpml": cp, "tshape": (nt_s, 2, nrec)}

    def upload_shots(self, plan, descs):
        plan.set_shots_elastic(descs)

    def snapshot_layout(self):
        """ela_models.jl:390-410: "ucur" = [ux, uz], "σ" = [σxx, σzz, σxz]."""
        nx, nz = self.gridsize
        return [("ucur", [(nx - 1, nz), (nx, nz - 1)], True), ("σ", [(nx, nz), (nx, nz), (nx - 1, nz - 1)], True)]

    def grad_fields_shapes(self):
        nx, nz = self.gridsize
        return [(nx - 1, nz), (nx, nz - 1), (nx, nz), (nx, nz), (nx - 1, nz - 1)]

    def finish_gradient(self, raw, shots_in_batch, per_shot):
        """Chain rule through the staggered averages (ela_gradient.jl:155-165)
        and muting (:166-178).  Gradients are "rho", "lambda", "mu"."""
        T = self.dtype
        mp = self.matprop
        rho = np.asarray(mp.rho, dtype=T)
        mu = np.asarray(mp.mu, dtype=T)
        gp = self.gradparams
        groups = raw if per_shot else [raw]
        tot = {k: np.zeros(self.gridsize, dtype=T) for k in ("rho", "lambda", "mu")}
        for k, (grih, grjh, glam, gmu, gmhh) in enumerate(groups):
            g = {"rho": (back_interp(mp.interp_method_rho, rho, grih, 1)
                         + back_interp(mp.interp_method_rho, rho, grjh, 2)).astype(T),
                 "lambda": glam.copy(),
                 "mu": (gmu + back_interp(mp.interp_method_mu, mu, gmhh, [1, 2])).astype(T)}
            if per_shot:
                shot = shots_in_batch[k]
                for arr in g.values():
                    mute_around_points(arr, shot.srcs.positions, self.spacing, gp.mute_radius_src)
                    mute_around_points(arr, shot.recs.positions, self.spacing, gp.mute_radius_rec)
            for key in tot:
                tot[key] += g[key]
        return tot


# ---------------------------------------------------------------------------
# build / run (apis/build.jl, forward.jl, misfit.jl, gradient.jl)
# ---------------------------------------------------------------------------
def build_wavesim(params, matprop, *, runparams: RunParameters, gradparams: GradParameters | None = None,
                  gradient: bool = False, **kwargs):
    """build_wavesim (apis/build.jl:18-40); extra keywords go to the wave
    simulation constructor (elastic: sincinterp)."""
    if gradparams is None and gradient:
        gradparams = GradParameters()
    elif gradparams is not None:
        assert gradient, "build_wavesim(...) specifies GradParameters, however, gradient keyword argument is set to false."
    if isinstance(matprop, VpAcousticCDMaterialProperties):
        cls = AcousticCDCPMLWaveSimulation
    elif isinstance(matprop, VpRhoAcousticVDMaterialProperties):
        cls = AcousticVDStaggeredCPMLWaveSimulation
    elif isinstance(matprop, ElasticIsoMaterialProperties):
        cls = ElasticIsoCPMLWaveSimulation
    else:
        raise TypeError(f"no wave simulation for material properties {type(matprop).__name__}")
    if not isinstance(params.boundcond, CPMLBoundaryConditionParameters):
        raise NotImplementedError("the mi355x backend implements the C-PML boundary models")
    return cls(params, matprop, params.boundcond, runparams=runparams, gradparams=gradparams, gradient=gradient,
               **kwargs)


def check_sim_consistency(wavesim, matprop, shots):
    """checks.jl:2-41."""
    t0 = (type(shots[0].srcs), type(shots[0].recs))
    for sh in shots:
        if (type(sh.srcs), type(sh.recs)) != t0:
            raise ValueError("Types of shots are inconsistent.")
    if isinstance(wavesim, ElasticIsoCPMLWaveSimulation):
        ok = (isinstance(matprop, ElasticIsoMaterialProperties)
              and all(isinstance(sh, (MomentTensorShot, ExternalForceShot)) for sh in shots))
    else:
        ok = (((isinstance(wavesim, AcousticCDCPMLWaveSimulation)
                and isinstance(matprop, VpAcousticCDMaterialProperties))
               or (isinstance(wavesim, AcousticVDStaggeredCPMLWaveSimulation)
                   and isinstance(matprop, VpRhoAcousticVDMaterialProperties)))
              and all(isinstance(sh, ScalarShot) for sh in shots))
    if not ok:
        raise ValueError("Types of WaveSimulation, MaterialProperties and Sources/Receivers are inconsistent")


def run_swforward(wavesim, matprop, shots):
    """run_swforward! (forward.jl:71-115), batched.  With
    RunParameters.snapevery set, returns the snapshots of every shot as the
    reference does: a list (one entry per shot) of {it: {name: field}},
    a field being an array or, for several arrays ("vcur", "ucur", "σ"), a
    list of arrays."""
    check_sim_consistency(wavesim, matprop, shots)
    wavesim.set_matprop(matprop)
    every = wavesim.runparams.snapevery
    snaps = []
    for batch, plan in wavesim.batches(len(shots)):
        descs = [wavesim.shot_desc(shots[s]) for s in batch]
        wavesim.upload_shots(plan, descs)
        traces = np.zeros(sum(int(np.prod(d["tshape"])) for d in descs), dtype=wavesim.dtype)
        if every:
            layout = wavesim.snapshot_layout()
            one = sum(int(np.prod(sh)) for _, shapes, _ in layout for sh in shapes)
            buf = np.zeros(sum(d["nt"] // every for d in descs) * one, dtype=wavesim.dtype)
            plan.set_snapshots(every, buf)
            try:
                plan.forward(traces, save_state=False)
            finally:
                plan.set_snapshots(0, None)
            snaps += _split_snapshots(buf, descs, every, layout)
        else:
            plan.forward(traces, save_state=False)
        _scatter_traces(traces, descs, [shots[s] for s in batch])
    return snaps if every else None


def _split_snapshots(buf, descs, every, layout):
    out, off = [], 0
    for d in descs:
        shot = {}
        for k in range(1, d["nt"] // every + 1):
            fields = {}
            for name, shapes, several in layout:
                arrs = []
                for sh in shapes:
                    n = int(np.prod(sh))
                    arrs.append(buf[off:off + n].reshape(sh, order="F").copy())
                    off += n
                fields[name] = arrs if several else arrs[0]
            shot[k * every] = fields
        out.append(shot)
    return out


def _scatter_traces(traces, descs, shots):
    off = 0
    for d, sh in zip(descs, shots):
        n = int(np.prod(d["tshape"]))
        blk = traces[off:off + n].reshape(d["tshape"], order="F")
        sh.recs.seismograms[...] = blk.astype(sh.recs.seismograms.dtype)
        off += n


def swforward(params_or_wavesim, matprop, shots, *, runparams: RunParameters | None = None):
    """swforward! (forward.jl:24-66): fills shot.recs.seismograms; returns
    the snapshots when RunParameters.snapevery is set."""
    if isinstance(params_or_wavesim, (InputParametersAcoustic, InputParametersElastic)):
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


def _device_l2(m):
    """An L2 misfit the device can evaluate: identity or diagonal inverse
    covariance, any windows."""
    return isinstance(m, L2Misfit) and (m.invcov is None or np.asarray(m.invcov).ndim == 1)


def _l2_observed(misfits, descs, T):
    """Observed data of the batch in the traces layout and, when a misfit
    has windows or an inverse covariance, the float64 weights mask x C^-1
    per sample (None when every weight is 1)."""
    obs, ws = [], []
    weighted = any(m.windows or m.invcov is not None for m in misfits)
    for m, d in zip(misfits, descs):
        o = np.asarray(m.observed)
        assert o.shape == tuple(d["tshape"]), "observed data and seismograms must have the same shape"
        obs.append(np.asfortranarray(o, dtype=T).ravel(order="F"))
        if weighted:
            nt = o.shape[0]
            w = np.ones(nt)
            if m.windows:
                w = np.zeros(nt)
                for a, b in m.windows:
                    w[a - 1:b] = 1.0
            if m.invcov is not None:
                w = w * np.asarray(m.invcov, dtype=np.float64)
            full = np.broadcast_to(w.reshape((nt,) + (1,) * (o.ndim - 1)), o.shape)
            ws.append(np.asfortranarray(full).ravel(order="F"))
    return np.concatenate(obs), (np.concatenate(ws) if weighted else None)


def run_swgradient(wavesim, matprop, shots, misfit, *, per_shot_gradients: bool = False):
    """run_swgradient! (gradient.jl:93-136), batched: forward with device
    checkpoints, adjoint source (on the device for L2 misfits with an
    identity or diagonal inverse covariance, on the host otherwise), adjoint
    with replay, correlation.

    With per_shot_gradients the result is the list of each shot's gradient
    (muted and chain-ruled as the reference does per shot, gradient.jl:
    117-128) instead of their sum; the library then returns one raw
    gradient per shot (swx_adjoint with per_shot = 1)."""
    check_sim_consistency(wavesim, matprop, shots)
    assert len(misfit) == len(shots)
    wavesim.set_matprop(matprop)
    gp = wavesim.gradparams
    per_shot = per_shot_gradients or gp.mute_radius_src > 0 or gp.mute_radius_rec > 0
    totgrad = None
    shot_grads = []
    totmisfit = 0.0
    T = wavesim.dtype
    shapes = wavesim.grad_fields_shapes()
    for batch, plan in wavesim.batches(len(shots)):
        bshots = [shots[s] for s in batch]
        descs = [wavesim.shot_desc(sh) for sh in bshots]
        wavesim.upload_shots(plan, descs)
        traces = np.zeros(sum(int(np.prod(d["tshape"])) for d in descs), dtype=T)
        plan.forward(traces, save_state=True)
        _scatter_traces(traces, descs, bshots)
        bmis = [misfit[s] for s in batch]
        ncopy = len(batch) if per_shot else 1
        size1 = sum(int(np.prod(s)) for s in shapes)
        rawflat = np.zeros(ncopy * size1, dtype=T)
        if wavesim.device_misfit and all(_device_l2(m) for m in bmis):
            # L2 adjoint source and misfit formed on the device from the
            # traces still there (acou_gradient.jl:45-51, L2Misfit.jl:24-95)
            obs, w = _l2_observed(bmis, descs, T)
            plan.set_observed(obs, w)
            shot_mis = plan.adjoint_l2(rawflat, per_shot, len(batch))
        else:
            adj = []
            for sh, d in zip(bshots, bmis):
                adj.append(np.asfortranarray(-d.dchi_du(sh.recs), dtype=T).ravel(order="F"))  # acou_gradient.jl:48
            adjsrc = np.concatenate(adj).astype(T)
            plan.adjoint(adjsrc, rawflat, per_shot)
            shot_mis = None
        raw = []
        for k in range(ncopy):
            fields, off = [], k * size1
            for s in shapes:
                n = int(np.prod(s))
                fields.append(rawflat[off:off + n].reshape(s, order="F").copy())
                off += n
            raw.append(fields)
        if per_shot_gradients:
            shot_grads += [wavesim.finish_gradient([raw[k]], [bshots[k]], True) for k in range(len(batch))]
        else:
            g = wavesim.finish_gradient(raw if per_shot else raw[0], bshots, per_shot)
            if totgrad is None:
                totgrad = g
            else:
                for key in totgrad:
                    totgrad[key] += g[key]
        if gp.compute_misfit:
            if shot_mis is not None:
                totmisfit += float(np.sum(shot_mis))
            else:
                totmisfit += sum(misfit[s].calcmisfit(shots[s].recs) for s in batch)
    out = shot_grads if per_shot_gradients else totgrad
    return (out, totmisfit) if gp.compute_misfit else out


def swgradient(params_or_wavesim, matprop, shots, misfit, *, runparams: RunParameters | None = None,
               gradparams: GradParameters | None = None, per_shot_gradients: bool = False):
    """swgradient! (gradient.jl:31-88).  per_shot_gradients: return the
    list of per-shot gradients instead of their sum (an extension)."""
    if isinstance(params_or_wavesim, (InputParametersAcoustic, InputParametersElastic)):
        ws = build_wavesim(params_or_wavesim, matprop, runparams=runparams or RunParameters(),
                           gradparams=gradparams or GradParameters(), gradient=True)
    else:
        ws = params_or_wavesim
    return run_swgradient(ws, matprop, shots, misfit, per_shot_gradients=per_shot_gradients)


__all__ = [
    "build_wavesim", "swforward", "swmisfit", "swgradient", "run_swforward", "run_swgradient",
    "AcousticCDCPMLWaveSimulation", "AcousticVDStaggeredCPMLWaveSimulation", "ElasticIsoCPMLWaveSimulation",
    "L2Misfit",
    "find_nearest_grid_points", "interp_stag", "back_interp", "mute_around_points",
]
