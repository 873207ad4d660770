"""Parameter and material-property types, mirroring the reference's structs.

Reference: src/models/acoustic/acou_params.jl:9-35,
src/models/acoustic/acou_material_properties.jl:10-46,
src/models/elastic/ela_params.jl:8-34, ela_material_properties.jl:9-22,
src/models/bdc_params.jl:16-27, src/models/genparameters.jl:25-78.

`T` in the reference is the element type of the arrays; here it is the numpy
dtype of the material arrays (float32 or float64).
"""
from __future__ import annotations

import logging
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

log = logging.getLogger("seismicwaves")


@dataclass
class CPMLBoundaryConditionParameters:
    """CPMLBoundaryConditionParameters, bdc_params.jl:16-25."""
    halo: int = 20
    rcoef: float = 0.0001
    freeboundtop: bool = True
    vel_max: Optional[float] = None


@dataclass
class ReflectiveBoundaryConditionParameters:
    """ReflectiveBoundaryConditionParameters, bdc_params.jl:7 (not on the
    GPU path: the MI355X backend implements the C-PML models only)."""


class InputParametersAcoustic:
    """InputParametersAcoustic, acou_params.jl:9-35."""

    def __init__(self, ntimesteps: int, dt: float, gridsize, gridspacing, boundcond):
        N = len(gridsize)
        assert N <= 3, "Dimensionality must be less than or equal to 3!"
        assert len(gridspacing) == N
        assert all(n > 0 for n in gridsize), "All numbers of grid points must be positive!"
        assert all(h > 0 for h in gridspacing), "All grid spacings must be positive!"
        assert ntimesteps > 0, "Number of timesteps must be positive!"
        assert dt > 0, "Timestep size must be positive!"
        self.ntimesteps = int(ntimesteps)
        self.dt = dt
        self.gridsize = tuple(int(n) for n in gridsize)
        self.gridspacing = tuple(gridspacing)
        self.boundcond = boundcond

    @property
    def ndim(self):
        return len(self.gridsize)


class InputParametersElastic(InputParametersAcoustic):
    """InputParametersElastic, src/models/elastic/ela_params.jl:8-34 (same
    fields and checks as the acoustic parameters)."""


class VpAcousticCDMaterialProperties:
    """VpAcousticCDMaterialProperties, acou_material_properties.jl:10-13."""

    def __init__(self, vp):
        self.vp = np.asarray(vp)


ARITHMETIC = "arithmetic"
HARMONIC = "harmonic"


class VpRhoAcousticVDMaterialProperties:
    """VpRhoAcousticVDMaterialProperties, acou_material_properties.jl:22-46."""

    def __init__(self, vp, rho, interp_method: str = ARITHMETIC):
        self.vp = np.asarray(vp)
        self.rho = np.asarray(rho)
        assert interp_method in (ARITHMETIC, HARMONIC)
        self.interp_method = interp_method


@dataclass
class RunParameters:
    """RunParameters, genparameters.jl:25-46.

    `parall` selects the backend; this package provides "mi355x" (the
    reference's symbols :serial/:threads/:CUDA/:AMDGPU/:Metal live in the
    Julia package).  `max_batch` bounds how many shots run in one batched
    launch (None = all shots of the call); `device` is the HIP ordinal.
    """
    parall: str = "mi355x"
    snapevery: Optional[int] = None
    infoevery: Optional[int] = None
    logger: logging.Logger = field(default_factory=lambda: log)
    erroronCFL: bool = True
    minPPW: int = 10
    erroronPPW: bool = True
    max_batch: Optional[int] = None
    device: int = 0
    mem_budget: float = 0.0

    def __post_init__(self):
        assert self.minPPW >= 0


@dataclass
class GradParameters:
    """GradParameters, genparameters.jl:63-78."""
    mute_radius_src: int = 0
    mute_radius_rec: int = 0
    compute_misfit: bool = False
    check_freq: int = 1

    def __post_init__(self):
        assert self.check_freq > 0


class ElasticIsoMaterialProperties:
    """ElasticIsoMaterialProperties, src/models/elastic/ela_material_properties.jl:9-22:
    Lame parameters lam (λ) and mu (μ), density rho (ρ), and the averaging
    used for the staggered points (ρ arithmetic, λ harmonic, μ arithmetic by
    default; λ is never averaged by the 2D scheme)."""

    def __init__(self, lam, mu, rho, interp_method_rho: str = ARITHMETIC, interp_method_lam: str = HARMONIC,
                 interp_method_mu: str = ARITHMETIC):
        self.lam = np.asarray(lam)
        self.mu = np.asarray(mu)
        self.rho = np.asarray(rho)
        for m in (interp_method_rho, interp_method_lam, interp_method_mu):
            assert m in (ARITHMETIC, HARMONIC)
        self.interp_method_rho = interp_method_rho
        self.interp_method_lam = interp_method_lam
        self.interp_method_mu = interp_method_mu

    # the reference's field names
    λ = property(lambda self: self.lam)
    μ = property(lambda self: self.mu)
    ρ = property(lambda self: self.rho)
