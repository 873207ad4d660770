"""MI355X backend for the SeismicWaves.jl finite-difference forward and
adjoint-gradient path, with a Python mirror of the reference's public API.

    from seismicwaves import *
    params = InputParametersAcoustic(nt, dt, (nx, nz), (dx, dz),
                                     CPMLBoundaryConditionParameters(halo=20))
    swforward(params, VpAcousticCDMaterialProperties(vp), shots,
              runparams=RunParameters(parall="mi355x"))

All numerical work runs in libswx.so (HIP kernels for gfx950) through the C
ABI of include/swx.h; there is no CPU fallback.
"""
from .params import (ARITHMETIC, HARMONIC, CPMLBoundaryConditionParameters, ElasticIsoMaterialProperties,
                     GradParameters, InputParametersAcoustic, InputParametersElastic,
                     ReflectiveBoundaryConditionParameters, RunParameters, VpAcousticCDMaterialProperties,
                     VpRhoAcousticVDMaterialProperties)
from .shots import (ExternalForceShot, ExternalForceSources, MomentTensor2D, MomentTensorShot, MomentTensorSources,
                    ScalarReceivers, ScalarShot, ScalarSources, VectorReceivers, gaussderivstf, gaussstf, rickerstf)
from .misfits import CCTSMisfit, L2Misfit
from .points import coeffsinc1d, spread_positions
from .api import (AcousticCDCPMLWaveSimulation, AcousticVDStaggeredCPMLWaveSimulation, ElasticIsoCPMLWaveSimulation,
                  back_interp, build_wavesim, find_nearest_grid_points, interp_stag, mute_around_points,
                  run_swforward, run_swgradient, swforward, swgradient, swmisfit)
from . import _native

__all__ = [
    "ARITHMETIC", "HARMONIC", "CPMLBoundaryConditionParameters", "GradParameters", "InputParametersAcoustic",
    "ReflectiveBoundaryConditionParameters", "RunParameters", "VpAcousticCDMaterialProperties",
    "VpRhoAcousticVDMaterialProperties", "ScalarReceivers", "ScalarShot", "ScalarSources", "gaussderivstf",
    "gaussstf", "rickerstf", "L2Misfit", "CCTSMisfit", "AcousticCDCPMLWaveSimulation", "AcousticVDStaggeredCPMLWaveSimulation",
    "back_interp", "build_wavesim", "find_nearest_grid_points", "interp_stag", "mute_around_points",
    "run_swforward", "run_swgradient", "swforward", "swgradient", "swmisfit",
    "InputParametersElastic", "ElasticIsoMaterialProperties", "MomentTensor2D", "MomentTensorSources",
    "ExternalForceSources", "VectorReceivers", "MomentTensorShot", "ExternalForceShot",
    "ElasticIsoCPMLWaveSimulation", "coeffsinc1d", "spread_positions",
]
