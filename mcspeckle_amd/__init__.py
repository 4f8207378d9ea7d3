"""MI355X-native speckle Monte Carlo core (SCMC's MCcov / MCfield /
MCfieldFourierQuad, the illumination-wave variants, MCCovHetro and
MCsampleHetro) behind the reference's interface.

The compute runs in libscmc.so (HIP kernels for gfx950); this package is the
host-side mirror of the reference's MATLAB interface (scmc.m and the
interface/ builders) over its C-ABI.
"""
from .engines import (MCCovHetro, MCcov, MCfield, MCfieldFourierQuad, MCfieldFourierQuadOnWave, MCfieldOnWave,
                      MCsampleHetro, ScmcArgumentError, make_problem, mccov_spec, mcfield_spec, quad_grid)
from .interface import (HGScatter, boxArea, farFieldSource, isotropicScatter, nearFieldSource,
                        tabulatedAmplitudeScatter)
from .scmc import scmc

__all__ = [
    "MCcov", "MCfield", "MCfieldFourierQuad", "MCfieldOnWave", "MCfieldFourierQuadOnWave", "MCCovHetro", "MCsampleHetro", "scmc", "boxArea", "farFieldSource", "nearFieldSource",
    "isotropicScatter", "HGScatter", "tabulatedAmplitudeScatter", "ScmcArgumentError", "make_problem",
    "mccov_spec", "mcfield_spec", "quad_grid",
]
