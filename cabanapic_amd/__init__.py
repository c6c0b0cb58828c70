"""MI355X-native implementation of the CabanaPIC per-timestep hot path.

The compute path is the HIP library ``libcabanapic_amd.so`` built in-tree
from ``cabanapic_amd/csrc``; ``Simulation`` (in ``pic``) drives it through the
C-ABI declared in ``include/cabanapic_amd.h``.
"""
from ._capi import CpicError, load as load_library  # noqa: F401
from .pic import (FIELD_MEMBERS, INTERP_MEMBERS, PARTICLE_MEMBERS, Adjoint,  # noqa: F401
                  Simulation, TangentSimulation, make_step_params)
