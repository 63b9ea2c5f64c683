"""flowvpm_amd -- MI355X-native particle-to-particle U/J path of FLOWVPM.

The compute lives in a C-ABI shared library of hand-written HIP kernels for
gfx950 (csrc/, built into lib/libflowvpm_hip.so).  This Python package is the
host-side mirror of the reference's call surface for that path:

  ParticleField, add_particle, remove_particle     (src/FLOWVPM_particlefield.jl)
  UJ_hip  -- the UJ functor, UJ_direct semantics   (src/FLOWVPM_UJ.jl:21-50)
  p2p_ranges -- FastMultipole leaf-pair P2P        (src/FLOWVPM_fmm.jl:102-168)
  rungekutta3, relaxation, DynamicSFS, CoreSpreading (device-resident)
  uj_adjoint -- hand-written adjoint all-pairs kernels

The directory is named flowvpm.jl_amd; import it through
__graft_entry__.load_package(), which registers it as `flowvpm_amd`.
"""
from . import fields, metrics, workloads  # noqa: F401
from .fields import (NFIELDS, X_INDEX, GAMMA_INDEX, SIGMA_INDEX, VOL_INDEX,  # noqa: F401
                     CIRCULATION_INDEX, U_INDEX, VORTICITY_INDEX, J_INDEX, PSE_INDEX,
                     M_INDEX, C_INDEX, SFS_INDEX, STATIC_INDEX, U_PREV_INDEX)

__version__ = "0.1.0"


def __getattr__(name):
    # The HIP-backed API is imported on first use so that the pure-Python
    # parts (layout, workloads, metrics) work on machines without the library.
    import importlib
    api = importlib.import_module(__name__ + ".api")
    if name == "api":
        return api
    if hasattr(api, name):
        return getattr(api, name)
    raise AttributeError(name)
