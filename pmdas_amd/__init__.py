"""pmdas_amd -- MI355X-native MDAS sampler (multidimensional adaptive sampling).

The product is libmdas.so: hand-written HIP kernels for gfx950 behind the C ABI
of include/mdas.h.  This package holds its sources (csrc/), the ctypes binding
(_lib) and a Python mirror of the reference's mdas::KDGrid (kdgrid).
"""
from ._lib import (MdasError, load, SCENE_AO, SCENE_CONSTANT, SCENE_DISKS_MB,  # noqa: F401
                   SCENE_RAMP, SCENE_SPHERES_DOF, SCENE_SPHERES_MBDOF)


def __getattr__(name):
    if name in ("KDGrid", "Scene"):
        from . import kdgrid
        return getattr(kdgrid, name)
    raise AttributeError(name)
