"""MI355X-native force-and-integrate step for the TensorDEM 2-D bonded-disk ice solver.

`DEMConfig` and `IceDEM` are drop-ins for `tensordem.DEMConfig` /
`tensordem.IceDEM` (src/tensordem/dem.py); `PlateEngine` runs many plates in
one launch; `DEM3DConfig` and `IceDEM3D` are drop-ins for the 3-D
bonded-sphere solver (src/tensordem/dem3d.py).  The compute path is the HIP library built by
`tensordem_amd.build.build()`; importing the package does not need it, but
constructing an engine does, and fails loudly without it.
"""
from .config import DEMConfig
from .dem3dconfig import DEM3DConfig

__all__ = ["DEMConfig", "DEM3DConfig", "IceDEM", "IceDEM3D", "PlateEngine", "PlateSpec"]


def __getattr__(name):
    if name == "IceDEM":
        from .icedem import IceDEM
        return IceDEM
    if name == "IceDEM3D":
        from .dem3d import IceDEM3D
        return IceDEM3D
    if name in ("PlateEngine", "PlateSpec"):
        from . import engine
        return getattr(engine, name)
    raise AttributeError(name)
