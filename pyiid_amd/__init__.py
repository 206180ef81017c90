"""pyiid_amd: MI355X (gfx950) backend for pyIID's scattering hot path.

Debye F(Q), its position gradient, the F(Q) -> G(r) transform and the Rw /
chi^2 energy and forces run as hand-written HIP kernels in
``pyiid_amd/_lib/libpyiid_hip.so`` behind a C ABI (``include/pyiid_hip.h``),
called through ctypes.  The Python side mirrors pyIID's ``ElasticScatter``
and ``Calc1D`` interfaces.
"""
from . import _lib  # noqa: F401
from .atoms import Atoms  # noqa: F401

__version__ = "0.1.0"


def __getattr__(name):
    # lazy: importing the package must not touch the GPU
    if name == "ElasticScatter":
        from .elasticscatter import ElasticScatter
        return ElasticScatter
    if name == "Calc1D":
        from .calc import Calc1D
        return Calc1D
    raise AttributeError(name)
