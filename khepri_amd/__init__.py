"""khepri_amd -- MI355X-native RCWA solve path behind Khepri's Crystal API.

    from khepri_amd import Crystal, Drawing
    cl = Crystal((5, 5))
    cl.add_layer_pixmap("Scyl", eps2d, 0.55)
    ...
    cl.set_source(1 / freqs, 1.0, 0.0)   # arrays -> one batched solve
    cl.solve()
    R, T = cl.poynting_flux_end()

All numerics run in libkhx.so (hand-written HIP kernels for gfx950); the
package raises if the library or a HIP device is missing.
"""
from .draw import Drawing  # noqa: F401
from .expansion import Expansion  # noqa: F401

__version__ = "0.1.0"


def __getattr__(name):
    # lazy imports: torch-backed modules load on first use
    if name in ("Crystal", "Multilayer"):
        from . import crystal
        return getattr(crystal, name)
    if name == "Layer":
        from .layer import Layer
        return Layer
    if name == "ExtendedLayer":
        from .extension import ExtendedLayer
        return ExtendedLayer
    if name in ("save", "load"):
        from . import state
        return getattr(state, name)
    raise AttributeError(name)
