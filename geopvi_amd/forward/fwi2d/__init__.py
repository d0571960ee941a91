"""2-D acoustic FWI on MI355X: drop-in for geopvi.forward.fwi2d.

``aco2d`` mirrors the reference's Cython extension (fwi / forward),
``posterior`` its likelihood adapter; ``solver`` holds the device-level
plain, batched and autograd entry points.
"""
from . import aco2d  # noqa: F401
from .params import FWIParams, prepare_fwi_parameters  # noqa: F401
