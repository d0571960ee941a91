"""geopvi_amd: MI355X-native FWI-2D likelihood path of GeoPVI.

The 2-D acoustic time-domain finite-difference forward + adjoint solver that
returns each sample's waveform misfit and velocity gradient, as hand-written
gfx950 HIP kernels behind a C ABI (include/geopvi_fwi2d.h), with the
reference's Python interface on top (geopvi_amd.forward.fwi2d).
"""
__version__ = "0.1.0"
