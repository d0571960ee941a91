"""geopvi_amd: MI355X-native FWI-2D likelihood path of GeoPVI.

The 2-D acoustic time-domain finite-difference forward + adjoint solver that
returns each sample's waveform misfit and velocity gradient, as hand-written
gfx950 HIP kernels behind a C ABI (include/geopvi_fwi2d.h), with the
reference's Python interface on top (geopvi_amd.forward.fwi2d).
"""
import os as _os

# Hardware queues the HIP runtime gives this process (read when the runtime
# starts; HIP's own default is 4).  The batched solver and the server each
# run two lanes of step launches on two streams so that one lane fills the
# GPU while the other's launches drain; two streams that share a hardware
# queue run one after the other.  With 4 queues and the streams torch and
# the library create, the two lanes of a server were seen to share one; with
# 8 they get their own.  Set here unless the environment already sets it; no
# effect once the runtime has started.
_os.environ.setdefault("GPU_MAX_HW_QUEUES", "8")

__version__ = "0.1.0"
