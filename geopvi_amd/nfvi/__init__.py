"""Variational-inference caller chain of the FWI-2D likelihood, on the device.

Drop-in counterparts of geopvi.nfvi (the flows and models the 2-D FWI
example uses: Linear, Real2Constr, Constr2Real, FlowsBasedDistribution,
VariationalInversion) whose parameters, base samples, transformed samples
and gradients stay on the GPU, so one inversion iteration is flow forward
-> batched FWI misfit+gradient -> backward -> optimiser step without host
round trips.
"""
from .flows import Constr2Real, Linear, Real2Constr, psvi_mask_2D  # noqa: F401
from .models import FlowsBasedDistribution, VariationalInversion  # noqa: F401
