"""ddrs_amd — the differentiable Muskingum-Cunge routing timestep of ddrs on
MI355X (gfx950): HIP kernels behind the C ABI of include/ddrs_hip.h, with a
PyTorch host mirror of the reference's routing and sparse interfaces."""
from . import _native
from .config import AttributeMinimums, ParameterRanges, Params, DT_SECONDS
from .sparse import Network, SparseAdjacency, Workspace, triangular_csr_solve
from .routing import (MuskingumCunge, RoutingInputs, SpatialParameters, denormalize, hotstart,
                      route_window, route_window_gauged, timestep_forward)
from ._native import FORM_PERSISTENT, FORM_PLAIN

__all__ = [
    "AttributeMinimums", "ParameterRanges", "Params", "DT_SECONDS", "Network",
    "SparseAdjacency", "Workspace", "triangular_csr_solve", "MuskingumCunge", "RoutingInputs",
    "SpatialParameters", "denormalize", "hotstart", "route_window", "route_window_gauged",
    "timestep_forward",
    "FORM_PERSISTENT", "FORM_PLAIN",
]
