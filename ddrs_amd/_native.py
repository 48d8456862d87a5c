"""ctypes binding of the in-tree HIP library ddrs_amd/libddrs_hip.so.

There is no fallback: if the library is missing, importing this module raises
and every routing entry point is unavailable.
"""
from __future__ import annotations

import ctypes
import os

# torch first: PyTorch-ROCm ships its own HIP runtime, and the library must
# bind to the one torch's tensors and streams live in. Loaded the other way
# round, the process holds two runtimes and the library's calls find no device.
import torch  # noqa: F401

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libddrs_hip.so")

c_int = ctypes.c_int32
c_float = ctypes.c_float
c_size = ctypes.c_size_t
c_u64 = ctypes.c_uint64
P = ctypes.c_void_p


class DdrsParams(ctypes.Structure):
    """ddrs_params of include/ddrs_hip.h."""

    _fields_ = [
        ("dt", c_float),
        ("discharge_lb", c_float),
        ("depth_lb", c_float),
        ("velocity_lb", c_float),
        ("bottom_width_lb", c_float),
        ("slope_lb", c_float),
        ("ddr_match", c_int),
        ("enforce_positivity", c_int),
    ]


FORM_PLAIN = 0
FORM_PERSISTENT = 1
WS_NO_DENSE_GRAD = 1
WS_SCHEDULE_FORCING = 2
LOSS_L1, LOSS_NSE_BATCH, LOSS_NNSE_KGE, LOSS_KGE = 0, 1, 2, 3


class LossConfig(ctypes.Structure):
    """ddrs_loss_config (the reference's LossConfig, src/config.rs:220-262)."""
    _fields_ = [("kind", ctypes.c_int32), ("eps", ctypes.c_float),
                ("nnse_weight", ctypes.c_float), ("kge_weight", ctypes.c_float),
                ("r_weight", ctypes.c_float), ("alpha_weight", ctypes.c_float),
                ("beta_weight", ctypes.c_float), ("kge_clamp", ctypes.c_float)]

# name -> (restype, argtypes); restype None means ddrs_status (int32)
_SIGS = {
    "ddrs_last_error": (ctypes.c_char_p, []),
    "ddrs_version": (ctypes.c_char_p, []),
    "ddrs_plan_create": (c_int, [c_int, c_int, P, P, P, ctypes.POINTER(P)]),
    "ddrs_plan_create_ex": (c_int, [c_int, c_int, P, P, P, c_int, c_int, ctypes.POINTER(P)]),
    "ddrs_plan_destroy": (None, [P]),
    "ddrs_plan_info": (c_int, [P, P, P, P, P]),
    "ddrs_plan_get_csr": (c_int, [P, P, P, P, P, P, P, P, P]),
    "ddrs_plan_get_schedule": (c_int, [P, P, P]),
    "ddrs_plan_positions": (c_int, [P, P]),
    "ddrs_plan_permute_forcing": (c_int, [P, c_int, P, P, P]),
    "ddrs_workspace_create": (c_int, [P, c_int, c_int, ctypes.POINTER(P)]),
    "ddrs_workspace_create_ex": (c_int, [P, c_int, c_int, c_int, ctypes.POINTER(P)]),
    "ddrs_workspace_destroy": (None, [P]),
    "ddrs_workspace_bytes": (c_int, [P, ctypes.POINTER(c_size)]),
    "ddrs_workspace_query_bytes": (c_int, [P, c_int, c_int, c_int, ctypes.POINTER(c_size)]),
    "ddrs_workspace_check": (c_int, [P, P]),
    "ddrs_mc_bind_statics": (c_int, [P, ctypes.POINTER(DdrsParams), P, P, P, P, P, P, P]),
    "ddrs_mc_step_fwd": (c_int, [P, ctypes.POINTER(DdrsParams), P, P, P, P, P]),
    "ddrs_mc_step_bwd": (c_int, [P, ctypes.POINTER(DdrsParams), P, P, P, P, P, P, P, P, P, P]),
    "ddrs_mc_window_fwd": (c_int, [P, ctypes.POINTER(DdrsParams), c_int, P, P, P, c_int, P]),
    "ddrs_mc_window_adj": (c_int, [P, ctypes.POINTER(DdrsParams), c_int, P, P, P, P, P, P, P, P,
                                   c_int, P]),
    "ddrs_workspace_set_gauges": (c_int, [P, c_int, P]),
    "ddrs_mc_window_fwd_gauged": (c_int, [P, ctypes.POINTER(DdrsParams), c_int, P, P, P, c_int,
                                          P]),
    "ddrs_mc_window_adj_gauged": (c_int, [P, ctypes.POINTER(DdrsParams), c_int, P, P, P, P, P,
                                          P, P, P, c_int, P]),
    "ddrs_tri_solve_fwd": (c_int, [P, P, P, P, P]),
    "ddrs_tri_solve_adj": (c_int, [P, P, P, P, P, P, P]),
    "ddrs_hotstart": (c_int, [P, ctypes.POINTER(DdrsParams), P, P, P]),
    "ddrs_hotstart_noalloc": (c_int, [P, ctypes.POINTER(DdrsParams), P, P, P, P]),
    "ddrs_hotstart_adj": (c_int, [P, ctypes.POINTER(DdrsParams), P, P, P, P]),
    "ddrs_set_timing": (c_int, [P, c_int]),
    "ddrs_get_timing": (c_int, [P, P, P, P, P]),
    "ddrs_get_adj_phase_timing": (c_int, [P, P, P, P]),
    "ddrs_workspace_layout": (c_int, [P, P, P, P, P]),
    "ddrs_workspace_set_chunk_width": (c_int, [P, c_int, c_int]),
    "ddrs_workspace_set_linearize_in_forward": (c_int, [P, c_int, c_int]),
    "ddrs_set_trace": (c_int, [P, c_int]),
    "ddrs_get_trace": (c_int, [P, c_int, P, c_int, P]),
    "ddrs_build_topology": (c_int, [c_int, P, P, c_int, P, P, P, c_int, P, P, c_int, P]),
    "ddrs_compress": (c_int, [c_int, P, P, c_int, P, P, c_int, P, P, P, P, P]),
    "ddrs_compress_ex": (c_int, [c_int, P, P, c_int, P, c_int, c_int, P, P, c_int, P, P, P, P,
                                 P, P, P, P, P, c_int, P]),
    "ddrs_gen_tree": (c_int, [c_int, c_int, c_u64, P, P]),
    "ddrs_gen_attributes": (c_int, [c_int, c_u64, P, P, P, P]),
    "ddrs_gen_forcing": (c_int, [c_int, c_int, c_u64, c_int, P]),
    "ddrs_gen_forcing_device": (c_int, [c_int, c_int, c_u64, c_int, P, P]),
    "ddrs_kan_workspace_bytes": (c_int, [P, c_int, ctypes.POINTER(c_size),
                                         ctypes.POINTER(c_size)]),
    "ddrs_kan_head_fwd": (c_int, [P, P, c_int, P, P, P, c_size, P]),
    "ddrs_kan_head_bwd": (c_int, [P, P, c_int, P, P, P, P, P, c_size, P]),
    "ddrs_zarr_array_info": (c_int, [ctypes.c_char_p, P, P, P, c_int, P]),
    "ddrs_zarr_read_array": (c_int, [ctypes.c_char_p, c_int, P, ctypes.c_int64]),
    "ddrs_zarr_group_attr": (c_int, [ctypes.c_char_p, ctypes.c_char_p, P, P, c_int]),
    "ddrs_gauge_plan_create": (c_int, [c_int, c_int, c_int, P, P, c_int, ctypes.POINTER(P)]),
    "ddrs_gauge_plan_destroy": (None, [P]),
    "ddrs_scatter_add_by_group": (c_int, [P, P, c_int, ctypes.c_int64, ctypes.c_int64, P, P]),
    "ddrs_scatter_add_by_group_bwd": (c_int, [P, P, c_int, ctypes.c_int64, ctypes.c_int64, P,
                                              P]),
    "ddrs_tau_trim_downsample": (c_int, [P, c_int, c_int, c_int, P, P]),
    "ddrs_tau_trim_downsample_bwd": (c_int, [P, c_int, c_int, c_int, P, P]),
    "ddrs_nse_batch_loss": (c_int, [P, P, P, P, c_int, c_int, c_float, P, P]),
    "ddrs_nse_batch_loss_bwd": (c_int, [P, P, P, c_int, c_int, c_float, P, P, P]),
    "ddrs_gauge_nse_loss": (c_int, [P, P, c_int, ctypes.c_int64, ctypes.c_int64, c_int, c_int,
                                    P, P, c_float, P, P, P, P, P]),
    "ddrs_gauge_batch_loss": (c_int, [P, ctypes.POINTER(LossConfig), P, c_int, ctypes.c_int64,
                                      ctypes.c_int64, c_int, c_int, P, P, P, P, P, P, P]),
}

EXPORTED = tuple(_SIGS)


class DdrsError(RuntimeError):
    pass


def _load() -> ctypes.CDLL:
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is not built; run `python -m ddrs_amd.build_native` "
            "(or __graft_entry__.build()) first. There is no CPU fallback.")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


def check(status: int, what: str = "") -> None:
    if status != 0:
        msg = lib.ddrs_last_error().decode(errors="replace")
        raise DdrsError(f"{what or 'ddrs'} failed (status {status}): {msg}")
