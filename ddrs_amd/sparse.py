"""The reference's sparse seam on the HIP library.

SparseAdjacency  src/sparse/mod.rs:44-105
Network          CsrPattern (src/sparse/mod.rs:115-304) + the device plan and
                 workspace that replace the cuSPARSE pattern cache
triangular_csr_solve  src/sparse/mod.rs:483-531 with the CsrSolveOp backward
                 (425-472) as a torch.autograd.Function
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Optional

import numpy as np
import torch

from . import _native as N


def _ptr(t) -> Optional[int]:
    if t is None:
        return None
    if isinstance(t, np.ndarray):
        return t.ctypes.data
    return t.data_ptr()


def _stream() -> int:
    return torch.cuda.current_stream().cuda_stream


def _dev_f32(t: torch.Tensor, n: Optional[int] = None) -> torch.Tensor:
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise TypeError("expected a CUDA (HIP) tensor")
    t = t.detach().to(torch.float32).contiguous()
    if n is not None and t.numel() != n:
        raise ValueError(f"expected {n} elements, got {t.numel()}")
    return t


@dataclass
class SparseAdjacency:
    """COO of the routing network N: rows = downstream, cols = upstream,
    rows > cols (topological order), plus per-reach length [m] and slope."""

    n: int
    rows: np.ndarray
    cols: np.ndarray
    values: np.ndarray
    length_m: np.ndarray
    slope: np.ndarray

    @classmethod
    def from_coo(cls, n, rows, cols, values=None, length_m=None, slope=None):
        rows = np.ascontiguousarray(rows, np.int32)
        cols = np.ascontiguousarray(cols, np.int32)
        values = np.ones(rows.shape[0], np.float32) if values is None else np.ascontiguousarray(values, np.float32)
        length_m = np.ones(n, np.float32) if length_m is None else np.ascontiguousarray(length_m, np.float32)
        slope = np.ones(n, np.float32) if slope is None else np.ascontiguousarray(slope, np.float32)
        return cls(int(n), rows, cols, values, length_m, slope)

    @property
    def nnz(self) -> int:
        return int(self.rows.shape[0])


class Network:
    """A routing network on the GPU: the bit-exact CsrPattern on the host plus
    the level-major device schedule (ddrs_plan_create)."""

    def __init__(self, adj: SparseAdjacency, max_sub_levels: int = 0, block: int = 0):
        """max_sub_levels: longest chain of links one block of the device
        schedule may hold; block: positions per block, 64 or 128 (0: library
        defaults). Results do not depend on either."""
        self.adj = adj
        self.max_sub_levels, self.block = int(max_sub_levels), int(block)
        h = ctypes.c_void_p()
        N.check(N.lib.ddrs_plan_create_ex(adj.n, adj.nnz, _ptr(adj.rows), _ptr(adj.cols),
                                          _ptr(adj.values), int(max_sub_levels), int(block),
                                          ctypes.byref(h)), "ddrs_plan_create_ex")
        self.handle = h
        info = [ctypes.c_int32() for _ in range(4)]
        N.check(N.lib.ddrs_plan_info(h, *[ctypes.byref(x) for x in info]))
        self.n, self.nnz, self.n_chunks, self.n_levels = (int(x.value) for x in info)
        npos = ctypes.c_int32()
        N.check(N.lib.ddrs_plan_positions(h, ctypes.byref(npos)))
        self.positions = int(npos.value)
        self._ws = {}

    def __del__(self):
        for ws in getattr(self, "_ws", {}).values():
            ws.close()
        # at interpreter exit the library binding may already be gone
        if getattr(self, "handle", None) and getattr(N, "lib", None) is not None:
            N.lib.ddrs_plan_destroy(self.handle)
            self.handle = None

    def csr(self) -> dict:
        n, nnz = self.n, self.nnz
        out = dict(crow=np.zeros(n + 1, np.int32), col=np.zeros(nnz, np.int32),
                   row_for_nnz=np.zeros(nnz, np.int32), adj_values=np.zeros(nnz, np.float32),
                   diag_mask=np.zeros(nnz, np.float32), trans_crow=np.zeros(n + 1, np.int32),
                   trans_col=np.zeros(nnz, np.int32), trans_to_orig=np.zeros(nnz, np.int32))
        N.check(N.lib.ddrs_plan_get_csr(self.handle, *[_ptr(v) for v in out.values()]))
        return out

    def schedule(self):
        perm = np.zeros(self.n, np.int32)
        level = np.zeros(self.n, np.int32)
        N.check(N.lib.ddrs_plan_get_schedule(self.handle, _ptr(perm), _ptr(level)))
        return perm, level

    def permute_forcing(self, q_prime: torch.Tensor) -> torch.Tensor:
        """q' [rows, n] (reach order, float32 on the GPU) in schedule order,
        [rows, positions] (ddrs_plan_permute_forcing), for workspaces made with
        schedule_forcing=True: the data side lays out a batch's forcing once."""
        q = _dev_f32(q_prime, None)
        if q.dim() != 2 or q.shape[1] != self.n:
            raise ValueError(f"q_prime must be [rows, {self.n}], got {tuple(q.shape)}")
        out = torch.empty((q.shape[0], self.positions), dtype=torch.float32, device=q.device)
        N.check(N.lib.ddrs_plan_permute_forcing(self.handle, int(q.shape[0]), q.data_ptr(),
                                                out.data_ptr(), _stream()),
                "ddrs_plan_permute_forcing")
        return out

    def workspace_bytes(self, window: int, max_steps: int, dense_grad: bool = True,
                        schedule_forcing: bool = False) -> int:
        """Device bytes a Workspace(self, window, max_steps, dense_grad,
        schedule_forcing) would hold (without its gauge buffers); allocates
        nothing."""
        b = N.c_size()
        flags = _ws_flags(dense_grad, schedule_forcing)
        N.check(N.lib.ddrs_workspace_query_bytes(self.handle, int(window), int(max_steps), flags,
                                                 ctypes.byref(b)), "ddrs_workspace_query_bytes")
        return int(b.value)

    def workspace(self, window: int = 1, max_steps: int = 1) -> "Workspace":
        key = (int(window), int(max_steps))
        for k, ws in self._ws.items():
            # skip the gauge-readout workspaces (no dense dL/dQ buffer)
            if len(k) == 2 and k[0] >= key[0] and k[1] >= key[1]:
                return ws
        ws = Workspace(self, *key)
        self._ws[key] = ws
        return ws


def _ws_flags(dense_grad: bool, schedule_forcing: bool) -> int:
    return ((0 if dense_grad else N.WS_NO_DENSE_GRAD)
            | (N.WS_SCHEDULE_FORCING if schedule_forcing else 0))


class Workspace:
    def __init__(self, net: Network, window: int, max_steps: int, dense_grad: bool = True,
                 schedule_forcing: bool = False):
        """dense_grad=False leaves out the [max_steps][n] dense dL/dQ buffer
        (only the gauged adjoint works then). schedule_forcing=True: the window
        calls take q' in schedule order, [T, positions] from
        Network.permute_forcing, and read it in place."""
        self.net = net
        self.window = window
        self.max_steps = max_steps
        self.schedule_forcing = bool(schedule_forcing)
        h = ctypes.c_void_p()
        flags = _ws_flags(dense_grad, schedule_forcing)
        N.check(N.lib.ddrs_workspace_create_ex(net.handle, window, max_steps, flags,
                                               ctypes.byref(h)), "ddrs_workspace_create_ex")
        self.handle = h
        self.fwd_token = None

    def close(self):
        # at interpreter exit the library binding may already be gone
        if getattr(self, "handle", None) and getattr(N, "lib", None) is not None:
            N.lib.ddrs_workspace_destroy(self.handle)
            self.handle = None

    def __del__(self):
        self.close()

    def nbytes(self) -> int:
        b = ctypes.c_size_t()
        N.check(N.lib.ddrs_workspace_bytes(self.handle, ctypes.byref(b)))
        return int(b.value)

    def check(self):
        N.check(N.lib.ddrs_workspace_check(self.handle, _stream()), "ddrs_workspace_check")

    def layout(self) -> dict:
        """Chunk width (reaches per lane) and workgroup count of the persistent
        forward and adjoint-sweep launches."""
        v = [ctypes.c_int32() for _ in range(4)]
        N.check(N.lib.ddrs_workspace_layout(self.handle, *[ctypes.byref(x) for x in v]))
        return dict(fwd_reaches_per_lane=v[0].value, fwd_workgroups=v[1].value,
                    sweep_reaches_per_lane=v[2].value, sweep_workgroups=v[3].value)

    def set_chunk_width(self, fwd: int = 0, sweep: int = 0):
        """Override the chunk width (1 or 2 reaches per lane; 0 keeps the
        automatic choice). Results do not depend on it."""
        N.check(N.lib.ddrs_workspace_set_chunk_width(self.handle, int(fwd), int(sweep)),
                "ddrs_workspace_set_chunk_width")

    def set_gauges(self, reaches) -> None:
        """Gauge readout of the _gauged window calls: distinct reach ids
        (reference order); their forward writes Q only at these reaches,
        [T, K], and their adjoint takes dL/dQ only there. Empty clears."""
        r = np.ascontiguousarray(reaches, np.int32)
        N.check(N.lib.ddrs_workspace_set_gauges(self.handle, int(r.size),
                                                r.ctypes.data if r.size else None),
                "ddrs_workspace_set_gauges")
        self.gauges = r

    def set_linearize_in_forward(self, enable: bool, grid: int = 0):
        """Linearise the adjoint while a persistent forward window runs, so
        that the adjoint that follows skips that phase (bit-identical results;
        worth it only when an adjoint follows). grid: workgroups of the
        concurrent launch, 0 for the default. Currently disabled in the
        library: enable=True raises (a 350000-reach window did not finish)."""
        N.check(N.lib.ddrs_workspace_set_linearize_in_forward(self.handle, int(bool(enable)),
                                                              int(grid)),
                "ddrs_workspace_set_linearize_in_forward")


class _TriSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, a_values, b, net):
        ws = net.workspace()
        a = _dev_f32(a_values, net.nnz)
        bb = _dev_f32(b, net.n)
        x = torch.empty(net.n, dtype=torch.float32, device=bb.device)
        N.check(N.lib.ddrs_tri_solve_fwd(ws.handle, _ptr(a), _ptr(bb), _ptr(x), _stream()),
                "ddrs_tri_solve_fwd")
        ws.check()
        ctx.net = net
        ctx.save_for_backward(a, x)
        return x

    @staticmethod
    def backward(ctx, gx):
        a, x = ctx.saved_tensors
        net = ctx.net
        ws = net.workspace()
        g = _dev_f32(gx, net.n)
        gb = torch.empty_like(g)
        ga = torch.empty(net.nnz, dtype=torch.float32, device=g.device) if ctx.needs_input_grad[0] else None
        N.check(N.lib.ddrs_tri_solve_adj(ws.handle, _ptr(a), _ptr(x), _ptr(g), _ptr(gb), _ptr(ga),
                                         _stream()), "ddrs_tri_solve_adj")
        ws.check()
        return ga, gb, None


def triangular_csr_solve(net: Network, a_values: torch.Tensor, b: torch.Tensor) -> torch.Tensor:
    """Solve A x = b for A in CsrPattern order (sparse/mod.rs:483-531)."""
    return _TriSolve.apply(a_values, b, net)
