"""KAN parameter head on the GPU: per-reach attributes -> routing parameters.

Host mirror of the reference's `KanHeadConfig` / `KanHead` (src/nn/kan_head.rs:
51-105 for the configuration, :187-246 for the forward): input Linear(F, H),
`num_hidden_layers` KanLayer(H, H) blocks, output Linear(H, P), sigmoid, and
the result split into one [N] tensor per learnable parameter. Forward and
backward run in the HIP kernels of ddrs_amd/csrc/ddrs_kan.hip through
`ddrs_kan_head_fwd` / `ddrs_kan_head_bwd`; there is no CPU path.

KanLayer is pykan's KANLayer as restated by the crate rskan v0.1.3:
    y[n, o] = sum_i mask[i, o] (scale_base[i, o] silu(x[n, i])
                                + scale_sp[i, o] sum_c coef[i, o, c] B_c(x[n, i]))
with B_c the order-k B-spline basis on input i's knots.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass
from typing import Dict, List

import numpy as np
import torch

from . import _native as N
from .sparse import _ptr, _stream

# pykan MultKAN's noise_scale, which the reference passes to every KanLayer
# (src/nn/kan_head.rs:48).
KAN_NOISE_SCALE = 0.3
# Xavier gain of the output layer (src/nn/kan_head.rs:46).
XAVIER_GAIN_OUTPUT = 0.1


class KanDims(ctypes.Structure):
    """ddrs_kan_dims of include/ddrs_hip.h."""

    _fields_ = [(n, N.c_int) for n in ("F", "H", "P", "L", "G", "k")]


class KanParams(ctypes.Structure):
    """ddrs_kan_params of include/ddrs_hip.h (device pointers)."""

    _fields_ = [(n, N.P) for n in ("input_weight", "input_bias", "grid", "coef", "scale_base",
                                   "scale_sp", "mask", "output_weight", "output_bias")]


class KanGrads(ctypes.Structure):
    """ddrs_kan_grads of include/ddrs_hip.h (device pointers)."""

    _fields_ = [(n, N.P) for n in ("input_weight", "input_bias", "coef", "scale_base",
                                   "scale_sp", "output_weight", "output_bias")]


# trained tensors, in the order of the autograd function's arguments
_TRAINED = ("input_weight", "input_bias", "coef", "scale_base", "scale_sp", "output_weight",
            "output_bias")


@dataclass
class KanHeadConfig:
    """Mirror of the reference's KanHeadConfig (src/nn/kan_head.rs:51-105);
    the optional disaggregation head is not part of this build."""

    input_var_names: List[str]
    learnable_parameters: List[str]
    seed: int
    hidden_size: int = 21
    num_hidden_layers: int = 2
    grid: int = 5
    k: int = 3

    def init(self, device="cuda") -> "KanHead":
        """Parameters drawn from the distributions of the reference's init
        (src/nn/kan_head.rs:117-185): Kaiming-normal input weight, Xavier-normal
        output weight (gain 0.1), zero biases, and pykan's KANLayer init for
        every block with the same seed (uniform extended grid on [-1, 1],
        coefficients fitted by least squares to noise of scale 0.3 / grid,
        scale_base = U(-1, 1) / sqrt(H), scale_sp = 1 / sqrt(H), mask 1).
        The draws come from torch's generator, not the reference's StdRng, so
        the values differ from the reference's at the same seed."""
        if not self.input_var_names or not self.learnable_parameters:
            raise ValueError("input_var_names and learnable_parameters must be non-empty")
        F, H, P = len(self.input_var_names), self.hidden_size, len(self.learnable_parameters)
        L, num, k = self.num_hidden_layers, self.grid, self.k
        gen = torch.Generator().manual_seed(self.seed)
        w_in = torch.randn(H, F, generator=gen, dtype=torch.float64) * math.sqrt(2.0 / F)
        w_out = torch.randn(P, H, generator=gen, dtype=torch.float64) * (
            XAVIER_GAIN_OUTPUT * math.sqrt(2.0 / (H + P)))
        h = 2.0 / num
        knots = torch.arange(-k, num + k + 1, dtype=torch.float64) * h - 1.0
        G, C = knots.numel(), knots.numel() - 1 - k
        grids, coefs, sbs, ssps, masks = [], [], [], [], []
        for _ in range(L):
            lg = torch.Generator().manual_seed(self.seed)
            grid = knots.expand(H, G).clone()
            x = torch.linspace(-1.0, 1.0, num + 1, dtype=torch.float64)
            noise = (torch.rand(num + 1, H, H, generator=lg, dtype=torch.float64) - 0.5) * (
                KAN_NOISE_SCALE / num)
            B = _basis_f64(x, knots, k)                       # [num+1, C]
            coef = torch.linalg.lstsq(B, noise.reshape(num + 1, H * H)).solution  # [C, H*H]
            grids.append(grid)
            coefs.append(coef.T.reshape(H, H, C))
            sbs.append((torch.rand(H, H, generator=lg, dtype=torch.float64) * 2 - 1)
                       / math.sqrt(H))
            ssps.append(torch.full((H, H), 1.0 / math.sqrt(H), dtype=torch.float64))
            masks.append(torch.ones(H, H, dtype=torch.float64))

        def stack(ts, shape):
            return torch.stack(ts) if ts else torch.zeros(shape, dtype=torch.float64)

        t = {
            "input_weight": w_in, "input_bias": torch.zeros(H, dtype=torch.float64),
            "grid": stack(grids, (0, H, G)), "coef": stack(coefs, (0, H, H, C)),
            "scale_base": stack(sbs, (0, H, H)), "scale_sp": stack(ssps, (0, H, H)),
            "mask": stack(masks, (0, H, H)),
            "output_weight": w_out, "output_bias": torch.zeros(P, dtype=torch.float64),
        }
        return KanHead(self.learnable_parameters, k, t, device)


def _basis_f64(x: torch.Tensor, knots: torch.Tensor, k: int) -> torch.Tensor:
    """Order-k B-spline basis of x on one knot row (Cox-de Boor), for the
    least-squares fit of the initial coefficients only."""
    x = x[:, None]
    B = ((x >= knots[:-1]) & (x < knots[1:])).to(torch.float64)
    for kk in range(1, k + 1):
        B = ((x - knots[:-(kk + 1)]) / (knots[kk:-1] - knots[:-(kk + 1)]) * B[:, :-1]
             + (knots[kk + 1:] - x) / (knots[kk + 1:] - knots[1:-kk]) * B[:, 1:])
        B = torch.nan_to_num(B)
    return B


class KanHead(torch.nn.Module):
    """The head's parameters on the device, stacked over blocks:
    input_weight [H, F], input_bias [H], grid [L, H, G], coef [L, H, H, C],
    scale_base / scale_sp / mask [L, H, H], output_weight [P, H],
    output_bias [P]. grid and mask are buffers (not trained)."""

    def __init__(self, learnable_parameters: List[str], k: int, tensors: Dict[str, torch.Tensor],
                 device="cuda"):
        super().__init__()
        self.learnable = list(learnable_parameters)
        self.k = int(k)
        dev = torch.device(device)
        for name in _TRAINED:
            self.register_parameter(name, torch.nn.Parameter(
                tensors[name].to(device=dev, dtype=torch.float32).contiguous()))
        for name in ("grid", "mask"):
            self.register_buffer(name, tensors[name].to(device=dev, dtype=torch.float32)
                                 .contiguous())
        H, F = self.input_weight.shape
        P = self.output_weight.shape[0]
        L, G = self.grid.shape[0], (self.grid.shape[2] if self.grid.dim() == 3 else 0)
        if P != len(self.learnable):
            raise ValueError(f"output width {P} != {len(self.learnable)} learnable parameters")
        if L and tuple(self.coef.shape) != (L, H, H, G - 1 - self.k):
            raise ValueError(f"coef shape {tuple(self.coef.shape)} does not match grid/k")
        self.dims = KanDims(F, H, P, L, G if L else self.k + 2, self.k)

    @classmethod
    def from_npz(cls, path: str, learnable_parameters: List[str], k: int,
                 device="cuda") -> "KanHead":
        """Load the reference's fixture layout (src/nn/kan_head.rs:299-360):
        input_weight/bias, output_weight/bias and block_{l}_{grid, coef,
        scale_base, scale_sp, mask} for l = 0, 1, ... Numeric arrays only
        (np.load without pickle)."""
        with np.load(path, allow_pickle=False) as d:
            names = set(d.files)
            L = 0
            while f"block_{L}_coef" in names:
                L += 1
            t = {n: torch.from_numpy(d[n].astype(np.float64))
                 for n in ("input_weight", "input_bias", "output_weight", "output_bias")}
            for n in ("grid", "coef", "scale_base", "scale_sp", "mask"):
                t[n] = torch.from_numpy(np.stack([d[f"block_{l}_{n}"].astype(np.float64)
                                                  for l in range(L)]))
        return cls(learnable_parameters, k, t, device)

    def learnable_parameters(self) -> List[str]:
        return list(self.learnable)

    def forward_packed(self, inputs: torch.Tensor) -> torch.Tensor:
        """[N, F] -> [P, N] (the rows of the reference's transposed output)."""
        return _KanHeadFn.apply(inputs, self.dims, self.grid, self.mask,
                                *[getattr(self, n) for n in _TRAINED])

    def forward(self, inputs: torch.Tensor) -> Dict[str, torch.Tensor]:
        """{name: [N]} for every learnable parameter (src/nn/kan_head.rs:215)."""
        out = self.forward_packed(inputs)
        return {name: out[i] for i, name in enumerate(self.learnable)}


def _params(dims: KanDims, grid, mask, t: Dict[str, torch.Tensor]) -> KanParams:
    return KanParams(_ptr(t["input_weight"]), _ptr(t["input_bias"]), _ptr(grid), _ptr(t["coef"]),
                     _ptr(t["scale_base"]), _ptr(t["scale_sp"]), _ptr(mask),
                     _ptr(t["output_weight"]), _ptr(t["output_bias"]))


def _workspace(dims: KanDims, n: int, device, backward: bool) -> torch.Tensor:
    fwd, bwd = N.c_size(), N.c_size()
    N.check(N.lib.ddrs_kan_workspace_bytes(ctypes.byref(dims), n, ctypes.byref(fwd),
                                           ctypes.byref(bwd)), "ddrs_kan_workspace_bytes")
    return torch.empty((bwd if backward else fwd).value, dtype=torch.uint8, device=device)


class _KanHeadFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, inputs, dims, grid, mask, *trained):
        if not inputs.is_cuda:
            raise TypeError("KanHead runs on the GPU: inputs must be a CUDA (HIP) tensor")
        x = inputs.detach().to(torch.float32).contiguous()
        if x.dim() != 2 or x.shape[1] != dims.F:
            raise ValueError(f"inputs must be [N, {dims.F}], got {tuple(x.shape)}")
        t = {n: v.detach().contiguous() for n, v in zip(_TRAINED, trained)}
        n = x.shape[0]
        out = torch.empty(dims.P, n, device=x.device, dtype=torch.float32)
        prm = _params(dims, grid, mask, t)
        ws = _workspace(dims, n, x.device, backward=False)
        N.check(N.lib.ddrs_kan_head_fwd(ctypes.byref(dims), ctypes.byref(prm), n, _ptr(x),
                                        _ptr(out), _ptr(ws), ws.numel(), _stream()),
                "ddrs_kan_head_fwd")
        ctx.dims = dims
        ctx.save_for_backward(x, grid, mask, *[t[n] for n in _TRAINED])
        ctx.in_dtype = inputs.dtype
        return out

    @staticmethod
    def backward(ctx, gout):
        x, grid, mask, *trained = ctx.saved_tensors
        t = dict(zip(_TRAINED, trained))
        dims = ctx.dims
        g = gout.detach().to(torch.float32).contiguous()
        grads = {n: torch.empty_like(v) for n, v in t.items()}
        gx = torch.empty_like(x) if ctx.needs_input_grad[0] else None
        prm = _params(dims, grid, mask, t)
        gr = KanGrads(*[_ptr(grads[n]) for n in _TRAINED])
        ws = _workspace(dims, x.shape[0], x.device, backward=True)
        N.check(N.lib.ddrs_kan_head_bwd(ctypes.byref(dims), ctypes.byref(prm), x.shape[0],
                                        _ptr(x), _ptr(g), ctypes.byref(gr), _ptr(gx), _ptr(ws),
                                        ws.numel(), _stream()), "ddrs_kan_head_bwd")
        if gx is not None:
            gx = gx.to(ctx.in_dtype)
        return (gx, None, None, None, *[grads[n] for n in _TRAINED])
