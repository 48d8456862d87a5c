"""CPU oracle of the KAN parameter head (test infrastructure only).

Only tests/, the smoke check and bench.py's CPU baseline use this module, and
only to check the HIP kernels; the product path never calls it.

The head (reference src/nn/kan_head.rs:187-215, `KanHead::forward`):
    h0 = inputs @ W_in^T + b_in                   Linear(F, H)
    h_{l+1} = KanLayer_l(h_l), l < L              KanLayer(H, H), no activation between
    out = sigmoid(h_L @ W_out^T + b_out)          Linear(H, P) -> sigmoid, [N, P]
Weights are kept in the fixture's torch layout, [out, in]
(src/nn/kan_head.rs:262-276 transposes them into BURN's [in, out]).

KanLayer comes from the crate rskan v0.1.3 (git tag v0.1.3, commit 44a46ebe,
not vendored in the reference). It restates pykan's KANLayer (Liu et al.,
"KAN: Kolmogorov-Arnold Networks", 2024; pykan KANLayer.forward, B_batch,
coef2curve), which is what this module follows:
    y[n, o] = sum_i mask[i, o] * (scale_base[i, o] * silu(x[n, i])
                                  + scale_sp[i, o] * sum_c coef[i, o, c] * B_c(x[n, i]; grid[i]))
with B_c the order-k B-spline basis on the extended grid of input i (G points,
G - 1 - k basis functions), built by the Cox-de Boor recursion
    B^0_j = [g_j <= x < g_{j+1}]
    B^k_j = (x - g_j) / (g_{j+k} - g_j) * B^{k-1}_j
            + (g_{j+k+1} - x) / (g_{j+k+1} - g_{j+1}) * B^{k-1}_{j+1}
with 0/0 replaced by 0 (torch.nan_to_num), and silu(x) = x * sigmoid(x).

Parity is pinned against the reference fixture tests/fixtures/kan_head_init_seed42.npz
(copied numerically to tests/golden/kan_head_seed42.npz): outputs to 1e-6 and
the gradients of sum(all outputs) to 1e-5, the tolerances of the reference's
tests/kan_head_fixture_forward.rs:69 and tests/kan_head_fixture_backward.rs:77.
The reverse mode below is derived by hand and checked there and against
central differences.
"""
from __future__ import annotations

import numpy as np


def sigmoid(z):
    return 1.0 / (1.0 + np.exp(-z))


def silu(x):
    return x * sigmoid(x)


def dsilu(x):
    s = sigmoid(x)
    return s * (1.0 + x * (1.0 - s))


def bspline_basis(x, grid, k):
    """Basis values and their x-derivatives: x [N, I], grid [I, G] ->
    (B [N, I, G-1-k], dB/dx [N, I, G-1-k]); the Cox-de Boor recursion with
    0/0 -> 0 at every order, the derivative carried along the recursion."""
    x = x[:, :, None]
    g = grid[None, :, :]
    B = ((x >= g[:, :, :-1]) & (x < g[:, :, 1:])).astype(x.dtype)
    dB = np.zeros_like(B)
    for kk in range(1, k + 1):
        d1 = g[:, :, kk:-1] - g[:, :, :-(kk + 1)]
        d2 = g[:, :, kk + 1:] - g[:, :, 1:-kk]
        with np.errstate(divide="ignore", invalid="ignore"):
            a = (x - g[:, :, :-(kk + 1)]) / d1
            c = (g[:, :, kk + 1:] - x) / d2
            val = a * B[:, :, :-1] + c * B[:, :, 1:]
            dval = (B[:, :, :-1] / d1 + a * dB[:, :, :-1]
                    - B[:, :, 1:] / d2 + c * dB[:, :, 1:])
        B = np.nan_to_num(val, nan=0.0, posinf=0.0, neginf=0.0)
        dB = np.nan_to_num(dval, nan=0.0, posinf=0.0, neginf=0.0)
    return B, dB


def kan_layer_fwd(x, layer, k):
    """One KanLayer: x [N, I] -> y [N, O] and the values the backward needs."""
    B, dB = bspline_basis(x, layer["grid"], k)
    spline = np.einsum("nic,ioc->nio", B, layer["coef"])          # [N, I, O]
    base = silu(x)                                                 # [N, I]
    m = layer["mask"][None]
    y = (m * (layer["scale_base"][None] * base[:, :, None]
              + layer["scale_sp"][None] * spline)).sum(axis=1)
    return y, dict(x=x, B=B, dB=dB, spline=spline, base=base)


def kan_layer_bwd(gy, layer, cache):
    """Gradients of one KanLayer for dL/dy = gy [N, O]."""
    x, B, dB, spline, base = (cache[k] for k in ("x", "B", "dB", "spline", "base"))
    m = layer["mask"]
    gym = gy[:, None, :] * m[None]                                 # [N, I, O]
    g_sb = (gym * base[:, :, None]).sum(axis=0)
    g_sp = (gym * spline).sum(axis=0)
    g_coef = np.einsum("nio,nic->ioc", gym * layer["scale_sp"][None], B)
    dspline = np.einsum("nic,ioc->nio", dB, layer["coef"])         # d spline / dx
    gx = (gym * (layer["scale_base"][None] * dsilu(x)[:, :, None]
                 + layer["scale_sp"][None] * dspline)).sum(axis=2)
    return gx, dict(coef=g_coef, scale_base=g_sb, scale_sp=g_sp)


def params_from_npz(d, n_layers=2):
    """The head's parameters from arrays named as in the reference fixture."""
    p = dict(input_weight=d["input_weight"], input_bias=d["input_bias"],
             output_weight=d["output_weight"], output_bias=d["output_bias"], blocks=[])
    for l in range(n_layers):
        p["blocks"].append({k: d[f"block_{l}_{k}"] for k in
                            ("grid", "coef", "scale_base", "scale_sp", "mask")})
    return p


def kan_head_fwd(inputs, p, k, dtype=np.float64):
    """Outputs [N, P] (sigmoid) and the forward cache."""
    cast = lambda a: np.asarray(a, dtype)
    h = cast(inputs) @ cast(p["input_weight"]).T + cast(p["input_bias"])
    caches = []
    for blk in p["blocks"]:
        layer = {kk: cast(v) for kk, v in blk.items()}
        h, c = kan_layer_fwd(h, layer, k)
        caches.append((layer, c))
    z = h @ cast(p["output_weight"]).T + cast(p["output_bias"])
    out = sigmoid(z)
    return out, dict(inputs=cast(inputs), hL=h, out=out, caches=caches)


def kan_head_bwd(gout, p, cache, dtype=np.float64):
    """Gradients of every trained parameter (and of the inputs) for
    dL/dout = gout [N, P]."""
    cast = lambda a: np.asarray(a, dtype)
    out = cache["out"]
    gz = cast(gout) * out * (1.0 - out)
    g = dict(output_weight=gz.T @ cache["hL"], output_bias=gz.sum(axis=0))
    gh = gz @ cast(p["output_weight"])
    blocks = []
    for layer, c in reversed(cache["caches"]):
        gh, gb = kan_layer_bwd(gh, layer, c)
        blocks.append(gb)
    g["blocks"] = blocks[::-1]
    g["input_weight"] = gh.T @ cache["inputs"]
    g["input_bias"] = gh.sum(axis=0)
    g["inputs"] = gh @ cast(p["input_weight"])
    return g
