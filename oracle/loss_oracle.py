"""CPU oracle of the gauge readout and training loss (test infrastructure only).

Only tests/, the smoke check and bench.py's CPU baseline use this module, and
only to check the HIP kernels; the product path never calls it.

Restates, in numpy:
- scatter_add_by_group (reference src/training/forward.rs:18-34):
  out[g, t] = sum over k with group_ids[k] == g of runoff[flat_indices[k], t],
  added in increasing k;
- area_pool_weights / tau_trim_and_downsample (src/training/loss.rs:33-95):
  hours [tau, T - (24 - tau)) pooled into T_days = (T - 24) // 24 bins of width
  s = L / M with fractional overlaps W[i, j] = (min(j+1, (i+1)s) - max(j, i s)) / s,
  all in float32 as the reference computes them;
- the driver's warmup slice and NaN-gauge filter (src/training/driver.rs:150-195)
  and the batch losses of batch_loss (src/training/loss.rs:179-208):
  nse-batch (loss.rs:226-237), the mean over kept gauges and post-warmup days
  of (p - o)^2 / (sigma + eps)^2, and L1 (loss.rs:186), the mean of |p - o|
  (its gradient sign(p - o) / n, 0 where p == o), and the per-gauge
  objectives nnse_kge_loss (loss.rs:257-318) and kge_component_loss
  (loss.rs:320-373): from each gauge's means, population moments with eps, the
  KGE parts r, alpha, beta and NNSE = 1 / (1 + sse / sso), averaged over
  gauges (their gradients derived by hand).

Pinned by the reference's own tests (tests/test_loss_oracle.py): the
scatter fixture of src/training/forward.rs:1030-1058, the hand-computed
nse_batch_loss of loss.rs:491-507, the pooling cases of loss.rs:609-727, and
torch's adaptive_avg_pool1d where the window is a whole number of days.
"""
from __future__ import annotations

import numpy as np


def scatter_add_by_group(runoff, flat_indices, group_ids, num_gauges, dtype=np.float32):
    """runoff [N, T] -> [G, T], sums in increasing k (float32: the reference's
    rounding)."""
    runoff = np.asarray(runoff, dtype)
    out = np.zeros((num_gauges, runoff.shape[1]), dtype)
    for k in range(len(flat_indices)):
        out[group_ids[k]] += runoff[flat_indices[k]]
    return out


def scatter_add_by_group_bwd(gout, flat_indices, group_ids, n_reaches):
    """dL/d runoff [N, T] for dL/d out = gout [G, T]."""
    gout = np.asarray(gout, np.float32)
    g = np.zeros((n_reaches, gout.shape[1]), np.float32)
    for k in range(len(flat_indices)):
        g[flat_indices[k]] += gout[group_ids[k]]
    return g


def area_pool_weights(l, m):
    """W [M, L] of loss.rs:70-95 in float32."""
    assert 0 < m <= l
    s = np.float32(l) / np.float32(m)
    W = np.zeros((m, l), np.float32)
    for i in range(m):
        left = np.float32(i) * s
        right = np.float32(i + 1) * s
        j_lo = int(np.floor(left))
        j_hi = min(int(np.ceil(right)), l)
        for j in range(j_lo, j_hi):
            cl = max(np.float32(j), left)
            cr = min(np.float32(j + 1), right)
            W[i, j] = np.float32(cr - cl) / s
    return W


def trim_window(t_hours, tau):
    assert 0 <= tau < 24
    start, end = tau, t_hours - (24 - tau)
    assert start < end
    L = end - start
    M = L // 24
    assert M > 0
    return start, end, L, M


def tau_trim_and_downsample(pred_hourly, tau, dtype=np.float64):
    """[G, T_hours] -> [G, T_days]: W applied in `dtype` (float64 for the
    reference value of the float32 weights, float32 to time it)."""
    x = np.asarray(pred_hourly)
    if x.dtype != np.float64:
        x = x.astype(np.float32)
    start, end, L, M = trim_window(x.shape[1], tau)
    W = area_pool_weights(L, M).astype(dtype)
    return x[:, start:end].astype(dtype) @ W.T


def tau_trim_and_downsample_bwd(gout, t_hours, tau):
    """dL/d pred_hourly [G, T_hours] (zero outside the trimmed window)."""
    gout = np.asarray(gout, np.float64)
    start, end, L, M = trim_window(t_hours, tau)
    W = area_pool_weights(L, M).astype(np.float64)
    g = np.zeros((gout.shape[0], t_hours))
    g[:, start:end] = gout @ W
    return g


def kept_gauges(obs, warmup):
    """Gauges whose post-warmup observations hold no NaN (driver.rs:168-176)."""
    obs = np.asarray(obs)
    return ~np.isnan(obs[:, warmup:]).any(axis=1)


def nse_batch_loss(p, o, sigma, eps):
    """Mean over (gauge, day) of (p - o)^2 / (sigma + eps)^2, in float64."""
    p, o = np.asarray(p, np.float64), np.asarray(o, np.float64)
    d = (np.asarray(sigma, np.float64) + eps)[:, None]
    return float((((p - o) ** 2) / (d * d)).mean())


def _kge_parts(p, o, eps):
    """Per-gauge statistics of nnse_kge_loss / kge_component_loss
    (loss.rs:257-373), float64: means, population variances, covariance, the
    std's with eps, r, alpha, beta, and 1 - NNSE with its x = sse / sso."""
    p, o = np.asarray(p, np.float64), np.asarray(o, np.float64)
    n = p.shape[1]
    mp, mo = p.mean(axis=1), o.mean(axis=1)
    pc, oc = p - mp[:, None], o - mo[:, None]
    var_p, var_o = (pc * pc).mean(axis=1), (oc * oc).mean(axis=1)
    sp, so = np.sqrt(var_p + eps), np.sqrt(var_o + eps)
    cov = (pc * oc).mean(axis=1)
    r = cov / (sp * so)
    alpha = sp / so
    beta = mp / (mo + eps)
    sse = ((p - o) ** 2).sum(axis=1)
    sso = (oc * oc).sum(axis=1) + eps
    nnse = 1.0 / (1.0 + sse / sso)
    return dict(n=n, mp=mp, mo=mo, pc=pc, oc=oc, resid=p - o, sp=sp, so=so, r=r,
                alpha=alpha, beta=beta, sso=sso, nnse=nnse)


def kge_batch_loss(p, o, kind, eps, nnse_weight=1.0, kge_weight=1.0, r_weight=1.0,
                   alpha_weight=1.0, beta_weight=1.0, clamp=10.0):
    """(loss, dL/dp) of nnse_kge_loss (kind "nnse-kge", loss.rs:257-318) or
    kge_component_loss (kind "kge", loss.rs:320-373) over p, o [G, T], both
    means over gauges; the gradient derived by hand (derivatives of the means,
    population moments, r, alpha, beta and NNSE), the clamp passing gradient
    where the component sum is at most the clamp."""
    k = _kge_parts(p, o, eps)
    G, n = len(k["r"]), k["n"]
    dr, da, db = k["r"] - 1.0, k["alpha"] - 1.0, k["beta"] - 1.0
    sp, so = k["sp"], k["so"]
    # d r / dp_t = oc_t / (n sp so) - r pc_t / (n sp^2); d alpha / dp_t = pc_t / (n sp so);
    # d beta / dp_t = 1 / (n (mo + eps))
    jr_pc, jr_oc = -k["r"] / (n * sp * sp), 1.0 / (n * sp * so)
    ja_pc = 1.0 / (n * sp * so)
    jb = 1.0 / (n * (k["mo"] + eps))
    nnse_term = 1.0 - k["nnse"]
    g_resid = 2.0 * k["nnse"] ** 2 / k["sso"]          # d(1 - nnse)/d resid_t per unit resid
    if kind == "nnse-kge":
        dist = np.sqrt(dr * dr + da * da + db * db)
        loss_g = nnse_weight * nnse_term + kge_weight * dist
        with np.errstate(divide="ignore", invalid="ignore"):
            w = kge_weight / dist
        A = w * (dr * jr_pc + da * ja_pc)
        B = w * dr * jr_oc
        D0 = w * db * jb
        C = nnse_weight * g_resid
    elif kind == "kge":
        term = r_weight * dr * dr + alpha_weight * da * da + beta_weight * db * db
        m = (term <= clamp).astype(np.float64)
        loss_g = np.minimum(term, clamp)
        A = 2.0 * m * (r_weight * dr * jr_pc + alpha_weight * da * ja_pc)
        B = 2.0 * m * r_weight * dr * jr_oc
        D0 = 2.0 * m * beta_weight * db * jb
        C = np.zeros(G)
        if nnse_weight != 0.0:
            loss_g = loss_g + nnse_weight * nnse_term
            C = nnse_weight * g_resid
    else:
        raise ValueError(kind)
    loss = float(loss_g.mean())
    grad = (A[:, None] * k["pc"] + B[:, None] * k["oc"] + C[:, None] * k["resid"]
            + D0[:, None]) / G
    return loss, grad


def l1_batch_loss(p, o):
    """Mean of |p - o| (loss.rs:186), in float64."""
    return float(np.abs(np.asarray(p, np.float64) - np.asarray(o, np.float64)).mean())


def gauge_nse_loss(runoff, flat_indices, group_ids, num_gauges, tau, warmup, obs, sigma, eps,
                   dtype=np.float64, kind="nse-batch", **weights):
    """The training objective from the routed discharge: (loss, n_valid,
    dL/d runoff). runoff [N, T_hours]; obs [G, T_days] with NaN for missing
    days; sigma [G] (nse-batch only). With no gauge kept the loss is 0 and
    n_valid 0. The gauge sums run in `dtype`, the rest in float64. kind:
    "nse-batch", "l1", "nnse-kge" or "kge" (weights: kge_batch_loss's);
    n_valid is loss_denominator (loss.rs:164-169): kept gauges x days for
    the first two, kept gauges for the others."""
    if kind not in ("nse-batch", "l1", "nnse-kge", "kge"):
        raise ValueError(kind)
    hourly = scatter_add_by_group(runoff, flat_indices, group_ids, num_gauges, dtype)
    daily = tau_trim_and_downsample(hourly, tau)
    keep = kept_gauges(obs, warmup)
    t_post = daily.shape[1] - warmup
    n_valid = int(keep.sum()) * (t_post if kind in ("nse-batch", "l1") else 1)
    g_daily = np.zeros_like(daily)
    if n_valid == 0:
        loss = 0.0
    elif kind in ("nnse-kge", "kge"):
        sel = np.ix_(keep, np.arange(warmup, daily.shape[1]))
        loss, g = kge_batch_loss(daily[keep, warmup:], np.asarray(obs, np.float64)[keep, warmup:],
                                 kind, eps, **weights)
        g_daily[sel] = g
    else:
        p = daily[keep, warmup:]
        o = np.asarray(obs, np.float64)[keep, warmup:]
        sel = np.ix_(keep, np.arange(warmup, daily.shape[1]))
        if kind == "l1":
            loss = float(np.abs(p - o).sum() / n_valid)
            g_daily[sel] = np.sign(p - o) / n_valid
        else:
            d = (np.asarray(sigma, np.float64)[keep] + eps)[:, None]
            loss = float((((p - o) ** 2) / (d * d)).sum() / n_valid)
            g_daily[sel] = 2.0 * (p - o) / (d * d) / n_valid
    g_hourly = tau_trim_and_downsample_bwd(g_daily, hourly.shape[1], tau)
    g_runoff = np.zeros(np.asarray(runoff).shape)
    for k in range(len(flat_indices)):
        g_runoff[flat_indices[k]] += g_hourly[group_ids[k]]
    return loss, n_valid, g_runoff
