"""Benchmark: reach-timesteps/s of the differentiable Muskingum-Cunge window,
forward + reverse-time adjoint, on one MI355X.

Workload (default, --config c2): a 350,000-reach synthetic river tree (each
reach drains to one drawn from the next 256, about 2.7k levels deep), 720
hourly steps, fp32 lateral inflow LogNormal(ln 0.5, 1) with storm pulses,
per-reach Manning n and q_spatial as leaves. One bench step = the batched
persistent forward window, the loss gradient (MSE against a fixed target at 64
gauge reaches) and the persistent adjoint window, producing dL/dn and
dL/dq_spatial for every reach. The window calls read out the 64 gauges in the
kernels (the forward writes Q only there, the adjoint takes dL/dQ only there);
results are bit-identical to the dense calls. Inputs are resident in HBM
before timing.

Also reported: the same step in the plain per-timestep-launch form, the
roofline of the dominant kernel (the persistent forward window) from HIP events
on the stream it runs on, gradient checks, and the CPU oracle timed on a bounded
sample of the same network.

"kan_head" reports the KAN parameter head (the producer of n and q_spatial in
the batched and training workloads): 10 attributes ~ N(0, 1) -> 21 hidden ->
two KAN blocks (grid 5, order 3) -> 2 sigmoid outputs, forward + backward over
350,000 reaches, against the fp32 peak and beside its CPU oracle.
--kan-only runs just that part.

"gauge_loss" reports the fused micro-batch objective of the training workload
(gauge readout, tau trim and daily pooling, warmup slice, NaN-gauge filter,
batch NSE, and its gradient into the routed discharge) for 2,000 gauges over a
time-major 8,760-hour x 350,000-reach discharge. --loss-only runs just that.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from ddrs_amd import workloads as W  # noqa: E402

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
# KAN head of the batched and training workloads, evaluated over 350,000 reaches
KAN_SHAPE = dict(F=10, H=21, P=2, L=2, grid=5, k=3, N=350_000)
# gauge readout + loss of the training workload
LOSS_SHAPE = dict(R=350_000, T=8760, G=2000, tau=9, warmup=30)
# Algorithmic bytes per reach-timestep, the per-step working set with no reuse
# across steps: forward 40 (q' 4, Q_t 4, the five statics n, q_spatial,
# p_spatial, length, slope 20, the CSR row pointer and one column index 8,
# Q_{t+1} written 4); adjoint 72 (lambda_{t+1} 4, the solve x_{t+1} and Q_t of
# the reach 4 + 4 and of its upstream 4 + 4, q' 4, statics 20, the downstream
# index 4, gb of the downstream reach 4, lambda_t written 4, read-modify-write of
# the n and q_spatial gradients 16). The roofline's achieved bandwidth is the
# forward's 40 B times its reach-steps over its launch time.
ALGO_BYTES = {"fwd": 40, "adjoint": 72}
# Compulsory HBM bytes per reach-timestep of each window kernel of this
# implementation (statics stay in registers), with the bench's gauge readout
# (Q written and dL/dQ read only at the gauges):
#   forward window: q' read 4, x history written 8 and read back 8 by the
#     downstream chunks (one link per reach in a tree), the upstream sums
#     (float2) written 8 for the linearisation; a dense Q output would add 4
#     written;
#   linearise: the reach's x history 8 and its upstream sums 8, q' 4;
#     records written 16 + 16;
#   sweep: records read 16, (P, G) published 16 and read 16 by the upstream
#     chunk, gb written 4; a dense dL/dQ would add 4 read;
#   finish: three record planes read 12, gb read 4 (this bench's loss asks
#     for no dL/dq', which would add the fourth plane, 4 read, and 4 written).
KERNEL_BYTES = {"fwd": 28, "lin": 52, "sweep": 52, "fin": 16}
KERNEL_NAMES = {"fwd": "k_fwd_window (persistent forward)", "lin": "k_adj_lin (linearise)",
                "sweep": "k_adj_sweep (persistent reverse sweep)", "fin": "k_adj_fin"}
def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", default="c2", choices=["c1", "c2", "c3", "c4", "c5"])
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-parity", action="store_true",
                    help="skip the oracle check of the timed step's results (c1, c2: the whole "
                         "network, about 90 s of CPU work at c2, and FP64 differences on 64 "
                         "reaches)")
    ap.add_argument("--cpu-steps", type=int, default=240,
                    help="timesteps of the single-thread oracle sample (about 12 s on C2)")
    ap.add_argument("--sublevels", type=int, default=0,
                    help="largest chain of links in one schedule block (0: library default)")
    ap.add_argument("--block", type=int, default=0, help="schedule block, 64 or 128 (0: default)")
    ap.add_argument("--graph", action="store_true",
                    help="also time the step captured in a HIP graph and replayed (and the plain "
                         "form as one graph unless --no-plain)")
    ap.add_argument("--tune", action="store_true",
                    help="choose the schedule's sub-level cap and the checkpoint window with the "
                         "autotuner (always on for c5 unless --sublevels is given)")
    ap.add_argument("--kan-only", action="store_true", help="only the KAN head measurement")
    ap.add_argument("--no-kan", action="store_true", help="skip the KAN head measurement")
    ap.add_argument("--loss-only", action="store_true", help="only the gauge loss measurement")
    ap.add_argument("--no-loss", action="store_true", help="skip the gauge loss measurement")
    ap.add_argument("--pmc-calibration", action="store_true",
                    help="first copy 1 GiB device to device once, as the known-size reference "
                         "for FETCH_SIZE / WRITE_SIZE counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(basins would shard across GPUs with no exchange during routing)"}))
        sys.exit(2)

    import torch
    from ddrs_amd import _native as NAT

    if args.kan_only:
        kan = kan_bench(torch, args.steps, args.warmup, not args.no_cpu)
        print(json.dumps({"metric": "reaches/s KAN head forward+backward", **kan}), flush=True)
        return
    if args.loss_only:
        gl = loss_bench(torch, args.steps, args.warmup, not args.no_cpu)
        print(json.dumps({"metric": "gauge-hours/s fused readout + loss + gradient", **gl}),
              flush=True)
        return

    if args.config == "c5":
        print(json.dumps(train_bench(args, torch)), flush=True)
        return

    if args.pmc_calibration:
        a = torch.empty(2**28, dtype=torch.float32, device="cuda:0")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    wl = W.build(args.config)
    R, T = wl.R, wl.T
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    log(f"[bench] building network R={R} T={T}")
    t0 = time.time()
    qprime = wl.forcing_device(dev)
    net, window, schedule = schedule_for(args, wl.adjacency(), T, qprime, wl.gauges, wl.ckpt,
                                         tune=args.tune)
    log(f"[bench] plan {time.time() - t0:.2f}s: {net.n_chunks} chunks, {net.n_levels} levels; "
        f"schedule {schedule}")
    gw = W.GaugedWindow(net, wl, qprime, window=window)
    ws = gw.ws
    log(f"[bench] workspace {ws.nbytes() / 2**30:.2f} GiB")
    # fixed observation target: the routed discharge at the gauges scaled by 0.9
    obs = (0.9 * gw.forward()).clone()
    gw.check()

    def step(form, stream=None):
        gw.step(obs, form, stream)

    def graph_of(form):
        """The step captured in a HIP graph (torch.cuda.graph's capture stream)."""
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            step(form, torch.cuda.current_stream().cuda_stream)
        return g

    for _ in range(args.warmup):
        step(NAT.FORM_PERSISTENT)
    gw.check()
    # HIP events around every kernel launch on the launch stream, summed over
    # the timed steps and read once after them (no host synchronisation inside)
    NAT.check(NAT.lib.ddrs_set_timing(ws.handle, 2))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        step(NAT.FORM_PERSISTENT)
        log(f"[bench] step {k + 1}/{args.steps} enqueued")
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    tm = [np.zeros(1, np.int32), np.zeros(1, np.float32), np.zeros(1, np.int32),
          np.zeros(1, np.float32)]
    NAT.check(NAT.lib.ddrs_get_timing(ws.handle, *[x.ctypes.data for x in tm]))
    ph = [np.zeros(1, np.float32) for _ in range(3)]
    NAT.check(NAT.lib.ddrs_get_adj_phase_timing(ws.handle, *[x.ctypes.data for x in ph]))
    fwd_ms = float(tm[1][0]) / args.steps
    phase_ms = np.array([float(x[0]) for x in ph]) / args.steps
    NAT.check(NAT.lib.ddrs_set_timing(ws.handle, 0))
    free_b, total_b = torch.cuda.mem_get_info()
    hbm = {"device_used_GiB": (total_b - free_b) / 2**30, "device_total_GiB": total_b / 2**30,
           "workspace_GiB": ws.nbytes() / 2**30,
           "note": "device memory in use after the timed steps (workspace, q', outputs, "
                   "runtime)"}
    log(f"[bench] HBM in use {hbm['device_used_GiB']:.1f} of {hbm['device_total_GiB']:.1f} GiB")
    gw.check()
    rts = R * (T - 1)
    value = rts * args.steps / elapsed
    ms_per_step = 1000.0 * elapsed / args.steps
    layout = ws.layout()
    log(f"[bench] persistent: {ms_per_step:.2f} ms/step, {value / 1e9:.3f} G reach-ts/s; "
        f"fwd kernel {fwd_ms:.2f} ms, adjoint {float(tm[3][0]) / args.steps:.2f} ms "
        f"(lin {phase_ms[0]:.2f}, sweep {phase_ms[1]:.2f}, fin {phase_ms[2]:.2f}); {layout}")
    gn_pers, gqsp_pers, Q_pers = gw.gn.clone(), gw.gqsp.clone(), gw.Q.clone()

    def same_as():
        return bool(torch.equal(gw.gn, gn_pers) and torch.equal(gw.gqsp, gqsp_pers)
                    and torch.equal(gw.Q, Q_pers))

    graph = None
    if args.graph:
        # the same step captured once and replayed: its replays must give the
        # eager step's bits
        gp = graph_of(NAT.FORM_PERSISTENT)
        gp.replay()
        torch.cuda.synchronize()
        gw.check()
        tg0 = time.perf_counter()
        for _ in range(args.steps):
            gp.replay()
        torch.cuda.synchronize()
        tg = (time.perf_counter() - tg0) / args.steps
        gw.check()
        graph = dict(value=rts / tg, ms_per_step=1000 * tg, bit_identical_to_eager=same_as())
        del gp
        log(f"[bench] graph replay: {1000 * tg:.2f} ms/step, "
            f"identical={graph['bit_identical_to_eager']}")
        if not args.no_plain:
            gpl = graph_of(NAT.FORM_PLAIN)
            torch.cuda.synchronize()
            tq0 = time.perf_counter()
            gpl.replay()
            torch.cuda.synchronize()
            tq = time.perf_counter() - tq0
            gw.check()
            graph["plain"] = dict(value=rts / tq, ms_per_step=1000 * tq,
                                  bit_identical_to_eager=same_as())
            del gpl
            log(f"[bench] plain form as one graph: {1000 * tq:.1f} ms/step")

    plain = None
    if not args.no_plain:
        NAT.check(NAT.lib.ddrs_set_timing(ws.handle, 1))
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        step(NAT.FORM_PLAIN)
        torch.cuda.synchronize()
        tp = time.perf_counter() - t1
        gw.check()
        a = [np.zeros(1, np.int32), np.zeros(1, np.float32), np.zeros(1, np.int32),
             np.zeros(1, np.float32)]
        NAT.check(NAT.lib.ddrs_get_timing(ws.handle, *[x.ctypes.data for x in a]))
        NAT.check(NAT.lib.ddrs_set_timing(ws.handle, 0))
        plain = dict(value=rts / tp, ms_per_step=1000 * tp, fwd_launches=int(a[0][0]),
                     fwd_kernel_ms=float(a[1][0]), adj_launches=int(a[2][0]),
                     adj_kernel_ms=float(a[3][0]), bit_identical_to_persistent=same_as())
        log(f"[bench] plain: {1000 * tp:.1f} ms/step ({value * tp / rts:.1f}x slower), "
            f"identical={plain['bit_identical_to_persistent']}")

    # per-launch times of the persistent form (one launch of each kernel per step)
    launch_ms = {"fwd": fwd_ms, "lin": phase_ms[0], "sweep": phase_ms[1], "fin": phase_ms[2]}
    traffic, traffic_note = stored_traffic(args.config)

    # the parity leg: the timed step's own results against the oracle
    results = {"Q": gw.Q.cpu().numpy(), "gQ": gw.gQ.cpu().numpy(), "q0": gw.q0.cpu().numpy(),
               "n": gw.gn.cpu().numpy(), "q_spatial": gw.gqsp.cpu().numpy()}
    dense = not args.no_parity and wl.sizes is None
    Qdense = gw.dense_forward().cpu().numpy() if dense else None
    qp_host = qprime.cpu().numpy() if dense else None
    gw.check()
    gw.close()
    del gw, ws, qprime, obs, gn_pers, gqsp_pers, Q_pers
    torch.cuda.empty_cache()
    copy = copy_peak(torch)
    log(f"[bench] device-to-device copy: {copy['GBs']:.0f} GB/s")

    cpu = None
    if not args.no_cpu:
        # bounded sample: about 80 M reach-steps of single-thread oracle work
        cpu_steps = args.cpu_steps if R * args.cpu_steps <= 1e8 else max(1, int(8e7 // R))
        cpu = cpu_baseline(wl, cpu_steps)

    parity = grad_check = None
    if dense:
        parity, grad_check = parity_leg(wl, qp_host, Qdense, results)
        del Qdense, qp_host
    elif not args.no_parity:
        parity = {"note": "forests are checked basin by basin at full size in "
                          "tests/test_gpu_scale.py"}

    out = {
        "metric": "reach-timesteps/s fwd+adjoint (batched vs plain, 1 GPU); grad rel err vs FD",
        "value": value,
        "unit": "reach-timesteps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded generator; no network access)",
        "config": {"workload": workload_name(args.config, wl, window), "reaches": R,
                   "timesteps": T, "levels": net.n_levels, "schedule": schedule,
                   "form": "persistent", "parallelism": "single GPU"},
        "roofline": roofline_of(launch_ms, rts, traffic, layout=layout, copy_peak=copy,
                                traffic_note=traffic_note,
                                note="latency- and VALU-bound persistent pipeline"),
        "hbm": hbm,
        "plain": plain,
        "batched_over_plain": (value / plain["value"]) if plain else None,
        "graph": graph,
        "grad_check": grad_check,
        "parity": parity,
        "cpu_baseline": cpu,
    }
    if not args.no_kan:
        out["kan_head"] = kan_bench(torch, args.steps, args.warmup, not args.no_cpu)
    if not args.no_loss:
        out["gauge_loss"] = loss_bench(torch, args.steps, args.warmup, not args.no_cpu)
    print(json.dumps(out), flush=True)


def workload_name(config, wl, window):
    cfg = W.CONFIGS[config]
    T = wl.T
    if wl.sizes is not None:
        sizes = (f"Zipf({cfg['zipf']}) capped at {cfg['size_hi']} and scaled to {cfg['total']}"
                 if "zipf" in cfg else f"log-uniform {cfg['size_lo']}-{cfg['size_hi']}")
        return (f"{config}: {cfg['basins']} basins, {wl.R} reaches (sizes {sizes}, generator "
                f"window {cfg['window']}), {T} hourly steps, checkpoint window "
                f"{window or (T - 1)}, forward + adjoint, MSE at {len(wl.gauges)} basin outlets")
    return (f"{config}: {wl.R}-reach tree (window {cfg['window']}), {T} hourly steps, "
            f"forward + adjoint, MSE at {len(wl.gauges)} gauges")


def copy_peak(torch, nbytes=4 << 30, reps=10):
    """Device-to-device copy bandwidth (bytes read + written per second) of a
    plain torch copy of nbytes, HIP events on the current stream: the
    attainable HBM rate beside the 8 TB/s published peak."""
    a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    a.fill_(1.0)
    b.copy_(a)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    ev[0].record()
    for _ in range(reps):
        b.copy_(a)
    ev[1].record()
    torch.cuda.synchronize()
    ms = ev[0].elapsed_time(ev[1]) / reps
    del a, b
    torch.cuda.empty_cache()
    gbs = 2 * nbytes / (ms * 1e-3) / 1e9
    return {"GBs": gbs, "bytes_copied": nbytes, "ms": ms, "frac_of_peak": gbs / HBM_PEAK_GBS}


def kernel_source_hash() -> str:
    """sha256 (16 hex digits) of the window kernels' sources."""
    import hashlib
    h = hashlib.sha256()
    for f in ("ddrs_kernels.hip", "mc_physics.h", "mc_math.h", "mc_ln_table.h", "ddrs_internal.h"):
        with open(os.path.join(ROOT, "ddrs_amd", "csrc", f), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()[:16]


def stored_traffic(config):
    """The forward window's HBM bytes per launch from the committed rocprofv3
    counter profile (profiles/traffic_fwd_window_<config>.json), only when it
    was measured on the current kernel sources: (bytes or None, note)."""
    path = os.path.join(ROOT, "profiles", f"traffic_fwd_window_{config}.json")
    if not os.path.exists(path):
        return None, "no counter profile for this workload"
    with open(path) as fh:
        prof = json.load(fh)
    if prof.get("kernels_sha16") != kernel_source_hash():
        return None, (f"the counter profile ({prof.get('source')}) was taken on other kernel "
                      f"sources; re-measure")
    return prof["bytes_per_launch"], (f"rocprofv3 FETCH_SIZE + WRITE_SIZE per launch from "
                                      f"{prof.get('source')} (kernel sources "
                                      f"{prof['kernels_sha16']}), not measured in this run")


def parity_leg(wl, qp, Qdense, res):
    """The timed step's discharge (a dense forward at every reach-step, and the
    gauge readout) and gradients against the oracle's three variants on the
    whole network, and FP64 central differences on 64 seeded reaches
    (oracle/scale_check.py). Returns (parity, grad_check)."""
    from concurrent.futures import ThreadPoolExecutor
    from oracle import oracle as O
    from oracle import scale_check as SC
    t0 = time.time()
    R, T = wl.R, wl.T
    P = O.make_params()
    pat = O.pattern_from_coo(R, wl.rows, wl.cols)
    at = wl.attrs
    p = np.full(R, 21.0, np.float32)
    gQd = np.zeros((T, R), np.float32)
    gQd[:, wl.gauges] = res["gQ"]
    with ThreadPoolExecutor(SC.threads()) as pool:
        r = SC.route_pair(P, pat, at["n"], at["q_spatial"], p, at["length"], at["slope"], qp,
                          gQd, res["q0"], want_x64=True, pool=pool)
        log(f"[bench] oracle (f32, f64, device-math f32) done in {time.time() - t0:.0f}s")
        par = {"discharge_all_reach_steps": SC.discharge_errors(Qdense, r),
               "discharge_gauges": SC.discharge_errors(res["Q"], r, wl.gauges)}
        for k in SC.GRAD_KEYS:
            par[f"grad_{k}"] = SC.grad_errors(res[k], r, k)
        reaches = np.sort(np.random.default_rng(4102).choice(R, 64, replace=False))
        fd = SC.fd_check(P, pat, {"n": at["n"], "q_spatial": at["q_spatial"], "p_spatial": p},
                         at["length"], at["slope"], qp, r["f64"]["Q"], r["f64"]["X"], res["gQ"],
                         wl.gauges, reaches, pool=pool)
    fe = SC.fd_errors(fd, reaches, res, r)
    for k in fe:
        for key in ("resolved_mask", "gpu_rel_err_per_reach", "f32_oracle_rel_err_per_reach",
                    "fp64_adjoint_rel_err_per_reach"):
            fe[k].pop(key)
    par["seconds"] = time.time() - t0
    par["note"] = ("oracle/scale_check.py: f32 = the reference's arithmetic with glibc powf/"
                   "logf; device-math = the same with correctly rounded powers; the discharge "
                   "spread between the two is libm's rounding amplified by the recurrence")
    grad = {"max_rel_err_vs_fp64_fd": max(fe[k]["max_rel_err_gpu"] for k in fe),
            "reaches": 64, "params": ["n", "q_spatial"],
            "network": f"the benchmark network itself ({R} reaches x {T} steps); the loss is "
                       f"the timed step's gauge MSE linearised at its dL/dQ",
            "per_param": fe}
    log(f"[bench] parity: discharge bit-identical to the device-math oracle at "
        f"{par['discharge_all_reach_steps']['bit_identical_to_device_math_oracle']:.6f} of "
        f"reach-steps, FD max rel err {grad['max_rel_err_vs_fp64_fd']:.2e} "
        f"({par['seconds']:.0f}s)")
    return par, grad


def kan_flops(F, H, P, L, k):
    """Algorithmic fp32 flops per reach (a fused multiply-add counts 2) of the
    KAN head: (forward, forward + backward). The backward kernels recompute the
    forward; that recomputation is not counted."""
    basis = 3 * k * (k + 1)            # k(k+1)/2 Cox-de Boor terms of about 6 flops
    edge = 2 * (k + 1) + 4             # spline dot, scale_base*silu + scale_sp*spline, sum
    fwd = 2 * F * H + L * (H * (basis + 4) + H * H * edge) + 2 * H * P + 4 * P
    # reverse: sigmoid and output layer, per KAN layer the basis derivatives and
    # dy/dx (spline derivative dot and the two scale terms), the input layer
    rev = 2 * H * P + 4 * P + L * (H * (2 * basis + 6) + H * H * (2 * (k + 1) + 6)) + 2 * H * F
    # parameter sums over reaches: per layer and input, the basis again and
    # (k+1 basis values + silu) x H output gradients; the two linear layers
    sums = L * H * (basis + 4 + 2 * (k + 2) * H) + 2 * (H + 1) * P + 2 * (F + 1) * H
    return fwd, fwd + rev + sums


def kan_bench(torch, steps, warmup, with_cpu):
    """Forward and backward of the KAN head through the C ABI, HIP events on
    the stream they run on; the CPU oracle on a bounded sample."""
    import ctypes
    from ddrs_amd import _native as NAT
    from ddrs_amd import kan as K
    s = KAN_SHAPE
    F, H, P, L, k, N = s["F"], s["H"], s["P"], s["L"], s["k"], s["N"]
    dev = torch.device("cuda:0")
    head = K.KanHeadConfig([f"attr{i}" for i in range(F)], ["n", "q_spatial"], 42, H, L,
                           s["grid"], k).init(dev)
    gen = torch.Generator(device=dev).manual_seed(7)
    x = torch.randn(N, F, generator=gen, device=dev)
    gout = torch.randn(P, N, generator=gen, device=dev)
    out = torch.empty(P, N, device=dev)
    gx = torch.empty_like(x)
    grads = {n: torch.empty_like(getattr(head, n)) for n in K._TRAINED}
    fb, bb = NAT.c_size(), NAT.c_size()
    NAT.check(NAT.lib.ddrs_kan_workspace_bytes(ctypes.byref(head.dims), N, ctypes.byref(fb),
                                               ctypes.byref(bb)))
    ws = torch.empty(bb.value, dtype=torch.uint8, device=dev)
    prm = K._params(head.dims, head.grid, head.mask,
                    {n: getattr(head, n).detach() for n in K._TRAINED})
    gr = K.KanGrads(*[grads[n].data_ptr() for n in K._TRAINED])
    stream = torch.cuda.current_stream()
    sp = stream.cuda_stream

    def fwd():
        NAT.check(NAT.lib.ddrs_kan_head_fwd(ctypes.byref(head.dims), ctypes.byref(prm), N,
                                            x.data_ptr(), out.data_ptr(), ws.data_ptr(),
                                            ws.numel(), sp), "ddrs_kan_head_fwd")

    def bwd():
        NAT.check(NAT.lib.ddrs_kan_head_bwd(ctypes.byref(head.dims), ctypes.byref(prm), N,
                                            x.data_ptr(), gout.data_ptr(), ctypes.byref(gr),
                                            gx.data_ptr(), ws.data_ptr(), ws.numel(), sp),
                  "ddrs_kan_head_bwd")

    for _ in range(max(warmup, 1)):
        fwd()
        bwd()
    torch.cuda.synchronize()
    reps = max(steps, 1) * 4
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3 * reps)]
    for r in range(reps):
        ev[3 * r].record(stream)
        fwd()
        ev[3 * r + 1].record(stream)
        bwd()
        ev[3 * r + 2].record(stream)
    torch.cuda.synchronize()
    fwd_ms = sum(ev[3 * r].elapsed_time(ev[3 * r + 1]) for r in range(reps)) / reps
    bwd_ms = sum(ev[3 * r + 1].elapsed_time(ev[3 * r + 2]) for r in range(reps)) / reps
    ff, fb_flops = kan_flops(F, H, P, L, k)
    tflops = fb_flops * N / ((fwd_ms + bwd_ms) * 1e-3) / 1e12
    log(f"[bench] kan head N={N}: fwd {fwd_ms:.3f} ms, bwd {bwd_ms:.3f} ms, {tflops:.1f} TFLOP/s")
    res = {
        "value": N / ((fwd_ms + bwd_ms) * 1e-3), "unit": "reaches/s",
        "fwd_ms": fwd_ms, "bwd_ms": bwd_ms, "calls_timed": reps, "dtype": "f32",
        "config": {"workload": f"KAN head {F} attributes -> {H} -> {L} KAN blocks (grid "
                               f"{s['grid']}, order {k}) -> {P} sigmoid outputs, forward + "
                               f"backward over {N} reaches", **s},
        "roofline": {"bound": "valu", "achieved": tflops, "peak": FP32_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": tflops / FP32_PEAK_TFLOPS, "traffic": None,
                     "flops_per_reach_fwd": ff, "flops_per_reach_fwd_bwd": fb_flops,
                     "fwd_TFLOPs": ff * N / (fwd_ms * 1e-3) / 1e12,
                     "note": "fp32 vector arithmetic, no packed math; the peak counts "
                             "packed/dual-issue fp32"},
    }
    if with_cpu:
        res["cpu_baseline"] = kan_cpu_baseline(F, H, P, L, s["grid"], k)
    return res


def kan_cpu_baseline(F, H, P, L, grid, k, n_sample=40_000):
    """The numpy oracle (f32, one BLAS thread) on a sample of reaches."""
    from threadpoolctl import threadpool_limits
    from oracle import kan_oracle as ko
    rng = np.random.default_rng(11)
    G = grid + 2 * k + 1
    knots = (np.arange(-k, grid + k + 1) * (2.0 / grid) - 1.0).astype(np.float32)
    p = dict(input_weight=rng.normal(0, 0.45, (H, F)).astype(np.float32),
             input_bias=np.zeros(H, np.float32),
             output_weight=rng.normal(0, 0.03, (P, H)).astype(np.float32),
             output_bias=np.zeros(P, np.float32),
             blocks=[dict(grid=np.tile(knots, (H, 1)),
                          coef=rng.normal(0, 0.01, (H, H, G - 1 - k)).astype(np.float32),
                          scale_base=rng.uniform(-0.2, 0.2, (H, H)).astype(np.float32),
                          scale_sp=np.full((H, H), 1 / np.sqrt(H), np.float32),
                          mask=np.ones((H, H), np.float32)) for _ in range(L)])
    x = rng.normal(size=(n_sample, F)).astype(np.float32)
    g = rng.normal(size=(n_sample, P)).astype(np.float32)
    with threadpool_limits(1):
        t0 = time.perf_counter()
        out, cache = ko.kan_head_fwd(x, p, k, np.float32)
        ko.kan_head_bwd(g, p, cache, np.float32)
        dt = time.perf_counter() - t0
    return {"value": n_sample / dt, "unit": "reaches/s", "cores": 1, "kind": "port",
            "sample": f"oracle/kan_oracle.py forward + backward in f32 over {n_sample} "
                      f"reaches of the same head shape, one thread ({dt:.1f} s)"}


def loss_bench(torch, steps, warmup, with_cpu):
    """The fused objective (ddrs_gauge_nse_loss) through the C ABI, HIP events
    on its stream; the numpy oracle on a bounded sample of gauges."""
    from ddrs_amd import _native as NAT
    from ddrs_amd import loss as LS
    s = LOSS_SHAPE
    R, T, G, tau, wu = s["R"], s["T"], s["G"], s["tau"], s["warmup"]
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(21)
    # one outflow reach per gauge, every tenth gauge with a second one
    first = rng.choice(R, G, replace=False)
    extra = rng.choice(R, G // 10, replace=False)
    flat = np.concatenate([first, extra]).astype(np.int32)
    group = np.concatenate([np.arange(G), np.arange(0, G, 10)]).astype(np.int32)
    plan = LS.GaugePlan(R, flat, group, G, T)
    gen = torch.Generator(device=dev).manual_seed(5)
    Q = torch.rand(T, R, generator=gen, device=dev) * 10 + 0.1      # time-major [T][R]
    gQ = torch.zeros_like(Q)
    days = (T - 24) // 24
    obs = torch.rand(G, days, generator=gen, device=dev) * 10 + 0.1
    obs[::7, days // 2] = float("nan")                                 # every 7th gauge dropped
    sigma = torch.rand(G, generator=gen, device=dev) + 0.5
    loss = torch.empty((), device=dev)
    nv = torch.empty((), dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream()
    sp = stream.cuda_stream

    def call():
        NAT.check(NAT.lib.ddrs_gauge_nse_loss(plan.handle, Q.data_ptr(), T, 1, R, tau, wu,
                                              obs.data_ptr(), sigma.data_ptr(), 0.1,
                                              loss.data_ptr(), nv.data_ptr(), gQ.data_ptr(),
                                              None, sp), "ddrs_gauge_nse_loss")

    for _ in range(max(warmup, 1)):
        call()
    torch.cuda.synchronize()
    reps = max(steps, 1) * 4
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2 * reps)]
    for r in range(reps):
        ev[2 * r].record(stream)
        call()
        ev[2 * r + 1].record(stream)
    torch.cuda.synchronize()
    ms = sum(ev[2 * r].elapsed_time(ev[2 * r + 1]) for r in range(reps)) / reps
    K = flat.size
    # algorithmic HBM bytes: discharge of the map's reaches read and their gradient
    # written (K x T x 4 each), the hourly readout written and read back and its
    # gradient the same (G x T x 4 x 4), observations read (G x days x 4)
    alg = 2 * K * T * 4 + 4 * G * T * 4 + G * days * 4
    gbs = alg / (ms * 1e-3) / 1e9
    log(f"[bench] gauge loss G={G} T={T}: {ms:.3f} ms, {gbs:.0f} GB/s algorithmic")
    res = {
        "value": G * T / (ms * 1e-3), "unit": "gauge-hours/s", "ms_per_call": ms,
        "calls_timed": reps, "loss": float(loss), "n_valid": int(nv), "dtype": "f32",
        "config": {"workload": f"fused readout + tau trim/daily pooling + warmup + NaN filter "
                               f"+ batch NSE + gradient: {G} gauges ({K} outflow entries) over "
                               f"a time-major {T} x {R} discharge", **s},
        "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": gbs / HBM_PEAK_GBS, "traffic": None, "bytes_per_call": alg,
                     "note": "4-byte reads and writes at scattered reaches of a time-major "
                             "series: each touches its own cache line"},
    }
    if with_cpu:
        from oracle import loss_oracle as lo
        gs = 200
        sel = np.isin(group, np.arange(gs))
        Qs = Q[:, torch.as_tensor(np.unique(flat[sel]).astype(np.int64), device=dev)]
        remap = {r: i for i, r in enumerate(np.unique(flat[sel]))}
        f2 = [remap[r] for r in flat[sel]]
        runoff = Qs.T.contiguous().cpu().numpy()
        o_np, s_np = obs[:gs].cpu().numpy(), sigma[:gs].cpu().numpy()
        t0 = time.perf_counter()
        lo.gauge_nse_loss(runoff, f2, group[sel], gs, tau, wu, o_np, s_np, 0.1, np.float32)
        dt = time.perf_counter() - t0
        res["cpu_baseline"] = {"value": gs * T / dt, "unit": "gauge-hours/s", "cores": 1,
                               "kind": "port",
                               "sample": f"oracle/loss_oracle.py gauge_nse_loss on the first "
                                         f"{gs} gauges of the same workload ({dt:.2f} s)"}
    del Q, gQ
    torch.cuda.empty_cache()
    return res


def roofline_of(launch_ms, rs, traffic=None, **extra):
    """The roofline object of the forward window (the dominant kernel) at the
    algorithmic 40 B per reach-step, with the adjoint's three kernels at 72 B
    together and every kernel at this implementation's compulsory bytes."""
    kernels = {k: {"launch_ms": v, "compulsory_bytes_per_reach_step": KERNEL_BYTES[k],
                   "compulsory_GBs": KERNEL_BYTES[k] * rs / (v * 1e-3) / 1e9}
               for k, v in launch_ms.items()}
    fwd = launch_ms["fwd"]
    adj_ms = launch_ms["lin"] + launch_ms["sweep"] + launch_ms["fin"]
    roof = ALGO_BYTES["fwd"] * rs / (fwd * 1e-3) / 1e9
    adj = ALGO_BYTES["adjoint"] * rs / (adj_ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": roof, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": roof / HBM_PEAK_GBS, "traffic": traffic, "kernel": KERNEL_NAMES["fwd"],
            "launch_ms": fwd, "bytes_per_reach_step": ALGO_BYTES["fwd"],
            "reach_steps_per_launch": rs,
            "adjoint": {"launch_ms": adj_ms, "bytes_per_reach_step": ALGO_BYTES["adjoint"],
                        "achieved_GBs": adj, "frac": adj / HBM_PEAK_GBS},
            "kernels": kernels, **extra}


def schedule_for(args, adj, T, qprime, gauges, ckpt, tune=True):
    """(network, checkpoint window, description): the autotuner's choice, or
    the fixed options (--sublevels / --block, the config's checkpoint window)."""
    from ddrs_amd import tune as TU
    from ddrs_amd.sparse import Network
    if args.sublevels or args.block or not tune:
        net = Network(adj, args.sublevels, args.block)
        return net, ckpt, {"tuned": False, "sub_levels": args.sublevels or 8,
                           "block": args.block or 64, "window": ckpt or (T - 1)}
    t0 = time.time()
    ts = TU.tune_schedule(adj, T, q_prime=qprime, gauges=gauges)
    desc = {"tuned": True, "sub_levels": ts.max_sub_levels, "block": ts.block,
            "window": ts.window, "predicted_ms": ts.predicted_ms,
            "tuning_s": time.time() - t0,
            "probes": [{k: v for k, v in pr.items() if k != "probe_lengths"} for pr in ts.probes],
            "probe_lengths": ts.probes[0]["probe_lengths"]}
    window = ts.window if ts.window < T - 1 else None
    return ts.network, window, desc


def train_bench(args, torch):
    """The training workload: one optimizer step of ddrs_amd.train.Trainer per
    bench step on a fixed micro-batch (HIP KAN head, gauged routing window,
    fused daily NSE, adjoint, KAN backward, clipping and AdaDelta on the
    device). Reach-timesteps/s of the whole step; the routing kernels' share
    from HIP events; the plain per-timestep-launch form of the same step once,
    with its gradients checked bitwise against the persistent form's."""
    from ddrs_amd import _native as NAT
    from ddrs_amd import train as TR

    cfg = W.CONFIGS["c5"]
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    t0 = time.time()
    wl = W.build_c5()
    T, R, R0 = wl.T, wl.R, wl.extra["R0"]
    basin = wl.extra["basin_sizes"]
    cg = wl.gauges
    qprime = wl.forcing_device(dev)
    net, window, schedule = schedule_for(args, wl.adjacency(), T, qprime, cg, cfg["ckpt"])
    log(f"[bench] c5 batch: {R} of {R0} reaches, {cg.size} gauges (basins {basin.min()}-"
        f"{basin.max()}), {net.n_levels} levels; setup {time.time() - t0:.1f}s; "
        f"schedule {schedule}")
    tau, wu, eps = cfg["tau"], cfg["warmup"], cfg["eps"]
    # observations: a hidden head's routed discharge at the gauges, daily
    # pooled, times LogNormal(0, 0.1) noise; sigma is each gauge's std
    batch = W.c5_batch(wl, qprime, net, window, dev)
    head = W.kan_head(42, dev)
    tr = TR.Trainer(head, tau=tau, warmup=wu, eps=eps, window=window, network=net)
    log(f"[bench] c5 observations ready ({time.time() - t0:.1f}s)")
    first = tr.step(batch)
    for _ in range(max(args.warmup - 1, 0)):
        tr.step(batch)
    ws = next(v for k, v in net._ws.items() if k[0] == "gauged")
    log(f"[bench] c5 workspace {ws.nbytes() / 2**30:.1f} GiB, first loss {first.loss:.5f}")
    # launch times summed over the timed steps, read once after them
    NAT.check(NAT.lib.ddrs_set_timing(ws.handle, 2))
    losses = []
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for k in range(args.steps):
        out = tr.step(batch)   # the step reads its loss back, as the training loop does
        losses.append(out.loss)
        log(f"[bench] c5 step {k + 1}/{args.steps}: loss {out.loss:.5f}, grad norm "
            f"{out.grad_norm:.4g}")
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t1
    a = [np.zeros(1, np.int32), np.zeros(1, np.float32), np.zeros(1, np.int32),
         np.zeros(1, np.float32)]
    NAT.check(NAT.lib.ddrs_get_timing(ws.handle, *[v.ctypes.data for v in a]))
    ph = [np.zeros(1, np.float32) for _ in range(3)]
    NAT.check(NAT.lib.ddrs_get_adj_phase_timing(ws.handle, *[v.ctypes.data for v in ph]))
    fwd_ms = float(a[1][0])
    phase_ms = np.array([float(v[0]) for v in ph])
    NAT.check(NAT.lib.ddrs_set_timing(ws.handle, 0))
    free_b, total_b = torch.cuda.mem_get_info()
    rs = R * (T - 1)
    value = rs * args.steps / elapsed
    ms_per_step = 1000.0 * elapsed / args.steps
    phase_ms /= args.steps
    launch_ms = {"fwd": fwd_ms / args.steps, "lin": phase_ms[0], "sweep": phase_ms[1],
                 "fin": phase_ms[2]}
    routing_ms = sum(launch_ms.values())
    log(f"[bench] c5 persistent: {ms_per_step:.1f} ms/step, {value / 1e9:.3f} G reach-ts/s; "
        f"routing kernels {routing_ms:.1f} ms (fwd {launch_ms['fwd']:.1f}, lin "
        f"{phase_ms[0]:.1f}, sweep {phase_ms[1]:.1f}, fin {phase_ms[2]:.1f})")

    def grads_of(form, step=False):
        tr.form = form
        for prm in head.parameters():
            prm.grad = None
        loss, _ = tr.loss(batch)
        loss.backward()
        g = [prm.grad.clone() for prm in head.parameters()]
        if step:
            TR.clip_grad_norm(head.parameters(), tr.grad_clip)
            tr.opt.step()
        return float(loss.detach()), g

    plain = None
    if not args.no_plain:
        lp, gp = grads_of(NAT.FORM_PERSISTENT)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        lq, gq = grads_of(NAT.FORM_PLAIN, step=True)
        torch.cuda.synchronize()
        tp = time.perf_counter() - t2
        same = lp == lq and all(torch.equal(u, v) for u, v in zip(gp, gq))
        plain = dict(value=rs / tp, ms_per_step=1000 * tp, bit_identical_to_persistent=same)
        log(f"[bench] c5 plain: {1000 * tp:.0f} ms/step ({value * tp / rs:.1f}x slower), "
            f"identical={same}")
        tr.form = NAT.FORM_PERSISTENT
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(wl, max(1, int(8e7 // R)))
        cpu["sample"] += " (routing forward + adjoint; the KAN head and loss add <1% there)"
    return {
        "metric": "reach-timesteps/s fwd+adjoint (batched vs plain, 1 GPU); grad rel err vs FD",
        "value": value,
        "unit": "reach-timesteps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded generators; observations from a hidden KAN head's routed "
                "discharge x LogNormal(0, 0.1); no network access)",
        "config": {"workload": (f"c5: training micro-batch, {cg.size} nested gauges (basins "
                                f"log-uniform {cfg['size_lo']}-{cfg['size_hi']} reaches) on a "
                                f"{R0}-reach tree (window {cfg['window']}), the union of their "
                                f"basins = {R} reaches, {T} hourly steps; KAN head (10 -> 21, "
                                f"2 blocks, grid 5, order 3 -> n, q_spatial) -> routing "
                                f"forward -> tau {tau} trim + daily pooling -> warmup {wu} d -> "
                                f"batch NSE -> adjoint -> KAN backward -> clip 1.0 -> AdaDelta"),
                   "reaches": R, "timesteps": T, "gauges": int(cg.size),
                   "levels": net.n_levels, "checkpoint_window": window or (T - 1),
                   "schedule": schedule, "form": "persistent", "parallelism": "single GPU"},
        "roofline": roofline_of(launch_ms, rs, None, routing_kernels_ms=routing_ms,
                                other_ms=ms_per_step - routing_ms,
                                note="other_ms: KAN head forward and backward, cold start, "
                                     "loss, optimizer, host synchronisation of the loss "
                                     "(q' is put in schedule order once per batch)"),
        "losses": losses,
        "hbm": {"device_used_GiB": (total_b - free_b) / 2**30,
                "device_total_GiB": total_b / 2**30, "workspace_GiB": ws.nbytes() / 2**30},
        "plain": plain,
        "batched_over_plain": (value / plain["value"]) if plain else None,
        "cpu_baseline": cpu,
    }


def cpu_baseline(wl, steps):
    """The oracle (the f32 restatement of the reference CPU path, one thread) on
    the workload's own network and forcing (the host generator, bit-identical
    to the device's) for `steps` timesteps, forward + adjoint."""
    from oracle import oracle as O
    T = steps + 1
    R = wl.R
    qp = wl.forcing_host(T)
    P = O.make_params()
    pat = O.pattern_from_coo(R, wl.rows, wl.cols)
    q0 = O.hotstart(P, pat, qp[0])
    p = np.full(R, 21.0, np.float32)
    gQ = np.zeros((T, R), np.float32)
    gQ[:, wl.gauges] = 1.0
    at = wl.attrs
    t0 = time.perf_counter()
    O.route_adj(P, pat, at["n"], at["q_spatial"], p, at["length"], at["slope"], qp, gQ, q0=q0)
    dt = time.perf_counter() - t0
    log(f"[bench] cpu oracle: {dt:.1f}s for {R}x{steps}")
    out = {"value": R * steps / dt, "unit": "reach-timesteps/s", "cores": 1, "kind": "port",
           "sample": f"same {R}-reach network and forcing, {steps} timesteps; oracle "
                     f"route_adj (forward, per-step recompute and adjoint), fp32, one thread "
                     f"({dt:.1f} s)",
           "cpu_model": _cpu_model(), "host_cores": os.cpu_count()}
    # the same sample with the per-reach loops of each step on every core the
    # job may use (the triangular solves and the scatters stay serial, as in
    # the reference); the reference's own CPU path is single-threaded
    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    if threads > 1:
        O.set_threads(threads)
        try:
            t0 = time.perf_counter()
            O.route_adj(P, pat, at["n"], at["q_spatial"], p, at["length"], at["slope"], qp, gQ,
                        q0=q0)
            dtp = time.perf_counter() - t0
        finally:
            O.set_threads(1)
        log(f"[bench] cpu oracle, {threads} threads: {dtp:.1f}s")
        out["parallel"] = {"value": R * steps / dtp, "unit": "reach-timesteps/s",
                           "cores": threads, "kind": "port",
                           "sample": "the same sample with each step's per-reach loops "
                                     f"(coefficients, local backward) on {threads} OpenMP "
                                     f"threads; solves and scatters serial ({dtp:.1f} s)"}
    return out


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


if __name__ == "__main__":
    main()
