#!/usr/bin/env python
"""Benchmark of the MI355X Richardson-Lucy deconvolution (one JSON line on stdout).

Metric: RL Mvoxel-iterations per second = voxels x iterations x steps / seconds / 1e6, with the
input already resident in HBM. A step is one complete deconvolution call (all iterations) on
the workload:

- ``plain3d`` (default): 128x512x512 float32 volume, 31x15x15 Gaussian PSF, 50 iterations
  (the largest single-array workload; the reference RL is 2-D only, this is its N-D form);
- ``plain2d``: 1024x1024, 25x25 PSF, 30 iterations (the reference's own 2-D entry point).

Alongside the throughput the line carries:

- ``roofline``: the dominant kernel's algorithmic HBM bytes per launch divided by its average
  launch time (HIP events on the library's stream over the timed region), against the MI355X
  HBM3E peak of 8 TB/s; ``traffic`` is the measured HBM bytes per launch of that kernel from the
  committed rocprofv3 counter summary under profiles/ (null when there is none);
- ``iteration``: the whole RL iteration's algorithmic bytes (24 n + 80 N_c in 2-D, 24 n + 144 N_c
  in 3-D, on the minimal linear-convolution domain n + k - 1) over the measured time per iteration;
- ``cpu_baseline``: the CPU restatement of the reference algorithm (direct correlation, FP32,
  OpenMP) timed on this host on a bounded sample of the same workload;
- ``batched``: the uneven-tile slide workload (300 tiles of 64..512 px per side, 5 Gaussian PSFs
  of 11..21 px, 20 iterations) through the batched form (one graph launch per call) against the
  plain 2-D entry point called once per tile, with both results compared; the achieved bandwidth
  from the algorithmic bytes and from the HBM bytes the committed rocprofv3 counters measured;
- ``grad``: the differentiable form on 64 tiles of 256x256, PSF 17x17, 10 iterations, checkpoint
  every 5: forward time, forward + backward (d/dx0 and d/dpsf) time and their ratio, also against the
  plain batched forward of the same tiles;
- ``plain2d`` (with the default ``plain3d`` headline): the same fields for the 1024x1024 workload,
  with its own roofline and CPU baseline;
- ``psf_fit``: milliseconds per Adam step of the PSF fit (16 volumes of 32..128 x 256x256, PSF
  31x15x15, 10 iterations, one batched forward + backward per step);
- ``wiener``: Wiener deconvolution of a 2048x2048 image with a 25x25 Gaussian PSF (7-smooth sizes:
  the R2C passes; 2039x2048: the general FFT with Bluestein columns), the unsupervised form on
  1024x1024 (30 iterations, burn-in 10), and the FP64 numpy restatement timed on this host;
- ``fft``: the general complex FFT over both axes of a 2048x2048 split re/im array, and of a
  2039x2039 one (Bluestein lines), against 32 bytes per element (two read+write passes);
- ``filter``: the correlation entry point on 2048x2048 with a 31x31 window (FFT passes) and a
  5x5 one (direct sum), phase correlation of a 1024x1024 pair at upsample 20, Butterworth and Gabor
  on 2048x2048, the device phantom renderers on the 2-D and 3-D workloads' truths, the whole 3-D
  workload generated on the device (render, blur, Poisson noise), and the oracle's CPU
  correlation and phase correlation timed on this host;
- ``tiled``: the zero-copy block walk (3 channels of 2048x2048, blocks of 480 with a 32-px halo,
  15x15 Gaussian PSF, 20 iterations), plan creation included, and the Poisson ``add_noise`` of a
  2048x2048 image against its 8 bytes per pixel.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBPS = 8000.0


def minimal_domain_bytes(shape, kshape) -> float:
    """Algorithmic bytes of one RL iteration on the minimal domain L = n + k - 1."""
    n = 1
    for s in shape:
        n *= s
    L = [a + b - 1 for a, b in zip(shape, kshape)]
    nc = 1
    for a in L[:-1]:
        nc *= a
    nc *= L[-1] // 2 + 1
    per_complex = 80 if len(shape) == 2 else 144
    return 24.0 * n + per_complex * nc


def load_traffic(workload: str, kernel: str):
    p = ROOT / "profiles" / f"traffic_{workload}.json"
    if not p.exists():
        return None
    try:
        d = json.loads(p.read_text())
        return d.get("kernels", {}).get(kernel, {}).get("hbm_bytes_per_launch")
    except Exception:
        return None


def cpu_baseline(prob, workload: str, want_seconds: float = 15.0):
    # one OpenMP thread per core, pinned (read when the oracle's OpenMP runtime starts)
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    from oracle import oracle as O

    O.build()
    y = prob.image
    if workload == "plain3d":
        iters = 6
        sample = f"full 128x512x512 volume, {iters} of the 50 iterations (cost per iteration is constant)"
    else:
        iters = prob.num_iter
        sample = f"full 1024x1024 image, all {iters} iterations"
    t0 = time.perf_counter()
    st, _ = O.richardson_lucy(y, prob.psf, iters, clip=False, nd=(y.ndim == 3))
    dt = time.perf_counter() - t0
    assert st == 0
    return {
        "value": round(y.size * iters / dt / 1e6, 3),
        "unit": "Mvoxel-iters/s",
        "cores": O.num_threads(),
        "kind": "port",
        "sample": sample + f"; FP32 direct correlation, {dt:.1f} s",
        "cpu_model": _cpu_model(),
    }


def _cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def device_info() -> dict:
    """The device's name and CU count, and its memory clock and bus width as the HIP runtime reports
    them (``hipDeviceGetAttribute``; attribute ids 60 and 59 of ROCm 7.2's ``hipDeviceAttribute_t``),
    with the peak they imply, beside the published 8 TB/s the roofline uses. HBM3 and HBM3E move
    4 bits per pin per reported memory clock (MI300X: 1.3 GHz, 5.2 Gb/s per pin)."""
    import ctypes

    import torch

    p = torch.cuda.get_device_properties(0)
    out = {"name": p.name, "gcn_arch": getattr(p, "gcnArchName", ""), "cus": p.multi_processor_count,
           "hbm_bytes": p.total_memory}
    try:
        hip = ctypes.CDLL("libamdhip64.so")
        v = ctypes.c_int()
        clk = bw = None
        if hip.hipDeviceGetAttribute(ctypes.byref(v), 60, 0) == 0:
            clk = v.value
        if hip.hipDeviceGetAttribute(ctypes.byref(v), 59, 0) == 0:
            bw = v.value
        out.update({"memory_clock_khz": clk, "memory_bus_bits": bw})
        if clk and bw:
            out["implied_peak_GBps"] = round(4 * clk * 1e3 * bw / 8 / 1e9, 1)
    except OSError:
        pass
    out["peak_used_GBps"] = HBM_PEAK_GBPS
    return out


def _timeit(fn, reps: int) -> float:
    import torch

    fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / reps


def batched_leg(steps: int) -> dict:
    import torch

    import lucem_amd as LA
    from lucem_amd import _native as N
    from lucem_amd import synth
    from lucem_amd.batched import RLBatch

    ts = synth.tiles()
    ys = [torch.from_numpy(t).cuda() for t in ts.tiles]
    psfs = [torch.from_numpy(p).cuda() for p in ts.psfs]
    settings = LA.RichardsonLucy(num_iter=ts.num_iter, clip=False)
    outs = [torch.empty_like(y) for y in ys]
    b = RLBatch([tuple(y.shape) for y in ys], [tuple(p.shape) for p in psfs], ts.psf_index, settings)
    tb = _timeit(lambda: b.execute(ys, outs, psfs), max(steps, 3))
    client = N.default_client(0)
    ws = N.Workspace(client)
    outs_p = [torch.empty_like(y) for y in ys]

    def loop():
        for y, o, i in zip(ys, outs_p, ts.psf_index):
            LA.richardson_lucy_into(y, psfs[i], o, settings, ws, client)

    tp = _timeit(loop, 2)
    diff = max(float(((a - c).norm() / c.norm()).item()) for a, c in zip(outs, outs_p))
    # the same loop with the plain call on its whole-array passes (four launches per iteration),
    # the per-tile baseline before the plain call took the one-launch tiled iteration
    os.environ["LUCEM_RL_TILED"] = "0"
    try:
        tw = _timeit(loop, 2)
    finally:
        del os.environ["LUCEM_RL_TILED"]
    nvox = sum(y.numel() for y in ys)
    alg = sum(minimal_domain_bytes(tuple(y.shape), tuple(psfs[i].shape)) for y, i in zip(ys, ts.psf_index))
    gbps = alg * ts.num_iter / tb / 1e9
    # HBM bytes the counters measured per launch of each pass (profiles/traffic_batched.json) times
    # the launches of one call (the tiled batch: one launch per iteration; the per-pass batch: 2
    # correlations, 1 ratio and 1 update pass per iteration)
    tiled = b.launches == 1 + 4 * len(psfs) + ts.num_iter
    counted = None
    if tiled:
        per_launch = load_traffic("batched", "tile_batch")
        counted = ts.num_iter * per_launch if per_launch is not None else None
    else:
        per = {k: load_traffic("batched", k) for k in ("line_fused_y", "row_ratio", "row_update")}
        if all(v is not None for v in per.values()):
            counted = ts.num_iter * (2 * per["line_fused_y"] + per["row_ratio"] + per["row_update"])
    return {
        "path": "tiled: one launch per iteration over the output tiles of every tile" if tiled else "per-pass",
        "counter_bytes_per_call": counted,
        "achieved_GBps_counters": round(counted / tb / 1e9, 1) if counted else None,
        "frac_counters": round(counted / tb / 1e9 / HBM_PEAK_GBPS, 4) if counted else None,
        "workload": f"{len(ys)} tiles (64..512 px per side), {len(psfs)} PSFs, {ts.num_iter} iterations",
        "ms_per_step": round(tb * 1e3, 3),
        "value": round(nvox * ts.num_iter / tb / 1e6, 2),
        "unit": "Mvoxel-iters/s",
        "achieved_GBps": round(gbps, 1),
        "frac": round(gbps / HBM_PEAK_GBPS, 4),
        "plain_loop_ms": round(tp * 1e3, 3),
        "plain_loop_value": round(nvox * ts.num_iter / tp / 1e6, 2),
        "plain_loop_achieved_GBps": round(alg * ts.num_iter / tp / 1e9, 1),
        "speedup_vs_plain_loop": round(tp / tb, 3),
        "plain_loop_whole_array_ms": round(tw * 1e3, 3),
        "speedup_vs_whole_array_plain_loop": round(tw / tb, 3),
        "max_rel_l2_batched_vs_plain": diff,
        "device_bytes": b.device_bytes,
    }


def grad_leg(steps: int) -> dict:
    import torch

    import lucem_amd as LA
    from lucem_amd import synth
    from lucem_amd.autograd import RLGradPlan
    from lucem_amd.batched import RLBatch

    ys, psf, _, T = synth.grad2d()
    y = torch.from_numpy(ys).cuda()
    p = torch.from_numpy(psf).cuda()
    n = y.shape[0]
    plan = RLGradPlan([tuple(y.shape[1:])] * n, psf.shape, T, 5)
    xs = list(torch.full_like(y, 0.5).unbind(0))
    yl = list(y.unbind(0))
    gx = [torch.empty_like(a) for a in xs]
    gp = torch.empty_like(p)

    def step():
        plan.forward(xs, yl, p)
        plan.backward(None, gx, gp)

    # the plain batched forward of the same tiles (RL alone, no checkpoints, loss or saved spectra)
    fwd = RLBatch([tuple(y.shape[1:])] * n, [tuple(psf.shape)], [0] * n, LA.RichardsonLucy(num_iter=T, clip=False))
    outs = [torch.empty_like(a) for a in yl]
    reps = max(steps, 5)
    tf = _timeit(lambda: plan.forward(xs, yl, p), reps)
    tb = _timeit(lambda: fwd.execute(yl, outs, [p]), reps)
    tg = _timeit(step, reps)
    return {
        "workload": f"{n} tiles of 256x256, psf 17x17, {T} iterations, checkpoint every 5",
        "forward_ms": round(tf * 1e3, 4),
        "batched_forward_ms": round(tb * 1e3, 4),
        "grad_step_ms": round(tg * 1e3, 4),
        "ratio": round(tg / tf, 3),
        "ratio_vs_batched_forward": round(tg / tb, 3),
        "device_bytes": plan.device_bytes,
    }


def psf_fit_leg(steps: int) -> dict:
    """The PSF-fitting workload: 16 volumes of 32..128 x 256 x 256 (generated on the device from the
    workload's geometry), sigma of a 31 x 15 x 15 Gaussian PSF, one Adam step = one batched forward +
    backward through 10 RL iterations."""
    import torch

    from lucem_amd import phantoms
    from lucem_amd.psf_fit import fit_gaussian_sigma

    vols, _ = phantoms.psffit3d()
    fit_gaussian_sigma(vols, (31, 15, 15), (2.5, 1.0, 1.0), steps=1)  # plans, tuning
    torch.cuda.synchronize()
    n = max(steps, 3)
    res = fit_gaussian_sigma(vols, (31, 15, 15), (2.5, 1.0, 1.0), steps=n)
    return {"workload": "16 volumes of 32..128 x 256 x 256, psf 31x15x15, 10 iterations, Adam",
            "depths": [int(v.shape[0]) for v in vols], "ms_per_adam_step": round(res.ms_per_step, 3),
            "steps_timed": n}


def plain_leg(workload: str, steps: int, warmup: int, with_cpu: bool) -> dict:
    """One plain-form workload: throughput, the dominant kernel's roofline, the iteration's
    algorithmic bandwidth, per-kernel times and (optionally) the CPU baseline."""
    import numpy as np
    import torch

    import lucem_amd as LA
    from lucem_amd import _native as N
    from lucem_amd import synth

    prob = synth.plain3d() if workload == "plain3d" else synth.plain2d()
    y = torch.from_numpy(prob.image).cuda()
    psf = torch.from_numpy(prob.psf).cuda()
    out = torch.empty_like(y)
    client = N.default_client(0)
    ws = N.Workspace(client)
    settings = LA.RichardsonLucy(num_iter=prob.num_iter, clip=False)
    run = LA.richardson_lucy_nd_into if y.dim() == 3 else LA.richardson_lucy_into

    for _ in range(warmup):
        run(y, psf, out, settings, ws, client)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        run(y, psf, out, settings, ws, client)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    # per-kernel times in separate calls: the events would otherwise sit in the timed region
    client.profile(True)
    for _ in range(max(2, min(steps, 5))):
        run(y, psf, out, settings, ws, client)
    stats = client.profile_read()
    client.profile(False)
    assert bool(torch.isfinite(out).all()), "non-finite output"
    whole_ms = None
    if "tile_iter" in stats:
        # the same call on the whole-array passes (four launches per iteration), for comparison
        os.environ["LUCEM_RL_TILED"] = "0"
        try:
            ws_whole = N.Workspace(client)
            run(y, psf, out, settings, ws_whole, client)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            for _ in range(steps):
                run(y, psf, out, settings, ws_whole, client)
            torch.cuda.synchronize()
            whole_ms = (time.perf_counter() - t1) * 1e3 / steps
            ws_whole.close()
        finally:
            del os.environ["LUCEM_RL_TILED"]

    nvox = int(np.prod(prob.image.shape))
    iter_kernels = {k: v for k, v in stats.items() if not k.startswith("otf") and k != "psf_prep"}
    dom_name, dom = max(iter_kernels.items(), key=lambda kv: kv[1]["total_ms"])
    avg_ms = dom["total_ms"] / dom["launches"]
    bytes_per_launch = dom["bytes"] / dom["launches"]
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    kernel_ms = sum(v["total_ms"] for v in stats.values())
    iter_bytes = minimal_domain_bytes(prob.image.shape, prob.psf.shape)
    iter_s = elapsed / (steps * prob.num_iter)
    leg = {
        "workload": prob.meta["workload"] + f", {prob.num_iter} iterations, Reflect, clip off",
        "shape": list(prob.image.shape),
        "psf": list(prob.psf.shape),
        "num_iter": prob.num_iter,
        "ms_per_step": round(elapsed * 1e3 / steps, 3),
        "value": round(nvox * prob.num_iter * steps / elapsed / 1e6, 2),
        "unit": "Mvoxel-iters/s",
        "roofline": {
            "bound": "hbm",
            "kernel": dom_name,
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBPS, 4),
            "traffic": load_traffic(workload, dom_name),
            "algorithmic_bytes_per_launch": bytes_per_launch,
            "avg_launch_ms": round(avg_ms, 5),
        },
        "iteration": {
            "algorithmic_bytes": iter_bytes,
            "ms": round(iter_s * 1e3, 4),
            "achieved_GBps": round(iter_bytes / iter_s / 1e9, 1),
            "frac": round(iter_bytes / iter_s / 1e9 / HBM_PEAK_GBPS, 4),
        },
        "kernels": {k: {"launches": v["launches"], "avg_ms": round(v["total_ms"] / v["launches"], 5),
                        "share": round(v["total_ms"] / kernel_ms, 4),
                        "GBps": round(v["bytes"] / (v["total_ms"] * 1e-3) / 1e9, 1)}
                    for k, v in stats.items()},
        "cpu_baseline": cpu_baseline(prob, workload) if with_cpu else None,
    }
    if whole_ms is not None:
        leg["path"] = ("tiled: one launch per iteration, each workgroup an output tile with its halo "
                       "(overlap-save FFT correlations in LDS); roofline bytes are the whole-array "
                       "iteration's canonical bytes, the tile kernel's HBM traffic is `traffic`")
        leg["whole_array_ms_per_step"] = round(whole_ms, 3)
    return leg


def _gaussian_psf(k: int, sigma: float):
    import numpy as np

    r = np.arange(k) - k // 2
    g = np.exp(-(r[:, None] ** 2 + r[None, :] ** 2) / (2 * sigma * sigma))
    return (g / g.sum()).astype(np.float32)


def wiener_leg(steps: int, with_cpu: bool) -> dict:
    import numpy as np
    import torch

    import lucem_amd as LA
    from lucem_amd import _native as N

    rng = np.random.default_rng(7)
    client = N.default_client(0)
    ws = N.Workspace(client)
    psf_np = _gaussian_psf(25, 2.0)
    psf = torch.from_numpy(psf_np).cuda()
    s = LA.WienerDeconvolve(balance=0.1, clip=False)
    out = {}
    for name, shape in (("r2c", (2048, 2048)), ("general", (2039, 2048))):
        y = torch.from_numpy(rng.uniform(0, 1, shape).astype(np.float32)).cuda()
        o = torch.empty_like(y)
        t = _timeit(lambda: LA.wiener_deconvolve_into(y, psf, o, s, ws, client), max(steps, 5))
        n = shape[0] * shape[1]
        out[name] = {"shape": list(shape), "ms": round(t * 1e3, 4), "Mpixel_per_s": round(n / t / 1e6, 1)}
    # the R2C route: image read and output write (4 B per pixel each); the half spectrum written,
    # read and written by the fused line pass, read again (4 transfers of 8 B per bin); the
    # filter written once and read once (2 more)
    H, W = 2048, 2048
    bins = H * (W // 2 + 1)
    alg = 8 * H * W + 6 * 8 * bins
    out["r2c"]["algorithmic_bytes"] = alg
    out["r2c"]["achieved_GBps"] = round(alg / (out["r2c"]["ms"] * 1e-3) / 1e9, 1)
    y1 = torch.from_numpy(rng.uniform(0, 1, (1024, 1024)).astype(np.float32)).cuda()
    us = LA.UnsupervisedWiener(num_iter=30, burnin=10, clip=False, seed=1)
    tu = _timeit(lambda: LA.unsupervised_wiener(y1, psf, us, client, ws), max(steps // 2, 3))
    out["unsupervised"] = {"shape": [1024, 1024], "num_iter": 30, "ms": round(tu * 1e3, 3)}
    if with_cpu:
        out["cpu_baseline"] = cpu_baseline_wiener(rng.uniform(0, 1, (2048, 2048)).astype(np.float32), psf_np)
        out["speedup_vs_cpu"] = round(out["cpu_baseline"]["ms"] / out["r2c"]["ms"], 1)
    return out


def cpu_baseline_wiener(y, psf) -> dict:
    """The oracle's FP64 numpy Wiener filter, timed once on this host (one core)."""
    from oracle import oracle as O

    t0 = time.perf_counter()
    O.wiener64(y, psf, 0.1, False)
    tc = time.perf_counter() - t0
    return {"ms": round(tc * 1e3, 1), "kind": "port", "cores": 1,
            "sample": f"the FP64 numpy restatement on the same {y.shape[0]}x{y.shape[1]} case"}


def fft_leg(steps: int) -> dict:
    import torch

    from lucem_amd import fft

    out = {}
    for name, shape in (("smooth", (2048, 2048)), ("bluestein", (2039, 2039))):
        re = torch.randn(shape, device="cuda")
        im = torch.randn(shape, device="cuda")
        ro, io = torch.empty_like(re), torch.empty_like(im)
        t = _timeit(lambda: fft.fft_nd_into(re, im, ro, io), max(steps, 5))
        alg = 32 * shape[0] * shape[1]
        gbps = alg / t / 1e9
        out[name] = {"shape": list(shape), "ms": round(t * 1e3, 4), "achieved_GBps": round(gbps, 1),
                     "frac": round(gbps / HBM_PEAK_GBPS, 4)}
    return out


def tiled_leg(steps: int) -> dict:
    import torch

    import lucem_amd as LA
    from lucem_amd import tiled

    g = torch.Generator(device="cuda").manual_seed(5)
    img = torch.rand((3, 2048, 2048), device="cuda", generator=g) * 100 + 50
    psf = torch.from_numpy(_gaussian_psf(15, 2.0)).cuda()
    op = tiled.RichardsonLucyOp(num_iter=20, clip=False)
    t = _timeit(lambda: tiled.deconvolve_blocks(img, psf, op, block=(480, 480), halo=32), max(steps // 2, 3))
    x = torch.rand((2048, 2048), device="cuda", generator=g)
    o = torch.empty_like(x)
    model = LA.NoiseModel.Poisson(20.0)
    tn = _timeit(lambda: LA.add_noise_into(x, o, model, 7, False), max(steps, 10))
    return {
        "workload": "3 x 2048x2048, blocks 480 + 32-px halo, psf 15x15, 20 iterations",
        "ms": round(t * 1e3, 3),
        "Mpixel_iters_per_s": round(img.numel() * 20 / t / 1e6, 1),
        "add_noise_poisson_2048_ms": round(tn * 1e3, 4),
        "add_noise_GBps": round(8 * x.numel() / tn / 1e9, 1),
    }


def slide_leg() -> dict:
    """The on-disk slide walk: a 5-channel 8192x8192 OME-Zarr image (1024-px chunks, 1.34 GB) in a
    temporary directory, deconvolved block by block into a new store, timed once after a small
    warm-up slide, with the in-memory block walk of the same blocks beside it."""
    import shutil
    import tempfile

    import numpy as np
    import torch

    from lucem_amd import slide, tiled
    from lucem_amd.batched import schedule

    tmp = Path(tempfile.mkdtemp(prefix="lucem-slide-"))
    try:
        rng = np.random.default_rng(9)
        C, n = 5, 8192
        shape, chunks = (1, n, n), (1, 1024, 1024)
        img = rng.random((C, n, n), dtype=np.float32)
        img *= 100
        img += 50
        w = slide.ZarrWriter.create(tmp / "in", slide.WriterOptions("bench", [f"c{c}" for c in range(C)], shape, chunks,
                                                                    0.5, 0.5))
        for b in schedule(shape, chunks, (0, 0, 0)):
            o, sz = b.origin, b.size
            for c in range(C):
                w.write(c, o, sz, img[c:c + 1, o[1]:o[1] + sz[1], o[2]:o[2] + sz[2]])
        src = w.finish()
        psf = _gaussian_psf(15, 2.0)
        op = tiled.RichardsonLucyOp(num_iter=20, clip=False)
        small = slide.ZarrWriter.create(tmp / "small", slide.WriterOptions("w", ["a"], (1, 512, 512), (1, 256, 256)))
        for b in schedule((1, 512, 512), (1, 256, 256), (0, 0, 0)):
            small.write(0, b.origin, b.size, img[:1, :256, :256])
        slide.deconvolve_slide(small.finish(), tmp / "small_out", psf, op, block=256, halo=32)
        st = slide.SlideStats()
        t0 = time.perf_counter()
        slide.deconvolve_slide(src, tmp / "out", psf, op, block=1024, halo=32, read_ahead=8, threads=8, stats=st)
        t = time.perf_counter() - t0
        dev = torch.from_numpy(img).cuda()
        pt = torch.from_numpy(psf).cuda()
        tm = _timeit(lambda: tiled.deconvolve_blocks(dev, pt, op, block=(1024, 1024), halo=32), 2)
        owned = C * n * n
        return {
            "workload": "5 x 8192x8192 OME-Zarr (1024-px chunks) on local disk, blocks 1024 + 32-px halo, "
                        "psf 15x15, 20 iterations, written to a new store with its coarser levels",
            "s": round(t, 3),
            "Mpixel_iters_per_s": round(owned * 20 / t / 1e6, 1),
            "blocks": st.blocks, "groups": st.groups, "plans_built": st.plans_built,
            "wait_read_s": round(st.wait_read_s, 3), "device_s": round(st.device_s, 3),
            "wait_write_s": round(st.wait_write_s, 3), "finish_s": round(st.finish_s, 3),
            "in_memory_walk_s": round(tm, 3),
        }
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def cpu_baseline_filters(x, w, ref, mov) -> dict:
    """The oracle's CPU forms timed once on this host: the direct FP32 correlation (OpenMP, all
    cores) on a 256-row strip of the 2048x2048 image, scaled to the whole image, and the FP64
    numpy phase correlation (one core) on the same 1024x1024 pair."""
    import numpy as np

    from oracle import oracle as O

    strip = np.ascontiguousarray(x[:256])
    t0 = time.perf_counter()
    O.correlate(strip, w, 0, 0.0)
    tc = (time.perf_counter() - t0) * x.shape[0] / strip.shape[0]
    t0 = time.perf_counter()
    O.phase_correlation64(ref, mov, 20, True)
    tp = time.perf_counter() - t0
    return {"correlate_ms": round(tc * 1e3, 1), "correlate_cores": O.num_threads(),
            "phase_correlation_ms": round(tp * 1e3, 1), "phase_correlation_cores": 1, "kind": "port",
            "sample": "correlate: 256 of 2048 rows, scaled; phase correlation: the same 1024x1024 pair"}


def filter_leg(steps: int, with_cpu: bool) -> dict:
    import numpy as np
    import torch

    import lucem_amd as LA
    from lucem_amd import _native as N
    from lucem_amd import phantoms, synth

    g = torch.Generator(device="cuda").manual_seed(6)
    x = torch.rand((2048, 2048), device="cuda", generator=g)
    w = torch.from_numpy(_gaussian_psf(31, 5.0)).cuda()
    o = torch.empty_like(x)
    ws = N.Workspace(N.default_client(0))
    s = LA.Correlate2d()
    tc = _timeit(lambda: LA.correlate_2d_into(x, w, o, s, ws), max(steps, 10))
    p2 = synth.ellipse_params((1024, 1024), 200, np.random.default_rng(1000))
    t2 = _timeit(lambda: phantoms.render_ellipses((1024, 1024), p2), max(steps, 10))
    p3 = synth.sphere_params((128, 512, 512), 1500, np.random.default_rng(1001))
    t3 = _timeit(lambda: phantoms.render_spheres((128, 512, 512), p3), max(steps, 3))
    tg = _timeit(lambda: phantoms.plain3d(), 2)
    rng = np.random.default_rng(8)
    ref = rng.uniform(size=(1024, 1024)).astype(np.float32)
    mov = np.roll(ref, (17, -40), axis=(0, 1))
    rt, mt = torch.from_numpy(ref).cuda(), torch.from_numpy(mov).cuda()
    tpc = _timeit(lambda: LA.phase_correlation(rt, mt, LA.PhaseCorrelation(upsample_factor=20)), 3)
    tbw = _timeit(lambda: LA.butterworth(x, 0.2, False, 2.0), max(steps, 5))
    tgb = _timeit(lambda: LA.gabor_2d(x, 0.1, 0.5, 3.0, 3.0), max(steps, 5))
    w5 = torch.from_numpy(_gaussian_psf(5, 1.0)).cuda()
    t5 = _timeit(lambda: LA.correlate_2d_into(x, w5, o, s, ws), max(steps, 10))
    gs = LA.GaussianBlur()
    tgs = _timeit(lambda: LA.gaussian_blur_into(x, [2.0], o, gs), max(steps, 10))
    out = {
        "correlate_2048_w31_ms": round(tc * 1e3, 4),
        "correlate_note": "2048x2048 f32, 31x31 window, Reflect: transfer function built in the call",
        "correlate_2048_w5_ms": round(t5 * 1e3, 4),
        "correlate_w5_GBps": round(8 * x.numel() / t5 / 1e9, 1),
        "gaussian_blur_2048_s2_ms": round(tgs * 1e3, 4),
        "gaussian_blur_s2_GBps": round(16 * x.numel() / tgs / 1e9, 1),
        "gaussian_note": "sigma 2 (17 taps) per axis, Reflect: two 1-D passes, 8 B per element each",
        "phase_correlation_1024_u20_ms": round(tpc * 1e3, 3),
        "butterworth_2048_ms": round(tbw * 1e3, 4),
        "gabor_2048_ms": round(tgb * 1e3, 4),
        "render_ellipses_1024_ms": round(t2 * 1e3, 4),
        "render_spheres_128x512x512_ms": round(t3 * 1e3, 3),
        "device_workload_plain3d_ms": round(tg * 1e3, 2),
    }
    if with_cpu:
        out["cpu_baseline"] = cpu_baseline_filters(x.cpu().numpy(), _gaussian_psf(31, 5.0), ref, mov)
    return out


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="plain3d", choices=["plain3d", "plain2d"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="skip the batched and grad legs")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py runs on exactly one GPU (asked for {max(args.gpus, world)}): the RL path is one "
              "array per GPU; tiles and volumes shard across GPUs without exchange", file=sys.stderr)
        return 2

    import torch

    if not torch.cuda.is_available():
        print("no HIP device visible", file=sys.stderr)
        return 1
    head = plain_leg(args.workload, args.steps, args.warmup, not args.no_cpu_baseline)
    line = {
        "metric": "RL Mvoxel-iters/s",
        "value": head["value"],
        "unit": "Mvoxel-iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": head["workload"],
            "shape": head["shape"],
            "psf": head["psf"],
            "num_iter": head["num_iter"],
        },
        "roofline": head["roofline"],
        "iteration": head["iteration"],
        "kernels": head["kernels"],
        "cpu_baseline": head["cpu_baseline"],
        "device": device_info(),
    }
    for key in ("path", "whole_array_ms_per_step"):
        if key in head:
            line[key] = head[key]
    if not args.no_extras:
        other = "plain2d" if args.workload == "plain3d" else "plain3d"
        line[other] = plain_leg(other, max(args.steps, 10) if other == "plain2d" else args.steps, 2,
                                not args.no_cpu_baseline)
        line["batched"] = batched_leg(args.steps)
        line["grad"] = grad_leg(args.steps)
        line["psf_fit"] = psf_fit_leg(args.steps)
        line["wiener"] = wiener_leg(args.steps, not args.no_cpu_baseline)
        line["fft"] = fft_leg(args.steps)
        line["tiled"] = tiled_leg(args.steps)
        line["slide"] = slide_leg()
        line["filter"] = filter_leg(args.steps, not args.no_cpu_baseline)
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
