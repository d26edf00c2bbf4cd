"""HBM traffic per kernel launch from rocprofv3 counter runs.

Two helpers for the ``traffic`` figure of ``bench.py``:

``python -m lucem_amd.profiling copy``
    a plain device-to-device copy of 1 GiB, three times: the calibration program. Its
    ``FETCH_SIZE`` / ``WRITE_SIZE`` readings against the known 1 GiB each way give the factor that
    turns each counter into bytes on this GPU.

``python -m lucem_amd.profiling run plain3d|plain2d|batched|grad|wiener|fft [REPS]``
    a short run of one workload for a kernel trace: the plain call on 128x512x512 (PSF 31x15x15,
    50 iterations) or 1024x1024 (PSF 25x25, 30 iterations), the uneven-tile slide batch (300
    tiles, 20 iterations), the differentiable form (64 tiles of 256x256, forward + backward), the Wiener
    filter (2048x2048, 25x25 Gaussian PSF, balance 0.1) or the complex FFT over both axes of a
    2048x2048 split re/im array; REPS calls (default 3) after one warm-up call.

``python -m lucem_amd.profiling summarize WORKLOAD FETCH_DIR WRITE_DIR FETCH_CAL_DIR WRITE_CAL_DIR``
    reads the ``run_counter_collection.csv`` files of four ``rocprofv3 --pmc FETCH_SIZE`` /
    ``--pmc WRITE_SIZE`` runs (the bench and the copy, each counter in a run of its own) and writes
    ``profiles/traffic_<WORKLOAD>.json``: per pass, the median calibrated HBM bytes per launch.

rocprofv3 names kernels by template (``col_pass_kernel<256, 16, 1>``); the op id is the last
template argument, and each RL pass of a workload is one (kernel family, op) pair. The tiled 2-D
iteration is one kernel family (``tile_iter_kernel<L0, L1>``, and ``tile_batch_kernel<L0, L1>`` for
the batched form) whatever its lengths, and so is the compile-time build of the fused line pass
(``col_fused_static_kernel<L, B>`` and ``col_fused_inplace_kernel<L, B>``, counted as ``COL_FUSED``); the compile-time row passes
(``row_static_kernel<L, P, OP>``) and y passes (``col_line_static_kernel<L, B, OP>``) count as their
``row_pass_kernel`` / ``col_pass_kernel`` op.
"""

from __future__ import annotations

import csv
import json
import re
import statistics
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent

# pass name used by the library's own profiling -> (kernel family, op id)
PASS_KERNELS = {
    "plain3d": {
        "line_fwd_y": ("col_pass_kernel", 0),
        "line_fused_z": ("col_pass_kernel", 1),
        "line_inv_y": ("col_pass_kernel", 2),
        "row_ratio": ("row_pass_kernel", 2),
        "row_update": ("row_pass_kernel", 3),
    },
    "plain2d": {
        "tile_iter": ("tile_iter_kernel", None),
        "line_fused_y": ("col_pass_kernel", 1),
        "row_ratio": ("row_pass_kernel", 2),
        "row_update": ("row_pass_kernel", 3),
    },
    "batched": {
        "tile_batch": ("tile_batch_kernel", None),
        "line_fused_y": ("col_batch_kernel", 1),
        "row_ratio": ("row_batch_kernel", 2),
        "row_update": ("row_batch_kernel", 3),
    },
}

_NAME = re.compile(r"lucem::(\w+)<(\d+), (\d+), (\d+)>")
_TILE = re.compile(r"(tile_iter_kernel|tile_batch_kernel)<\d+, \d+>")
# the compile-time build of the fused line pass (COL_FUSED, op 1), named by its length and columns
_COL_STATIC = re.compile(r"col_fused_(static|inplace)_kernel<\d+, \d+>")
CAL_BYTES = 1 << 30


def copy_calibration() -> None:
    import torch

    a = torch.empty(CAL_BYTES // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(f"copied {CAL_BYTES} bytes 3 times", flush=True)


def run_workload(name: str, reps: int = 3) -> None:
    import torch

    from . import synth
    from .restore import RichardsonLucy

    if name in ("plain3d", "plain2d"):
        from . import _native as N
        from .restore import richardson_lucy_into, richardson_lucy_nd_into

        prob = synth.plain3d() if name == "plain3d" else synth.plain2d()
        y = torch.from_numpy(prob.image).cuda()
        p = torch.from_numpy(prob.psf).cuda()
        o = torch.empty_like(y)
        client = N.default_client(0)
        ws = N.Workspace(client)
        s = RichardsonLucy(num_iter=prob.num_iter, clip=False)
        run = richardson_lucy_nd_into if y.dim() == 3 else richardson_lucy_into
        fn = lambda: run(y, p, o, s, ws, client)  # noqa: E731
    elif name == "batched":
        from .batched import RLBatch

        ts = synth.tiles()
        ys = [torch.from_numpy(t).cuda() for t in ts.tiles]
        psfs = [torch.from_numpy(p).cuda() for p in ts.psfs]
        outs = [torch.empty_like(y) for y in ys]
        b = RLBatch([tuple(y.shape) for y in ys], [tuple(p.shape) for p in psfs], ts.psf_index,
                    RichardsonLucy(num_iter=ts.num_iter, clip=False))
        fn = lambda: b.execute(ys, outs, psfs)  # noqa: E731
    elif name == "grad":
        from .autograd import RLGradPlan

        ys, psf, _, T = synth.grad2d()
        y = torch.from_numpy(ys).cuda()
        p = torch.from_numpy(psf).cuda()
        plan = RLGradPlan([tuple(y.shape[1:])] * y.shape[0], psf.shape, T, 5)
        xs = list(torch.full_like(y, 0.5).unbind(0))
        yl = list(y.unbind(0))
        gx = [torch.empty_like(a) for a in xs]
        gp = torch.empty_like(p)

        def fn():
            plan.forward(xs, yl, p)
            plan.backward(None, gx, gp)
    elif name == "wiener":
        import numpy as np

        from .wiener import WienerDeconvolve, wiener_deconvolve_into

        r = np.arange(25) - 12
        g = np.exp(-(r[:, None] ** 2 + r[None, :] ** 2) / 8.0)
        psf = torch.from_numpy((g / g.sum()).astype(np.float32)).cuda()
        y = torch.rand((2048, 2048), device="cuda")
        o = torch.empty_like(y)
        s = WienerDeconvolve(balance=0.1, clip=False)
        fn = lambda: wiener_deconvolve_into(y, psf, o, s)  # noqa: E731
    elif name == "fft":
        from .fft import fft_nd_into

        re, im = torch.randn((2048, 2048), device="cuda"), torch.randn((2048, 2048), device="cuda")
        ro, io = torch.empty_like(re), torch.empty_like(im)
        fn = lambda: fft_nd_into(re, im, ro, io)  # noqa: E731
    else:
        raise ValueError(f"unknown workload {name!r}")
    fn()
    torch.cuda.synchronize()
    for _ in range(reps):
        fn()
    torch.cuda.synchronize()
    print(f"{name}: {reps} calls done", flush=True)


def _rows(d: Path, counter: str):
    files = sorted(Path(d).rglob("*counter_collection.csv"))
    if not files:
        raise FileNotFoundError(f"no counter_collection.csv under {d}")
    for f in files:
        for r in csv.DictReader(open(f)):
            if r["Counter_Name"] == counter:
                yield r["Kernel_Name"], float(r["Counter_Value"])


def _calibration(d: Path, counter: str) -> float:
    vals = [v for name, v in _rows(d, counter) if "copyBuffer" in name and v > 0.5 * CAL_BYTES / 1024 / 4]
    if not vals:
        raise ValueError(f"no copy kernel of 1 GiB in {d}")
    return CAL_BYTES / (statistics.median(vals) * 1024.0)  # bytes moved / bytes counted


def _per_pass(d: Path, counter: str, workload: str):
    groups: dict = {}
    for name, v in _rows(d, counter):
        m = _NAME.search(name)
        if m:
            # the compile-time builds (row_static_kernel<L, P, OP>, col_line_static_kernel<L, B, OP>)
            # count as their run-time pass of the same op
            fam = {"row_static_kernel": "row_pass_kernel", "col_line_static_kernel": "col_pass_kernel"}.get(
                m.group(1), m.group(1))
            groups.setdefault((fam, int(m.group(4))), []).append(v)
        elif (t := _TILE.search(name)):
            groups.setdefault((t.group(1), None), []).append(v)
        elif _COL_STATIC.search(name):
            groups.setdefault(("col_pass_kernel", 1), []).append(v)
    out = {}
    for pname, key in PASS_KERNELS[workload].items():
        if key in groups:
            out[pname] = (statistics.median(groups[key]), len(groups[key]))
    return out


def summarize(workload: str, fetch_dir, write_dir, fetch_cal, write_cal) -> dict:
    kf = _calibration(Path(fetch_cal), "FETCH_SIZE")
    kw = _calibration(Path(write_cal), "WRITE_SIZE")
    f = _per_pass(Path(fetch_dir), "FETCH_SIZE", workload)
    w = _per_pass(Path(write_dir), "WRITE_SIZE", workload)
    kernels = {}
    for p in PASS_KERNELS[workload]:
        if p in f and p in w:
            rb, wb = f[p][0] * 1024.0 * kf, w[p][0] * 1024.0 * kw
            kernels[p] = {"read_bytes": round(rb), "write_bytes": round(wb), "hbm_bytes_per_launch": round(rb + wb),
                          "dispatches": f[p][1]}
    return {"workload": workload, "fetch_scale": round(kf, 4), "write_scale": round(kw, 4),
            "scale_note": "bytes moved by a 1 GiB device copy / counter reading of that copy", "kernels": kernels}


def main(argv) -> int:
    if argv[:1] == ["copy"]:
        copy_calibration()
        return 0
    if argv[:1] == ["run"] and len(argv) in (2, 3):
        run_workload(argv[1], int(argv[2]) if len(argv) == 3 else 3)
        return 0
    if argv[:1] == ["summarize"] and len(argv) == 6:
        res = summarize(*argv[1:])
        dst = ROOT / "profiles" / f"traffic_{argv[1]}.json"
        dst.write_text(json.dumps(res, indent=1) + "\n")
        print(json.dumps(res, indent=1))
        return 0
    print(__doc__, file=sys.stderr)
    return 2


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
