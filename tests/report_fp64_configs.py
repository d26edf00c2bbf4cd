"""Relative L2 error of the HIP path against the FP64 oracle for every benchmark workload (run on a
GPU machine from the repository root: ``python tests/report_fp64_configs.py``; prints one JSON line).

- plain2d: 1024x1024, 25x25 PSF, 30 iterations, against the FFT-FP64 restatement of the
  reference algorithm (oracle.rl_fft64);
- plain3d: 128x512x512, 31x15x15 PSF, 50 iterations, the same at full size;
- tiles: the uneven-tile slide workload through the batched form, every 10th of the 300 tiles;
- grad2d: the differentiable form's forward estimate x_10 of 8 of the 64 tiles, and the loss and
  gradients (d/dx0, d/dpsf) of 2 tiles against the FP64 reverse-mode oracle;
- psffit3d: 2 of the 16 volumes, 10 iterations with the true PSF.

The FP32 oracle (the reference's own FP32 arithmetic) is within 1e-4 of FP64 on these workloads as
well; the numbers here are the HIP path's distance to FP64.
"""

from __future__ import annotations

import json
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def main() -> None:
    import torch

    import lucem_amd as LA
    from lucem_amd import synth
    from lucem_amd.autograd import RLGradPlan
    from lucem_amd.batched import RLBatch
    from oracle import oracle as O

    out = {}
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()  # noqa: E731

    p = synth.plain2d()
    got = LA.richardson_lucy(dev(p.image), dev(p.psf), LA.RichardsonLucy(num_iter=p.num_iter, clip=False))
    out["plain2d"] = rel(got.cpu().numpy(), O.rl_fft64(p.image, p.psf, p.num_iter))
    print("plain2d", out["plain2d"], flush=True)

    p = synth.plain3d()
    y = dev(p.image)
    o = torch.empty_like(y)
    LA.richardson_lucy_nd_into(y, dev(p.psf), o, LA.RichardsonLucy(num_iter=p.num_iter, clip=False))
    x = None
    for done in range(0, p.num_iter, 10):  # in steps of 10 iterations, reporting progress
        x = O.rl_fft64(p.image, p.psf, min(10, p.num_iter - done), x0=x)
        print(f"plain3d FP64: {done + 10} of {p.num_iter} iterations", flush=True)
    out["plain3d"] = rel(o.cpu().numpy(), x)
    print("plain3d", out["plain3d"], flush=True)

    ts = synth.tiles()
    ys = [dev(t) for t in ts.tiles]
    psfs = [dev(q) for q in ts.psfs]
    outs = [torch.empty_like(t) for t in ys]
    b = RLBatch([tuple(t.shape) for t in ys], [tuple(q.shape) for q in psfs], ts.psf_index,
                LA.RichardsonLucy(num_iter=ts.num_iter, clip=False))
    b.execute(ys, outs, psfs)
    torch.cuda.synchronize()
    errs = [rel(outs[i].cpu().numpy(), O.rl_fft64(ts.tiles[i], ts.psfs[ts.psf_index[i]], ts.num_iter))
            for i in range(0, len(ys), 10)]
    out["tiles_max"] = max(errs)
    out["tiles_median"] = float(np.median(errs))
    print("tiles", out["tiles_max"], flush=True)

    gy, gpsf, _, T = synth.grad2d()
    sel = list(range(8))
    plan = RLGradPlan([gy.shape[1:]] * len(sel), gpsf.shape, T, checkpoint_every=5)
    x0 = [torch.full(gy.shape[1:], 0.5, device="cuda") for _ in sel]
    yt = [dev(gy[i]) for i in sel]
    xo = [torch.empty_like(t) for t in yt]
    plan.forward(x0, yt, dev(gpsf), xo)
    torch.cuda.synchronize()
    out["grad2d_forward_max"] = max(rel(xo[j].cpu().numpy(), O.rl_fft64(gy[i], gpsf, T)) for j, i in enumerate(sel))
    sel2 = [0, 1]
    plan2 = RLGradPlan([gy.shape[1:]] * len(sel2), gpsf.shape, T, checkpoint_every=5)
    x0n = [np.random.default_rng(i).uniform(0.5, 2.0, gy.shape[1:]).astype(np.float32) for i in sel2]
    loss = plan2.forward([dev(x) for x in x0n], [dev(gy[i]) for i in sel2], dev(gpsf))
    gx = [torch.empty(gy.shape[1:], device="cuda") for _ in sel2]
    gp = torch.zeros(gpsf.shape, device="cuda")
    plan2.backward(None, gx, gp)
    torch.cuda.synchronize()
    e_loss, e_gx, gp_want = [], [], np.zeros(gpsf.shape)
    for j, i in enumerate(sel2):
        l64, gx64, gp64 = O.rl_nll_grad64(gy[i], x0n[j], gpsf, T)
        e_loss.append(abs(float(loss[j]) - l64) / abs(l64))
        e_gx.append(rel(gx[j].cpu().numpy(), gx64))
        gp_want += gp64
    out["grad2d_loss_max"] = max(e_loss)
    out["grad2d_dx0_max"] = max(e_gx)
    out["grad2d_dpsf"] = rel(gp.cpu().numpy(), gp_want)
    print("grad2d", out["grad2d_dx0_max"], out["grad2d_dpsf"], flush=True)

    vols, vpsf = synth.psffit3d(count=2)
    e = []
    for v in vols:
        o = torch.empty(v.shape, device="cuda")
        LA.richardson_lucy_nd_into(dev(v), dev(vpsf), o, LA.RichardsonLucy(num_iter=10, clip=False))
        e.append(rel(o.cpu().numpy(), O.rl_fft64(v, vpsf, 10)))
    out["psffit3d_max"] = max(e)
    print(json.dumps({k: float(f"{v:.3g}") for k, v in out.items()}))


if __name__ == "__main__":
    main()
