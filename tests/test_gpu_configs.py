"""Parity at the benchmark workloads' own sizes, against the FP64 oracle.

- Uneven-tile slide (300 tiles of 64..512 px per side, 5 Gaussian PSFs, 20 iterations): the whole
  plan runs in one batched call; every 10th tile, the three smallest edge tiles and two 512 x 512
  tiles of the channel with the largest PSF are compared with the FFT-FP64 RL
  (``oracle.rl_fft64``) at 1e-4 relative L2. The persistent-grid schedules (``stride``, ``queue``)
  run the same plan and agree with the default schedule.
- Differentiable form (64 tiles of 256 x 256, PSF sigma 1.8 on 17 x 17, 10 iterations, checkpoint
  every 5): losses of all 64 tiles, d/dx0 of two tiles and d/dpsf against the FP64 reverse-mode
  oracle; d/dx0 at 16 seeded pixels and d/dsigma against FP64 central finite differences
  (steps 1e-3 max(1, |x|) and 1e-4), all at 1e-3.
- PSF fitting (volumes of uneven depth x 256 x 256, PSF 31 x 15 x 15): two volumes of different
  depth in one plan against the FP64 reverse-mode oracle at the starting sigma (2.5, 1.0, 1.0),
  and three Adam steps of the fit.

The FP64 side is ``oracle.rl_nll_grad_fft64`` / ``oracle.rl_nll_fft64``, pinned against the
direct FP64 oracle in tests/test_oracle.py. Reference: crates/lucem-ops/src/restore/deconvolve.rs:45-149.
"""

from __future__ import annotations

import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def _dev(a):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


@pytest.fixture(scope="module")
def slide():
    from lucem_amd import synth

    return synth.tiles()


def _run_plan(ts, schedule=None):
    import torch

    import lucem_amd as LA
    from lucem_amd.batched import RLBatch

    old = os.environ.get("LUCEM_BATCH_SCHEDULE")
    if schedule:
        os.environ["LUCEM_BATCH_SCHEDULE"] = schedule
    try:
        ys = [_dev(t) for t in ts.tiles]
        psfs = [_dev(q) for q in ts.psfs]
        outs = [torch.empty_like(t) for t in ys]
        plan = RLBatch([tuple(t.shape) for t in ys], [tuple(q.shape) for q in psfs], ts.psf_index,
                       LA.RichardsonLucy(num_iter=ts.num_iter, clip=False))
    finally:
        if schedule:
            if old is None:
                del os.environ["LUCEM_BATCH_SCHEDULE"]
            else:
                os.environ["LUCEM_BATCH_SCHEDULE"] = old
    plan.execute(ys, outs, psfs)
    torch.cuda.synchronize()
    res = [o.cpu().numpy() for o in outs]
    plan.close()
    return res


def test_slide_plan_against_fp64(gpu, slide, oracle):
    ts = slide
    print(f"slide: {len(ts.tiles)} tiles, PSFs {[p.shape for p in ts.psfs]}, {ts.num_iter} iterations", flush=True)
    got = _run_plan(ts)
    areas = [t.size for t in ts.tiles]
    big_psf = int(np.argmax([p.shape[0] for p in ts.psfs]))
    pick = set(range(0, len(ts.tiles), 10))
    pick |= set(int(i) for i in np.argsort(areas)[:3])
    pick |= set([i for i, (t, c) in enumerate(zip(ts.tiles, ts.psf_index)) if c == big_psf and t.shape == (512, 512)][:2])
    errs = {}
    for i in sorted(pick):
        want = oracle.rl_fft64(ts.tiles[i], ts.psfs[ts.psf_index[i]], ts.num_iter)
        errs[i] = _rel(got[i], want)
        assert errs[i] < 1e-4, (i, ts.tiles[i].shape, errs[i])
    e = np.array(list(errs.values()))
    print(f"slide: {len(errs)} tiles against FP64: max {e.max():.2e}, median {np.median(e):.2e}; smallest "
          f"{[ts.tiles[i].shape for i in np.argsort(areas)[:3]]}, largest PSF {ts.psfs[big_psf].shape}")


@pytest.mark.parametrize("schedule", ["stride", "queue"])
def test_slide_plan_persistent_schedules(gpu, slide, schedule):
    """The persistent grids (a grid of occupancy x CUs workgroups striding over the work items, or
    taking them from an atomic counter) give the default schedule's results."""
    ts = slide
    print(f"slide: schedule {schedule}", flush=True)
    base = _run_plan(ts)
    got = _run_plan(ts, schedule)
    worst = max(_rel(a, b) for a, b in zip(got, base))
    print(f"slide: schedule {schedule} vs default: max rel-L2 {worst:.2e}")
    assert worst < 1e-6


@pytest.fixture(scope="module")
def grad_tiles():
    from lucem_amd import synth

    return synth.grad2d()


def test_grad_tiles_against_fp64(gpu, grad_tiles, oracle):
    import torch

    from lucem_amd.autograd import RLGradPlan

    ys, psf, _, T = grad_tiles
    n = ys.shape[0]
    print(f"grad: {n} tiles of {ys.shape[1:]}, psf {psf.shape}, {T} iterations, checkpoint 5", flush=True)
    plan = RLGradPlan([ys.shape[1:]] * n, psf.shape, T, 5)
    x0 = [torch.full(ys.shape[1:], 0.5, device="cuda") for _ in range(n)]
    yt = [_dev(y) for y in ys]
    loss = plan.forward(x0, yt, _dev(psf))
    gx = [torch.empty_like(x) for x in x0]
    gp = torch.empty(psf.shape, device="cuda")
    plan.backward(None, gx, gp)
    torch.cuda.synchronize()
    loss = loss.cpu().numpy()
    x0n = np.full(ys.shape[1:], 0.5)
    want_loss = np.array([oracle.rl_nll_fft64(ys[i], x0n, psf, T) for i in range(n)])
    e_loss = float(np.max(np.abs(loss - want_loss) / np.abs(want_loss)))
    assert e_loss < 1e-6
    for i in (0, 37):
        _, g64, _ = oracle.rl_nll_grad_fft64(ys[i], x0n, psf, T)
        e = _rel(gx[i].cpu().numpy(), g64)
        print(f"grad: tile {i}: d/dx0 rel-L2 {e:.2e}")
        assert e < 1e-3
    gp_want = sum(oracle.rl_nll_grad_fft64(ys[i], x0n, psf, T)[2] for i in range(n))
    e_p = _rel(gp.cpu().numpy(), gp_want)
    print(f"grad: losses max rel {e_loss:.2e}, d/dpsf over {n} tiles rel-L2 {e_p:.2e}")
    assert e_p < 1e-3
    # d/dx0 of tile 0 at 16 seeded pixels against FP64 central differences of the FP64 loss
    rng = np.random.default_rng(3003)
    pts = list(zip(rng.integers(0, ys.shape[1], 16), rng.integers(0, ys.shape[2], 16)))
    g0 = gx[0].cpu().numpy().astype(np.float64)
    fd, hip = [], []
    for (a, b) in pts:
        h = 1e-3 * max(1.0, abs(x0n[a, b]))
        xp, xm = x0n.copy(), x0n.copy()
        xp[a, b] += h
        xm[a, b] -= h
        fd.append((oracle.rl_nll_fft64(ys[0], xp, psf, T) - oracle.rl_nll_fft64(ys[0], xm, psf, T)) / (2 * h))
        hip.append(g0[a, b])
    e_fd = _rel(hip, fd)
    print(f"grad: 16 pixels against FP64 central differences rel-L2 {e_fd:.2e}")
    assert e_fd < 1e-3
    plan.close()


def test_grad_sigma_against_finite_differences(gpu, grad_tiles, oracle):
    import torch

    from lucem_amd import synth
    from lucem_amd.autograd import gaussian_psf, rl_nll

    ys, psf, _, T = grad_tiles
    sel = [0, 1]
    print(f"grad: d/dsigma over tiles {sel}", flush=True)
    sigma = torch.tensor(1.8, dtype=torch.float32, device="cuda", requires_grad=True)
    p = gaussian_psf(sigma, psf.shape).float().contiguous()
    x0 = [torch.full(ys.shape[1:], 0.5, device="cuda") for _ in sel]
    loss = rl_nll(x0, [_dev(ys[i]) for i in sel], p, T, checkpoint_every=5).sum()
    loss.backward()
    g_hip = float(sigma.grad)
    x0n = np.full(ys.shape[1:], 0.5)

    def f(s):
        q = synth.gaussian_psf(s, psf.shape)
        return sum(oracle.rl_nll_fft64(ys[i], x0n, q, T) for i in sel)

    h = 1e-4
    fd = (f(1.8 + h) - f(1.8 - h)) / (2 * h)
    e = abs(g_hip - fd) / abs(fd)
    print(f"grad: d/dsigma HIP {g_hip:.6e}, FP64 central difference {fd:.6e}, rel {e:.2e}")
    assert e < 1e-3


def test_psf_fit_volumes_against_fp64(gpu, oracle):
    import torch

    from lucem_amd import synth
    from lucem_amd.autograd import RLGradPlan
    from lucem_amd.psf_fit import fit_gaussian_sigma

    vols, _ = synth.psffit3d(count=6)
    sel = [1, 5]  # depths 32 and 46: uneven volumes in one plan
    psf0 = synth.gaussian_psf((2.5, 1.0, 1.0), (31, 15, 15)).astype(np.float32)
    shapes = [vols[i].shape for i in sel]
    print(f"psf fit: volumes {shapes}, psf {psf0.shape}", flush=True)
    plan = RLGradPlan(shapes, psf0.shape, 10, 5)
    x0 = [torch.full(s, 0.5, device="cuda") for s in shapes]
    loss = plan.forward(x0, [_dev(vols[i]) for i in sel], _dev(psf0))
    gx = [torch.empty_like(x) for x in x0]
    gp = torch.empty(psf0.shape, device="cuda")
    plan.backward(None, gx, gp)
    torch.cuda.synchronize()
    loss = loss.cpu().numpy()
    gp_want = np.zeros(psf0.shape)
    for j, i in enumerate(sel):
        l64, g64, p64 = oracle.rl_nll_grad_fft64(vols[i], np.full(shapes[j], 0.5), psf0, 10)
        gp_want += p64
        e_l = abs(loss[j] - l64) / abs(l64)
        e_g = _rel(gx[j].cpu().numpy(), g64)
        print(f"psf fit: volume {shapes[j]}: loss rel {e_l:.2e}, d/dx0 rel-L2 {e_g:.2e}")
        assert e_l < 1e-6 and e_g < 1e-3
    e_p = _rel(gp.cpu().numpy(), gp_want)
    print(f"psf fit: d/dpsf rel-L2 {e_p:.2e}")
    assert e_p < 1e-3
    plan.close()
    res = fit_gaussian_sigma([_dev(vols[i]) for i in sel], (31, 15, 15), (2.5, 1.0, 1.0), num_iter=10, steps=3)
    losses = [h[1] for h in res.history]
    print(f"psf fit: 3 Adam steps: sigma {res.sigma}, losses {losses}, {res.ms_per_step:.1f} ms per step")
    assert len(res.history) == 3 and all(np.isfinite(losses))
    assert all(np.isfinite(res.sigma)) and res.sigma != [2.5, 1.0, 1.0]
    assert losses[-1] < losses[0]
