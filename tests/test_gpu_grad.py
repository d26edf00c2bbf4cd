"""GPU tests of the differentiable form: loss, d/dx0 and d/dpsf against the FP64 oracle (exact
reverse-mode derivative of the FP64 direct restatement, itself pinned by finite differences in
tests/test_oracle.py), checkpointing invariance, 3-D and uneven tiles, autograd plumbing.

Tolerance: gradients within 1e-3 relative L2 of FP64; losses 1e-6 relative.
"""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GRAD_TOL = 1e-3


def _tiles(rng, n, shape, lo=50.0, hi=150.0):
    return [rng.uniform(lo, hi, shape).astype(np.float32) for _ in range(n)]


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-30))


def _run(x0s, ys, psf, num_iter, k=5, mode=0):
    import torch

    from lucem_amd.autograd import rl_nll

    x0 = [torch.from_numpy(x).cuda().requires_grad_(True) for x in x0s]
    y = [torch.from_numpy(v).cuda() for v in ys]
    p = torch.from_numpy(psf).cuda().requires_grad_(True)
    loss = rl_nll(x0, y, p, num_iter, checkpoint_every=k, mode=mode)
    loss.sum().backward()
    torch.cuda.synchronize()
    return loss.detach().cpu().numpy(), [x.grad.cpu().numpy() for x in x0], p.grad.cpu().numpy()


@pytest.mark.parametrize("mode", [0, 1, 2, 3, 4])
def test_grad_2d_matches_fp64_oracle(gpu, oracle, mode):
    rng = np.random.default_rng(30 + mode)
    shape, kshape, T = (24, 29), (5, 4), 4
    x0s = _tiles(rng, 3, shape, 0.5, 2.0)
    ys = _tiles(rng, 3, shape)
    psf = rng.uniform(0.1, 1.0, kshape).astype(np.float32)
    loss, gx, gp = _run(x0s, ys, psf, T, k=2, mode=mode)
    gp_ref = np.zeros(kshape)
    for i in range(3):
        L, gxr, gpr = oracle.rl_nll_grad64(ys[i], x0s[i], psf, T, mode=mode)
        assert abs(loss[i] - L) <= 1e-6 * abs(L)
        e = _rel(gx[i], gxr)
        print(f"mode {mode} tile {i}: dL/dx0 rel-L2 vs FP64 {e:.2e}")
        assert e < GRAD_TOL
        gp_ref += gpr
    e = _rel(gp, gp_ref)
    print(f"mode {mode}: dL/dpsf rel-L2 vs FP64 {e:.2e}")
    assert e < GRAD_TOL


@pytest.mark.parametrize("k", [1, 2, 3, 5, 7])
def test_checkpoint_interval_does_not_change_results(gpu, k):
    rng = np.random.default_rng(40)
    x0s = _tiles(rng, 2, (33, 20), 0.5, 2.0)
    ys = _tiles(rng, 2, (33, 20))
    psf = rng.uniform(0.1, 1.0, (7, 5)).astype(np.float32)
    ref = _run(x0s, ys, psf, 7, k=7)
    got = _run(x0s, ys, psf, 7, k=k)
    np.testing.assert_allclose(got[0], ref[0], rtol=1e-12)
    for a, b in zip(got[1], ref[1]):
        np.testing.assert_allclose(a, b, rtol=1e-6, atol=1e-9)
    np.testing.assert_allclose(got[2], ref[2], rtol=1e-5, atol=1e-9)


def test_grad_3d_uneven_depths(gpu, oracle):
    rng = np.random.default_rng(50)
    shapes = [(6, 14, 12), (9, 14, 12), (4, 10, 16)]
    psf = rng.uniform(0.1, 1.0, (3, 5, 3)).astype(np.float32)
    x0s = [rng.uniform(0.5, 2.0, s).astype(np.float32) for s in shapes]
    ys = [rng.uniform(50, 150, s).astype(np.float32) for s in shapes]
    loss, gx, gp = _run(x0s, ys, psf, 3, k=2)
    gp_ref = np.zeros(psf.shape)
    for i in range(3):
        L, gxr, gpr = oracle.rl_nll_grad64(ys[i], x0s[i], psf, 3)
        assert abs(loss[i] - L) <= 1e-6 * abs(L)
        assert _rel(gx[i], gxr) < GRAD_TOL
        gp_ref += gpr
    assert _rel(gp, gp_ref) < GRAD_TOL


def test_grad_loss_weights_and_autograd_chain(gpu, oracle):
    """Weighted sum of tile losses, and a Gaussian-sigma parameter chained by torch autograd."""
    import torch

    from lucem_amd.autograd import gaussian_psf, rl_nll

    rng = np.random.default_rng(60)
    shape = (20, 22)
    x0s = _tiles(rng, 2, shape, 0.5, 2.0)
    ys = _tiles(rng, 2, shape)
    sigma = torch.tensor(1.3, device="cuda", dtype=torch.float32, requires_grad=True)
    psf = gaussian_psf(sigma, (7, 7))
    x0 = [torch.from_numpy(x).cuda() for x in x0s]
    y = [torch.from_numpy(v).cuda() for v in ys]
    loss = rl_nll(x0, y, psf, 3, checkpoint_every=2)
    (2.0 * loss[0] + 0.5 * loss[1]).backward()
    g = float(sigma.grad)
    # FP64 central difference on sigma through the oracle
    def total(s):
        p = gaussian_psf(torch.tensor(s, dtype=torch.float64), (7, 7)).numpy()
        return (2.0 * oracle.rl_nll64(ys[0], x0s[0], p, 3) + 0.5 * oracle.rl_nll64(ys[1], x0s[1], p, 3))
    h = 1e-4
    fd = (total(1.3 + h) - total(1.3 - h)) / (2 * h)
    print(f"d/dsigma: {g:.6e} vs FP64 FD {fd:.6e}")
    assert abs(g - fd) <= GRAD_TOL * abs(fd)


def test_grad_fd_on_pixels(gpu, oracle):
    """d loss / d x0 at seeded pixels against FP64 central finite differences (config law,
    reduced size), step 1e-3 * max(1, |x|)."""
    from lucem_amd import synth

    ys_all, psf, truths, T = synth.grad2d(count=2, side=48, num_iter=10)
    x0s = [np.full(ys_all[i].shape, 0.5 * float(ys_all[i].mean()), np.float32) for i in range(2)]
    ys = [ys_all[i] for i in range(2)]
    loss, gx, gp = _run(x0s, ys, psf, T, k=5)
    rng = np.random.default_rng(70)
    for i in range(2):
        for _ in range(8):
            p = tuple(int(v) for v in rng.integers(0, 48, 2))
            x = x0s[i].astype(np.float64)
            step = 1e-3 * max(1.0, abs(x[p]))
            xp, xm = x.copy(), x.copy()
            xp[p] += step
            xm[p] -= step
            fd = (oracle.rl_nll64(ys[i], xp, psf, T) - oracle.rl_nll64(ys[i], xm, psf, T)) / (2 * step)
            assert abs(gx[i][p] - fd) <= GRAD_TOL * max(abs(fd), 1e-3 * np.abs(gx[i]).max())


def test_backward_after_other_forward_recomputes(gpu):
    import torch

    from lucem_amd.autograd import rl_nll

    rng = np.random.default_rng(80)
    x0 = [torch.from_numpy(rng.uniform(0.5, 2.0, (16, 16)).astype(np.float32)).cuda().requires_grad_(True)]
    y1 = [torch.from_numpy(rng.uniform(50, 150, (16, 16)).astype(np.float32)).cuda()]
    y2 = [torch.from_numpy(rng.uniform(50, 150, (16, 16)).astype(np.float32)).cuda()]
    psf = torch.from_numpy(rng.uniform(0.1, 1.0, (3, 3)).astype(np.float32)).cuda()
    l1 = rl_nll(x0, y1, psf, 3)
    l1.sum().backward()
    g_ref = x0[0].grad.clone()
    x0[0].grad = None
    l1 = rl_nll(x0, y1, psf, 3)
    l2 = rl_nll(x0, y2, psf, 3)  # same plan, later forward
    l1.sum().backward()
    torch.cuda.synchronize()
    np.testing.assert_allclose(x0[0].grad.cpu().numpy(), g_ref.cpu().numpy(), rtol=1e-6)
    del l2


# Axes whose FFT is longer than 4096 take the 512- and 1024-thread kernels in every pass of the
# forward, the adjoint and the PSF-gradient finish.
LONG_GRAD_CASES = [
    ((2, 6000), (1, 5), 0),     # rows of L = 6048: 512 threads
    ((5000, 2), (5, 1), 3),     # columns of L = 5000: 512-thread line passes (transposed ones too)
    ((2, 12000), (3, 3), 4),    # rows of L = 12005: 1024 threads
    ((3, 9000, 2), (3, 3, 1), 0),  # 3-D: y lines of L = 9002 -> 9072, 1024 threads
]


@pytest.mark.parametrize("shape,kshape,mode", LONG_GRAD_CASES)
def test_grad_long_axes_match_fp64_oracle(gpu, oracle, shape, kshape, mode):
    rng = np.random.default_rng(sum(shape) + mode)
    x0s = _tiles(rng, 2, shape, 0.5, 2.0)
    ys = _tiles(rng, 2, shape)
    psf = rng.uniform(0.1, 1.0, kshape).astype(np.float32)
    T = 3
    loss, gx, gp = _run(x0s, ys, psf, T, k=2, mode=mode)
    gp_ref = np.zeros(kshape)
    for i in range(2):
        L, gxr, gpr = oracle.rl_nll_grad64(ys[i], x0s[i], psf, T, mode=mode)
        assert abs(loss[i] - L) <= 1e-6 * abs(L)
        e = _rel(gx[i], gxr)
        print(f"{shape} tile {i}: dL/dx0 rel-L2 vs FP64 {e:.2e}")
        assert e < GRAD_TOL
        gp_ref += gpr
    e = _rel(gp, gp_ref)
    print(f"{shape}: dL/dpsf rel-L2 vs FP64 {e:.2e}")
    assert e < GRAD_TOL


def test_grad_plan_replays_match_every_time(gpu):
    """One plan's forward and backward replayed four times with eager work between: losses and
    gradients equal the first replay's bit for bit."""
    import torch

    from lucem_amd.autograd import RLGradPlan

    rng = np.random.default_rng(31)
    shapes = [(48, 40), (48, 40)]
    ys = [torch.from_numpy(y).cuda() for y in _tiles(rng, 2, shapes[0])]
    x0 = [torch.from_numpy(x).cuda() for x in _tiles(rng, 2, shapes[0], 0.5, 2.0)]
    psf = torch.from_numpy(rng.uniform(0.1, 1.0, (7, 5)).astype(np.float32)).cuda()
    plan = RLGradPlan(shapes, (7, 5), 6, checkpoint_every=3)
    ref = None
    for rep in range(4):
        junk = torch.randn(1 << 18, device="cuda").sum()
        loss = plan.forward(x0, ys, psf)
        gx = [torch.full(shapes[0], float(rep), device="cuda") for _ in shapes]
        gp = torch.full((7, 5), float(rep), device="cuda")
        plan.backward(None, gx, gp)
        torch.cuda.synchronize()
        assert bool(torch.isfinite(junk))
        got = [loss.cpu().numpy()] + [g.cpu().numpy() for g in gx] + [gp.cpu().numpy()]
        if ref is None:
            ref = got
        else:
            for a, b in zip(got, ref):
                np.testing.assert_array_equal(a, b)
    plan.close()


def test_plan_cache_is_bounded(gpu):
    """rl_nll keeps at most the configured number of plans (least recently used first out); an
    evicted plan whose backward is still pending stays alive through its graph and gives the same
    gradient as a fresh one."""
    import gc
    import weakref

    import torch

    from lucem_amd import autograd as AG

    AG.clear_plan_cache()
    AG.set_plan_cache_size(2)
    try:
        psf = torch.rand((3, 3), device="cuda") + 0.1
        y = torch.rand((20, 24), device="cuda") * 50 + 10
        x0 = torch.full_like(y, 0.5, requires_grad=True)
        pending = AG.rl_nll([x0], [y], psf, 3).sum()  # its plan is about to be evicted
        refs = []
        for h in (21, 22, 23):
            yy = torch.rand((h, 24), device="cuda") * 50 + 10
            AG.rl_nll([torch.full_like(yy, 0.5)], [yy], psf, 3)
            refs.append(weakref.ref(AG._plans[next(reversed(AG._plans))]))
        assert len(AG._plans) == 2
        gc.collect()
        assert refs[0]() is None  # evicted and unreferenced: freed
        pending.backward()
        g1 = x0.grad.clone()
        x1 = torch.full_like(y, 0.5, requires_grad=True)
        AG.rl_nll([x1], [y], psf, 3).sum().backward()
        torch.cuda.synchronize()
        assert torch.equal(g1, x1.grad)
    finally:
        AG.set_plan_cache_size(4)
        AG.clear_plan_cache()


def test_loss_against_a_separate_target(gpu, oracle):
    """RL on y, the NLL against another array of counts (the held-out half of a thinning): loss,
    d/dx0 and d/dpsf against the FP64 oracle, 2-D and 3-D tiles of different shapes in one plan."""
    import torch

    from lucem_amd.autograd import RLGradPlan

    rng = np.random.default_rng(44)
    for shapes, k in ((((30, 28), (25, 33)), (5, 5)), (((6, 20, 18), (9, 16, 20)), (3, 5, 5))):
        ys = [rng.uniform(10, 100, s).astype(np.float32) for s in shapes]
        ts = [rng.poisson(y).astype(np.float32) for y in ys]
        psf = rng.uniform(0.1, 1.0, k).astype(np.float32)
        print(f"target loss: tiles {shapes}, psf {k}", flush=True)
        plan = RLGradPlan(shapes, k, 4, 2)
        x0 = [torch.full(s, 0.5, device="cuda") for s in shapes]
        loss = plan.forward(x0, [torch.from_numpy(y).cuda() for y in ys], torch.from_numpy(psf).cuda(),
                            targets=[torch.from_numpy(t).cuda() for t in ts])
        gx = [torch.empty_like(x) for x in x0]
        gp = torch.empty(k, device="cuda")
        plan.backward(None, gx, gp)
        torch.cuda.synchronize()
        gp_want = np.zeros(k)
        for j in range(len(shapes)):
            l64, g64, p64 = oracle.rl_nll_grad_fft64(ys[j], np.full(shapes[j], 0.5), psf, 4, target=ts[j])
            gp_want += p64
            assert abs(float(loss[j]) - l64) <= 1e-6 * abs(l64)
            assert oracle.rel_l2(gx[j].cpu().numpy(), g64) < 1e-3
        assert oracle.rel_l2(gp.cpu().numpy(), gp_want) < 1e-3
        plan.close()


def test_thinning_splits_counts(gpu):
    """Binomial thinning: the halves add up to the counts, each half is integral and within the
    counts, and each takes about half of them."""
    import torch

    from lucem_amd.psf_fit import thin

    y = torch.poisson(torch.full((64, 256, 256), 40.0, device="cuda"))
    y1, y2 = thin([y], seed=3)
    assert torch.equal(y1[0] + y2[0], y)
    assert bool((y1[0] >= 0).all()) and bool((y2[0] >= 0).all()) and bool((y1[0] == y1[0].round()).all())
    frac = float(y1[0].sum() / y.sum())
    assert abs(frac - 0.5) < 2e-3


def test_known_object_fit_recovers_sigma(gpu):
    """The known-object objective (held-out likelihood, RL started from the true object) identifies
    the PSF: 4 volumes of the PSF-fitting law (31x15x15 Gaussian of sigma (3.5, 1.4, 1.6)), 2
    unrolled iterations, 100 Adam steps from (2.5, 1.0, 1.0) end within 3% on every axis."""
    import torch

    from lucem_amd import synth
    from lucem_amd.psf_fit import fit_gaussian_sigma

    vols, _, objs = synth.psffit3d(count=4, return_truth=True)
    print(f"known object: volumes {[v.shape for v in vols]}", flush=True)
    vt = [torch.from_numpy(v).cuda() for v in vols]
    ot = [torch.from_numpy(o).cuda() for o in objs]
    res = fit_gaussian_sigma(vt, (31, 15, 15), (2.5, 1.0, 1.0), num_iter=2, steps=100, lr=0.05,
                             objective="known", objects=ot)
    err = [abs(a - b) / b for a, b in zip(res.sigma, (3.5, 1.4, 1.6))]
    print(f"known object: sigma {res.sigma}, relative errors {err}, {res.ms_per_step:.1f} ms per step")
    assert max(err) < 0.03
