"""Seeded random sweeps over shapes against the oracles: the plain RL form in 2-D and 3-D, the
batched form, the Wiener filter, the complex FFT, the correlation and phase correlation, at sizes
from 1 up to a few hundred (odd, prime and 7-smooth), PSFs and windows from 1 to 31 taps, every
pad mode. Each case prints what it ran."""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _size(rng, lo, hi):
    return int(round(np.exp(rng.uniform(np.log(lo), np.log(hi)))))


def _dev(a):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


@pytest.mark.parametrize("seed", range(4))
def test_rl_2d_random_shapes(gpu, oracle, seed):
    import torch

    import lucem_amd as LA

    rng = np.random.default_rng(700 + seed)
    for case in range(8):
        shape = (_size(rng, 1, 400), _size(rng, 1, 400))
        kshape = (int(rng.integers(1, 32)), int(rng.integers(1, 32)))
        mode, clip, it = int(rng.integers(0, 5)), bool(rng.integers(0, 2)), int(rng.integers(1, 4))
        y = rng.uniform(0.05, 0.95, shape) if clip else rng.uniform(10, 200, shape)
        y = y.astype(np.float32)
        p = rng.uniform(0.05, 1.0, kshape).astype(np.float32)
        out = torch.empty(shape, device="cuda")
        LA.richardson_lucy_into(_dev(y), _dev(p), out, LA.RichardsonLucy(num_iter=it, clip=clip, mode=mode))
        torch.cuda.synchronize()
        st, ref = oracle.richardson_lucy(y, p, it, clip=clip, mode=mode)
        e = oracle.rel_l2(out.cpu().numpy(), ref)
        print(f"seed {seed} case {case}: {shape} psf {kshape} mode {mode} clip {clip} iters {it}: {e:.2e}")
        assert st == 0 and e < 1e-4


@pytest.mark.parametrize("seed", range(2))
def test_rl_3d_random_shapes(gpu, oracle, seed):
    import torch

    import lucem_amd as LA

    rng = np.random.default_rng(800 + seed)
    for case in range(5):
        shape = (_size(rng, 1, 40), _size(rng, 1, 70), _size(rng, 1, 70))
        kshape = tuple(int(rng.integers(1, 8)) for _ in range(3))
        mode, it = int(rng.integers(0, 5)), int(rng.integers(1, 3))
        y = rng.uniform(10, 200, shape).astype(np.float32)
        p = rng.uniform(0.05, 1.0, kshape).astype(np.float32)
        out = torch.empty(shape, device="cuda")
        LA.richardson_lucy_nd_into(_dev(y), _dev(p), out, LA.RichardsonLucy(num_iter=it, clip=False, mode=mode))
        torch.cuda.synchronize()
        st, ref = oracle.richardson_lucy(y, p, it, clip=False, mode=mode, nd=True)
        e = oracle.rel_l2(out.cpu().numpy(), ref)
        print(f"seed {seed} case {case}: {shape} psf {kshape} mode {mode} iters {it}: {e:.2e}")
        assert st == 0 and e < 1e-4


@pytest.mark.parametrize("seed", range(2))
def test_batched_random_tiles(gpu, oracle, seed):
    import lucem_amd as LA
    from lucem_amd.batched import richardson_lucy_batched

    rng = np.random.default_rng(900 + seed)
    psfs = [rng.uniform(0.05, 1.0, (int(rng.integers(1, 16)), int(rng.integers(1, 16)))).astype(np.float32)
            for _ in range(3)]
    shapes = [(_size(rng, 1, 300), _size(rng, 1, 300)) for _ in range(12)]
    index = [int(rng.integers(0, 3)) for _ in shapes]
    ys = [rng.uniform(10, 200, s).astype(np.float32) for s in shapes]
    mode = int(rng.integers(0, 5))
    outs = richardson_lucy_batched([_dev(y) for y in ys], [_dev(p) for p in psfs], index,
                                   LA.RichardsonLucy(num_iter=3, clip=False, mode=mode))
    for y, o, i in zip(ys, outs, index):
        st, ref = oracle.richardson_lucy(y, psfs[i], 3, clip=False, mode=mode)
        e = oracle.rel_l2(o.cpu().numpy(), ref)
        print(f"seed {seed}: tile {y.shape} psf {psfs[i].shape} mode {mode}: {e:.2e}")
        assert st == 0 and e < 1e-4


@pytest.mark.parametrize("seed", range(2))
def test_wiener_random_shapes(gpu, oracle, seed):
    import lucem_amd as LA

    rng = np.random.default_rng(1000 + seed)
    for case in range(8):
        shape = (_size(rng, 1, 300), _size(rng, 1, 300))
        kshape = (int(rng.integers(1, min(shape[0], 25) + 1)), int(rng.integers(1, min(shape[1], 25) + 1)))
        balance = float(np.exp(rng.uniform(np.log(1e-3), np.log(1.0))))
        y = rng.uniform(0, 1, shape).astype(np.float32)
        p = rng.uniform(0.05, 1.0, kshape).astype(np.float32)
        out = LA.wiener_deconvolve(_dev(y), _dev(p), LA.WienerDeconvolve(balance=balance, clip=False))
        ref = oracle.wiener64(y, p, balance, False)
        e = oracle.rel_l2(out.cpu().numpy(), ref)
        print(f"seed {seed} case {case}: {shape} psf {kshape} balance {balance:.3g}: {e:.2e}")
        assert e < 1e-4


@pytest.mark.parametrize("seed", range(2))
def test_fft_random_shapes(gpu, oracle, seed):
    import torch

    from lucem_amd import fft

    rng = np.random.default_rng(1100 + seed)
    for case in range(8):
        nd = int(rng.integers(1, 4))
        shape = tuple(_size(rng, 1, 3000 if nd == 1 else (300 if nd == 2 else 40)) for _ in range(nd))
        re, im = rng.normal(size=shape).astype(np.float32), rng.normal(size=shape).astype(np.float32)
        inverse = bool(rng.integers(0, 2))
        fr, fi = fft.fft_nd(_dev(re), _dev(im), inverse=inverse)
        torch.cuda.synchronize()
        wr, wi = oracle.fft_axis64(re, im, inverse=inverse)
        got = fr.cpu().numpy().astype(np.float64) + 1j * fi.cpu().numpy()
        want = wr + 1j * wi
        e = float(np.linalg.norm((got - want).ravel()) / max(np.linalg.norm(want.ravel()), 1e-30))
        print(f"seed {seed} case {case}: {shape} inverse {inverse}: {e:.2e}")
        assert e < 2e-6


def test_grad_random_shapes(gpu, oracle):
    import torch

    from lucem_amd.autograd import rl_nll

    rng = np.random.default_rng(1200)
    for case in range(6):
        nd = 2 if case < 4 else 3
        shape = tuple(_size(rng, 2, 120 if nd == 2 else 24) for _ in range(nd))
        kshape = tuple(int(rng.integers(1, 10 if nd == 2 else 5)) for _ in range(nd))
        mode, T = int(rng.integers(0, 5)), int(rng.integers(1, 4))
        x0 = rng.uniform(0.5, 2.0, shape).astype(np.float32)
        y = rng.uniform(10, 200, shape).astype(np.float32)
        p = rng.uniform(0.1, 1.0, kshape).astype(np.float32)
        xt = _dev(x0).requires_grad_(True)
        pt = _dev(p).requires_grad_(True)
        loss = rl_nll([xt], [_dev(y)], pt, T, checkpoint_every=2, mode=mode)
        loss.sum().backward()
        torch.cuda.synchronize()
        loss = loss.detach()
        L, gx, gp = oracle.rl_nll_grad64(y, x0, p, T, mode=mode)
        ex = oracle.rel_l2(xt.grad.cpu().numpy(), gx)
        ep = oracle.rel_l2(pt.grad.cpu().numpy(), gp)
        print(f"case {case}: {shape} psf {kshape} mode {mode} T {T}: loss {abs(float(loss[0]) - L) / abs(L):.1e}, "
              f"dx0 {ex:.2e}, dpsf {ep:.2e}")
        assert abs(float(loss[0]) - L) <= 1e-6 * abs(L) and ex < 1e-3 and ep < 1e-3


@pytest.mark.parametrize("seed", range(3))
def test_correlate_random_shapes(gpu, oracle, seed):
    """The correlation in 2-D (direct sum up to 13 x 13, FFT passes above) and 3-D, random
    windows, origins, modes and cval, against the FP64 oracle at the reference's tolerance."""
    import lucem_amd as LA

    rng = np.random.default_rng(900 + seed)
    for case in range(10):
        nd = 2 if case < 7 else 3
        shape = tuple(_size(rng, 1, 300 if nd == 2 else 40) for _ in range(nd))
        kshape = tuple(int(rng.integers(1, 26 if nd == 2 else 9)) for _ in range(nd))
        origin = tuple(int(rng.integers(-(k // 2), (k - 1) // 2 + 1)) for k in kshape)
        mode = int(rng.integers(0, 5))
        cval = float(rng.uniform(-2, 2)) if mode == 1 else 0.0
        x = rng.uniform(-1, 1, shape).astype(np.float32)
        w = rng.standard_normal(kshape).astype(np.float32)
        print(f"correlate {shape} window {kshape} origin {origin} mode {mode} cval {cval:.3f}", flush=True)
        s = LA.Correlate2d(mode=mode, cval=cval, origin=origin)
        got = (LA.correlate_2d if nd == 2 else LA.correlate_nd)(_dev(x), _dev(w), s).cpu().numpy()
        want = oracle.correlate(x.astype(np.float64), w.astype(np.float64), mode, cval, origin=origin)
        scale = float(np.abs(want).max())
        assert np.all(np.abs(got - want) <= 1e-5 * np.abs(want) + 1e-5 * scale)


@pytest.mark.parametrize("seed", range(2))
def test_phase_correlation_random_shifts(gpu, oracle, seed):
    """Random sizes (1-D to 3-D), shifts and upsampling factors: the same shift as the oracle."""
    import lucem_amd as LA

    rng = np.random.default_rng(950 + seed)
    for case in range(6):
        nd = 1 + case % 3
        shape = tuple(_size(rng, 16, 200 if nd < 3 else 40) for _ in range(nd))
        u = int(rng.choice([1, 2, 4, 10]))
        shift = tuple(float(np.round(rng.uniform(-n / 4, n / 4) * u) / u) for n in shape)
        x = rng.uniform(size=shape)
        X = np.fft.fftn(x)
        ph = np.zeros(shape)
        for ax, (n, s_) in enumerate(zip(shape, shift)):
            k = np.arange(n)
            f = np.where(k <= n // 2, k, k - n) / n
            ph = ph + (f * s_).reshape([n if a == ax else 1 for a in range(nd)])
        y = np.fft.ifftn(X * np.exp(-2j * np.pi * ph)).real + rng.normal(scale=1e-3, size=shape)
        x32, y32 = x.astype(np.float32), y.astype(np.float32)
        print(f"phase_correlation {shape} shift {shift} u {u}", flush=True)
        got = LA.phase_correlation(_dev(x32), _dev(y32), LA.PhaseCorrelation(upsample_factor=u))
        want, _, _ = oracle.phase_correlation64(x32, y32, u, True)
        assert got.shift == want
