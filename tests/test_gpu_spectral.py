"""GPU tests of the frequency-domain filters (filter/spectral.rs:20-212) against the FP64 oracle:
the Gaussian, box and phase-shift transfer functions on 1-D to 3-D spectra (one value or one per
axis, even and odd sizes), Butterworth low- and high-pass filters on the device FFT (including a
long Bluestein axis), and the reference's checks."""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _dev(a):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


def _rel(a, b):
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


SPECTRAL_CASES = [
    ("gaussian", (64, 80), [1.5]),
    ("gaussian", (9, 16, 11), [0.5, 2.0, 3.0]),
    ("box", (64, 80), [3.0, 7.0]),
    ("box", (101,), [4.5]),
    ("shift", (64, 80), [2.25, -1.5]),
    ("shift", (15, 21, 8), [0.3, -7.0, 2.5]),
]


@pytest.mark.parametrize("kind,shape,vals", SPECTRAL_CASES)
def test_spectral_filters_match_oracle(gpu, oracle, kind, shape, vals):
    import lucem_amd as LA

    rng = np.random.default_rng(len(shape) + len(vals))
    re = rng.normal(size=shape).astype(np.float32)
    im = rng.normal(size=shape).astype(np.float32)
    r, i = _dev(re), _dev(im)
    fn = {"gaussian": LA.fourier_gaussian_into, "box": LA.fourier_box_into, "shift": LA.fourier_shift_into}[kind]
    fn(r, i, vals)
    wr, wi = oracle.spectral_axes64(re, im, kind, vals)
    e = _rel(r.cpu().numpy() + 1j * i.cpu().numpy(), wr + 1j * wi)
    print(f"{kind} {shape} {vals}: rel-L2 {e:.2e}")
    assert e < 1e-6


@pytest.mark.parametrize("shape,cutoff,high,order", [
    ((128, 96), 0.1, False, 2.0), ((128, 96), 0.25, True, 3.0), ((255,), 0.5, False, 1.0),
    ((20, 32, 24), 0.2, True, 2.5), ((16, 8209), 0.15, False, 2.0),
])
def test_butterworth_matches_oracle(gpu, oracle, shape, cutoff, high, order):
    import lucem_amd as LA

    x = np.random.default_rng(sum(shape)).normal(size=shape).astype(np.float32)
    got = LA.butterworth(_dev(x), cutoff, high, order).cpu().numpy()
    want = oracle.butterworth64(x, cutoff, high, order)
    e = _rel(got, want)
    print(f"butterworth {shape} cutoff {cutoff} high {high} order {order}: rel-L2 {e:.2e}")
    assert e < 1e-5


def test_spectral_errors(gpu):
    import torch

    import lucem_amd as LA

    x = torch.ones((8, 8), device="cuda")
    with pytest.raises(LA.LucemError) as e:
        LA.fourier_gaussian_into(x, torch.ones((8, 9), device="cuda"), [1.0])
    assert e.value.kind == "SHAPE_MISMATCH"
    with pytest.raises(LA.LucemError) as e:
        LA.fourier_box_into(x, x.clone(), [1.0, 2.0, 3.0])
    assert e.value.kind == "INVALID_PARAMETER"
    for c in (0.0, 0.6, float("nan")):
        with pytest.raises(LA.LucemError) as e:
            LA.butterworth(x, c, False, 2.0)
        assert e.value.kind == "INVALID_PARAMETER"
    e0 = torch.empty((0, 4), device="cuda")
    LA.fourier_shift_into(e0, e0.clone(), [1.0])
    assert LA.butterworth(e0, 0.2, False, 2.0).shape == (0, 4)


@pytest.mark.parametrize("params", [
    (0.1, 0.0, 1.0, 1.0, 3.0, 0.0, 0, 0.0),
    (0.25, 0.6, 2.0, 1.5, 3.0, 0.4, 1, 0.5),
    (0.05, 1.2, 4.0, 2.0, 2.5, 0.0, 3, 0.0),
])
def test_gabor_matches_oracle(gpu, oracle, params):
    """gabor_2d (spectral.rs:249-268): the two correlations of the kernel's parts."""
    import lucem_amd as LA

    f, th, sx, sy, ns, off, mode, cval = params
    x = np.random.default_rng(int(f * 100)).uniform(size=(96, 128)).astype(np.float32)
    gr, gi = LA.gabor_2d(_dev(x), f, th, sx, sy, ns, off, mode, cval)
    kr, ki = oracle.gabor_kernel64(f, th, sx, sy, ns, off)
    wr = oracle.correlate(x.astype(np.float64), kr.astype(np.float32).astype(np.float64), mode, cval)
    wi = oracle.correlate(x.astype(np.float64), ki.astype(np.float32).astype(np.float64), mode, cval)
    e = max(_rel(gr.cpu().numpy(), wr), _rel(gi.cpu().numpy(), wi))
    print(f"gabor {params}: rel-L2 {e:.2e}")
    assert e < 1e-5
