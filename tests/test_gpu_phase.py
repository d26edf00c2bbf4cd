"""GPU tests of phase correlation (register/phase.rs:52-250) against the FP64 oracle: integer and
subpixel shifts in 1-D, 2-D and 3-D, with and without phase normalisation, several upsampling
factors, the reference's own case, and the reference's checks. The shift is compared exactly
(both land on the same 1/u grid point), the squared error and the phase difference to 1e-5."""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _dev(a):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


def _fourier_shift(x, shift):
    X = np.fft.fftn(x)
    ph = np.zeros(x.shape)
    for ax, (n, s) in enumerate(zip(x.shape, shift)):
        k = np.arange(n)
        f = np.where(k <= n // 2, k, k - n) / n
        shp = [1] * x.ndim
        shp[ax] = n
        ph = ph + (f * s).reshape(shp)
    return np.fft.ifftn(X * np.exp(-2j * np.pi * ph)).real


CASES = [
    ((64, 80), (3.0, -7.0), 1, True),
    ((64, 80), (2.3, -1.7), 10, True),
    ((63, 75), (-4.25, 6.5), 4, True),
    ((50, 50), (1.6, 2.2), 5, False),
    ((257,), (-11.3,), 10, True),
    ((16, 24, 20), (1.5, -2.25, 3.75), 4, True),
    ((512, 512), (17.35, -40.6), 20, True),
    ((100, 8209), (2.5, -3.5), 2, True),   # a long Bluestein axis in passes
]


@pytest.mark.parametrize("shape,shift,u,normalize", CASES)
def test_phase_correlation_matches_oracle(gpu, oracle, shape, shift, u, normalize):
    import lucem_amd as LA

    rng = np.random.default_rng(sum(shape) + u)
    # a white image (every frequency well above FP32 rounding, so that the normalised cross-power
    # spectrum is well conditioned), moved and with a little noise
    x = rng.uniform(0.0, 1.0, shape)
    y = _fourier_shift(x, shift) + rng.normal(scale=1e-3, size=shape)
    x32, y32 = x.astype(np.float32), y.astype(np.float32)
    print(f"phase_correlation {shape} shift {shift} u {u} normalize {normalize}", flush=True)
    got = LA.phase_correlation(_dev(x32), _dev(y32), LA.PhaseCorrelation(upsample_factor=u, normalize_phase=normalize))
    want_shift, want_err, want_ph = oracle.phase_correlation64(x32, y32, u, normalize)
    print(f"  shift {got.shift} vs {want_shift}; error {got.error:.6f} vs {want_err:.6f}; "
          f"phase {got.phasediff:.3e} vs {want_ph:.3e}")
    assert got.shift == want_shift
    # error = sqrt(1 - q): compare 1 - q, which FP32 rounding moves by ~1e-7 however small it is
    assert abs(got.error ** 2 - want_err ** 2) < 1e-5
    assert abs(got.phasediff - want_ph) < 1e-5


def test_phase_correlation_reference_case(gpu):
    """register/phase.rs:262-293: moved by (2.25, -1.5), upsample 20, shift (-2.25, 1.5) within
    assert_close at rtol 0.1."""
    import lucem_amd as LA

    base = np.array([((i * 6151 + 49999) % 1013) * 0.1 + 0.1 for i in range(256)]).reshape(16, 16)
    moved = _fourier_shift(base, (2.25, -1.5))
    r = LA.phase_correlation(_dev(base), _dev(moved), LA.PhaseCorrelation(upsample_factor=20))
    want = np.array([-2.25, 1.5])
    assert np.all(np.abs(np.array(r.shift) - want) <= 0.1 * np.abs(want) + 0.1 * np.abs(want).max())


def test_phase_correlation_errors(gpu):
    import torch

    import lucem_amd as LA

    x = torch.ones((8, 8), device="cuda")
    with pytest.raises(LA.LucemError) as e:
        LA.phase_correlation(x, torch.ones((8, 9), device="cuda"))
    assert e.value.kind == "SHAPE_MISMATCH"
    with pytest.raises(LA.LucemError) as e:
        LA.phase_correlation(torch.ones((0, 8), device="cuda"), torch.ones((0, 8), device="cuda"))
    assert e.value.kind == "EMPTY_INPUT"
    with pytest.raises(LA.LucemError) as e:
        LA.phase_correlation(x, x, LA.PhaseCorrelation(upsample_factor=0))
    assert e.value.kind == "INVALID_PARAMETER"
    r = LA.phase_correlation(x, x)
    assert r.shift == [0.0, 0.0] and r.error < 1e-3
