"""GPU tests of add_noise (intensity/noise.rs:35-143) against the float32 oracle: impulse models
bit for bit (integer hash and exact comparisons), Gaussian and speckle to the rounding of log/cos,
Poisson counts equal but for rare draws at a rounding boundary; the reference's own check and
the empty-input error."""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _run(x, model, seed, clip):
    import torch

    import lucem_amd as LA

    return LA.add_noise(torch.from_numpy(x).cuda(), model, seed, clip).cpu().numpy()


@pytest.mark.parametrize("kind,a,b", [(2, 0.1, 0.0), (3, 0.2, 0.0), (4, 0.3, 0.4)])
def test_impulse_noise_is_bit_exact(gpu, oracle, kind, a, b):
    from lucem_amd.noise import NoiseModel

    x = np.random.default_rng(kind).uniform(0, 1, (300, 217)).astype(np.float32)
    got = _run(x, NoiseModel(kind, a, b), 99 + (5 << 32), False)
    want = oracle.add_noise32(x, kind, a, b, seed=99 + (5 << 32), clip=False)
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("kind,clip", [(0, False), (0, True), (1, False)])
def test_gaussian_and_speckle_match_to_rounding(gpu, oracle, kind, clip):
    from lucem_amd.noise import NoiseModel

    x = np.random.default_rng(10 + kind).uniform(0, 1, (512, 384)).astype(np.float32)
    got = _run(x, NoiseModel(kind, 0.01, 0.08), 2024, clip)
    want = oracle.add_noise32(x, kind, 0.01, 0.08, seed=2024, clip=clip)
    err = float(np.max(np.abs(got - want)))
    print(f"kind {kind} clip {clip}: max abs diff {err:.2e}")
    assert err < 4e-6


@pytest.mark.parametrize("scale", [1.0, 10.0, 100.0])
def test_poisson_counts_match(gpu, oracle, scale):
    from lucem_amd.noise import NoiseModel

    x = np.random.default_rng(int(scale)).uniform(0, 1, (400, 400)).astype(np.float32)
    got = _run(x, NoiseModel.Poisson(scale), 5, False)
    want = oracle.add_noise32(x, 5, scale, seed=5, clip=False)
    diff = np.abs(got - want) * scale
    frac = float((diff > 1e-3).mean())
    print(f"scale {scale}: fraction of differing draws {frac:.2e}")
    assert frac < 1e-4
    assert np.allclose(diff, np.round(diff), atol=1e-3) and diff.max() <= 1.0 + 1e-3


def test_reference_check_and_in_place(gpu, oracle):
    """intensity/mod.rs:52-63: Gaussian (0, 0.05), seed 42, clip on the 8x8 ramp stays in [0, 1]."""
    import torch

    import lucem_amd as LA
    from lucem_amd.noise import NoiseModel, add_noise_into

    ramp = (np.arange(64, dtype=np.float32) / 63.0).reshape(8, 8)
    y = _run(ramp, NoiseModel.Gaussian(0.0, 0.05), 42, True)
    assert ((y >= 0) & (y <= 1)).all()
    t = torch.from_numpy(ramp).cuda()
    add_noise_into(t, t, NoiseModel.Salt(0.5), 1, False)  # in place
    want = oracle.add_noise32(ramp, 2, 0.5, seed=1, clip=False)
    np.testing.assert_array_equal(t.cpu().numpy(), want)
    with pytest.raises(LA.LucemError) as e:
        LA.add_noise(torch.empty((0, 4), device="cuda"), NoiseModel.Poisson(10.0))
    assert e.value.kind == "EMPTY_INPUT"
