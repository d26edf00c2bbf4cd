"""GPU tests of the device-side synthetic workloads: the rendered phantoms equal the host
generator's float64 images rounded to float32 bit for bit (same drawn parameters), the device
blur matches the host FP64 blur, and the device Poisson noise has Poisson moments around the
blurred image."""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _ellipse_case(shape, count, seed, axes=(3.0, 20.0)):
    from lucem_amd import synth as S

    return S.ellipse_params(shape, count, np.random.default_rng(seed), axes)


@pytest.mark.parametrize("shape,count,seed", [
    ((1024, 1024), 200, 1000),   # the 2-D workload's truth
    ((511, 300), 96, 7),         # an edge tile of the slide workload
    ((37, 1000), 40, 8),
    ((1, 64), 5, 9),
    ((300, 2), 12, 10),
    ((64, 64), 0, 11),           # background only
    ((16, 16), 300, 12),         # more objects than one chunk of 256, most overlapping
])
def test_render_ellipses_matches_host_bit_for_bit(gpu, shape, count, seed):
    from lucem_amd import phantoms as P
    from lucem_amd import synth as S

    p = _ellipse_case(shape, count, seed)
    print(f"ellipses {shape} count {count}", flush=True)
    got = P.render_ellipses(shape, p, 100.0).cpu().numpy()
    want = S.render_ellipses(shape, p, 100.0).astype(np.float32)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


@pytest.mark.parametrize("shape,count,seed", [
    ((128, 512, 512), 1500, 1001),   # the 3-D workload's truth
    ((40, 60, 50), 200, 3),
    ((5, 9, 13), 40, 4),
    ((1, 32, 17), 10, 5),
    ((8, 8, 8), 600, 6),              # several chunks, heavy overlap
])
def test_render_spheres_matches_host_bit_for_bit(gpu, shape, count, seed):
    from lucem_amd import phantoms as P
    from lucem_amd import synth as S

    p = S.sphere_params(shape, count, np.random.default_rng(seed))
    print(f"spheres {shape} count {count}", flush=True)
    got = P.render_spheres(shape, p, 100.0).cpu().numpy()
    want = S.render_spheres(shape, p, 100.0).astype(np.float32)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def _poisson_moments(y, b):
    """Mean and variance of y - b against Poisson's (mean 0, variance E[b] + 1/12 with the
    rounded normal draw used for lambda >= 20)."""
    d = (y - b).astype(np.float64)
    n = d.size
    lam = float(b.mean())
    assert abs(d.mean()) < 6.0 * np.sqrt(lam / n) + 1e-3
    assert abs(d.var() / (lam + 1.0 / 12.0) - 1.0) < 0.02


def test_device_plain2d_matches_host_generator(gpu, oracle):
    from lucem_amd import phantoms as P
    from lucem_amd import synth as S

    prob = P.plain2d()
    geo = np.random.default_rng(1000)
    truth = S.ellipses((1024, 1024), 200, geo)
    assert np.array_equal(prob.truth.cpu().numpy(), truth.astype(np.float32))
    b_host = S.blur(truth.astype(np.float32), S.gaussian_psf(2.0, (25, 25)))
    b_dev = P.blur(prob.truth, S.gaussian_psf(2.0, (25, 25))).cpu().numpy()
    e = oracle.rel_l2(b_dev, b_host)
    print(f"device blur vs host FP64 blur: rel-L2 {e:.2e}")
    assert e < 1e-6
    y = prob.image.cpu().numpy()
    assert np.all(y >= 0) and np.array_equal(y, np.round(y))
    _poisson_moments(y, b_dev)


def test_device_plain3d_matches_host_generator(gpu, oracle):
    from lucem_amd import phantoms as P
    from lucem_amd import synth as S

    prob = P.plain3d()
    geo = np.random.default_rng(1001)
    truth = S.spheres((128, 512, 512), 1500, geo)
    assert np.array_equal(prob.truth.cpu().numpy(), truth.astype(np.float32))
    psf = S.gaussian_psf((4.0, 1.5, 1.5), (31, 15, 15))
    b_host = S.blur(truth.astype(np.float32), psf)
    b_host *= 1000.0 / b_host.max()
    b_dev = P.blur(prob.truth, psf)
    b_dev.mul_(1000.0 / b_dev.max())
    b_dev = b_dev.cpu().numpy()
    e = oracle.rel_l2(b_dev, b_host)
    print(f"device 3-D blur vs host FP64 blur: rel-L2 {e:.2e}")
    assert e < 1e-6
    y = prob.image.cpu().numpy()
    assert np.all(y >= 0) and np.array_equal(y, np.round(y))
    _poisson_moments(y, b_dev)
