"""GPU tests of the Gaussian filters (filter/gaussian.rs:88-252) against scipy.ndimage's
gaussian_filter1d applied axis by axis in FP64 (the same sampled kernels, boundary rules and pass
order): every mode, derivative orders, radii, axes and channel axes, f32 to 1e-5 rel L2 and f64 to
1e-12, plus the reference's own constant-image case."""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SCIPY_MODE = {"reflect": "reflect", "constant": "constant", "nearest": "nearest", "mirror": "mirror",
              "wrap": "grid-wrap"}


def _want(x, sigmas, orders, radii, axes, mode, cval, truncate):
    from scipy import ndimage as ndi

    y = x.astype(np.float64)
    for ax, s, od, r in zip(axes, sigmas, orders, radii):
        if s > 0:
            y = ndi.gaussian_filter1d(y, s, axis=ax, order=od, mode=SCIPY_MODE[mode], cval=cval,
                                      truncate=truncate, radius=r)
    return y


def _rel(a, b):
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-30))


CASES = [
    # shape, sigmas, orders, radii, axes, channel_axis, mode
    ((17, 21), [1.5], [0], None, None, None, "reflect"),
    ((64, 80), [2.0, 0.8], [0], None, None, None, "constant"),
    ((64, 80), [2.0], [1, 0], None, None, None, "nearest"),
    ((33, 40, 28), [1.2, 2.5, 0.0], [0, 2, 0], None, None, None, "mirror"),
    ((33, 40, 28), [3.0], [0], [2], [2, 0], None, "wrap"),
    ((48, 50, 3), [2.2], [0], None, None, -1, "reflect"),
    ((256, 300), [4.0], [3], None, None, None, "reflect"),
    ((1, 5000), [6.0], [4], None, None, None, "constant"),
]


@pytest.mark.parametrize("shape,sigmas,orders,radii,axes,channel_axis,mode", CASES)
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_gaussian_blur_matches_scipy(gpu, shape, sigmas, orders, radii, axes, channel_axis, mode, dtype):
    import torch

    import lucem_amd as LA

    rng = np.random.default_rng(sum(shape))
    x = rng.uniform(-1.0, 2.0, shape).astype(dtype)
    nd = len(shape)
    ax = list(range(nd)) if axes is None else [a % nd for a in axes]
    if channel_axis is not None:
        ax = [a for a in ax if a != channel_axis % nd]
    k = len(ax)
    sig = sigmas * k if len(sigmas) == 1 else sigmas
    od = orders * k if len(orders) == 1 else orders
    rd = [None] * k if radii is None else (radii * k if len(radii) == 1 else radii)
    print(f"gaussian_blur {shape} {dtype.__name__} sigmas {sig} orders {od} mode {mode}", flush=True)
    s = LA.GaussianBlur(orders=orders, mode=mode, cval=0.75, radii=radii, axes=axes, channel_axis=channel_axis)
    got = LA.gaussian_blur(torch.from_numpy(x).cuda(), sigmas, s).cpu().numpy()
    want = _want(x, sig, od, rd, ax, mode, 0.75, 4.0)
    err = _rel(got.astype(np.float64), want)
    print(f"  rel L2 {err:.3e}")
    assert err < (1e-5 if dtype == np.float32 else 1e-12)


def test_gaussian_1d_and_zero_sigma(gpu):
    import torch

    import lucem_amd as LA

    x = np.random.default_rng(3).normal(size=(40, 70)).astype(np.float32)
    xd = torch.from_numpy(x).cuda()
    got = LA.gaussian_1d(xd, 1.7, LA.Gaussian1d(axis=0, order=2, mode="nearest", truncate=3.0)).cpu().numpy()
    want = _want(x, [1.7], [2], [None], [0], "nearest", 0.0, 3.0)
    assert _rel(got.astype(np.float64), want) < 1e-5
    # sigma 0 copies; all-zero sigmas in the blur copy too
    assert np.array_equal(LA.gaussian_1d(xd, 0.0).cpu().numpy(), x)
    assert np.array_equal(LA.gaussian_blur(xd, [0.0, 0.0]).cpu().numpy(), x)


def test_gaussian_blur_reference_case(gpu):
    """gaussian.rs:448-457: the shape is kept, and a constant image (7.25) stays constant under
    Reflect within 1e-5."""
    import torch

    import lucem_amd as LA

    x = torch.full((17, 21), 7.25, device="cuda")
    out = LA.gaussian_blur(x, [2.0]).cpu().numpy()
    assert out.shape == (17, 21)
    np.testing.assert_allclose(out, 7.25, rtol=1e-5, atol=1e-5)


def test_gaussian_blur_on_own_stream_client(gpu):
    """A client on its own HIP stream: the scratch arrays and copies follow the library's stream,
    so four chained passes read what the previous one wrote."""
    import torch

    import lucem_amd as LA

    x = np.random.default_rng(5).normal(size=(8, 30, 31, 32)).astype(np.float32)
    c = LA.Client(0, own_stream=True)
    got = LA.gaussian_blur(torch.from_numpy(x).cuda(), [1.0, 1.5, 2.0, 2.5], client=c)
    c.sync()
    want = _want(x, [1.0, 1.5, 2.0, 2.5], [0] * 4, [None] * 4, [0, 1, 2, 3], "reflect", 0.0, 4.0)
    assert _rel(got.cpu().numpy().astype(np.float64), want) < 1e-5
