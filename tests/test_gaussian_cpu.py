"""CPU tests of the Gaussian filter family's host side (filter/gaussian.rs:18-252): the FP64 taps
of gaussian_kernel_1d against the reference's own assertions and against scipy's Gaussian
derivative kernels (an independent construction), and the argument checks that run before any
device call."""

from __future__ import annotations

import numpy as np
import pytest


def test_kernel_reference_assertions():
    """gaussian.rs:441-446: order 0 sums to 1, order 1 to 0 (sigma 2, radius 8)."""
    from lucem_amd.filter import gaussian_kernel_1d

    assert abs(gaussian_kernel_1d(2.0, 0, 8).sum() - 1.0) < 1e-12
    assert abs(gaussian_kernel_1d(2.0, 1, 8).sum()) < 1e-12
    assert list(gaussian_kernel_1d(1.0, 0, 0)) == [1.0]
    assert list(gaussian_kernel_1d(1.0, 2, 0)) == [0.0]


@pytest.mark.parametrize("sigma,radius", [(0.7, 3), (2.0, 8), (3.3, 13)])
@pytest.mark.parametrize("order", [0, 1, 2, 3, 4, 5])
def test_kernel_matches_scipy_derivatives(sigma, radius, order):
    """scipy's gaussian_filter1d correlates with the sampled normalised Gaussian's order-th
    derivative (its own polynomial recurrence): a unit impulse read back through it gives the
    correlation taps, which must equal ours (odd orders negated for correlation)."""
    from scipy import ndimage as ndi

    from lucem_amd.filter import gaussian_kernel_1d

    n = 2 * radius + 1
    x = np.zeros(3 * n)
    x[n + radius] = 1.0
    resp = ndi.gaussian_filter1d(x, sigma, order=order, mode="constant", radius=radius)
    # correlation with w puts w[t] at output index (centre - t + radius): reverse the window
    taps = resp[n:2 * n][::-1]
    np.testing.assert_allclose(gaussian_kernel_1d(sigma, order, radius), taps, rtol=1e-12,
                               atol=1e-14 * np.abs(taps).max())


def test_resolved_radius():
    """Gaussian1d::resolved_radius (gaussian.rs:83-86): floor(truncate * sigma + 0.5) unless set."""
    import lucem_amd as LA

    assert LA.Gaussian1d().resolved_radius(1.5) == 6
    assert LA.Gaussian1d(truncate=2.0).resolved_radius(1.2) == 2
    assert LA.Gaussian1d(radius=3).resolved_radius(10.0) == 3


def test_argument_checks_before_device_work():
    """The checks of gaussian_1d_into / gaussian_blur_into in the reference's order, on host
    tensors (no device call is reached)."""
    import torch

    import lucem_amd as LA

    x = torch.zeros((4, 5, 6))
    for bad in (-1.0, float("nan"), float("inf")):
        with pytest.raises(LA.LucemError) as e:
            LA.gaussian_1d_into(x, bad, torch.zeros_like(x))
        assert e.value.kind == "INVALID_PARAMETER"
    cases = [
        (1.0, LA.GaussianBlur(axes=[0, 3]), "AXIS_OUT_OF_BOUNDS"),
        (1.0, LA.GaussianBlur(axes=[0, -3]), "INVALID_PARAMETER"),      # repeated axis
        (1.0, LA.GaussianBlur(channel_axis=5), "AXIS_OUT_OF_BOUNDS"),
        ([1.0, 2.0], LA.GaussianBlur(), "INVALID_PARAMETER"),           # 2 sigmas for 3 axes
        ([1.0, -2.0], LA.GaussianBlur(channel_axis=0), "INVALID_PARAMETER"),
        (1.0, LA.GaussianBlur(orders=[0, 1]), "INVALID_PARAMETER"),
        (1.0, LA.GaussianBlur(radii=[1, 2]), "INVALID_PARAMETER"),
    ]
    for sig, s, kind in cases:
        with pytest.raises(LA.LucemError) as e:
            LA.gaussian_blur_into(x, sig, torch.zeros_like(x), s)
        assert e.value.kind == kind, (sig, s, e.value)
    with pytest.raises(LA.LucemError) as e:
        LA.gaussian_blur_into(x, 1.0, torch.zeros((4, 5, 7)))
    assert e.value.kind == "SHAPE_MISMATCH"
    with pytest.raises(LA.LucemError) as e:
        LA.gaussian_blur_into(x, 1.0, x)
    assert e.value.kind == "ALIASED_BUFFERS"
    # empty arrays: nothing to do after the checks
    LA.gaussian_blur_into(torch.zeros((0, 3)), 1.0, torch.zeros((0, 3)))
