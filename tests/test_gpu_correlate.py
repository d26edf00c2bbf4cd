"""GPU tests of the correlation entry points (correlate_2d_into, filter/convolve.rs:126-166)
against the oracle's direct correlation: every boundary mode, window origins, Constant mode with
a nonzero cval, weights larger than the input, 3-D arrays, the FP64 direct form, and the
reference's checks.

Tolerances: the reference's own correlation test bound (``assert_close`` with rtol 1e-5 and
atol = rtol * max|want|, crates/lucem-core/src/test_support.rs:42-50, used at
filter/convolve.rs:211-228) and 1e-6 relative L2 for the FFT form in FP32; 1e-12 for FP64."""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MODES = [0, 1, 2, 3, 4]
SCIPY_MODES = {0: "reflect", 1: "constant", 2: "nearest", 3: "mirror", 4: "wrap"}


def _dev(a, dtype=np.float32):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a, dtype=dtype)).cuda()


def assert_close(got, want, rtol):
    got = np.asarray(got, dtype=np.float64)
    want = np.asarray(want, dtype=np.float64)
    atol = rtol * float(np.max(np.abs(want))) if want.size else 0.0
    bad = np.abs(got - want) > rtol * np.abs(want) + atol
    assert not bad.any(), f"{int(bad.sum())} values differ, worst {float(np.max(np.abs(got - want)))}"


def _corr(x, w, mode, cval=0.0, origin=(0, 0), nd=False):
    import lucem_amd as LA

    s = LA.Correlate2d(mode=mode, cval=cval, origin=origin)
    fn = LA.correlate_nd if nd else LA.correlate_2d
    return fn(_dev(x), _dev(w), s).cpu().numpy()


CASES_2D = [
    ((64, 64), (5, 5), (0, 0)),
    ((37, 53), (4, 6), (-2, 2)),      # even windows, origins at both ends of the range
    ((48, 80), (7, 3), (3, -1)),
    ((9, 11), (15, 21), (0, 0)),      # weights larger than the input: folded pads
    ((1, 64), (1, 7), (0, 3)),
    ((64, 1), (9, 1), (-4, 0)),
    ((3, 3), (3, 3), (1, -1)),
    ((125, 98), (25, 25), (-12, 12)),
    ((1024, 1024), (25, 25), (0, 0)),
]


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("shape,kshape,origin", CASES_2D)
def test_correlate_2d_matches_oracle(gpu, oracle, mode, shape, kshape, origin):
    rng = np.random.default_rng(shape[0] * 31 + shape[1] + kshape[0] * 7 + mode)
    x = rng.uniform(-1.0, 1.0, shape).astype(np.float32)
    w = rng.standard_normal(kshape).astype(np.float32)
    cval = 1.5 if mode == 1 else 0.0
    print(f"correlate_2d {shape} weights {kshape} origin {origin} mode {mode} cval {cval}", flush=True)
    got = _corr(x, w, mode, cval, origin)
    want = oracle.correlate(x.astype(np.float64), w.astype(np.float64), mode, cval, origin=origin)
    e = oracle.rel_l2(got, want)
    print(f"  rel-L2 vs FP64 {e:.2e}")
    assert_close(got, want, 1e-5)
    assert e < 1e-6


CASES_3D = [
    ((16, 20, 24), (3, 5, 7), (0, 0, 0)),
    ((9, 12, 10), (4, 3, 6), (-2, 1, 2)),
    ((5, 7, 6), (7, 9, 11), (0, 0, 0)),    # weights larger than the input
    ((64, 96, 80), (15, 7, 7), (3, -3, 0)),
]


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("shape,kshape,origin", CASES_3D)
def test_correlate_nd_matches_oracle(gpu, oracle, mode, shape, kshape, origin):
    rng = np.random.default_rng(sum(shape) * 13 + sum(kshape) + mode)
    x = rng.uniform(-1.0, 1.0, shape).astype(np.float32)
    w = rng.standard_normal(kshape).astype(np.float32)
    cval = -0.75 if mode == 1 else 0.0
    print(f"correlate_nd {shape} weights {kshape} origin {origin} mode {mode} cval {cval}", flush=True)
    got = _corr(x, w, mode, cval, origin, nd=True)
    want = oracle.correlate(x.astype(np.float64), w.astype(np.float64), mode, cval, origin=origin)
    e = oracle.rel_l2(got, want)
    print(f"  rel-L2 vs FP64 {e:.2e}")
    assert_close(got, want, 1e-5)
    assert e < 1e-6


@pytest.mark.parametrize("mode", MODES)
def test_correlate_2d_f64_matches_oracle(gpu, oracle, mode):
    import lucem_amd as LA

    for shape, kshape, origin in [((37, 53), (4, 6), (-2, 2)), ((9, 11), (15, 21), (0, 0)),
                                  ((40, 40), (81, 81), (40, -1))]:  # 81 x 81: the unstaged kernel
        rng = np.random.default_rng(shape[0] + kshape[1] + mode)
        x = rng.uniform(-1.0, 1.0, shape)
        w = rng.standard_normal(kshape)
        cval = 1.5 if mode == 1 else 0.0
        print(f"correlate_2d f64 {shape} weights {kshape} origin {origin} mode {mode}", flush=True)
        got = LA.correlate_2d(_dev(x, np.float64), _dev(w, np.float64),
                              LA.Correlate2d(mode=mode, cval=cval, origin=origin)).cpu().numpy()
        want = oracle.correlate(x, w, mode, cval, origin=origin)
        assert oracle.rel_l2(got, want) < 1e-12


@pytest.mark.parametrize("shape,kshape,origin,mode", [
    ((3, 16400), (3, 7), (0, 2), 1),
    ((16390, 2), (5, 1), (-2, 0), 0),
    ((2, 16380), (1, 9), (0, -4), 4),   # n fits the FFT limit, n + k - 1 does not
])
def test_correlate_longer_than_16384_takes_the_direct_sum(gpu, oracle, shape, kshape, origin, mode):
    rng = np.random.default_rng(shape[0] + shape[1])
    x = rng.uniform(-1.0, 1.0, shape).astype(np.float32)
    w = rng.standard_normal(kshape).astype(np.float32)
    cval = 1.5 if mode == 1 else 0.0
    print(f"correlate_2d direct {shape} weights {kshape} origin {origin} mode {mode}", flush=True)
    got = _corr(x, w, mode, cval, origin)
    want = oracle.correlate(x.astype(np.float64), w.astype(np.float64), mode, cval, origin=origin)
    assert_close(got, want, 1e-5)
    assert oracle.rel_l2(got, want) < 1e-6


def test_correlate_matches_rl_blur(gpu, oracle):
    """The forward blur of the RL iteration (deconvolve.rs:120): correlate(x, flip(psf / sum))."""
    rng = np.random.default_rng(5)
    x = rng.uniform(0.0, 2.0, (200, 150)).astype(np.float32)
    psf = rng.uniform(0.0, 1.0, (11, 9))
    h = psf / psf.sum()
    rev = h[::-1, ::-1].astype(np.float32)
    got = _corr(x, rev, 0)
    want = oracle.correlate_fft64(x, rev.astype(np.float64), 0)
    assert oracle.rel_l2(got, want) < 1e-6


def test_correlate_delta_shifts_exactly(gpu):
    """A delta window at (dy, dx) moves the image by (dy - k // 2, dx - k // 2) (Wrap: circular)."""
    rng = np.random.default_rng(6)
    x = rng.uniform(0.0, 1.0, (256, 320)).astype(np.float32)
    w = np.zeros((9, 7), np.float32)
    w[7, 1] = 1.0
    got = _corr(x, w, 4)
    want = np.roll(x, (-(7 - 4), -(1 - 3)), axis=(0, 1))
    assert np.max(np.abs(got - want)) < 2e-6


def test_correlate_empty_and_errors(gpu):
    import torch

    import lucem_amd as LA

    e = torch.empty((0, 8), device="cuda")
    LA.correlate_2d_into(e, _dev(np.ones((3, 3))), torch.empty((0, 8), device="cuda"))
    x = _dev(np.ones((8, 8)))
    with pytest.raises(LA.LucemError) as ei:
        LA.correlate_2d(x, _dev(np.ones((3, 4))), LA.Correlate2d(origin=(2, 0)))
    assert ei.value.kind == "INVALID_ORIGIN"
    with pytest.raises(LA.LucemError) as ei:
        LA.correlate_2d_into(x, _dev(np.ones((3, 3))), x)
    assert ei.value.kind == "ALIASED_BUFFERS"
    with pytest.raises(LA.LucemError) as ei:
        LA.correlate_2d(_dev(np.ones((2, 8, 8))), _dev(np.ones((3, 3, 3))))
    assert ei.value.kind == "INVALID_DIMENSION"


@pytest.mark.parametrize("mode", MODES)
def test_correlate_small_windows_through_the_fft_passes(gpu, oracle, mode, monkeypatch):
    """Windows up to 13 x 13 take the direct sum by default; LUCEM_CORRELATE_DIRECT_MAX=0 sends
    them through the FFT passes, which must agree as well."""
    monkeypatch.setenv("LUCEM_CORRELATE_DIRECT_MAX", "0")
    for shape, kshape, origin in [((64, 64), (5, 5), (0, 0)), ((37, 53), (4, 6), (-2, 2)), ((3, 3), (3, 3), (1, -1))]:
        rng = np.random.default_rng(shape[0] + kshape[0] + mode)
        x = rng.uniform(-1.0, 1.0, shape).astype(np.float32)
        w = rng.standard_normal(kshape).astype(np.float32)
        cval = 1.5 if mode == 1 else 0.0
        got = _corr(x, w, mode, cval, origin)
        want = oracle.correlate(x.astype(np.float64), w.astype(np.float64), mode, cval, origin=origin)
        assert_close(got, want, 1e-5)
        assert oracle.rel_l2(got, want) < 1e-6


def _host_correlate_1d(x, dims, w, axis, mode, cval, origin):
    """The reference test's host_correlate_1d (filter/convolve.rs:181-209), in FP64."""
    x = np.asarray(x, np.float64).reshape(dims)
    out = np.zeros_like(x)
    n = dims[axis]
    left = len(w) // 2 + origin
    xm = np.moveaxis(x, axis, -1)
    om = np.moveaxis(out, axis, -1)
    for i in range(n):
        acc = np.zeros(xm.shape[:-1])
        for t, wt in enumerate(w):
            p = i - left + t
            from oracle import oracle as O

            q = O.pad_index(p, n, mode)
            acc = acc + wt * (cval if q < 0 else xm[..., q])
        om[..., i] = acc
    return out


@pytest.mark.parametrize("mode", MODES)
def test_correlate_1d_reference_case(gpu, mode):
    """filter/convolve.rs:211-228: dims [1], [7], [32, 33], [5, 1, 9]; every axis; weights
    ([1, -2, 3], origin 0) and ([0.5] * 4, origin -1); cval 1.5; assert_close at 1e-5; in f32 and
    f64 (f64 to 1e-12)."""
    import torch

    import lucem_amd as LA

    for dims in [(1,), (7,), (32, 33), (5, 1, 9)]:
        n = int(np.prod(dims))
        x = np.array([((i * 7919 + 104729) % 1013) * 0.1 - 50.0 for i in range(n)]).reshape(dims)
        for axis in range(len(dims)):
            for w, origin in [([1.0, -2.0, 3.0], 0), ([0.5] * 4, -1)]:
                s = LA.Correlate1d(axis=axis, mode=mode, cval=1.5, origin=origin)
                want = _host_correlate_1d(x, dims, w, axis, mode, 1.5, origin)
                got = LA.correlate_1d(torch.from_numpy(x.astype(np.float32)).cuda(), w, s).cpu().numpy()
                assert_close(got, want, 1e-5)
                got64 = LA.correlate_1d(torch.from_numpy(x).cuda(), w, s).cpu().numpy()
                np.testing.assert_allclose(got64, want, rtol=1e-12, atol=1e-12)


def test_convolve_1d_is_the_mirrored_correlation(gpu, oracle):
    """convolve_1d (convolve.rs:94-111) equals scipy's convolve1d (the same reversal and origin
    rule) on even and odd windows."""
    import torch
    from scipy import ndimage as ndi

    import lucem_amd as LA

    x = np.random.default_rng(3).normal(size=(20, 30))
    for w, origin in [([1.0, 2.0, -1.0], 0), ([0.25, 0.5, 0.75, 1.0], 1), ([1.0, -1.0], 0)]:
        got = LA.convolve_1d(torch.from_numpy(x).cuda(), w, LA.Correlate1d(axis=1, origin=origin)).cpu().numpy()
        want = ndi.convolve1d(x, w, axis=1, mode="reflect", origin=origin)
        np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12)
    with pytest.raises(LA.LucemError):
        LA.correlate_1d(torch.from_numpy(x).cuda(), [1.0, float("nan")])
    with pytest.raises(LA.LucemError) as e:
        LA.correlate_1d(torch.from_numpy(x).cuda(), [1.0], LA.Correlate1d(axis=2))
    assert e.value.kind == "AXIS_OUT_OF_BOUNDS"


@pytest.mark.parametrize("mode", MODES)
def test_correlate_1d_tiled_forms_match_scipy(gpu, mode):
    """The LDS-tiled forms of the 1-D correlation: the last axis with lines longer than 64
    (segments of 256 and a ragged end), strided axes with inner extents of 64 and more (64-wide
    column blocks, ragged in both directions), windows up to and past what the tiles hold (the
    general form takes those), origins and cval, against scipy.ndimage.correlate1d in FP64
    (same origin and mode meanings): f32 to 1e-5 rel L2, f64 to 1e-12."""
    import torch
    from scipy import ndimage as ndi

    import lucem_amd as LA

    rng = np.random.default_rng(11)
    for dims in [(3, 300, 130), (70, 1000), (4, 2500), (5, 300), (2, 65, 64, 3)]:
        x = rng.uniform(-1.0, 1.0, dims)
        for axis in range(len(dims)):
            for size, origin in [(5, 0), (17, -3), (70, 20), (200, 0)]:
                if size > dims[axis] * 4:
                    continue
                w = rng.normal(size=size)
                print(f"correlate_1d {dims} axis {axis} size {size} origin {origin} {mode}", flush=True)
                want = ndi.correlate1d(x, w, axis=axis, mode=SCIPY_MODES[mode], cval=0.5, origin=origin)
                s = LA.Correlate1d(axis=axis, mode=mode, cval=0.5, origin=origin)
                got32 = LA.correlate_1d(torch.from_numpy(x.astype(np.float32)).cuda(),
                                        w.astype(np.float32).tolist(), s).cpu().numpy()
                err = np.linalg.norm(got32 - want) / np.linalg.norm(want)
                assert err < 1e-5, (dims, axis, size, origin, err)
                got64 = LA.correlate_1d(torch.from_numpy(x).cuda(), w.tolist(), s).cpu().numpy()
                np.testing.assert_allclose(got64, want, rtol=1e-12, atol=1e-12 * np.abs(want).max())


@pytest.mark.parametrize("mode", [0, 4])
def test_correlate_1d_large_strided_axis(gpu, mode):
    """A strided axis with enough blocks for the 64-row column tiles (4100 x 2100 along axis 0,
    ragged in both directions), against scipy.ndimage.correlate1d in FP64 to 1e-5 rel L2."""
    import torch
    from scipy import ndimage as ndi

    import lucem_amd as LA

    rng = np.random.default_rng(12)
    x = rng.uniform(-1.0, 1.0, (4100, 2100))
    w = rng.normal(size=17)
    print(f"correlate_1d (4100, 2100) axis 0 size 17 origin -3 mode {mode}", flush=True)
    want = ndi.correlate1d(x, w, axis=0, mode=SCIPY_MODES[mode], cval=0.0, origin=-3)
    s = LA.Correlate1d(axis=0, mode=mode, origin=-3)
    got = LA.correlate_1d(torch.from_numpy(x.astype(np.float32)).cuda(), w.astype(np.float32).tolist(),
                          s).cpu().numpy()
    assert np.linalg.norm(got - want) / np.linalg.norm(want) < 1e-5


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("shape,kshape,mode", [((125, 98), (5, 7), 0), ((48, 80), (3, 3), 3), ((70, 130), (9, 5), 4),
                                               ((17, 64), (4, 6), 1), ((33, 200), (11, 13), 2)])
def test_direct_wide_tiles_equal_square_tiles(gpu, monkeypatch, dtype, shape, kshape, mode):
    """The 64x16 output tiles (four outputs per thread) and the 16x16 ones sum every output in the
    same order: equal bit for bit, on ragged widths, in f32 and f64 (widths of 64 and more take the
    wide tiles; LUCEM_DIRECT_TILE=16 forces the square ones)."""
    import lucem_amd as LA

    rng = np.random.default_rng(sum(shape))
    x = _dev(rng.uniform(-1, 1, shape), dtype)
    w = _dev(rng.uniform(-1, 1, kshape), dtype)
    s = LA.Correlate2d(mode=mode, cval=0.25 if mode == 1 else 0.0)
    print(f"direct tiles: {shape} window {kshape} mode {mode} {np.dtype(dtype).name}", flush=True)
    wide = LA.correlate_2d(x, w, s).cpu().numpy()
    monkeypatch.setenv("LUCEM_DIRECT_TILE", "16")
    square = LA.correlate_2d(x, w, s).cpu().numpy()
    np.testing.assert_array_equal(wide, square)


def test_direct_grid_strides_past_a_million_tiles(gpu, oracle):
    """A 20,000,000 x 1 column has 1.25 million 16x16 output tiles, more than one launch's blocks
    (2^20): the blocks stride over the rest. Checked against the oracle's correlation."""
    import lucem_amd as LA

    rng = np.random.default_rng(9)
    x = rng.uniform(0, 1, (20_000_000, 1))
    w = rng.uniform(0, 1, (5, 1))
    print("direct grid: 20000000 x 1, window 5 x 1, f64", flush=True)
    got = LA.correlate_2d(_dev(x, np.float64), _dev(w, np.float64), LA.Correlate2d()).cpu().numpy()
    want = oracle.correlate(x, w, 0)
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12)
