"""Pins the CPU oracle before anything is compared with it.

Each check is chosen so that a plausible mistake in the oracle (a wrong pad period, a missing
flip, an off-by-one centre, a wrong accumulation order, a dropped term in the gradient) fails:

- pad_index: the reference's own host restatement (crates/lucem-ops/src/test_support.rs:21-45,
  rewritten below) by brute force, and numpy.pad's definition of each mode;
- correlation: scipy.ndimage.correlate (same mode names and centring, k//2 + origin) and the
  reference's correlate_1d KAT (crates/lucem-ops/src/filter/convolve.rs:211-228, 1e-5);
- RL: the reference's identity and box-PSF KATs (crates/lucem-ops/src/restore/deconvolve.rs:517-566),
  the closed-form result for a shifted delta PSF, flat fixed points, FP32 vs FP64, the FFT form;
- gradient: FP64 central finite differences.
"""

from __future__ import annotations

import numpy as np
import pytest
import scipy.ndimage as ndi

MODES = {0: "reflect", 1: "constant", 2: "nearest", 3: "mirror", 4: "wrap"}


def host_pad_index(p: int, n: int, mode: int):
    """The reference's f64 host restatement (test_support.rs:21-45), in Python."""
    if 0 <= p < n:
        return p
    if mode == 1:
        return None
    if mode == 2:
        return min(max(p, 0), n - 1)
    if mode == 4:
        return p % n
    if mode == 0:
        period = 2 * n
        r = p % period
        return r if r < n else period - 1 - r
    if n <= 1:
        return 0
    period = 2 * n - 2
    r = p % period
    return r if r < n else period - r


def sample_f64(n: int) -> np.ndarray:
    """crates/lucem-ops/src/test_support.rs:10-12."""
    i = np.arange(n, dtype=np.int64)
    return ((i * 7919 + 104_729) % 1013).astype(np.float64) * 0.1 - 50.0


def assert_close(got, want, rtol):
    """crates/lucem-core/src/test_support.rs:42-50: |g - w| <= rtol*|w| + rtol*max|want|."""
    got = np.asarray(got, dtype=np.float64).ravel()
    want = np.asarray(want, dtype=np.float64).ravel()
    atol = rtol * np.abs(want).max()
    bad = np.abs(got - want) > rtol * np.abs(want) + atol
    assert not bad.any(), f"index {np.argmax(bad)}: got {got[bad][0]}, expected {want[bad][0]}"


@pytest.mark.parametrize("mode", list(MODES))
def test_pad_index_matches_reference_host_restatement(oracle, mode):
    for n in range(1, 9):
        for p in range(-40, 41):
            want = host_pad_index(p, n, mode)
            assert oracle.pad_index(p, n, mode) == (-1 if want is None else want), (p, n, mode)


@pytest.mark.parametrize("mode,np_mode", [(0, "symmetric"), (2, "edge"), (3, "reflect"), (4, "wrap")])
def test_pad_index_matches_numpy_pad(oracle, mode, np_mode):
    for n in range(2, 7):
        x = np.arange(n)
        padded = np.pad(x, 3 * n, mode=np_mode)
        for i, v in enumerate(padded):
            assert oracle.pad_index(i - 3 * n, n, mode) == v


@pytest.mark.parametrize("mode", list(MODES))
def test_correlate_matches_scipy(oracle, mode):
    rng = np.random.default_rng(10 + mode)
    cases = [((17, 23), (5, 3)), ((7, 9), (4, 6)), ((3, 2), (9, 11)), ((6, 5, 9), (3, 4, 5)), ((1, 1), (3, 3))]
    for shp, ks in cases:
        x = rng.standard_normal(shp)
        w = rng.standard_normal(ks)
        got = oracle.correlate(x, w, mode, 0.0)
        want = ndi.correlate(x, w, mode=MODES[mode], cval=0.0)
        np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12)
        got32 = oracle.correlate(x.astype(np.float32), w.astype(np.float32), mode, 0.0)
        assert_close(got32, want, 1e-5)


@pytest.mark.parametrize("mode", list(MODES))
def test_correlate_origin_and_cval_match_scipy(oracle, mode):
    """The window origin (left = k // 2 + origin, convolve.rs:146-147) and cval in 2-D and 3-D,
    against scipy.ndimage.correlate, whose origin and cval have the same meaning."""
    rng = np.random.default_rng(40 + mode)
    cases = [((9, 11), (4, 5), (-2, 1)), ((9, 11), (4, 5), (1, -2)), ((12, 7), (3, 6), (-1, 2)),
             ((5, 6, 7), (3, 4, 5), (1, -2, 2)), ((4, 9, 6), (2, 3, 4), (-1, 0, -2))]
    for shp, ks, org in cases:
        x = rng.standard_normal(shp)
        w = rng.standard_normal(ks)
        got = oracle.correlate(x, w, mode, 1.5, origin=org)
        want = ndi.correlate(x, w, mode=MODES[mode], cval=1.5, origin=org)
        np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12)


def host_correlate_1d(x, dims, w, axis, mode, cval, origin):
    """filter/convolve.rs:185-209 restated: f64 sums over host_pad_index."""
    x = np.asarray(x, dtype=np.float64).reshape(dims)
    left = len(w) // 2 + origin
    out = np.zeros_like(x)
    n = dims[axis]
    for idx in np.ndindex(*dims):
        s = 0.0
        for t, wt in enumerate(w):
            q = host_pad_index(idx[axis] - left + t, n, mode)
            if q is None:
                v = cval
            else:
                j = list(idx)
                j[axis] = q
                v = x[tuple(j)]
            s += wt * v
        out[idx] = s
    return out


def test_correlate_1d_reference_kat(oracle):
    """The reference KAT at filter/convolve.rs:211-228 (dims, weight sets, origins, cval 1.5)."""
    for dims in [(1,), (7,), (32, 33), (5, 1, 9)]:
        x = sample_f64(int(np.prod(dims)))
        for axis in range(len(dims)):
            for w, origin in [([1.0, -2.0, 3.0], 0), ([0.5] * 4, -1)]:
                for mode in MODES:
                    kshape = [1] * len(dims)
                    kshape[axis] = len(w)
                    org = [0] * len(dims)
                    org[axis] = origin
                    got = oracle.correlate(x.astype(np.float32).reshape(dims),
                                           np.array(w, np.float32).reshape(kshape), mode, 1.5, org)
                    want = host_correlate_1d(x, dims, w, axis, mode, 1.5, origin)
                    assert_close(got, want, 1e-5)


def test_rl_identity_kat(oracle):
    """deconvolve.rs:519-539: delta PSF, 15 iterations, no clip -> the input within 1e-4."""
    img = ((np.arange(64, dtype=np.float32) / 64.0) * 0.8 + 0.1).reshape(8, 8).astype(np.float32)
    delta = np.zeros((3, 3), np.float32)
    delta[1, 1] = 1.0
    st, out = oracle.richardson_lucy(img, delta, 15, clip=False)
    assert st == 0
    assert_close(out, img.astype(np.float64), 1e-4)


def test_rl_box_sharpening_kat(oracle):
    """deconvolve.rs:551-566: RL raises the peak of a box-blurred spot by more than 0.15."""
    spot = np.full((8, 8), 0.1, np.float32)
    spot[4, 4] = 0.9
    box = np.full((3, 3), 1.0 / 9.0, np.float32)
    blurred = oracle.correlate(spot, box, 0, 0.0)
    st, restored = oracle.richardson_lucy(blurred, box, 25, clip=True)
    assert st == 0
    assert restored[4, 4] > blurred[4, 4] + 0.15


@pytest.mark.parametrize("mode", [0, 2, 3, 4])
def test_rl_shifted_delta_closed_form(oracle, mode):
    """A delta PSF at (0, 1) of a 3x3 support: correlate(x, flip(h)) reads x[p + (1, 0)] and
    correlate(r, h) reads r[p - (1, 0)], so RL returns y shifted down one row (row 0 is the
    padded row -1) after any number of iterations. A missing flip shifts the other way."""
    rng = np.random.default_rng(3)
    y = rng.uniform(1.0, 2.0, (9, 7))
    h = np.zeros((3, 3))
    h[0, 1] = 1.0
    for it in (1, 4):
        st, out = oracle.richardson_lucy(y, h, it, clip=False, mode=mode)
        assert st == 0
        want = np.empty_like(y)
        want[1:] = y[:-1]
        want[0] = y[oracle.pad_index(-1, 9, mode)]
        np.testing.assert_allclose(out, want, rtol=1e-12)


@pytest.mark.parametrize("mode", [0, 2, 3, 4])
def test_rl_flat_fixed_point(oracle, mode):
    rng = np.random.default_rng(4)
    psf = rng.uniform(0.1, 1.0, (5, 4))
    y = np.full((11, 13), 37.5)
    st, out = oracle.richardson_lucy(y, psf, 6, clip=False, mode=mode)
    np.testing.assert_allclose(out, y, rtol=1e-12)


def test_rl_f32_close_to_f64_and_fft64(oracle):
    rng = np.random.default_rng(5)
    y = rng.uniform(50, 150, (23, 31))
    psf = rng.uniform(0.0, 1.0, (7, 6))
    for mode in MODES:
        _, a64 = oracle.richardson_lucy(y, psf, 8, clip=False, mode=mode)
        _, a32 = oracle.richardson_lucy(y.astype(np.float32), psf.astype(np.float32), 8, clip=False, mode=mode)
        f64 = oracle.rl_fft64(y, psf, 8, mode=mode)
        assert oracle.rel_l2(f64, a64) < 1e-12
        assert oracle.rel_l2(a32, a64) < 1e-5


def test_rl_3d_fft64_matches_direct(oracle):
    rng = np.random.default_rng(6)
    y = rng.uniform(50, 150, (6, 9, 8))
    psf = rng.uniform(0.0, 1.0, (3, 4, 5))
    for mode in MODES:
        st, a = oracle.richardson_lucy(y, psf, 4, clip=False, mode=mode, nd=True)
        assert st == 0
        assert oracle.rel_l2(oracle.rl_fft64(y, psf, 4, mode=mode), a) < 1e-12


def test_rl_error_order(oracle):
    """The reference check order (deconvolve.rs:53-71)."""
    y = np.ones((4, 5), np.float32)
    psf = np.ones((3, 3), np.float32)
    assert oracle.richardson_lucy(y, psf, 3, out=np.zeros((4, 6), np.float32))[0] == 1
    assert oracle.richardson_lucy(np.ones((2, 3, 4), np.float32), psf, 3)[0] == 3
    assert oracle.richardson_lucy(y, np.ones((3,), np.float32), 3)[0] == 3
    assert oracle.richardson_lucy(y, np.ones((0, 3), np.float32), 3)[0] == 4
    assert oracle.richardson_lucy(y, psf, 0)[0] == 5
    assert oracle.richardson_lucy(np.ones((0, 5), np.float32), psf, 3)[0] == 0
    # shape mismatch is reported before the dimension check
    assert oracle.richardson_lucy(np.ones((2, 3, 4), np.float32), psf, 0, out=np.zeros((2, 3), np.float32))[0] == 1


@pytest.mark.parametrize("mode", list(MODES))
def test_gradient_matches_finite_differences(oracle, mode):
    rng = np.random.default_rng(7 + mode)
    x0 = rng.uniform(0.5, 2.0, (9, 10))
    y = rng.uniform(50, 150, (9, 10))
    psf = rng.uniform(0.1, 1.0, (3, 5))
    _, gx, gp = oracle.rl_nll_grad64(y, x0, psf, 4, mode=mode)
    for idx in [(0, 0), (4, 5), (8, 9), (2, 7)]:
        h = 1e-5
        xp, xm = x0.copy(), x0.copy()
        xp[idx] += h
        xm[idx] -= h
        fd = (oracle.rl_nll64(y, xp, psf, 4, mode=mode) - oracle.rl_nll64(y, xm, psf, 4, mode=mode)) / (2 * h)
        assert abs(fd - gx[idx]) <= 1e-4 * abs(fd) + 1e-7
    for idx in [(0, 0), (1, 2), (2, 4)]:
        h = 1e-6
        pp, pm = psf.copy(), psf.copy()
        pp[idx] += h
        pm[idx] -= h
        fd = (oracle.rl_nll64(y, x0, pp, 4, mode=mode) - oracle.rl_nll64(y, x0, pm, 4, mode=mode)) / (2 * h)
        assert abs(fd - gp[idx]) <= 1e-4 * max(abs(fd), 1e-3)


def test_gradient_3d_matches_finite_differences(oracle):
    rng = np.random.default_rng(9)
    x0 = rng.uniform(0.5, 2.0, (4, 5, 6))
    y = rng.uniform(50, 150, (4, 5, 6))
    psf = rng.uniform(0.1, 1.0, (3, 3, 2))
    _, gx, gp = oracle.rl_nll_grad64(y, x0, psf, 3)
    for idx in [(0, 0, 0), (2, 3, 4), (3, 4, 5)]:
        h = 1e-5
        xp, xm = x0.copy(), x0.copy()
        xp[idx] += h
        xm[idx] -= h
        fd = (oracle.rl_nll64(y, xp, psf, 3) - oracle.rl_nll64(y, xm, psf, 3)) / (2 * h)
        assert abs(fd - gx[idx]) <= 1e-4 * abs(fd) + 1e-7
    idx = (1, 2, 1)
    pp, pm = psf.copy(), psf.copy()
    pp[idx] += 1e-6
    pm[idx] -= 1e-6
    fd = (oracle.rl_nll64(y, x0, pp, 3) - oracle.rl_nll64(y, x0, pm, 3)) / 2e-6
    assert abs(fd - gp[idx]) <= 1e-4 * max(abs(fd), 1e-3)


@pytest.mark.parametrize("mode", range(5))
@pytest.mark.parametrize("shape,k", [((23, 17), (5, 7)), ((9, 11, 8), (3, 5, 3)), ((6, 5), (9, 11))])
def test_fft_gradient_matches_the_direct_gradient(oracle, mode, shape, k):
    """The FFT form of the FP64 gradient oracle (config-size volumes) against the direct form,
    which the finite differences above pin: every mode, 2-D and 3-D, a PSF larger than the image."""
    rng = np.random.default_rng(len(shape) * 10 + mode)
    y = rng.uniform(10, 100, shape)
    x0 = rng.uniform(0.5, 2.0, shape)
    psf = rng.uniform(0.1, 1.0, k)
    l1, g1, p1 = oracle.rl_nll_grad64(y, x0, psf, 4, mode=mode)
    l2, g2, p2 = oracle.rl_nll_grad_fft64(y, x0, psf, 4, mode=mode)
    assert abs(l2 - l1) <= 1e-12 * abs(l1)
    assert abs(oracle.rl_nll_fft64(y, x0, psf, 4, mode=mode) - l1) <= 1e-12 * abs(l1)
    assert oracle.rel_l2(g2, g1) < 1e-12
    assert oracle.rel_l2(p2, p1) < 1e-12


def test_fft_gradient_with_a_target_matches_finite_differences(oracle):
    """The loss against a separate target (the held-out half of thinned counts): d/dx0 and d/dpsf
    of the FFT oracle against FP64 central differences of its own loss, 2-D and 3-D."""
    rng = np.random.default_rng(21)
    for shape, k in (((11, 9), (3, 5)), ((5, 6, 7), (3, 3, 3))):
        y = rng.uniform(10, 100, shape)
        t = rng.uniform(10, 100, shape)
        x0 = rng.uniform(0.5, 2.0, shape)
        psf = rng.uniform(0.1, 1.0, k)
        loss, gx, gp = oracle.rl_nll_grad_fft64(y, x0, psf, 3, target=t)
        assert abs(loss - oracle.rl_nll_fft64(y, x0, psf, 3, target=t)) <= 1e-12 * abs(loss)
        for idx in [(0,) * len(shape), tuple(s // 2 for s in shape)]:
            h = 1e-5
            xp, xm = x0.copy(), x0.copy()
            xp[idx] += h
            xm[idx] -= h
            fd = (oracle.rl_nll_fft64(y, xp, psf, 3, target=t) - oracle.rl_nll_fft64(y, xm, psf, 3, target=t)) / (2 * h)
            assert abs(fd - gx[idx]) <= 1e-5 * abs(fd) + 1e-7
        idx = tuple(s // 2 for s in k)
        pp, pm = psf.copy(), psf.copy()
        pp[idx] += 1e-6
        pm[idx] -= 1e-6
        fd = (oracle.rl_nll_fft64(y, x0, pp, 3, target=t) - oracle.rl_nll_fft64(y, x0, pm, 3, target=t)) / 2e-6
        assert abs(fd - gp[idx]) <= 1e-4 * max(abs(fd), 1e-3)


def test_fft_gradient_planted_mistake_is_caught(oracle):
    """Dropping the fold of the padded positions (the transpose of the boundary gather) in the
    FFT gradient, a plausible mistake, moves d/dx0 far beyond the tolerance above."""
    rng = np.random.default_rng(5)
    y = rng.uniform(10, 100, (12, 10))
    x0 = rng.uniform(0.5, 2.0, (12, 10))
    psf = rng.uniform(0.1, 1.0, (5, 5))
    _, g1, _ = oracle.rl_nll_grad64(y, x0, psf, 3)
    real_fold = oracle._fold64
    try:  # the planted mistake: keep only the centre of the padded array
        oracle._fold64 = lambda q, ns, ks, mode: q[tuple(slice(k // 2, k // 2 + n) for n, k in zip(ns, ks))]
        _, g2, _ = oracle.rl_nll_grad_fft64(y, x0, psf, 3)
    finally:
        oracle._fold64 = real_fold
    assert oracle.rel_l2(g2, g1) > 1e-6


def _two_pixel_closed_form(y, a, b, iters, clip_each=False):
    """RL on a 1 x 2 image, PSF [[a, b]] (a + b = 1), Wrap: with left = 1, correlate(x, w) is
    (w0 x1 + w1 x0, w0 x0 + w1 x1), so u = (b x1 + a x0, b x0 + a x1) for rev = (b, a) and
    c = (a r1 + b r0, a r0 + b r1) for h = (a, b), written out by hand."""
    x0, x1 = 0.5, 0.5
    for _ in range(iters):
        u0, u1 = b * x1 + a * x0, b * x0 + a * x1
        r0, r1 = y[0] / u0, y[1] / u1
        x0, x1 = x0 * (a * r1 + b * r0), x1 * (a * r0 + b * r1)
        if clip_each:
            x0, x1 = min(max(x0, -1.0), 1.0), min(max(x1, -1.0), 1.0)
    return np.array([[x0, x1]])


def test_clip_applies_on_the_last_iteration_only(oracle):
    """deconvolve.rs:134-145 clamps to [-1, 1] only when the iteration is the last. On a two-pixel
    image whose first iterate exceeds 1, clamping every iteration (the plausible mistake) gives
    another first pixel after two iterations; the oracle matches the last-only closed form."""
    y = np.array([[3.0, 0.2]])
    psf = np.array([[0.75, 0.25]])
    want = _two_pixel_closed_form(y[0], 0.75, 0.25, 2)
    st, free = oracle.richardson_lucy(y, psf, 2, clip=False, mode=4)
    assert st == 0
    np.testing.assert_allclose(free, want, rtol=1e-13)
    st, out = oracle.richardson_lucy(y, psf, 2, clip=True, mode=4)
    np.testing.assert_allclose(out, np.clip(want, -1.0, 1.0), rtol=1e-13)
    wrong = _two_pixel_closed_form(y[0], 0.75, 0.25, 2, clip_each=True)
    assert abs(wrong[0, 0] - out[0, 0]) > 0.2  # 0.868 against 0.609


def test_planted_flip_mistake_is_caught(oracle):
    """A restatement that forgets the 180-degree flip (correlating with h in both steps) fails
    the shifted-delta closed form: the pin above is sensitive to that mistake."""
    rng = np.random.default_rng(11)
    y = rng.uniform(1.0, 2.0, (9, 7))
    h = np.zeros((3, 3))
    h[0, 1] = 1.0
    x = np.full_like(y, 0.5)
    for _ in range(2):  # the planted mistake: both correlations use h
        u = oracle.correlate(x, h, 0)
        x = x * oracle.correlate(y / u, h, 0)
    want = np.empty_like(y)
    want[1:] = y[:-1]
    want[0] = y[0]
    assert not np.allclose(x, want, rtol=1e-6)


# ---- Wiener forms (restore/deconvolve.rs:184-509) ------------------------------------------


def circular_blur(x, psf):
    """Brute force: y[i, j] = sum psf[ky, kx] x[(i - (ky - kh//2)) mod H, (j - (kx - kw//2)) mod W],
    the image whose spectrum is H X for the centred PSF."""
    H, W = x.shape
    kh, kw = psf.shape
    y = np.zeros((H, W))
    for ky in range(kh):
        for kx in range(kw):
            y += psf[ky, kx] * np.roll(x, (ky - kh // 2, kx - kw // 2), axis=(0, 1))
    return y


def test_wiener_identity_kat(oracle):
    """deconvolve.rs:542-550: a delta PSF with balance 1e-7 returns the image within 1e-3."""
    img = (np.arange(64) / 64.0 * 0.8 + 0.1).reshape(8, 8)
    delta = np.zeros((3, 3))
    delta[1, 1] = 1.0
    out = oracle.wiener64(img, delta, balance=1e-7, clip=False)
    assert np.max(np.abs(out - img)) < 1e-3
    assert np.max(np.abs(out - img)) < 1e-6  # FP64: only the regulariser's 1e-7 remains


@pytest.mark.parametrize("shape,kshape", [((16, 20), (5, 3)), ((9, 12), (4, 4)), ((7, 5), (7, 5))])
def test_wiener_inverts_a_circular_blur_at_balance_zero(oracle, shape, kshape):
    """With balance 0 the filter is 1 / H: a circularly blurred image comes back exactly. A wrong
    centre, a missing conjugate or a transposed axis breaks this."""
    rng = np.random.default_rng(sum(shape) + kshape[0])
    x = rng.uniform(0.0, 1.0, shape)
    # a dominant centre keeps every |H| >= 1 - (the other taps' sum) > 0: a well-posed inverse
    psf = rng.uniform(0.0, 0.02, kshape)
    psf[kshape[0] // 2, kshape[1] // 2] += 1.0
    assert np.min(np.abs(np.fft.fft2(oracle.centred_kernel64(psf, shape)))) > 0.3
    y = circular_blur(x, psf)
    out = oracle.wiener64(y, psf, balance=0.0, clip=False)
    np.testing.assert_allclose(out, x, atol=1e-9)


def dft_matrix(n, inverse=False):
    k = np.arange(n)
    sign = 1.0 if inverse else -1.0
    m = np.exp(sign * 2j * np.pi * np.outer(k, k) / n)
    return m / n if inverse else m


def test_wiener_matches_dft_matrix_restatement(oracle):
    """The reference's pipeline with explicit DFT matrices per axis (the dft_axis_kernel sum,
    core/kernels/fft.rs:11-47) and its Laplacian built through the same centring."""
    rng = np.random.default_rng(5)
    y = rng.uniform(0, 1, (6, 7))
    psf = rng.uniform(0, 1, (3, 4))
    F0, F1 = dft_matrix(6), dft_matrix(7)
    I0, I1 = dft_matrix(6, True), dft_matrix(7, True)
    fft2 = lambda a: F0 @ a @ F1.T  # noqa: E731
    Hf = fft2(oracle.centred_kernel64(psf, (6, 7)))
    lap = np.array([[0, -1, 0], [-1, 4, -1], [0, -1, 0]], dtype=np.float64)
    Lf = fft2(oracle.centred_kernel64(lap, (6, 7)))
    for balance in (0.1, 3.0):
        X = np.conj(Hf) * fft2(y) / np.maximum(np.abs(Hf) ** 2 + balance * np.abs(Lf) ** 2, 1e-12)
        want = (I0 @ X @ I1.T).real
        np.testing.assert_allclose(oracle.wiener64(y, psf, balance, clip=False), want, atol=1e-12)
    # the closed form of the Laplacian's spectrum: 4 - 2 cos(2 pi u / H) - 2 cos(2 pi v / W)
    u, v = np.meshgrid(np.arange(6), np.arange(7), indexing="ij")
    closed = (4 - 2 * np.cos(2 * np.pi * u / 6) - 2 * np.cos(2 * np.pi * v / 7)) ** 2
    np.testing.assert_allclose(oracle.laplacian_spectrum64((6, 7)), closed, atol=1e-12)
    np.testing.assert_allclose(oracle.laplacian_spectrum64((2, 7)), np.ones((2, 7)))


def conv_matrix(psf, shape):
    """The circular blur by a centred kernel as an N x N matrix (row-major pixels)."""
    n = shape[0] * shape[1]
    A = np.zeros((n, n))
    for j in range(n):
        e = np.zeros(n)
        e[j] = 1.0
        A[:, j] = circular_blur(e.reshape(shape), psf).ravel()
    return A


def test_unsupervised_first_balance_from_image_domain_sums(oracle):
    """The first chain value from sums taken in the image domain: by Parseval, rss = |y - A x|^2
    and reg = |L x|^2 for x the Wiener estimate at balance 0.05, and the trace term is
    tr(A'A (A'A + b L'L)^-1) for the blur and Laplacian matrices."""
    rng = np.random.default_rng(8)
    shape = (6, 5)
    y = rng.uniform(0, 1, shape)
    psf = rng.uniform(0.1, 1.0, (3, 3))
    lap = np.array([[0, -1, 0], [-1, 4, -1], [0, -1, 0]], dtype=np.float64)
    A, Lm = conv_matrix(psf, shape), conv_matrix(lap, shape)
    b0, n = 0.05, float(y.size)
    M = A.T @ A + b0 * Lm.T @ Lm
    x = np.linalg.solve(M, A.T @ y.ravel())
    np.testing.assert_allclose(x.reshape(shape), oracle.wiener64(y, psf, b0, clip=False), atol=1e-10)
    rss = float(np.sum((y.ravel() - A @ x) ** 2))
    reg = float(np.sum((Lm @ x) ** 2))
    trace = float(np.trace(A.T @ A @ np.linalg.inv(M)))
    sn2 = max(rss / max(n - trace, 1.0), 1e-8 * max(float(np.var(y)), 1e-6))
    sx2 = max(reg / max(trace, 1.0), 1e-8)
    jit = oracle.unsupervised_wiener_jitter(3, 4, n)
    want = min(max(sn2 / sx2 * jit[0], 1e-6), 1e3)
    img, chain = oracle.unsupervised_wiener64(y, psf, num_iter=4, burnin=3, clip=False, seed=3)
    assert len(chain) == 4
    assert abs(chain[0] - want) <= 1e-9 * want
    # with one post-burn-in iteration the image is the Wiener estimate at that iteration's balance
    np.testing.assert_allclose(img, oracle.wiener64(y, psf, chain[2], clip=False), atol=1e-10)
    assert all(np.exp(-1 / np.sqrt(n)) <= j <= np.exp(1 / np.sqrt(n)) for j in jit)


@pytest.mark.parametrize("n", [1, 2, 7, 11, 12, 97])
def test_fft_axis64_matches_dft_sum(oracle, n):
    """The reference's DFT (core/kernels/fft.rs:11-47): X_k = sum_m x_m exp(-+2 pi i k m / n),
    1/n on the inverse, evaluated term by term."""
    rng = np.random.default_rng(n)
    re, im = rng.normal(size=(3, n)), rng.normal(size=(3, n))
    for inverse in (False, True):
        sign = 1.0 if inverse else -1.0
        m = np.arange(n)
        W = np.exp(sign * 2j * np.pi * np.outer(m, m) / n) / (n if inverse else 1)
        want = (re + 1j * im) @ W.T
        gr, gi = oracle.fft_axis64(re, im, axes=[1], inverse=inverse)
        np.testing.assert_allclose(gr + 1j * gi, want, atol=1e-10)
    # a round trip and the N-D form as consecutive axes
    x = rng.normal(size=(4, n, 3))
    fr, fi = oracle.fft_axis64(x, np.zeros_like(x))
    br, bi = oracle.fft_axis64(fr, fi, inverse=True)
    np.testing.assert_allclose(br, x, atol=1e-10)
    np.testing.assert_allclose(bi, 0.0, atol=1e-10)


# ---- add_noise (intensity/noise.rs:35-143) -----------------------------------------------


def test_noise_uniform_stream_is_uniform(oracle):
    """The counter hash's uniforms: mean, variance and a 64-bin chi-square over 2^20 counters."""
    u = oracle._uniform32(np.uint32(12345), np.arange(1 << 20, dtype=np.uint32)).astype(np.float64)
    n = u.size
    assert abs(u.mean() - 0.5) < 4 / np.sqrt(12 * n)
    assert abs(u.var() - 1 / 12) < 2e-3
    counts = np.bincount(np.minimum((u * 64).astype(int), 63), minlength=64)
    chi2 = float(((counts - n / 64) ** 2 / (n / 64)).sum())
    assert chi2 < 120  # 63 degrees of freedom: the 99.99th percentile is about 113
    # a permutation of the 32-bit counters: no collisions
    h = oracle.pcg_hash32(np.arange(1 << 20, dtype=np.uint32))
    assert np.unique(h).size == h.size


def test_noise_gaussian_and_speckle_moments(oracle):
    x = np.full((512, 512), 0.5, np.float32)
    g = oracle.add_noise32(x, 0, 0.1, 0.05, seed=7, clip=False) - x
    assert abs(g.mean() - 0.1) < 5e-4 and abs(g.std() - 0.05) < 5e-4
    s = oracle.add_noise32(x, 1, 0.0, 0.2, seed=7, clip=False)
    assert abs((s / x - 1).std() - 0.2) < 2e-3
    # the reference's own check (intensity/mod.rs:52-63): clipped Gaussian noise stays in [0, 1]
    ramp = (np.arange(64, dtype=np.float32) / 63.0).reshape(8, 8)
    y = oracle.add_noise32(ramp, 0, 0.0, 0.05, seed=42, clip=True)
    assert ((y >= 0) & (y <= 1)).all()


@pytest.mark.parametrize("lam", [0.5, 3.0, 12.0, 50.0])
def test_noise_poisson_moments(oracle, lam):
    """Poisson draws at lambda = x * scale: integer photon counts with mean and variance lambda,
    from Knuth's product method below 20 and the rounded normal approximation above."""
    scale = 10.0
    x = np.full(1 << 18, lam / scale, np.float32)
    y = oracle.add_noise32(x, 5, scale, seed=3, clip=False) * np.float32(scale)
    assert np.allclose(y, np.round(y), atol=1e-3)
    assert abs(y.mean() - lam) < 5 * np.sqrt(lam / y.size) + 1e-3
    assert abs(y.var() / lam - 1) < 0.02


def test_noise_impulse_fractions(oracle):
    x = np.full((1024, 1024), 0.5, np.float32)
    salt = oracle.add_noise32(x, 2, 0.1)
    assert set(np.unique(salt)) == {0.5, 1.0} and abs((salt == 1).mean() - 0.1) < 2e-3
    pep = oracle.add_noise32(x, 3, 0.2)
    assert set(np.unique(pep)) == {0.0, 0.5} and abs((pep == 0).mean() - 0.2) < 2e-3
    sp = oracle.add_noise32(x, 4, 0.3, 0.25)
    hit = sp != 0.5
    assert abs(hit.mean() - 0.3) < 2e-3 and abs((sp[hit] == 1).mean() - 0.25) < 5e-3
    # the 64-bit seed is folded: a change in its high half changes the stream
    assert (oracle.add_noise32(x, 2, 0.1, seed=1 << 32) != salt).any()


def _fourier_shift(x, shift):
    """x moved by ``shift`` (pixels per axis) through its spectrum, with the signed frequencies
    of phase.rs:145-158."""
    X = np.fft.fftn(x)
    ph = np.zeros(x.shape)
    for ax, (n, s) in enumerate(zip(x.shape, shift)):
        k = np.arange(n)
        f = np.where(k <= n // 2, k, k - n) / n
        shp = [1] * x.ndim
        shp[ax] = n
        ph = ph + (f * s).reshape(shp)
    return np.fft.ifftn(X * np.exp(-2j * np.pi * ph)).real


def test_phase_correlation_reference_kat(oracle):
    """The reference test (register/phase.rs:262-293): a 16x16 sample_positive_f64 image moved by
    (2.25, -1.5) in the Fourier domain, upsample 20: the shift (-2.25, 1.5) within assert_close
    at rtol 0.1."""
    base = np.array([((i * 6151 + 49999) % 1013) * 0.1 + 0.1 for i in range(256)]).reshape(16, 16)
    moved = _fourier_shift(base, (2.25, -1.5))
    shift, _, _ = oracle.phase_correlation64(base.astype(np.float32), moved.astype(np.float32), 20, True)
    assert_close(np.array(shift), np.array([-2.25, 1.5]), 0.1)


@pytest.mark.parametrize("shape,shift", [((40, 50), (3, -7)), ((33,), (-5,)), ((6, 9, 10), (1, -4, 5))])
def test_phase_correlation_integer_circular_shifts(oracle, shape, shift):
    """A circular shift is recovered exactly (negated, as the signed index k <= n / 2 of
    phase.rs:123-131, so -5 on an axis of 10 reads +5), with zero error and zero phase."""
    x = np.random.default_rng(len(shape)).uniform(size=shape)
    want = []
    for s_, n in zip(shift, shape):
        k = (-s_) % n
        want.append(float(k if k <= n // 2 else k - n))
    for normalize in (True, False):
        got, err, ph = oracle.phase_correlation64(x, np.roll(x, shift, axis=tuple(range(x.ndim))), 1, normalize)
        assert got == want
        assert err < 1e-6 and abs(ph) < 1e-9


@pytest.mark.parametrize("shape,shift,u", [((33, 35), (1.25, -2.75), 4), ((51,), (-3.5,), 2),
                                           ((21, 25), (0.4, -1.6), 5)])
def test_phase_correlation_subpixel_shift_on_the_grid(oracle, shape, shift, u):
    """Odd sizes (no Nyquist bin) and a shift on the 1/u grid: the normalised cross-power spectrum
    is an exact phase ramp, so the upsampled peak is exactly the shift (negated) with |peak| = 1."""
    x = np.random.default_rng(7).uniform(size=shape)
    got, err, ph = oracle.phase_correlation64(x, _fourier_shift(x, shift), u, True)
    np.testing.assert_allclose(got, [-s for s in shift], atol=1e-12)
    assert abs(ph) < 1e-9


@pytest.mark.parametrize("kind", ["gaussian", "box", "shift"])
def test_spectral_filters_match_scipy_fourier(oracle, kind):
    """fourier_gaussian / fourier_uniform / fourier_shift of scipy.ndimage (complex input, the
    same transfer functions) on odd sizes, where the Nyquist sign convention does not arise."""
    rng = np.random.default_rng(3)
    z = rng.normal(size=(15, 21)) + 1j * rng.normal(size=(15, 21))
    vals = {"gaussian": [1.5, 2.5], "box": [3.0, 5.0], "shift": [2.25, -1.5]}[kind]
    fn = {"gaussian": ndi.fourier_gaussian, "box": ndi.fourier_uniform, "shift": ndi.fourier_shift}[kind]
    want = fn(z, vals)
    gr, gi = oracle.spectral_axes64(z.real, z.imag, kind, vals)
    np.testing.assert_allclose(gr + 1j * gi, want, rtol=1e-12, atol=1e-12)


def test_spectral_shift_by_integers_is_a_circular_roll(oracle):
    x = np.random.default_rng(4).normal(size=(16, 20))
    X = np.fft.fft2(x)
    gr, gi = oracle.spectral_axes64(X.real, X.imag, "shift", [3, -5])
    np.testing.assert_allclose(np.fft.ifft2(gr + 1j * gi).real, np.roll(x, (3, -5), axis=(0, 1)), atol=1e-12)


def test_butterworth_closed_forms(oracle):
    """A cosine at frequency f0 is scaled by exactly 1 / (1 + (f0 / c)^(2 n)); the low and high
    passes add up to the input; the low pass keeps the mean."""
    n = 64
    t = np.arange(n)
    for f0, c, order in [(5 / 64, 0.1, 2.0), (12 / 64, 0.25, 3.5), (20 / 64, 0.2, 1.0)]:
        x = np.cos(2 * np.pi * f0 * t)[:, None] * np.ones((1, 8))
        got = oracle.butterworth64(x, c, False, order)
        np.testing.assert_allclose(got, x / (1.0 + (f0 / c) ** (2 * order)), atol=1e-12)
    x = np.random.default_rng(5).normal(size=(30, 41))
    lo = oracle.butterworth64(x, 0.2, False, 2.0)
    hi = oracle.butterworth64(x, 0.2, True, 2.0)
    np.testing.assert_allclose(lo + hi, x, atol=1e-12)
    assert abs(lo.mean() - x.mean()) < 1e-12


def test_gabor_kernel_closed_forms(oracle):
    """gabor_kernel_2d: at frequency 0 and offset 0 the real part is the normalised sampled
    Gaussian (sum near 1 for 4 sigma) and the imaginary part is 0; with offset 0 the real part is
    even and the imaginary part odd; the centre is norm * (cos offset, sin offset); theta = pi/2
    swaps the axes."""
    re, im = oracle.gabor_kernel64(0.0, 0.0, 2.0, 3.0, 4.0, 0.0)
    assert re.shape == (25, 17) and np.all(im == 0.0)
    assert abs(re.sum() - 1.0) < 1e-3
    re, im = oracle.gabor_kernel64(0.15, 0.7, 2.0, 1.5, 3.0, 0.0)
    np.testing.assert_allclose(re, re[::-1, ::-1], atol=1e-15)
    np.testing.assert_allclose(im, -im[::-1, ::-1], atol=1e-15)
    re, im = oracle.gabor_kernel64(0.2, 0.3, 1.0, 2.0, 3.0, 0.9)
    c = (re.shape[0] // 2, re.shape[1] // 2)
    norm = 1.0 / (2.0 * np.pi * 2.0)
    assert abs(re[c] - norm * np.cos(0.9)) < 1e-15 and abs(im[c] - norm * np.sin(0.9)) < 1e-15
    a, _ = oracle.gabor_kernel64(0.0, 0.0, 1.0, 2.5, 3.0, 0.0)
    b, _ = oracle.gabor_kernel64(0.0, np.pi / 2, 1.0, 2.5, 3.0, 0.0)
    np.testing.assert_allclose(a, b.T, atol=1e-12)
