"""The overlap-save decomposition of the one-launch tiled iteration, checked in FP64 on the CPU.

Each output tile [t0, t0 + T) (fewer on an axis's last tile) correlates P(x) on
[t0 - 2 left, t0 + T - 1 + 2 right] with the flipped PSF on an L-point circular domain
(L >= T + 2 (k - 1)), forms rel = y / safe(u) on its own rel window, pads rel by gathering the pad
targets from that window, correlates again with the PSF and multiplies x on the tile. Restated here
in NumPy FFTs, the decomposition must reproduce the whole-array FP64 oracle of the reference's RL
(restore/deconvolve.rs:45-149) to rounding wherever the admissibility rule of the host accepts the
grid, in every boundary mode, for odd and even PSF extents; and the rule must refuse the grids
whose folds leave a tile (Wrap across several tiles).
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O


def _pad(q: int, n: int, mode: int) -> int:
    return O.pad_index(q, n, mode)


def _admits(n: int, k: int, T: int, mode: int) -> bool:
    """The host rule: every pad target of every tile's rel window lies in that window."""
    left = k // 2
    for t0 in range(0, n, T):
        m = min(T, n - t0) + k - 1
        q0 = t0 - left
        for i in range(m):
            s = _pad(q0 + i, n, mode)
            if s >= 0 and not (q0 <= s < q0 + m):
                return False
    return True


def _corr_circ(Q: np.ndarray, w: np.ndarray) -> np.ndarray:
    """out[p] = sum_j w[j] Q[(p + j) mod L], both axes."""
    W = np.zeros_like(Q)
    W[: w.shape[0], : w.shape[1]] = w
    return np.real(np.fft.ifft2(np.fft.fft2(Q) * np.conj(np.fft.fft2(W))))


def _gather(src: np.ndarray, y0: int, x0: int, h: int, w: int, mode: int) -> np.ndarray:
    n0, n1 = src.shape
    iy = np.array([_pad(y0 + r, n0, mode) for r in range(h)])
    ix = np.array([_pad(x0 + c, n1, mode) for c in range(w)])
    out = np.zeros((h, w))
    ok = (iy[:, None] >= 0) & (ix[None, :] >= 0)
    out[ok] = src[np.clip(iy, 0, None)[:, None].repeat(w, 1)[ok], np.clip(ix, 0, None)[None, :].repeat(h, 0)[ok]]
    return out


def tiled_rl64(y: np.ndarray, psf: np.ndarray, num_iter: int, mode: int, T, L, eps: float = 1e-12) -> np.ndarray:
    n0, n1 = y.shape
    k0, k1 = psf.shape
    h = psf / psf.sum()
    rev = h[::-1, ::-1]
    l0, l1 = k0 // 2, k1 // 2
    x = np.full(y.shape, 0.5)
    for _ in range(num_iter):
        xn = np.empty_like(x)
        for t0y in range(0, n0, T[0]):
            for t0x in range(0, n1, T[1]):
                u = _corr_circ(_gather(x, t0y - 2 * l0, t0x - 2 * l1, L[0], L[1], mode), rev)
                my, mx = min(T[0], n0 - t0y) + k0 - 1, min(T[1], n1 - t0x) + k1 - 1
                qy0, qx0 = t0y - l0, t0x - l1
                Q2 = np.zeros(L)
                for r in range(my):
                    sy = _pad(qy0 + r, n0, mode)
                    for c in range(mx):
                        sx = _pad(qx0 + c, n1, mode)
                        if sy >= 0 and sx >= 0:
                            uu = u[sy - qy0, sx - qx0]
                            Q2[r, c] = y[sy, sx] / (uu if abs(uu) >= eps else eps)
                cc = _corr_circ(Q2, h)
                hh, ww = min(T[0], n0 - t0y), min(T[1], n1 - t0x)
                xn[t0y:t0y + hh, t0x:t0x + ww] = x[t0y:t0y + hh, t0x:t0x + ww] * cc[:hh, :ww]
        x = xn
    return x


@pytest.mark.parametrize("mode", [0, 1, 2, 3])
@pytest.mark.parametrize("shape,kshape,T", [((33, 47), (5, 3), (14, 10)), ((40, 26), (4, 6), (9, 26)),
                                             ((21, 30), (7, 7), (21, 8)), ((12, 50), (3, 9), (5, 17))])
def test_decomposition_equals_whole_array(mode, shape, kshape, T):
    rng = np.random.default_rng(hash((shape, kshape)) % 2**32)
    y = rng.uniform(50, 150, shape)
    psf = rng.uniform(0.1, 1.0, kshape)
    assert _admits(shape[0], kshape[0], T[0], mode) and _admits(shape[1], kshape[1], T[1], mode)
    L = (T[0] + 2 * (kshape[0] - 1), T[1] + 2 * (kshape[1] - 1))
    got = tiled_rl64(y, psf, 3, mode, T, L)
    st, ref = O.richardson_lucy(y, psf, 3, clip=False, mode=mode)
    assert st == 0
    assert O.rel_l2(got, ref) < 1e-12


def test_longer_domain_changes_nothing():
    """Any L >= T + 2 (k - 1) gives the same values (no wrap reaches the kept outputs)."""
    rng = np.random.default_rng(3)
    y = rng.uniform(50, 150, (30, 30))
    psf = rng.uniform(0.1, 1.0, (5, 5))
    a = tiled_rl64(y, psf, 2, 0, (12, 12), (20, 20))
    b = tiled_rl64(y, psf, 2, 0, (12, 12), (32, 27))
    assert O.rel_l2(a, b) < 1e-13


def test_too_short_domain_wraps():
    """A domain one sample shorter than T + 2 (k - 1) lets the circular wrap reach the tile."""
    rng = np.random.default_rng(4)
    y = rng.uniform(50, 150, (30, 30))
    psf = rng.uniform(0.1, 1.0, (5, 5))
    _, ref = O.richardson_lucy(y, psf, 2, clip=False, mode=0)
    bad = tiled_rl64(y, psf, 2, 0, (12, 12), (19, 20))
    assert O.rel_l2(bad, ref) > 1e-6


def test_rule_refuses_wrap_across_tiles():
    """Wrap folds a left-edge window onto the right edge: refused on several tiles, admitted on one."""
    assert not _admits(40, 5, 10, 4)
    assert _admits(40, 5, 40, 4)
    for mode in (0, 1, 2, 3):
        assert _admits(512, 21, 88, mode)
