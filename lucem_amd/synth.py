"""Synthetic Richardson-Lucy workloads (phantoms, Gaussian PSFs, blur, Poisson noise).

Everything is generated on the host with NumPy's PCG64 so the same seed gives the same arrays
on every machine: workload ``c`` uses geometry seed ``1000 + c`` and noise seed ``2000 + c``.
Ground truth is additive (``background + sum of objects``); the blur is the reference forward
operator (correlation with the flipped, normalised PSF under Reflect padding, i.e.
crates/lucem-ops/src/restore/deconvolve.rs:120) evaluated in float64 with FFTs; the noise is
``Poisson(blurred)`` cast to float32 (photon counts).

Workloads (sizes as used by the benchmark and tests):

- ``plain2d``   1024x1024, 200 ellipses (semi-axes U(3,20), intensity U(200,2000)), background
                100, Gaussian PSF sigma 2.0 on 25x25, 30 iterations.
- ``plain3d``   128x512x512 ZYX, 1500 spheres (radius U(2,8), intensity U(100,1000)), background
                100, Gaussian PSF sigma (4.0,1.5,1.5) on 31x15x15, blurred volume scaled to a
                peak of 1000 photons, 50 iterations.
- ``tiles``     5 channels; per channel 48 read tiles of 512x512 and 12 edge tiles of U{64..511}^2
                (300 tiles), ellipse nuclei at about one per 40x40 px, channel PSF sigma
                U(1.2,2.5) with radius floor(4 sigma + 0.5) (crates/lucem-ops/src/filter/
                gaussian.rs:83-85), 20 iterations.
- ``grad2d``    64 tiles of 256x256 drawn from the plain2d law (about 13 ellipses each), PSF
                sigma 1.8 on 17x17, 10 iterations.
- ``psffit3d``  16 volumes of U{32..128} slices x 256x256, spheres, true PSF sigma
                (3.5,1.4,1.6) on 31x15x15.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


def gaussian_psf(sigma, shape) -> np.ndarray:
    """Sampled Gaussian centred at ``k//2`` on each axis, normalised to sum 1 (float64)."""
    sigma = np.broadcast_to(np.asarray(sigma, dtype=np.float64), (len(shape),))
    grids = np.meshgrid(*[np.arange(k, dtype=np.float64) - (k // 2) for k in shape], indexing="ij")
    r2 = sum((g / s) ** 2 for g, s in zip(grids, sigma))
    p = np.exp(-0.5 * r2)
    return p / p.sum()


def _reflect_index(n: int, k: int) -> np.ndarray:
    """Source index of every padded position i - k//2, i < n + k - 1, under Reflect (period 2n)."""
    p = np.arange(n + k - 1) - k // 2
    r = np.mod(p, 2 * n)
    return np.where(r < n, r, 2 * n - 1 - r)


def blur(x: np.ndarray, psf: np.ndarray) -> np.ndarray:
    """The reference forward blur in float64: correlate(x, flip(psf/sum)) with Reflect padding."""
    x = np.asarray(x, dtype=np.float64)
    h = np.asarray(psf, dtype=np.float64)
    h = h / h.sum()
    rev = h[tuple(slice(None, None, -1) for _ in range(h.ndim))]
    q = x
    for ax in range(x.ndim):
        q = np.take(q, _reflect_index(x.shape[ax], h.shape[ax]), axis=ax)
    axes = list(range(x.ndim))
    L = q.shape
    try:
        import scipy.fft as F

        kw = {"workers": -1}
    except ImportError:  # numpy's FFT gives the same result, more slowly
        F, kw = np.fft, {}
    spec = F.rfftn(q, L, axes=axes, **kw) * np.conj(F.rfftn(rev, L, axes=axes, **kw))
    c = F.irfftn(spec, L, axes=axes, **kw)
    return np.ascontiguousarray(c[tuple(slice(0, s) for s in x.shape)])


def ellipse_params(shape, count: int, rng: np.random.Generator, axes=(3.0, 20.0),
                   intensity=(200.0, 2000.0)) -> dict:
    """Random ellipses: centre, semi-axes, angle (with its cosine and sine) and intensity."""
    h, w = shape
    cy = rng.uniform(0, h, count)
    cx = rng.uniform(0, w, count)
    a = rng.uniform(axes[0], axes[1], count)
    b = rng.uniform(axes[0], axes[1], count)
    th = rng.uniform(0, np.pi, count)
    val = rng.uniform(intensity[0], intensity[1], count)
    cos = np.array([np.cos(t) for t in th], dtype=np.float64)
    sin = np.array([np.sin(t) for t in th], dtype=np.float64)
    return {"cy": cy, "cx": cx, "a": a, "b": b, "theta": th, "cos": cos, "sin": sin, "value": val}


def render_ellipses(shape, p: dict, background: float = 100.0) -> np.ndarray:
    """``background`` plus the filled ellipses of ``p`` (float64), in list order."""
    h, w = shape
    img = np.full((h, w), background, dtype=np.float64)
    cy, cx, a, b, val = p["cy"], p["cx"], p["a"], p["b"], p["value"]
    for i in range(len(cy)):
        r = int(np.ceil(max(a[i], b[i]))) + 1
        y0, y1 = max(int(cy[i]) - r, 0), min(int(cy[i]) + r + 1, h)
        x0, x1 = max(int(cx[i]) - r, 0), min(int(cx[i]) + r + 1, w)
        if y0 >= y1 or x0 >= x1:
            continue
        yy, xx = np.mgrid[y0:y1, x0:x1]
        dy, dx = yy - cy[i], xx - cx[i]
        c, s = p["cos"][i], p["sin"][i]
        u = (dx * c + dy * s) / a[i]
        v = (-dx * s + dy * c) / b[i]
        img[y0:y1, x0:x1] += np.where(u * u + v * v <= 1.0, val[i], 0.0)
    return img


def ellipses(shape, count: int, rng: np.random.Generator, axes=(3.0, 20.0), intensity=(200.0, 2000.0),
             background: float = 100.0) -> np.ndarray:
    """``background`` plus ``count`` filled ellipses (random centre, semi-axes, angle, intensity)."""
    return render_ellipses(shape, ellipse_params(shape, count, rng, axes, intensity), background)


def sphere_params(shape, count: int, rng: np.random.Generator, radius=(2.0, 8.0), intensity=(100.0, 1000.0)) -> dict:
    """Random spheres: centre, radius (and its square) and intensity."""
    d, h, w = shape
    cz = rng.uniform(0, d, count)
    cy = rng.uniform(0, h, count)
    cx = rng.uniform(0, w, count)
    rad = rng.uniform(radius[0], radius[1], count)
    val = rng.uniform(intensity[0], intensity[1], count)
    rad_sq = np.array([r ** 2 for r in rad], dtype=np.float64)
    return {"cz": cz, "cy": cy, "cx": cx, "radius": rad, "radius_sq": rad_sq, "value": val}


def render_spheres(shape, p: dict, background: float = 100.0) -> np.ndarray:
    """``background`` plus the filled spheres of ``p`` (float64), in list order."""
    d, h, w = shape
    vol = np.full((d, h, w), background, dtype=np.float64)
    cz, cy, cx, rad, val = p["cz"], p["cy"], p["cx"], p["radius"], p["value"]
    for i in range(len(cz)):
        r = int(np.ceil(rad[i])) + 1
        z0, z1 = max(int(cz[i]) - r, 0), min(int(cz[i]) + r + 1, d)
        y0, y1 = max(int(cy[i]) - r, 0), min(int(cy[i]) + r + 1, h)
        x0, x1 = max(int(cx[i]) - r, 0), min(int(cx[i]) + r + 1, w)
        if z0 >= z1 or y0 >= y1 or x0 >= x1:
            continue
        zz, yy, xx = np.mgrid[z0:z1, y0:y1, x0:x1]
        inside = (zz - cz[i]) ** 2 + (yy - cy[i]) ** 2 + (xx - cx[i]) ** 2 <= p["radius_sq"][i]
        vol[z0:z1, y0:y1, x0:x1] += np.where(inside, val[i], 0.0)
    return vol


def spheres(shape, count: int, rng: np.random.Generator, radius=(2.0, 8.0), intensity=(100.0, 1000.0),
            background: float = 100.0) -> np.ndarray:
    return render_spheres(shape, sphere_params(shape, count, rng, radius, intensity), background)


def poisson(blurred: np.ndarray, rng: np.random.Generator) -> np.ndarray:
    return rng.poisson(np.maximum(blurred, 0.0)).astype(np.float32)


@dataclass
class Problem:
    """One RL problem: observed image, PSF, truth and iteration count."""

    image: np.ndarray
    psf: np.ndarray
    num_iter: int
    truth: np.ndarray | None = None
    meta: dict = field(default_factory=dict)


def plain2d(size: int = 1024, count: int = 200, sigma: float = 2.0, k: int = 25, num_iter: int = 30) -> Problem:
    geo = np.random.default_rng(1000)
    noise = np.random.default_rng(2000)
    truth = ellipses((size, size), count, geo)
    psf = gaussian_psf(sigma, (k, k))
    y = poisson(blur(truth, psf), noise)
    return Problem(y, psf.astype(np.float32), num_iter, truth.astype(np.float32),
                   {"workload": f"plain2d {size}x{size} psf {k}x{k} sigma {sigma}"})


def plain3d(shape=(128, 512, 512), count: int = 1500, sigma=(4.0, 1.5, 1.5), k=(31, 15, 15), num_iter: int = 50,
            peak: float = 1000.0) -> Problem:
    geo = np.random.default_rng(1001)
    noise = np.random.default_rng(2001)
    truth = spheres(shape, count, geo)
    psf = gaussian_psf(sigma, k)
    b = blur(truth, psf)
    b *= peak / b.max()
    y = poisson(b, noise)
    return Problem(y, psf.astype(np.float32), num_iter, truth.astype(np.float32),
                   {"workload": f"plain3d {shape[0]}x{shape[1]}x{shape[2]} psf {k[0]}x{k[1]}x{k[2]} sigma {sigma}"})


@dataclass
class TileSet:
    """Tiles of uneven size with a per-tile PSF index (the batched workload)."""

    tiles: list
    psfs: list
    psf_index: list
    num_iter: int
    meta: dict = field(default_factory=dict)


def tiles(channels: int = 5, full: int = 48, edge: int = 12, side: int = 512, num_iter: int = 20,
          density: float = 1.0 / 1600.0, scale: float = 1.0) -> TileSet:
    """The uneven-tile slide workload. ``scale`` < 1 shrinks tile sides (for quick tests)."""
    geo = np.random.default_rng(1002)
    noise = np.random.default_rng(2002)
    psfs, idx, out = [], [], []
    for c in range(channels):
        sigma = float(geo.uniform(1.2, 2.5))
        r = int(np.floor(4.0 * sigma + 0.5))
        psfs.append(gaussian_psf(sigma, (2 * r + 1, 2 * r + 1)).astype(np.float32))
        shapes = [(side, side)] * full
        lo = max(int(64 * scale), 8)
        hi = max(int(511 * scale), lo + 1)
        shapes += [(int(geo.integers(lo, hi + 1)), int(geo.integers(lo, hi + 1))) for _ in range(edge)]
        for shp in shapes:
            shp = (max(int(shp[0] * (scale if shp == (side, side) else 1.0)), 8),
                   max(int(shp[1] * (scale if shp == (side, side) else 1.0)), 8))
            n = max(int(shp[0] * shp[1] * density), 1)
            truth = ellipses(shp, n, geo)
            out.append(poisson(blur(truth, psfs[c]), noise))
            idx.append(c)
    return TileSet(out, psfs, idx, num_iter, {"workload": f"tiles {len(out)} tiles, {channels} PSFs"})


def grad2d(count: int = 64, side: int = 256, sigma: float = 1.8, k: int = 17, num_iter: int = 10):
    geo = np.random.default_rng(1003)
    noise = np.random.default_rng(2003)
    psf = gaussian_psf(sigma, (k, k))
    ys, truths = [], []
    per = max(int(round(200 * side * side / (1024 * 1024))), 1)
    for _ in range(count):
        t = ellipses((side, side), per, geo)
        truths.append(t.astype(np.float32))
        ys.append(poisson(blur(t, psf), noise))
    return np.stack(ys), psf.astype(np.float32), np.stack(truths), num_iter


def psffit3d(count: int = 16, side: int = 256, depth=(32, 128), sigma=(3.5, 1.4, 1.6), k=(31, 15, 15),
             spheres_per_mvox: float = 1500 / 33.554432, return_truth: bool = False):
    geo = np.random.default_rng(1004)
    noise = np.random.default_rng(2004)
    psf = gaussian_psf(sigma, k)
    vols, truths = [], []
    for _ in range(count):
        d = int(geo.integers(depth[0], depth[1] + 1))
        n = max(int(spheres_per_mvox * d * side * side / 1e6), 1)
        t = spheres((d, side, side), n, geo)
        vols.append(poisson(blur(t, psf), noise))
        if return_truth:
            truths.append(t.astype(np.float32))
    if return_truth:
        return vols, psf.astype(np.float32), truths
    return vols, psf.astype(np.float32)
