"""Synthetic Richardson-Lucy workloads generated on the device.

The object parameters are drawn on the host with the same NumPy PCG64 streams as
:mod:`lucem_amd.synth` (a few kilobytes); everything image-sized stays on the device:

- the ground truth is rendered by ``lucem_render_ellipses_f32`` / ``lucem_render_spheres_f32``,
  equal bit for bit to :func:`lucem_amd.synth.render_ellipses` / ``render_spheres`` rounded to
  float32;
- the blur is the reference forward operator, correlate(x, flip(psf / sum(psf))) under Reflect
  (crates/lucem-ops/src/restore/deconvolve.rs:120), as one device correlation
  (:func:`lucem_amd.filter.correlate_nd`, FP32 FFT);
- the noise is the reference's Poisson ``add_noise`` (crates/lucem-ops/src/intensity/noise.rs:
  120-143) in photon units (scale 1), seeded with the workload's noise seed.

The truth matches the host generator exactly and the blur to FP32 rounding; the noise is the
reference's generator, so its draws differ from NumPy's Poisson sampler of the host workloads.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import _native as N
from . import synth as S


@dataclass
class DeviceProblem:
    """One RL problem on the device: observed image, PSF, truth and iteration count."""

    image: object
    psf: object
    num_iter: int
    truth: object
    meta: dict = field(default_factory=dict)


def _upload(rows: np.ndarray):
    import torch

    return torch.from_numpy(np.ascontiguousarray(rows, dtype=np.float64)).cuda()


def render_ellipses(shape, params: dict, background: float = 100.0, client: N.Client | None = None):
    """Background plus the filled ellipses of ``params`` (as drawn by
    :func:`lucem_amd.synth.ellipse_params`), a float32 device tensor of ``shape`` = (h, w)."""
    import torch

    if client is None:
        client = N.default_client(torch.cuda.current_device())
    keys = ("cy", "cx", "a", "b", "cos", "sin", "value")
    rows = np.stack([np.asarray(params[k], dtype=np.float64) for k in keys], axis=1) if len(params["cy"]) else \
        np.zeros((0, 7))
    obj = _upload(rows)
    out = torch.empty(tuple(shape), dtype=torch.float32, device="cuda")
    N.check(N.lib().lucem_render_ellipses_f32(
        client.handle, N._vp(out.data_ptr() if out.numel() else None), N.dims(out.shape), out.dim(),
        N._vp(obj.data_ptr() if obj.numel() else None), len(rows), float(background)))
    return out


def render_spheres(shape, params: dict, background: float = 100.0, client: N.Client | None = None):
    """Background plus the filled spheres of ``params`` (as drawn by
    :func:`lucem_amd.synth.sphere_params`), a float32 device tensor of ``shape`` = (d, h, w)."""
    import torch

    if client is None:
        client = N.default_client(torch.cuda.current_device())
    keys = ("cz", "cy", "cx", "radius", "radius_sq", "value")
    rows = np.stack([np.asarray(params[k], dtype=np.float64) for k in keys], axis=1) if len(params["cz"]) else \
        np.zeros((0, 6))
    obj = _upload(rows)
    out = torch.empty(tuple(shape), dtype=torch.float32, device="cuda")
    N.check(N.lib().lucem_render_spheres_f32(
        client.handle, N._vp(out.data_ptr() if out.numel() else None), N.dims(out.shape), out.dim(),
        N._vp(obj.data_ptr() if obj.numel() else None), len(rows), float(background)))
    return out


def blur(x, psf: np.ndarray):
    """The reference forward blur on the device: correlate(x, flip(psf / sum)) under Reflect."""
    import torch

    from .filter import Correlate2d, correlate_nd

    h = np.asarray(psf, dtype=np.float64)
    h = h / h.sum()
    rev = h[tuple(slice(None, None, -1) for _ in range(h.ndim))]
    w = torch.from_numpy(np.ascontiguousarray(rev, dtype=np.float32)).to(x.device)
    return correlate_nd(x, w, Correlate2d(mode=0, origin=(0,) * h.ndim))


def poisson(blurred, seed: int):
    """Poisson counts of ``blurred`` (photon units) with the reference's add_noise generator."""
    from .noise import NoiseModel, add_noise

    return add_noise(blurred, NoiseModel.Poisson(1.0), seed=seed, clip=False)


def plain2d(size: int = 1024, count: int = 200, sigma: float = 2.0, k: int = 25, num_iter: int = 30,
            client: N.Client | None = None) -> DeviceProblem:
    """The 2-D workload of :func:`lucem_amd.synth.plain2d` (geometry seed 1000, noise seed 2000)."""
    import torch

    geo = np.random.default_rng(1000)
    truth = render_ellipses((size, size), S.ellipse_params((size, size), count, geo), 100.0, client)
    psf = S.gaussian_psf(sigma, (k, k))
    y = poisson(blur(truth, psf), 2000)
    return DeviceProblem(y, torch.from_numpy(psf.astype(np.float32)).cuda(), num_iter, truth,
                         {"workload": f"plain2d {size}x{size} psf {k}x{k} sigma {sigma} (device)"})


def plain3d(shape=(128, 512, 512), count: int = 1500, sigma=(4.0, 1.5, 1.5), k=(31, 15, 15), num_iter: int = 50,
            peak: float = 1000.0, client: N.Client | None = None) -> DeviceProblem:
    """The 3-D workload of :func:`lucem_amd.synth.plain3d` (geometry seed 1001, noise seed 2001):
    the blurred volume is scaled to a peak of ``peak`` photons before the noise."""
    import torch

    geo = np.random.default_rng(1001)
    truth = render_spheres(shape, S.sphere_params(shape, count, geo), 100.0, client)
    psf = S.gaussian_psf(sigma, k)
    b = blur(truth, psf)
    b.mul_(peak / b.max())
    y = poisson(b, 2001)
    return DeviceProblem(y, torch.from_numpy(psf.astype(np.float32)).cuda(), num_iter, truth,
                         {"workload": f"plain3d {shape[0]}x{shape[1]}x{shape[2]} psf {k[0]}x{k[1]}x{k[2]} (device)"})


def psffit3d(count: int = 16, side: int = 256, depth=(32, 128), sigma=(3.5, 1.4, 1.6), k=(31, 15, 15),
             spheres_per_mvox: float = 1500 / 33.554432, client: N.Client | None = None):
    """The volumes of :func:`lucem_amd.synth.psffit3d` (geometry seed 1004: the same depths and
    spheres) rendered, blurred by the true PSF and given Poisson noise on the device (noise seed
    2004, the reference's generator). Returns (volumes, psf) as float32 device tensors."""
    import torch

    geo = np.random.default_rng(1004)
    psf = S.gaussian_psf(sigma, k)
    vols = []
    for i in range(count):
        d = int(geo.integers(depth[0], depth[1] + 1))
        n = max(int(spheres_per_mvox * d * side * side / 1e6), 1)
        truth = render_spheres((d, side, side), S.sphere_params((d, side, side), n, geo), 100.0, client)
        vols.append(poisson(blur(truth, psf), 2004 + i))
    return vols, torch.from_numpy(psf.astype(np.float32)).cuda()
