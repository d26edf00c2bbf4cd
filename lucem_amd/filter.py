"""Host-side mirror of lucem-ops' N-D correlation interface, over the C ABI.

Reference (neuro-kitchen/lucem):

- ``Correlate2d`` (defaults)          crates/lucem-ops/src/filter/convolve.rs:113-123
- ``correlate_2d_into``               crates/lucem-ops/src/filter/convolve.rs:126-166
- ``correlate_2d`` (allocating form)  crates/lucem-ops/src/filter/convolve.rs:169-177
- ``check::origin``                   crates/lucem-ops/src/core/check.rs:59-72
- ``Correlate1d``, ``correlate_1d_into``, ``correlate_1d``, ``convolve_1d``
                                      crates/lucem-ops/src/filter/convolve.rs:15-111
- ``fourier_gaussian_into``, ``fourier_box_into``, ``fourier_shift_into``, ``butterworth``
                                      crates/lucem-ops/src/filter/spectral.rs:20-212
- ``gabor_2d`` (``gabor_kernel_2d``)  crates/lucem-ops/src/filter/spectral.rs:216-268

``out[p] = sum_t w[t] * P(in)[p - (k // 2 + origin) + t]`` over the window ``t``, ``P`` the
boundary mode (``cval`` outside the array in Constant mode). float32 tensors run as an FFT
linear correlation through the passes of the Richardson-Lucy path (padded domain ``n + k - 1``
per axis; past 16384 a 2-D array takes the direct sum); float64 tensors run the direct sum in
the reference's order. Errors raise :class:`lucem_amd.LucemError` in the reference's check order.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass

from . import _native as N
from .restore import PadMode


@dataclass
class Correlate2d:
    """Settings of 2-D / N-D correlation (convolve.rs:113-123): boundary mode, cval and the
    window origin per axis of the input."""

    mode: PadMode | str = PadMode.Reflect
    cval: float = 0.0
    origin: tuple = (0, 0)

    def to_c(self) -> N.CorrelateSettings:
        s = N.CorrelateSettings()
        s.mode = int(PadMode.parse(self.mode)) if isinstance(self.mode, str) else int(PadMode(self.mode))
        org = tuple(int(o) for o in self.origin)
        if len(org) > 3:
            raise ValueError("origin has at most 3 entries")
        for i, o in enumerate(org):
            s.origin[i] = o
        s.cval = float(self.cval)
        return s


def _call(fn, input, weights, output, settings, ws, client, dtype=None, with_ws=True):
    x, w, o = N.as_tensor(input), N.as_tensor(weights), N.as_tensor(output)
    if client is None:
        client = N.default_client(x.device.index if x.is_cuda else None)
    if x.numel() and not (x.is_cuda and w.is_cuda and o.is_cuda):
        raise ValueError("input, weights and output must be device tensors")
    xp = N.device_ptr(x, "input", dtype) if x.numel() else 0
    wp = N.device_ptr(w, "weights", dtype) if w.numel() else 0
    op = N.device_ptr(o, "output", dtype) if o.numel() else 0
    st = (settings or Correlate2d()).to_c()
    args = [client.handle, N._vp(xp or None), N.dims(x.shape), x.dim(), N._vp(wp or None), N.dims(w.shape), w.dim(),
            N._vp(op or None), N.dims(o.shape), o.dim(), ctypes.byref(st)]
    if with_ws:
        args.append(ws.handle if ws is not None else None)
    N.check(fn(*args))


def correlate_2d_into(input, weights, output, settings: Correlate2d | None = None, ws: N.Workspace | None = None,
                      client: N.Client | None = None) -> None:
    """Correlates 2-D ``input`` with 2-D ``weights`` into ``output`` (convolve.rs:126-166).

    float32: FFT linear correlation (rounding error about 1e-7 of the output's largest value);
    float64: the direct sum in FP64 (``ws`` unused). Enqueues on the client's stream.
    """
    import torch

    if N.as_tensor(input).dtype == torch.float64:
        _call(N.lib().lucem_correlate_2d_into_f64, input, weights, output, settings, ws, client, torch.float64,
              with_ws=False)
        return
    _call(N.lib().lucem_correlate_2d_into_f32, input, weights, output, settings, ws, client)


def correlate_2d(input, weights, settings: Correlate2d | None = None, client: N.Client | None = None):
    """Allocates and returns the correlation (convolve.rs:169-177)."""
    import torch

    x = N.as_tensor(input)
    out = torch.empty_like(x)
    correlate_2d_into(x, weights, out, settings, None, client)
    return out


def correlate_nd_into(input, weights, output, settings: Correlate2d | None = None, ws: N.Workspace | None = None,
                      client: N.Client | None = None) -> None:
    """The same operator on 2-D or 3-D float32 arrays (weights of the input's rank, origin per
    axis)."""
    _call(N.lib().lucem_correlate_nd_into_f32, input, weights, output, settings, ws, client)


def correlate_nd(input, weights, settings: Correlate2d | None = None, client: N.Client | None = None):
    import torch

    x = N.as_tensor(input)
    out = torch.empty_like(x)
    correlate_nd_into(x, weights, out, settings, None, client)
    return out


def _spectral(fn, re, im, values, client):
    r, i = N.as_tensor(re), N.as_tensor(im)
    if client is None:
        client = N.default_client(r.device.index if r.is_cuda else None)
    vals = [float(v) for v in (values if hasattr(values, "__len__") else [values])]
    arr = (ctypes.c_double * max(len(vals), 1))(*vals)
    rp = N.device_ptr(r, "re") if r.numel() else 0
    ip = N.device_ptr(i, "im") if i.numel() else 0
    N.check(fn(client.handle, N._vp(rp or None), N.dims(r.shape), r.dim(), N._vp(ip or None), N.dims(i.shape), i.dim(),
               arr, len(vals)))


def fourier_gaussian_into(re, im, sigmas, client: N.Client | None = None) -> None:
    """Multiplies the spectrum (re, im) in place by exp(-2 pi^2 sigma_d^2 f_d^2) along each axis
    (spectral.rs:20-62); ``sigmas``: one value or one per axis."""
    _spectral(N.lib().lucem_fourier_gaussian_into_f32, re, im, sigmas, client)


def fourier_box_into(re, im, sizes, client: N.Client | None = None) -> None:
    """Multiplies the spectrum in place by sinc(size_d f_d) along each axis (spectral.rs:64-108)."""
    _spectral(N.lib().lucem_fourier_box_into_f32, re, im, sizes, client)


def fourier_shift_into(re, im, shifts, client: N.Client | None = None) -> None:
    """Multiplies the spectrum in place by exp(-2 pi i shift_d f_d) along each axis
    (spectral.rs:110-165): the image moves by ``shifts`` pixels."""
    _spectral(N.lib().lucem_fourier_shift_into_f32, re, im, shifts, client)


def butterworth(input, cutoff_ratio: float, high_pass: bool, order: float, ws: N.Workspace | None = None,
                client: N.Client | None = None):
    """Butterworth low- or high-pass filter with normalised cutoff in (0, 0.5] (spectral.rs:
    166-212): returns a new float32 device tensor."""
    import torch

    x = N.as_tensor(input)
    if client is None:
        client = N.default_client(x.device.index if x.is_cuda else None)
    out = torch.empty_like(x)
    xp = N.device_ptr(x, "input") if x.numel() else 0
    op = N.device_ptr(out, "output") if out.numel() else 0
    N.check(N.lib().lucem_butterworth_f32(client.handle, N._vp(xp or None), N.dims(x.shape), x.dim(),
                                          N._vp(op or None), N.dims(out.shape), out.dim(), float(cutoff_ratio),
                                          1 if high_pass else 0, float(order), ws.handle if ws is not None else None))
    return out


def gabor_2d(input, frequency: float, theta: float, sigma_x: float, sigma_y: float, n_stds: float = 3.0,
             offset: float = 0.0, mode: PadMode | str = PadMode.Reflect, cval: float = 0.0,
             ws: N.Workspace | None = None, client: N.Client | None = None):
    """The 2-D Gabor filter (spectral.rs:249-268): (real response, imaginary response), the
    correlations of ``input`` with the two parts of gabor_kernel_2d's kernel."""
    import torch

    x = N.as_tensor(input)
    if client is None:
        client = N.default_client(x.device.index if x.is_cuda else None)
    s = N.GaborSettings()
    N.lib().lucem_gabor_settings_default(ctypes.byref(s))
    s.frequency, s.theta, s.sigma_x, s.sigma_y = float(frequency), float(theta), float(sigma_x), float(sigma_y)
    s.n_stds, s.offset, s.cval = float(n_stds), float(offset), float(cval)
    s.mode = int(PadMode.parse(mode)) if isinstance(mode, str) else int(PadMode(mode))
    ore, oim = torch.empty_like(x), torch.empty_like(x)
    xp = N.device_ptr(x, "input") if x.numel() else 0
    N.check(N.lib().lucem_gabor_2d_f32(client.handle, N._vp(xp or None), N.dims(x.shape), x.dim(),
                                       N._vp(ore.data_ptr() if ore.numel() else None),
                                       N._vp(oim.data_ptr() if oim.numel() else None), ctypes.byref(s),
                                       ws.handle if ws is not None else None))
    return ore, oim


@dataclass
class Correlate1d:
    """Settings of 1-D correlation (convolve.rs:15-29): axis, boundary mode, cval, origin."""

    axis: int = -1
    mode: PadMode | str = PadMode.Reflect
    cval: float = 0.0
    origin: int = 0

    def to_c(self) -> N.Correlate1dSettings:
        s = N.Correlate1dSettings()
        N.lib().lucem_correlate1d_settings_default(ctypes.byref(s))
        s.axis = int(self.axis)
        s.mode = int(PadMode.parse(self.mode)) if isinstance(self.mode, str) else int(PadMode(self.mode))
        s.origin = int(self.origin)
        s.cval = float(self.cval)
        return s


def correlate_1d_into(input, weights, output, settings: Correlate1d | None = None,
                      client: N.Client | None = None) -> None:
    """1-D correlation along ``settings.axis`` with device-resident ``weights`` (convolve.rs:40-75);
    float32 or float64 (input, weights and output alike)."""
    import torch

    x, w, o = N.as_tensor(input), N.as_tensor(weights), N.as_tensor(output)
    if client is None:
        client = N.default_client(x.device.index if x.is_cuda else None)
    f64 = x.dtype == torch.float64
    dt = torch.float64 if f64 else None
    xp = N.device_ptr(x, "input", dt) if x.numel() else 0
    wp = N.device_ptr(w, "weights", dt) if w.numel() else 0
    op = N.device_ptr(o, "output", dt) if o.numel() else 0
    fn = N.lib().lucem_correlate_1d_into_f64 if f64 else N.lib().lucem_correlate_1d_into_f32
    st = (settings or Correlate1d()).to_c()
    N.check(fn(client.handle, N._vp(xp or None), N.dims(x.shape), x.dim(), N._vp(wp or None), w.numel(),
               N._vp(op or None), N.dims(o.shape), o.dim(), ctypes.byref(st)))


def _check_weights(w) -> list:
    """check::weights (core/check.rs:74-82): non-empty and finite."""
    import math

    w = [float(v) for v in w]
    if not w:
        raise N.LucemError(4, "weights: filter weights cannot be empty")
    if not all(math.isfinite(v) for v in w):
        raise N.LucemError(4, "weights: filter weights must be finite")
    return w


def correlate_1d(input, weights, settings: Correlate1d | None = None, client: N.Client | None = None):
    """correlate_1d (convolve.rs:77-92): host weights uploaded in the input's precision."""
    import torch

    x = N.as_tensor(input)
    w = torch.as_tensor(_check_weights(weights), dtype=x.dtype, device=x.device)
    out = torch.empty_like(x)
    correlate_1d_into(x, w, out, settings, client)
    return out


def convolve_1d(input, weights, settings: Correlate1d | None = None, client: N.Client | None = None):
    """convolve_1d (convolve.rs:94-111): correlation with the reversed weights and the mirrored
    origin -origin - (1 if the size is even)."""
    s = settings or Correlate1d()
    w = _check_weights(weights)[::-1]
    mirrored = Correlate1d(axis=s.axis, mode=s.mode, cval=s.cval, origin=-s.origin - (1 if len(w) % 2 == 0 else 0))
    return correlate_1d(input, w, mirrored, client)
