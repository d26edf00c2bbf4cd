"""Host-side mirror of lucem-ops' N-D correlation interface, over the C ABI.

Reference (neuro-kitchen/lucem):

- ``Correlate2d`` (defaults)          crates/lucem-ops/src/filter/convolve.rs:113-123
- ``correlate_2d_into``               crates/lucem-ops/src/filter/convolve.rs:126-166
- ``correlate_2d`` (allocating form)  crates/lucem-ops/src/filter/convolve.rs:169-177
- ``check::origin``                   crates/lucem-ops/src/core/check.rs:59-72
- ``Correlate1d``, ``correlate_1d_into``, ``correlate_1d``, ``convolve_1d``
                                      crates/lucem-ops/src/filter/convolve.rs:15-111
- ``gaussian_kernel_1d``, ``Gaussian1d``, ``gaussian_1d(_into)``, ``GaussianBlur``,
  ``gaussian_blur(_into)``            crates/lucem-ops/src/filter/gaussian.rs:18-252 (the
                                      separable passes of core/separable.rs:8-50 on the device's
                                      1-D correlation)
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
from dataclasses import dataclass, field

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
    taps = _check_weights(weights)
    client, ctx = _client_stream(x, client)
    with ctx:  # the weights are freed on the stream that reads them
        w = torch.as_tensor(taps, dtype=x.dtype, device=x.device)
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


def gaussian_kernel_1d(sigma: float, order: int, radius: int):
    """gaussian_kernel_1d (gaussian.rs:18-64) in FP64 on the host: the normalised sampled Gaussian
    of half-width ``radius``, or its ``order``-th derivative through the Hermite recurrence, with
    odd orders negated for the correlation convention."""
    import numpy as np

    if radius == 0:
        return np.array([1.0 if order == 0 else 0.0])
    s2 = sigma * sigma
    x = np.arange(-radius, radius + 1, dtype=np.float64)
    phi = np.exp(-0.5 * x * x / s2)
    phi = phi / phi.sum()
    if order == 0:
        return phi
    if order == 1:
        q = -x / s2
    elif order == 2:
        q = (x * x - s2) / (s2 * s2)
    elif order == 3:
        q = -(x ** 3 - 3.0 * x * s2) / s2 ** 3
    else:
        h_prev, h_cur = np.ones_like(x), -x / s2
        for k in range(1, order):
            h_prev, h_cur = h_cur, (-x / s2) * h_cur - (k / s2) * h_prev
        q = h_cur
    return (-1.0 if order % 2 else 1.0) * q * phi


_GAUSSIAN_TAPS: dict = {}


def _gaussian_taps(sigma: float, order: int, radius: int, like):
    """The taps on the device in ``like``'s precision, kept per device under the reference's key
    (gaussian.rs:106-110: sigma to 12 decimals, order, radius), so that a repeated filter uploads
    nothing and no call frees taps that queued work still reads."""
    import numpy as np
    import torch

    key = (str(like.device), like.dtype, f"gauss1d-{sigma:.12f}-o{order}-r{radius}")
    t = _GAUSSIAN_TAPS.get(key)
    if t is None:
        host = gaussian_kernel_1d(sigma, order, radius)
        np_dt = np.float64 if like.dtype == torch.float64 else np.float32
        t = torch.from_numpy(host.astype(np_dt)).to(like.device)
        _GAUSSIAN_TAPS[key] = t
    return t


@dataclass
class Gaussian1d:
    """Settings of the 1-D Gaussian filter (gaussian.rs:66-87)."""

    axis: int = -1
    order: int = 0
    mode: PadMode | str = PadMode.Reflect
    cval: float = 0.0
    truncate: float = 4.0
    radius: int | None = None

    def resolved_radius(self, sigma: float) -> int:
        import math

        return int(self.radius) if self.radius is not None else max(int(math.floor(self.truncate * sigma + 0.5)), 0)


def gaussian_1d_into(input, sigma: float, output, settings: Gaussian1d | None = None,
                     client: N.Client | None = None) -> None:
    """gaussian_1d_into (gaussian.rs:88-118): sigma finite and >= 0; 0 copies; otherwise the
    correlation with gaussian_kernel_1d along the axis."""
    import math

    import torch

    s = settings or Gaussian1d()
    if not math.isfinite(sigma) or sigma < 0.0:
        raise N.LucemError(5, f"invalid parameter `sigma`: sigma must be non-negative and finite, got {sigma}")
    x, o = N.as_tensor(input), N.as_tensor(output)
    if sigma == 0.0:
        _copy_into(x, o, client)
        return
    wt = _gaussian_taps(sigma, int(s.order), s.resolved_radius(sigma), x)
    correlate_1d_into(x, wt, o, Correlate1d(axis=s.axis, mode=s.mode, cval=s.cval, origin=0), client)


def gaussian_1d(input, sigma: float, settings: Gaussian1d | None = None, client: N.Client | None = None):
    import torch

    x = N.as_tensor(input)
    out = torch.empty_like(x)
    gaussian_1d_into(x, sigma, out, settings, client)
    return out


@dataclass
class GaussianBlur:
    """Settings of the N-D separable Gaussian blur (gaussian.rs:133-171)."""

    orders: list = field(default_factory=lambda: [0])
    mode: PadMode | str = PadMode.Reflect
    cval: float = 0.0
    truncate: float = 4.0
    radii: list | None = None
    axes: list | None = None
    channel_axis: int | None = None


def _client_stream(x, client):
    """A torch stream context on ``client``'s HIP stream, so that the copies and scratch arrays of
    the separable passes are ordered with the library's own work on that stream."""
    import contextlib

    import torch

    if not x.is_cuda:  # host arrays: the library reports them
        return client, contextlib.nullcontext()
    if client is None:
        client = N.default_client(x.device.index)
    if client.stream == torch.cuda.current_stream(x.device).cuda_stream:
        return client, contextlib.nullcontext()  # the default client: already torch's stream
    return client, torch.cuda.stream(torch.cuda.ExternalStream(client.stream, device=x.device))


def _copy_into(x, o, client) -> None:
    """elementwise::copy_into (core/elementwise.rs): same shape, then a device copy."""
    if tuple(x.shape) != tuple(o.shape):
        raise N.LucemError(1, f"input shape {list(x.shape)} does not match output shape {list(o.shape)}")
    if x.numel() == 0:
        return
    N.device_ptr(x, "input", x.dtype)
    N.device_ptr(o, "output", x.dtype)
    client, ctx = _client_stream(x, client)
    with ctx:
        o.copy_(x)


def _broadcast(name: str, values, n: int) -> list:
    """check::broadcast_per_axis (core/check.rs:92-108)."""
    values = list(values)
    if n == 0:
        return []
    if len(values) == 1:
        return values * n
    if len(values) == n:
        return values
    raise N.LucemError(5, f"invalid parameter `{name}`: expected 1 or {n} values, got {len(values)}")


def gaussian_blur_into(input, sigmas, output, settings: GaussianBlur | None = None,
                       client: N.Client | None = None) -> None:
    """gaussian_blur_into (gaussian.rs:174-238): one 1-D Gaussian pass per axis with sigma > 0,
    chained through two scratch arrays as core/separable.rs:8-50 does (0 passes: a copy)."""
    import math

    import torch

    s = settings or GaussianBlur()
    x, o = N.as_tensor(input), N.as_tensor(output)
    nd = x.dim()

    def norm_axis(a):
        # Shape::axis / Shape::axes (lucem-core/src/shape.rs:50-75), numpy's messages
        a = int(a)
        if a < -nd or a >= nd:
            raise N.LucemError(12, f"axis {a} is out of bounds for array of dimension {nd}")
        return a + nd if a < 0 else a

    axes = list(range(nd)) if s.axes is None else []
    if s.axes is not None:
        for a in s.axes:
            na = norm_axis(a)
            if na in axes:
                raise N.LucemError(5, f"repeated axis {a}")
            axes.append(na)
    if s.channel_axis is not None:
        ch = norm_axis(s.channel_axis)
        axes = [a for a in axes if a != ch]
    sig = _broadcast("sigmas", sigmas if hasattr(sigmas, "__len__") else [sigmas], len(axes))
    for v in sig:
        if not math.isfinite(float(v)) or float(v) < 0.0:
            raise N.LucemError(5, f"invalid parameter `sigmas`: sigma must be non-negative and finite, got {v}")
    orders = _broadcast("orders", s.orders, len(axes))
    radii = _broadcast("radii", s.radii, len(axes)) if s.radii is not None else [None] * len(axes)
    passes = [(a, float(v), int(od), r) for a, v, od, r in zip(axes, sig, orders, radii) if float(v) > 0.0]
    if tuple(x.shape) != tuple(o.shape):
        raise N.LucemError(1, f"input shape {list(x.shape)} does not match output shape {list(o.shape)}")
    if x.numel() and x.data_ptr() == o.data_ptr():
        raise N.LucemError(2, "input and output arrays must not share the same device buffer")
    if x.numel() == 0:
        return
    if not passes:
        _copy_into(x, o, client)
        return
    client, ctx = _client_stream(x, client)

    def step(src, dst, p):
        a, v, od, r = p
        gaussian_1d_into(src, v, dst, Gaussian1d(axis=a, order=od, mode=s.mode, cval=s.cval, truncate=s.truncate,
                                                 radius=r), client)

    with ctx:
        bufs = [torch.empty_like(x) for _ in range(min(len(passes) - 1, 2))]
    src = x
    for i, p in enumerate(passes):
        dst = o if i == len(passes) - 1 else bufs[i % 2]
        step(src, dst, p)
        src = dst


def gaussian_blur(input, sigmas, settings: GaussianBlur | None = None, client: N.Client | None = None):
    import torch

    x = N.as_tensor(input)
    out = torch.empty_like(x)
    gaussian_blur_into(x, sigmas, out, settings, client)
    return out
