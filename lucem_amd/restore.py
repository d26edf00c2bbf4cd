"""Host-side mirror of lucem-ops' Richardson-Lucy interface, over the C ABI.

Reference (neuro-kitchen/lucem):

- ``PadMode`` and ``PadMode::parse``        crates/lucem-ops/src/core/pad.rs:13-57
- ``RichardsonLucy`` (defaults)             crates/lucem-ops/src/restore/deconvolve.rs:19-41
- ``richardson_lucy_into``                  crates/lucem-ops/src/restore/deconvolve.rs:45-149
- ``richardson_lucy`` (allocating form)     crates/lucem-ops/src/restore/deconvolve.rs:152-162

Arrays are float32 device tensors (torch, or any DLPack exporter). Errors raise
:class:`lucem_amd.LucemError` with the status of the reference's check that failed, in the
reference's order (deconvolve.rs:53-71).
"""

from __future__ import annotations

import ctypes
import enum
from dataclasses import dataclass
from typing import Optional

from . import _native as N


class PadMode(enum.IntEnum):
    """Boundary modes; values are the reference's ``PadMode::id`` (core/pad.rs:29-45)."""

    Reflect = 0
    Constant = 1
    Nearest = 2
    Mirror = 3
    Wrap = 4

    @classmethod
    def parse(cls, name: str) -> "PadMode":
        """``PadMode::parse``: names and aliases of core/pad.rs:48-57 (case-insensitive)."""
        out = ctypes.c_uint32()
        N.check(N.lib().lucem_pad_mode_parse(str(name).encode(), ctypes.byref(out)))
        return cls(out.value)

    @property
    def label(self) -> str:
        return self.name.lower()


@dataclass
class RichardsonLucy:
    """Settings of iterative Richardson-Lucy deconvolution (deconvolve.rs:19-41)."""

    num_iter: int = 30
    clip: bool = True
    filter_epsilon: Optional[float] = None
    mode: PadMode = PadMode.Reflect

    def to_c(self) -> N.RlSettings:
        s = N.RlSettings()
        s.num_iter = int(self.num_iter)
        s.clip = 1 if self.clip else 0
        s.has_filter_epsilon = 0 if self.filter_epsilon is None else 1
        s.filter_epsilon = 0.0 if self.filter_epsilon is None else float(self.filter_epsilon)
        s.mode = int(PadMode(self.mode)) if not isinstance(self.mode, str) else int(PadMode.parse(self.mode))
        return s


def _call(fn, input, psf, output, settings: RichardsonLucy, ws, client, dtype=None):
    x = N.as_tensor(input)
    p = N.as_tensor(psf)
    o = N.as_tensor(output)
    if client is None:
        client = N.default_client(x.device.index if x.is_cuda else None)
    st = settings.to_c()
    args_ok = x.numel() == 0 or (x.is_cuda and p.is_cuda and o.is_cuda)
    xp = N.device_ptr(x, "input", dtype) if x.numel() else 0
    pp = N.device_ptr(p, "psf", dtype) if p.numel() else 0
    op = N.device_ptr(o, "output", dtype) if o.numel() else 0
    if not args_ok:
        raise ValueError("input, psf and output must be device tensors")
    N.check(fn(client.handle, N._vp(xp or None), N.dims(x.shape), x.dim(), N._vp(pp or None), N.dims(p.shape),
               p.dim(), N._vp(op or None), N.dims(o.shape), o.dim(), ctypes.byref(st),
               ws.handle if ws is not None else None))


def richardson_lucy_into(input, psf, output, settings: RichardsonLucy | None = None, ws: N.Workspace | None = None,
                         client: N.Client | None = None) -> None:
    """Restores 2-D ``input`` blurred by 2-D ``psf`` into ``output`` (deconvolve.rs:45-149).

    Enqueues on the client's stream (by default torch's current stream) and returns. float32
    tensors take the FFT path; float64 tensors (input, psf and output alike, as ``F`` is one type
    in the reference) take the FP64 direct path (``richardson_lucy_into::<f64>``).
    """
    import torch

    if N.as_tensor(input).dtype == torch.float64:
        _call(N.lib().lucem_richardson_lucy_into_f64, input, psf, output, settings or RichardsonLucy(), ws, client,
              torch.float64)
        return
    _call(N.lib().lucem_richardson_lucy_into_f32, input, psf, output, settings or RichardsonLucy(), ws, client)


def richardson_lucy(input, psf, settings: RichardsonLucy | None = None, client: N.Client | None = None):
    """Allocates and returns the Richardson-Lucy deconvolution of ``input`` (deconvolve.rs:152-162)."""
    import torch

    x = N.as_tensor(input)
    out = torch.empty_like(x)
    richardson_lucy_into(x, psf, out, settings, None, client)
    return out


def richardson_lucy_nd_into(input, psf, output, settings: RichardsonLucy | None = None,
                            ws: N.Workspace | None = None, client: N.Client | None = None) -> None:
    """The same operator on 2-D or 3-D arrays ``[z, y, x]`` (the PSF has the input's rank).

    float64 tensors take the FP64 form, which is 2-D as the reference's operator is: a 3-D float64
    input is refused with ``INVALID_DIMENSION``.
    """
    import torch

    if N.as_tensor(input).dtype == torch.float64:
        _call(N.lib().lucem_richardson_lucy_into_f64, input, psf, output, settings or RichardsonLucy(), ws, client,
              torch.float64)
        return
    _call(N.lib().lucem_richardson_lucy_nd_into_f32, input, psf, output, settings or RichardsonLucy(), ws, client)


def richardson_lucy_nd(input, psf, settings: RichardsonLucy | None = None, client: N.Client | None = None):
    import torch

    x = N.as_tensor(input)
    out = torch.empty_like(x)
    richardson_lucy_nd_into(x, psf, out, settings, None, client)
    return out
