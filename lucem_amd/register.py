"""Host-side mirror of lucem-ops' phase correlation interface, over the C ABI.

Reference (neuro-kitchen/lucem):

- ``PhaseCorrelation`` (defaults)        crates/lucem-ops/src/register/phase.rs:18-36
- ``PhaseCorrelationResult``             crates/lucem-ops/src/register/phase.rs:38-47
- ``phase_correlation``                  crates/lucem-ops/src/register/phase.rs:52-250

Subpixel translation registration by Fourier phase cross-correlation (Guizar-Sicairos, Thurman
and Fienup, 2008) on the device's general FFT. Arrays are float32 device tensors of 1 to 3
dimensions. Errors raise :class:`lucem_amd.LucemError` in the reference's check order.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass

from . import _native as N


@dataclass
class PhaseCorrelation:
    """Settings of phase_correlation (phase.rs:18-36)."""

    upsample_factor: int = 1
    normalize_phase: bool = True

    def to_c(self) -> N.PhaseCorrelationSettings:
        s = N.PhaseCorrelationSettings()
        N.lib().lucem_phase_correlation_settings_default(ctypes.byref(s))
        s.upsample_factor = int(self.upsample_factor)
        s.normalize_phase = 1 if self.normalize_phase else 0
        return s


@dataclass
class PhaseCorrelationResult:
    """The shift that registers ``moving`` onto ``reference``, the normalised RMS error and the
    global phase difference (phase.rs:38-47)."""

    shift: list
    error: float
    phasediff: float


def phase_correlation(reference, moving, settings: PhaseCorrelation | None = None, ws: N.Workspace | None = None,
                      client: N.Client | None = None) -> PhaseCorrelationResult:
    """Estimates the translation registering ``moving`` onto ``reference`` (phase.rs:52-250).
    Waits for the result, as the reference does."""
    if int((settings or PhaseCorrelation()).upsample_factor) < 0:
        raise ValueError("upsample_factor must be >= 0")
    r, m = N.as_tensor(reference), N.as_tensor(moving)
    if client is None:
        client = N.default_client(r.device.index if r.is_cuda else None)
    if r.numel() and not (r.is_cuda and m.is_cuda):
        raise ValueError("reference and moving must be device tensors")
    rp = N.device_ptr(r, "reference") if r.numel() else 0
    mp = N.device_ptr(m, "moving") if m.numel() else 0
    st = (settings or PhaseCorrelation()).to_c()
    out = N.PhaseCorrelationResult()
    N.check(N.lib().lucem_phase_correlation_f32(client.handle, N._vp(rp or None), N.dims(r.shape), r.dim(),
                                                N._vp(mp or None), N.dims(m.shape), m.dim(), ctypes.byref(st),
                                                ctypes.byref(out), ws.handle if ws is not None else None))
    return PhaseCorrelationResult([float(out.shift[i]) for i in range(out.ndim)], float(out.error),
                                  float(out.phasediff))
