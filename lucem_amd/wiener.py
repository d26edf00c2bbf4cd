"""Host-side mirror of lucem-ops' Wiener deconvolution interface, over the C ABI.

Reference (neuro-kitchen/lucem):

- ``WienerDeconvolve`` (defaults)           crates/lucem-ops/src/restore/deconvolve.rs:164-181
- ``wiener_deconvolve_into``                crates/lucem-ops/src/restore/deconvolve.rs:213-315
- ``wiener_deconvolve`` (allocating form)   crates/lucem-ops/src/restore/deconvolve.rs:318-327
- ``UnsupervisedWiener`` (defaults)         crates/lucem-ops/src/restore/deconvolve.rs:330-353
- ``unsupervised_wiener``                   crates/lucem-ops/src/restore/deconvolve.rs:357-509

Arrays are float32 device tensors. Both forms work on the circular domain of the image: the R2C
passes for 7-smooth sizes up to 16384, the general complex FFT for any other size. Errors raise
:class:`lucem_amd.LucemError` in the reference's check order.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass

from . import _native as N


@dataclass
class WienerDeconvolve:
    """Settings of regularised frequency-domain Wiener deconvolution (deconvolve.rs:164-181)."""

    balance: float = 0.1
    clip: bool = True

    def to_c(self) -> N.WienerSettings:
        s = N.WienerSettings()
        s.balance = float(self.balance)
        s.clip = 1 if self.clip else 0
        return s


@dataclass
class UnsupervisedWiener:
    """Settings of unsupervised Bayesian Wiener deconvolution (deconvolve.rs:330-353)."""

    num_iter: int = 30
    burnin: int = 10
    clip: bool = True
    seed: int = 0

    def to_c(self) -> N.UnsupervisedWienerSettings:
        s = N.UnsupervisedWienerSettings()
        s.num_iter = int(self.num_iter)
        s.burnin = int(self.burnin)
        s.clip = 1 if self.clip else 0
        s.seed = int(self.seed) & 0xFFFFFFFFFFFFFFFF
        return s


def _pointers(x, p, o):
    if x.numel() and not (x.is_cuda and p.is_cuda and o.is_cuda):
        raise ValueError("input, psf and output must be device tensors")
    xp = N.device_ptr(x, "input") if x.numel() else 0
    pp = N.device_ptr(p, "psf") if p.numel() else 0
    op = N.device_ptr(o, "output") if o.numel() else 0
    return N._vp(xp or None), N._vp(pp or None), N._vp(op or None)


def wiener_deconvolve_into(input, psf, output, settings: WienerDeconvolve | None = None,
                           ws: N.Workspace | None = None, client: N.Client | None = None) -> None:
    """Deconvolves 2-D ``input`` with ``psf`` into ``output`` (deconvolve.rs:213-315).

    ``X = conj(H) Y / max(|H|^2 + balance |L|^2, 1e-12)`` with ``L`` the centred 3x3 Laplacian,
    then the inverse transform and the optional clip. Enqueues on the client's stream.
    """
    x, p, o = N.as_tensor(input), N.as_tensor(psf), N.as_tensor(output)
    if client is None:
        client = N.default_client(x.device.index if x.is_cuda else None)
    st = (settings or WienerDeconvolve()).to_c()
    xp, pp, op = _pointers(x, p, o)
    N.check(N.lib().lucem_wiener_deconvolve_into_f32(
        client.handle, xp, N.dims(x.shape), x.dim(), pp, N.dims(p.shape), p.dim(), op, N.dims(o.shape), o.dim(),
        ctypes.byref(st), ws.handle if ws is not None else None))


def wiener_deconvolve(input, psf, settings: WienerDeconvolve | None = None, client: N.Client | None = None):
    """Allocates and returns Wiener-deconvolved ``input`` (deconvolve.rs:318-327)."""
    import torch

    x = N.as_tensor(input)
    out = torch.empty_like(x)
    wiener_deconvolve_into(x, psf, out, settings, None, client)
    return out


def unsupervised_wiener(input, psf, settings: UnsupervisedWiener | None = None, client: N.Client | None = None,
                        ws: N.Workspace | None = None):
    """Unsupervised Wiener deconvolution (deconvolve.rs:357-509): returns ``(image, balance_chain)``.

    The loop runs on the device; the call returns once the chain is on the host, as the
    reference's does.
    """
    import torch

    s = settings or UnsupervisedWiener()
    x, p = N.as_tensor(input), N.as_tensor(psf)
    out = torch.empty_like(x)
    if client is None:
        client = N.default_client(x.device.index if x.is_cuda else None)
    st = s.to_c()
    cap = max(int(s.num_iter), 0)
    chain = (ctypes.c_double * max(cap, 1))()
    got = ctypes.c_uint64(0)
    xp, pp, op = _pointers(x, p, out)
    N.check(N.lib().lucem_unsupervised_wiener_f32(
        client.handle, xp, N.dims(x.shape), x.dim(), pp, N.dims(p.shape), p.dim(), op, ctypes.byref(st), chain,
        cap, ctypes.byref(got), ws.handle if ws is not None else None))
    return out, [float(chain[i]) for i in range(got.value)]
