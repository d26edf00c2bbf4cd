"""Host-side mirror of lucem-ops' complex FFT interface (split re/im arrays), over the C ABI.

Reference (neuro-kitchen/lucem): crates/lucem-ops/src/core/fft.rs

- ``fft_1d_into``        core/fft.rs:13-51   the DFT along one axis (negative axes count from the end)
- ``fft_nd_into``        core/fft.rs:53-100  the DFT over every axis
- ``complex_mul_into``   core/fft.rs:102-135 pointwise complex product, optionally with conj(b)

Forward transforms use exp(-2 pi i k m / n) and no scaling; inverse transforms exp(+...) / n.
Any length up to 2^27: one pass for 7-smooth lengths up to 16384 and other lengths up to 8192
(Bluestein), four-step passes for longer lines.
"""

from __future__ import annotations

from . import _native as N


def _arr(t, what):
    t = N.as_tensor(t)
    p = N.device_ptr(t, what) if t.numel() else 0
    return [N._vp(p or None), N.dims(t.shape), t.dim()]


def _client(t, client):
    if client is not None:
        return client
    t = N.as_tensor(t)
    return N.default_client(t.device.index if t.is_cuda else None)


def fft_1d_into(re_in, im_in, re_out, im_out, axis: int = -1, inverse: bool = False,
                client: N.Client | None = None) -> None:
    """The DFT of (re_in, im_in) along ``axis`` into (re_out, im_out) (core/fft.rs:13-51)."""
    args = _arr(re_in, "re_in") + _arr(im_in, "im_in") + _arr(re_out, "re_out") + _arr(im_out, "im_out")
    N.check(N.lib().lucem_fft_1d_into_f32(_client(re_in, client).handle, *args, int(axis), 1 if inverse else 0))


def fft_nd_into(re_in, im_in, re_out, im_out, inverse: bool = False, ws: N.Workspace | None = None,
                client: N.Client | None = None) -> None:
    """The DFT over every axis (core/fft.rs:53-100); outputs may alias inputs for ndim >= 2."""
    args = _arr(re_in, "re_in") + _arr(im_in, "im_in") + _arr(re_out, "re_out") + _arr(im_out, "im_out")
    N.check(N.lib().lucem_fft_nd_into_f32(_client(re_in, client).handle, *args, 1 if inverse else 0,
                                          ws.handle if ws is not None else None))


def fft_nd(re, im, inverse: bool = False, client: N.Client | None = None):
    """Allocating form of :func:`fft_nd_into`: returns ``(re_out, im_out)``."""
    import torch

    re_out, im_out = torch.empty_like(N.as_tensor(re)), torch.empty_like(N.as_tensor(im))
    fft_nd_into(re, im, re_out, im_out, inverse, None, client)
    return re_out, im_out


def complex_mul_into(a_re, a_im, b_re, b_im, out_re, out_im, conj_b: bool = False,
                     client: N.Client | None = None) -> None:
    """out = a * b (or a * conj(b)) pointwise (core/fft.rs:102-135)."""
    a = _arr(a_re, "a_re")
    b = _arr(b_re, "b_re")
    o = _arr(out_re, "out_re")
    ai = _arr(a_im, "a_im")[0]
    bi = _arr(b_im, "b_im")[0]
    oi = _arr(out_im, "out_im")[0]
    N.check(N.lib().lucem_complex_mul_into_f32(_client(a_re, client).handle, *a, ai, *b, bi, *o, oi,
                                               1 if conj_b else 0))
