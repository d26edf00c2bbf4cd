"""CPU oracle for lucem-ops Richardson-Lucy deconvolution (ctypes over ``liblucem_oracle.so``).

TEST INFRASTRUCTURE ONLY: ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg
of ``bench.py`` load this module as the checker. The product path never imports it.

Functions and the reference code each one restates (paths relative to the reference tree):

- :func:`pad_index`           crates/lucem-ops/src/core/kernels/pad.rs:11-34
- :func:`correlate`           crates/lucem-ops/src/filter/kernels/convolve.rs:61-118 and
                              filter/convolve.rs:125-166 (N-D extension for 3-D)
- :func:`richardson_lucy`     crates/lucem-ops/src/restore/deconvolve.rs:45-149 with
                              restore/kernels/deconvolve.rs:7-43 (FP32 or FP64)
- :func:`rl_fft64`            the same RL in FP64 with each correlation done as an FFT linear
                              correlation of the padded array (for shapes too large for the direct
                              FP64 oracle); pinned against the direct FP64 oracle on small shapes
- :func:`rl_nll_grad64`       FP64 reverse-mode gradient of unrolled RL + Poisson NLL (no
                              reference counterpart); pinned by central finite differences
- :func:`rl_nll_grad_fft64`, :func:`rl_nll_fft64`
                              the same gradient and loss with FFT correlations (config-size
                              volumes); pinned against rl_nll_grad64 on small shapes
- :func:`fft_axis64`          crates/lucem-ops/src/core/fft.rs:13-100 (the DFT of
                              core/kernels/fft.rs:11-47) in FP64 through numpy's FFT
- :func:`add_noise32`         crates/lucem-ops/src/intensity/noise.rs:35-143 with
                              intensity/kernels/noise.rs:1-135 and core/kernels/random.rs:7-12 in
                              numpy float32 / uint32
- :func:`wiener64`            crates/lucem-ops/src/restore/deconvolve.rs:184-315 with
                              restore/kernels/deconvolve.rs:44-99 in FP64, numpy's FFT in place of
                              the reference's O(N^2) DFT (core/fft.rs:55-100)
- :func:`unsupervised_wiener64`
                              restore/deconvolve.rs:357-509 in FP64 over the full spectrum
- :func:`phase_correlation64` crates/lucem-ops/src/register/phase.rs:52-250 in FP64, numpy's FFT
                              in place of the reference's DFT
- :func:`spectral_axes64`, :func:`butterworth64`, :func:`gabor_kernel64`
                              crates/lucem-ops/src/filter/spectral.rs:20-246 in FP64

Parity pins (see tests/test_oracle.py): pad_index against the reference's own host restatement
(crates/lucem-ops/src/test_support.rs:21-45, restated in the test) and numpy.pad; correlation
against scipy.ndimage.correlate for every mode and the reference KAT pattern
(filter/convolve.rs:211-228); RL against the reference's identity and box-PSF KATs
(restore/deconvolve.rs:517-566), closed-form fixed points, brute-force FP64 and the FFT form.
Wiener against the reference's identity KAT (deconvolve.rs:542-550), exact inversion of a
brute-force circular blur at balance 0, and a literal restatement with DFT matrices; the
unsupervised chain's first value against Parseval-equivalent sums taken in the image domain.
"""

from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
_LIB_PATH = _HERE / "liblucem_oracle.so"
_lib = None

PAD_REFLECT, PAD_CONSTANT, PAD_NEAREST, PAD_MIRROR, PAD_WRAP = 0, 1, 2, 3, 4

_i64p = ctypes.POINTER(ctypes.c_int64)
_f32p = ctypes.POINTER(ctypes.c_float)
_f64p = ctypes.POINTER(ctypes.c_double)


def build(force: bool = False) -> Path:
    """Compiles the oracle with ``make`` (g++), in place."""
    if force or not _LIB_PATH.exists():
        subprocess.run(["make", "-s", "-C", str(_HERE), "liblucem_oracle.so"], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not _LIB_PATH.exists():
            build()
        L = ctypes.CDLL(str(_LIB_PATH))
        L.lo_pad_index.restype = ctypes.c_int64
        L.lo_pad_index.argtypes = [ctypes.c_int64, ctypes.c_int64, ctypes.c_int]
        L.lo_num_threads.restype = ctypes.c_int
        L.lo_set_num_threads.argtypes = [ctypes.c_int]
        L.lo_correlate_f32.argtypes = [_f32p, _i64p, ctypes.c_int, _f32p, _i64p, _i64p, ctypes.c_int,
                                       ctypes.c_float, _f32p]
        L.lo_correlate_f64.argtypes = [_f64p, _i64p, ctypes.c_int, _f64p, _i64p, _i64p, ctypes.c_int,
                                       ctypes.c_double, _f64p]
        rl_args = [_i64p, ctypes.c_int, None, _i64p, ctypes.c_int, None, _i64p, ctypes.c_int, ctypes.c_uint64,
                   ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int]
        L.lo_rl_f32.restype = ctypes.c_int
        L.lo_rl_f32.argtypes = [_f32p] + [(_f32p if a is None else a) for a in rl_args]
        L.lo_rl_f64.restype = ctypes.c_int
        L.lo_rl_f64.argtypes = [_f64p] + [(_f64p if a is None else a) for a in rl_args]
        L.lo_rl_nll_grad_f64.restype = ctypes.c_double
        L.lo_rl_nll_grad_f64.argtypes = [_f64p, _f64p, _i64p, ctypes.c_int, _f64p, _i64p, ctypes.c_uint64,
                                         ctypes.c_double, ctypes.c_int, _f64p, _f64p]
        L.lo_rl_nll_f64.restype = ctypes.c_double
        L.lo_rl_nll_f64.argtypes = [_f64p, _f64p, _i64p, ctypes.c_int, _f64p, _i64p, ctypes.c_uint64,
                                    ctypes.c_double, ctypes.c_int]
        _lib = L
    return _lib


def set_num_threads(n: int) -> None:
    lib().lo_set_num_threads(int(n))


def num_threads() -> int:
    return int(lib().lo_num_threads())


def _dims(shape) -> ctypes.Array:
    return (ctypes.c_int64 * max(len(shape), 1))(*[int(s) for s in shape])


def _ptr(a: np.ndarray, kind):
    return a.ctypes.data_as(kind)


def pad_index(p: int, n: int, mode: int) -> int:
    """Reference ``pad_index`` (core/kernels/pad.rs:11-34); -1 for Constant outside ``[0, n)``."""
    return int(lib().lo_pad_index(int(p), int(n), int(mode)))


def correlate(x: np.ndarray, w: np.ndarray, mode: int = PAD_REFLECT, cval: float = 0.0, origin=None) -> np.ndarray:
    """Direct correlation ``out[p] = sum_k w[k] * P_mode(x)[p - (k_dim//2 + origin) + k]``.

    FP32 when ``x`` is float32 (sequential FP32 accumulation, as the reference kernel), else FP64.
    """
    f64 = x.dtype == np.float64
    dt = np.float64 if f64 else np.float32
    x = np.ascontiguousarray(x, dtype=dt)
    w = np.ascontiguousarray(w, dtype=dt)
    assert x.ndim == w.ndim and 1 <= x.ndim <= 3
    out = np.empty_like(x)
    org = None if origin is None else _dims(origin)
    if f64:
        lib().lo_correlate_f64(_ptr(x, _f64p), _dims(x.shape), x.ndim, _ptr(w, _f64p), _dims(w.shape), org, mode,
                               float(cval), _ptr(out, _f64p))
    else:
        lib().lo_correlate_f32(_ptr(x, _f32p), _dims(x.shape), x.ndim, _ptr(w, _f32p), _dims(w.shape), org, mode,
                               float(cval), _ptr(out, _f32p))
    return out


def richardson_lucy(image: np.ndarray, psf: np.ndarray, num_iter: int = 30, clip: bool = True,
                    filter_epsilon: float | None = None, mode: int = PAD_REFLECT, nd: bool = False,
                    out: np.ndarray | None = None) -> tuple[int, np.ndarray]:
    """Reference ``richardson_lucy_into`` restated on the CPU. Returns ``(status, output)``.

    Status codes follow the reference's check order (deconvolve.rs:53-71). ``nd=True`` allows the
    3-D extension; otherwise ndim must be 2 as in the reference. Precision follows ``image.dtype``.
    """
    f64 = image.dtype == np.float64
    dt = np.float64 if f64 else np.float32
    image = np.ascontiguousarray(image, dtype=dt)
    psf = np.ascontiguousarray(psf, dtype=dt)
    if out is None:
        out = np.zeros_like(image)
    kind = _f64p if f64 else _f32p
    fn = lib().lo_rl_f64 if f64 else lib().lo_rl_f32
    st = fn(_ptr(image, kind), _dims(image.shape), image.ndim, _ptr(psf, kind), _dims(psf.shape), psf.ndim,
            _ptr(out, kind), _dims(out.shape), out.ndim, int(num_iter), int(bool(clip)),
            int(filter_epsilon is not None), float(filter_epsilon or 0.0), int(mode), 3 if nd else 2)
    return int(st), out


def _normalised(psf: np.ndarray):
    psf = np.asarray(psf, dtype=np.float64)
    s = float(np.sum(psf.ravel()))
    norm = 1.0 / s if abs(s) > 1e-12 else 1.0
    h = psf * norm
    rev = h[tuple(slice(None, None, -1) for _ in range(h.ndim))]
    return h, rev


def _pad_indices(n: int, k: int, mode: int) -> np.ndarray:
    return np.array([pad_index(i - k // 2, n, mode) for i in range(n + k - 1)], dtype=np.int64)


def correlate_fft64(x: np.ndarray, w: np.ndarray, mode: int = PAD_REFLECT) -> np.ndarray:
    """FP64 correlation (cval 0, origin 0) as an FFT linear correlation of the padded array."""
    x = np.asarray(x, dtype=np.float64)
    w = np.asarray(w, dtype=np.float64)
    q = x
    for ax in range(x.ndim):
        idx = _pad_indices(x.shape[ax], w.shape[ax], mode)
        valid = idx >= 0
        q = np.take(q, np.where(valid, idx, 0), axis=ax)
        if not valid.all():
            sl = [slice(None)] * x.ndim
            sl[ax] = ~valid
            q[tuple(sl)] = 0.0
    L = q.shape
    axes = list(range(x.ndim))
    try:  # scipy's FFT runs on every core; numpy's gives the same FP64 result on one
        import scipy.fft as F

        kw = {"workers": -1}
    except ImportError:
        F, kw = np.fft, {}
    Q = F.rfftn(q, L, axes=axes, **kw)
    W = F.rfftn(w, L, axes=axes, **kw)
    c = F.irfftn(Q * np.conj(W), L, axes=axes, **kw)
    return np.ascontiguousarray(c[tuple(slice(0, s) for s in x.shape)])


def rl_fft64(image: np.ndarray, psf: np.ndarray, num_iter: int, mode: int = PAD_REFLECT,
             filter_epsilon: float | None = None, x0: np.ndarray | None = None) -> np.ndarray:
    """FP64 RL with the reference's semantics (0.5 start, eps rule, no clip), FFT correlations."""
    y = np.asarray(image, dtype=np.float64)
    h, rev = _normalised(psf)
    eps = max(1e-12 if filter_epsilon is None else float(filter_epsilon), 1e-15)
    x = np.full_like(y, 0.5) if x0 is None else np.array(x0, dtype=np.float64)
    for _ in range(int(num_iter)):
        u = correlate_fft64(x, rev, mode)
        r = y / np.where(np.abs(u) < eps, eps, u)
        x = x * correlate_fft64(r, h, mode)
    return x


def _fft_mod():
    try:  # scipy's FFT runs on every core; numpy's gives the same FP64 result on one
        import scipy.fft as F

        return F, {"workers": -1}
    except ImportError:
        return np.fft, {}


def _gather64(x: np.ndarray, ks, mode: int) -> np.ndarray:
    """The padded array Q[i] = x[pad(i - k // 2)] on [0, n + k - 1) per axis (0 outside in Constant
    mode): the values correlate_2d_kernel reads (filter/kernels/convolve.rs:99-115)."""
    q = x
    for ax in range(x.ndim):
        idx = _pad_indices(x.shape[ax], ks[ax], mode)
        valid = idx >= 0
        q = np.take(q, np.where(valid, idx, 0), axis=ax)
        if not valid.all():
            sl = [slice(None)] * x.ndim
            sl[ax] = ~valid
            q[tuple(sl)] = 0.0
    return q


def _fold64(qbar: np.ndarray, ns, ks, mode: int) -> np.ndarray:
    """Transpose of :func:`_gather64`: every padded position adds into the source it was read from."""
    g = qbar
    for ax in range(g.ndim):
        idx = _pad_indices(ns[ax], ks[ax], mode)
        valid = np.nonzero(idx >= 0)[0]
        m = np.zeros((ns[ax], len(idx)))
        m[idx[valid], valid] = 1.0
        g = np.moveaxis(np.tensordot(m, g, axes=([1], [ax])), 0, ax)
    return g


class _Fft64:
    """FFT linear correlations of one geometry (array dims ns, window dims ks) on L >= n + k - 1."""

    def __init__(self, ns, ks):
        self.F, self.kw = _fft_mod()
        self.ns, self.ks = tuple(ns), tuple(ks)
        self.L = tuple(self.F.next_fast_len(n + k - 1) if hasattr(self.F, "next_fast_len") else n + k - 1
                       for n, k in zip(ns, ks))
        self.axes = list(range(len(ns)))

    def spec(self, a):
        return self.F.rfftn(a, self.L, axes=self.axes, **self.kw)

    def back(self, sp, shape):
        c = self.F.irfftn(sp, self.L, axes=self.axes, **self.kw)
        return np.ascontiguousarray(c[tuple(slice(0, s) for s in shape)])

    def corr(self, q, w):
        """out[p] = sum_j w[j] q[p + j], p < n (q padded, w of dims ks)."""
        return self.back(self.spec(q) * np.conj(self.spec(w)), self.ns)

    def corr_t(self, obar, w):
        """Transpose of corr in q: qbar[i] = sum_p obar[p] w[i - p], i < n + k - 1."""
        return self.back(self.spec(obar) * self.spec(w), tuple(n + k - 1 for n, k in zip(self.ns, self.ks)))

    def corr_w(self, q, obar):
        """Transpose of corr in w: wbar[j] = sum_p obar[p] q[p + j], j < k."""
        return self.back(self.spec(q) * np.conj(self.spec(obar)), self.ks)


def rl_nll_fft64(y: np.ndarray, x0: np.ndarray, psf: np.ndarray, num_iter: int, eps: float = 1e-12,
                 mode: int = PAD_REFLECT, target: np.ndarray | None = None) -> float:
    """FP64 loss ``sum(u - t log u)`` after ``num_iter`` RL iterations on ``y`` from ``x0`` (FFT
    correlations); ``t`` is ``target`` when given, else ``y``."""
    y = np.asarray(y, dtype=np.float64)
    t = y if target is None else np.asarray(target, dtype=np.float64)
    x = np.array(x0, dtype=np.float64)
    h, rev = _normalised(psf)
    f = _Fft64(y.shape, h.shape)
    for _ in range(int(num_iter)):
        u = f.corr(_gather64(x, h.shape, mode), rev)
        r = y / np.where(np.abs(u) < eps, eps, u)
        x = x * f.corr(_gather64(r, h.shape, mode), h)
    u = f.corr(_gather64(x, h.shape, mode), rev)
    return float(np.sum(u - t * np.log(u)))


def rl_nll_grad_fft64(y: np.ndarray, x0: np.ndarray, psf: np.ndarray, num_iter: int, eps: float = 1e-12,
                      mode: int = PAD_REFLECT, target: np.ndarray | None = None):
    """:func:`rl_nll_grad64` with FFT correlations (for volumes the direct FP64 oracle cannot
    finish): the forward keeps every iteration's padded estimate and ratio, the reverse sweep
    applies the transposes of the gather (fold), the correlation in its input (full convolution)
    and in its window, then the chain rule through ``h = psf / sum(psf)`` (deconvolve.rs:77-79).
    Pinned against the direct :func:`rl_nll_grad64` on small 2-D and 3-D shapes in every mode.
    ``target``: the counts the loss compares ``u`` with, in place of ``y`` (only the seed of the
    reverse sweep changes, ``1 - target / u``); pinned by central differences."""
    y = np.asarray(y, dtype=np.float64)
    tgt = y if target is None else np.asarray(target, dtype=np.float64)
    x = np.array(x0, dtype=np.float64)
    psf = np.asarray(psf, dtype=np.float64)
    ssum = float(np.sum(psf.ravel()))
    norm = 1.0 / ssum if abs(ssum) > 1e-12 else 1.0
    h, rev = _normalised(psf)
    ks, ns = h.shape, y.shape
    f = _Fft64(ns, ks)
    tape = []
    for _ in range(int(num_iter)):
        qx = _gather64(x, ks, mode)
        u = f.corr(qx, rev)
        small = np.abs(u) < eps
        safe = np.where(small, eps, u)
        r = y / safe
        qr = _gather64(r, ks, mode)
        c = f.corr(qr, h)
        tape.append((x, qx, safe, small, qr, c))
        x = x * c
    qx = _gather64(x, ks, mode)
    u = f.corr(qx, rev)
    loss = float(np.sum(u - tgt * np.log(u)))
    ubar = 1.0 - tgt / u
    grev = f.corr_w(qx, ubar)
    gh = np.zeros_like(h)
    xbar = _fold64(f.corr_t(ubar, rev), ns, ks, mode)
    for xt, qxt, safe, small, qrt, ct in reversed(tape):
        cbar = xbar * xt
        xnew = xbar * ct
        gh += f.corr_w(qrt, cbar)
        rbar = _fold64(f.corr_t(cbar, h), ns, ks, mode)
        ubar_t = np.where(small, 0.0, -rbar * y / (safe * safe))
        grev += f.corr_w(qxt, ubar_t)
        xbar = xnew + _fold64(f.corr_t(ubar_t, rev), ns, ks, mode)
    g = gh + grev[tuple(slice(None, None, -1) for _ in range(grev.ndim))]
    gpsf = (g - float(np.sum(g * h))) * norm if abs(ssum) > 1e-12 else g
    return loss, xbar, gpsf


def rl_nll_grad64(y: np.ndarray, x0: np.ndarray, psf: np.ndarray, num_iter: int, eps: float = 1e-12,
                  mode: int = PAD_REFLECT):
    """FP64 loss and gradients ``(loss, d/dx0, d/dpsf)`` of ``sum(u - y log u)``, ``u = A x_T``."""
    y = np.ascontiguousarray(y, dtype=np.float64)
    x0 = np.ascontiguousarray(x0, dtype=np.float64)
    psf = np.ascontiguousarray(psf, dtype=np.float64)
    gx = np.zeros_like(x0)
    gp = np.zeros_like(psf)
    loss = lib().lo_rl_nll_grad_f64(_ptr(y, _f64p), _ptr(x0, _f64p), _dims(y.shape), y.ndim, _ptr(psf, _f64p),
                                    _dims(psf.shape), int(num_iter), float(eps), int(mode), _ptr(gx, _f64p),
                                    _ptr(gp, _f64p))
    return float(loss), gx, gp


def rl_nll64(y: np.ndarray, x0: np.ndarray, psf: np.ndarray, num_iter: int, eps: float = 1e-12,
             mode: int = PAD_REFLECT) -> float:
    """FP64 loss ``sum(u - y log u)`` after ``num_iter`` RL iterations from ``x0``."""
    y = np.ascontiguousarray(y, dtype=np.float64)
    x0 = np.ascontiguousarray(x0, dtype=np.float64)
    psf = np.ascontiguousarray(psf, dtype=np.float64)
    return float(lib().lo_rl_nll_f64(_ptr(y, _f64p), _ptr(x0, _f64p), _dims(y.shape), y.ndim, _ptr(psf, _f64p),
                                     _dims(psf.shape), int(num_iter), float(eps), int(mode)))


def fft_axis64(re: np.ndarray, im: np.ndarray, axes=None, inverse: bool = False):
    """fft_1d_into along each of ``axes`` in turn (all axes: fft_nd_into), FP64: forward
    exp(-2 pi i k m / n) unscaled, inverse exp(+...) / n. Returns ``(re, im)``."""
    z = np.asarray(re, dtype=np.float64) + 1j * np.asarray(im, dtype=np.float64)
    for ax in (range(z.ndim) if axes is None else axes):
        z = np.fft.ifft(z, axis=ax) if inverse else np.fft.fft(z, axis=ax)
    return z.real.copy(), z.imag.copy()


def _signed_freqs(n: int) -> np.ndarray:
    """phase.rs:145-158: k / n for k <= n / 2, else (k - n) / n."""
    k = np.arange(n)
    return np.where(k <= n // 2, k, k - n).astype(np.float64) / n


def phase_correlation64(reference: np.ndarray, moving: np.ndarray, upsample_factor: int = 1,
                        normalize_phase: bool = True):
    """phase_correlation (register/phase.rs:52-250) in FP64. Returns ``(shift, error, phasediff)``.

    The cross-power spectrum A conj(B) (normalised by its magnitude above 1e-12, or 0), its
    inverse transform, the first index of the largest |cc|^2 as the integer shift (signed), then
    for ``upsample_factor > 1`` two coordinate sweeps of the local inverse DFT of the cross-power
    spectrum (radius ceil(1.5 u) at step 1/u, then radius 2 at step 0.25/u, each axis in turn,
    keeping the first strictly larger |sum|^2), the shift snapped to the 1/u grid; the phase of
    the peak; the normalised RMS error from the powers of both images and the unnormalised
    correlation at the peak index.
    """
    ref = np.asarray(reference, dtype=np.float64)
    mov = np.asarray(moving, dtype=np.float64)
    dims = ref.shape
    ndim = ref.ndim
    n_total = ref.size
    A = np.fft.fftn(ref)
    B = np.fft.fftn(mov)
    prod = A * np.conj(B)
    if normalize_phase:
        mag = np.abs(prod)
        prod = np.where(mag > 1e-12, prod / np.where(mag > 1e-12, mag, 1.0), 0.0)
    cc = np.fft.ifftn(prod)
    max_idx = int(np.argmax((cc.real ** 2 + cc.imag ** 2).ravel()))
    coords = np.unravel_index(max_idx, dims)
    shift = [float(k) if k <= n // 2 else float(k - n) for k, n in zip(coords, dims)]
    peak = complex(cc.ravel()[max_idx])
    if upsample_factor > 1:
        u = float(upsample_factor)
        half_span = int(np.ceil(1.5 * u))
        freqs = [_signed_freqs(n) for n in dims]
        grids = np.meshgrid(*freqs, indexing="ij")
        pr = prod.ravel()
        for pas in range(2):
            radius = half_span if pas == 0 else 2
            step = 1.0 / u if pas == 0 else 0.25 / u
            for ax in range(ndim):
                other = np.zeros(dims)
                for d in range(ndim):
                    if d != ax:
                        other = other + 2.0 * np.pi * grids[d] * shift[d]
                oc, os_ = np.cos(other).ravel(), np.sin(other).ravel()
                center = shift[ax]
                best = (-1.0, center, peak)
                for m in range(-radius, radius + 1):
                    cand = center + m * step
                    a = (2.0 * np.pi * grids[ax] * cand).ravel()
                    ac, as_ = np.cos(a), np.sin(a)
                    c = oc * ac - os_ * as_
                    sn = os_ * ac + oc * as_
                    sre = float(np.sum(pr.real * c - pr.imag * sn)) / n_total
                    sim = float(np.sum(pr.real * sn + pr.imag * c)) / n_total
                    mag2 = sre * sre + sim * sim
                    if mag2 > best[0]:
                        best = (mag2, cand, complex(sre, sim))
                shift[ax] = best[1]
                peak = best[2]
        shift = [round(s_ * u) / u for s_ in shift]
    phasediff = float(np.arctan2(peak.imag, peak.real))
    ref_power = float(np.sum(ref * ref))
    mov_power = float(np.sum(mov * mov))
    denom = ref_power * mov_power
    if denom <= np.finfo(np.float64).eps:
        error = 0.0
    else:
        if normalize_phase:
            un = np.fft.ifftn(A * np.conj(B)).ravel()[max_idx]
        else:
            un = peak
        amp2 = (un.real ** 2 + un.imag ** 2) * n_total * n_total
        error = float(np.sqrt(1.0 - min(max(amp2 / denom, 0.0), 1.0)))
    return shift, error, phasediff


def spectral_axes64(re: np.ndarray, im: np.ndarray, kind: str, values) -> tuple:
    """fourier_gaussian_into / fourier_box_into / fourier_shift_into (filter/spectral.rs:20-165)
    in FP64: per axis, the spectrum times exp(-2 pi^2 p^2 f^2) ("gaussian"), sinc(p f) ("box",
    1 where |pi p f| < 1e-12) or exp(-2 pi i p f) ("shift"), f the signed frequency."""
    z = np.asarray(re, dtype=np.float64) + 1j * np.asarray(im, dtype=np.float64)
    vals = list(values) if hasattr(values, "__len__") else [values]
    if len(vals) == 1:
        vals = vals * z.ndim
    for ax, p in enumerate(vals):
        n = z.shape[ax]
        f = _signed_freqs(n)
        if kind == "gaussian":
            w = np.exp(-2.0 * np.pi ** 2 * p * p * f * f)
        elif kind == "box":
            x = np.pi * p * f
            w = np.where(np.abs(x) < 1e-12, 1.0, np.sin(x) / np.where(np.abs(x) < 1e-12, 1.0, x))
        else:
            w = np.exp(-2j * np.pi * p * f)
        shp = [1] * z.ndim
        shp[ax] = n
        z = z * w.reshape(shp)
    return z.real.copy(), z.imag.copy()


def butterworth64(x: np.ndarray, cutoff_ratio: float, high_pass: bool, order: float) -> np.ndarray:
    """butterworth (filter/spectral.rs:166-212) in FP64: the real part of the inverse transform of
    the spectrum times 1 / (1 + |f / cutoff|^(2 order)) (or 1 minus it)."""
    x = np.asarray(x, dtype=np.float64)
    grids = np.meshgrid(*[_signed_freqs(n) for n in x.shape], indexing="ij")
    r2 = sum((g / cutoff_ratio) ** 2 for g in grids)
    lp = 1.0 / (1.0 + np.sqrt(r2) ** (2.0 * order))
    w = 1.0 - lp if high_pass else lp
    return np.fft.ifftn(np.fft.fftn(x) * w).real


def gabor_kernel64(frequency: float, theta: float, sigma_x: float, sigma_y: float, n_stds: float = 3.0,
                   offset: float = 0.0) -> tuple:
    """gabor_kernel_2d (filter/spectral.rs:216-246) in FP64: (real, imaginary) kernels of shape
    (2 ry + 1, 2 rx + 1)."""
    ct, st = np.cos(theta), np.sin(theta)
    rx = int(max(np.ceil(n_stds * np.hypot(sigma_x * ct, sigma_y * st)), 1.0))
    ry = int(max(np.ceil(n_stds * np.hypot(sigma_x * st, sigma_y * ct)), 1.0))
    y, x = np.meshgrid(np.arange(-ry, ry + 1, dtype=np.float64), np.arange(-rx, rx + 1, dtype=np.float64),
                       indexing="ij")
    xr = x * ct + y * st
    yr = -x * st + y * ct
    norm = 1.0 / (2.0 * np.pi * sigma_x * sigma_y)
    env = norm * np.exp(-0.5 * (xr * xr / (sigma_x * sigma_x) + yr * yr / (sigma_y * sigma_y)))
    phase = 2.0 * np.pi * frequency * xr + offset
    return env * np.cos(phase), env * np.sin(phase)


def pcg_hash32(x: np.ndarray) -> np.ndarray:
    """core/kernels/random.rs:7-12 on uint32 arrays (arithmetic wraps at 32 bits)."""
    x = np.asarray(x, dtype=np.uint32)
    state = x * np.uint32(747796405) + np.uint32(2891336453)
    word = ((state >> ((state >> np.uint32(28)) + np.uint32(4))) ^ state) * np.uint32(277803737)
    return (word >> np.uint32(22)) ^ word


def _uniform32(seed: np.uint32, c: np.ndarray) -> np.ndarray:
    return pcg_hash32(c ^ seed).astype(np.float32) * np.float32(1.0 / 4294967296.0)


def _normal32(seed: np.uint32, idx: np.ndarray) -> np.ndarray:
    u1 = np.maximum(_uniform32(seed, idx * np.uint32(2)), np.float32(1e-7))
    u2 = _uniform32(seed, idx * np.uint32(2) + np.uint32(1))
    return np.sqrt(np.float32(-2.0) * np.log(u1)) * np.cos(np.float32(6.2831855) * u2)


def add_noise32(image: np.ndarray, kind: int, p0: float, p1: float = 0.0, seed: int = 0,
                clip: bool = True) -> np.ndarray:
    """add_noise (intensity/noise.rs:35-143) in float32: kind 0 Gaussian (mean, std), 1 Speckle
    (mean, std), 2 Salt (amount), 3 Pepper (amount), 4 SaltAndPepper (amount, salt_vs_pepper),
    5 Poisson (scale)."""
    x = np.ascontiguousarray(image, dtype=np.float32).ravel()
    if x.size == 0:
        raise ValueError("empty input to add_noise")
    seed = int(seed) & ((1 << 64) - 1)
    s32 = np.uint32(((seed & 0xFFFFFFFF) ^ (((seed >> 32) * 0x9E3779B9) & 0xFFFFFFFF)) & 0xFFFFFFFF)
    idx = np.arange(x.size, dtype=np.uint32)
    a, b = np.float32(p0), np.float32(p1)
    if kind in (0, 1):
        z = _normal32(s32, idx)
        out = (x + a + b * z) if kind == 0 else (x + x * (a + b * z))
        out = out.astype(np.float32)
        if clip:
            out = np.clip(out, np.float32(0), np.float32(1))
    elif kind in (2, 3, 4):
        b = np.float32(1.0) if kind == 2 else (np.float32(0.0) if kind == 3 else b)
        u = _uniform32(s32, idx * np.uint32(2))
        v = _uniform32(s32, idx * np.uint32(2) + np.uint32(1))
        out = np.where(u < a, np.where(v < b, np.float32(1), np.float32(0)), x).astype(np.float32)
    elif kind == 5:
        scale = np.float32(max(float(p0), 1.0))
        lam = (np.maximum(x, np.float32(0)) * scale).astype(np.float32)
        out = np.zeros_like(x)
        small = (lam > 0) & (lam < 20)
        l = np.exp(-lam[small]).astype(np.float32)
        si = idx[small]
        p = np.ones(si.size, dtype=np.float32)
        k = np.zeros(si.size, dtype=np.uint32)
        alive = np.ones(si.size, dtype=bool)
        for step in range(1, 129):
            k = np.where(alive, np.uint32(step), k)
            u = np.maximum(_uniform32(s32, si * np.uint32(131) + np.uint32(step)), np.float32(1e-7))
            p = np.where(alive, (p * u).astype(np.float32), p)
            alive &= ~(p <= l)
        out[small] = (k - np.uint32(1)).astype(np.float32) / scale
        big = lam >= 20
        z = _normal32(s32, idx[big])
        lb = lam[big]
        out[big] = np.maximum(np.floor(lb + np.sqrt(lb) * z + np.float32(0.5)), np.float32(0)) / scale
        if clip:
            out = np.clip(out, np.float32(0), np.float32(1))
    else:
        raise ValueError(f"unknown noise model {kind}")
    return out.reshape(np.shape(image))


def centred_kernel64(kernel: np.ndarray, shape) -> np.ndarray:
    """fft_shift_2d_kernel (restore/kernels/deconvolve.rs:76-99): kernel[ky, kx] stored at
    ((ky - kh/2) mod H, (kx - kw/2) mod W) of a zero H x W image (later entries win)."""
    kernel = np.asarray(kernel, dtype=np.float64)
    H, W = shape
    kh, kw = kernel.shape
    out = np.zeros((H, W))
    for ky in range(kh):
        for kx in range(kw):
            out[(ky + H - kh // 2) % H, (kx + W - kw // 2) % W] = kernel[ky, kx]
    return out


def laplacian_spectrum64(shape) -> np.ndarray:
    """|L|^2 of the regulariser of deconvolve.rs:266-285: the centred 3x3 Laplacian, or a 1x1
    kernel of 1 when an axis is shorter than 3."""
    H, W = shape
    if min(H, 3) == 3 and min(W, 3) == 3:
        lap = np.array([[0.0, -1.0, 0.0], [-1.0, 4.0, -1.0], [0.0, -1.0, 0.0]])
    else:
        lap = np.ones((1, 1))
    L = np.fft.fft2(centred_kernel64(lap, shape))
    return (L * np.conj(L)).real


def wiener64(image: np.ndarray, psf: np.ndarray, balance: float = 0.1, clip: bool = True) -> np.ndarray:
    """wiener_deconvolve_into (deconvolve.rs:213-315) in FP64: the real part of
    ifft2(conj(H) Y / max(|H|^2 + balance |L|^2, 1e-12)), clipped to [-1, 1] when asked."""
    y = np.asarray(image, dtype=np.float64)
    Y = np.fft.fft2(y)
    Hf = np.fft.fft2(centred_kernel64(psf, y.shape))
    denom = np.maximum((Hf * np.conj(Hf)).real + float(balance) * laplacian_spectrum64(y.shape), 1e-12)
    out = np.fft.ifft2(np.conj(Hf) * Y / denom).real
    return np.clip(out, -1.0, 1.0) if clip else out


def unsupervised_wiener_jitter(seed: int, num_iter: int, n: float) -> list:
    """The hyperparameter jitter of deconvolve.rs:429-431, :472-475 (64-bit LCG)."""
    mask = (1 << 64) - 1
    state = (int(seed) + 0x9E3779B97F4A7C15) & mask
    out = []
    for _ in range(int(num_iter)):
        state = (state * 6364136223846793005 + 1) & mask
        u = (state >> 33) / float(1 << 31) - 0.5
        out.append(float(np.exp(u * (2.0 / np.sqrt(n)))))
    return out


def unsupervised_wiener64(image: np.ndarray, psf: np.ndarray, num_iter: int = 30, burnin: int = 10,
                          clip: bool = True, seed: int = 0):
    """unsupervised_wiener (deconvolve.rs:357-509) in FP64 over the full spectrum: returns
    ``(image, balance_chain)``."""
    y = np.asarray(image, dtype=np.float64)
    n = float(y.size)
    Y = np.fft.fft2(y)
    Hf = np.fft.fft2(centred_kernel64(psf, y.shape))
    hm2 = (Hf * np.conj(Hf)).real
    lm2 = laplacian_spectrum64(y.shape)
    hy = Y * np.conj(Hf)
    img_var = float(np.var(y))
    balance = 0.05
    chain = []
    mean_x = np.zeros_like(Y)
    count = 0
    jitter = unsupervised_wiener_jitter(seed, num_iter, n)
    for it in range(int(num_iter)):
        denom = np.maximum(hm2 + balance * lm2, 1e-12)
        X = hy / denom
        D = Y - Hf * X
        rss = float(np.sum((D.real ** 2 + D.imag ** 2) / n))
        reg = float(np.sum(lm2 * (X.real ** 2 + X.imag ** 2) / n))
        trace = float(np.sum(hm2 / denom))
        if it >= burnin:
            mean_x += X
            count += 1
        eff_dof = max(n - trace, 1.0)
        sn2 = max(rss / eff_dof, 1e-8 * max(img_var, 1e-6))
        sx2 = max(reg / max(trace, 1.0), 1e-8)
        balance = min(max(sn2 / sx2 * jitter[it], 1e-6), 1e3)
        chain.append(balance)
    out = np.fft.ifft2(mean_x / max(count, 1)).real
    return (np.clip(out, -1.0, 1.0) if clip else out), chain


def rel_l2(a: np.ndarray, b: np.ndarray) -> float:
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    d = np.linalg.norm((a - b).ravel())
    n = np.linalg.norm(b.ravel())
    return float(d / n) if n > 0 else float(d)


if os.environ.get("LUCEM_ORACLE_THREADS"):
    set_num_threads(int(os.environ["LUCEM_ORACLE_THREADS"]))
