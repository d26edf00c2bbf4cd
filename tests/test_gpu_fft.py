"""GPU tests of the general complex FFT (core/fft.rs:13-135) against the FP64 oracle: every axis
of 1-D to 3-D arrays, 7-smooth lengths and Bluestein lengths (primes up to 8191 in one pass,
longer lines in four-step passes), the N-D form in place and as a round trip, the complex
product, and the reference's checks."""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _dev(a):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


def _err(gr, gi, wr, wi):
    g = gr.cpu().numpy().astype(np.float64) + 1j * gi.cpu().numpy()
    w = wr + 1j * wi
    return float(np.linalg.norm((g - w).ravel()) / max(np.linalg.norm(w.ravel()), 1e-30))


AXIS_CASES = [
    ((6, 8, 12), 0), ((6, 8, 12), 1), ((6, 8, 12), 2), ((6, 8, 12), -1),
    ((5, 11, 3), 1),          # 11: Bluestein, strided
    ((3, 97), 1),             # 97: Bluestein, contiguous
    ((97, 40), 0),
    ((4, 1000), 1),           # 7-smooth, not a power of two
    ((2, 1021), 1),           # prime: Bluestein of length 2048
    ((4099, 2), 0),           # prime: Bluestein of length 8192 (512 threads), strided
    ((2, 8191), 1),           # prime: Bluestein of length 16384 (1024 threads)
    ((3, 16384), 1),          # 7-smooth at the 16384 limit
    ((1, 1), 0), ((7, 1), 1),
    ((16384, 3), 0),          # strided at the 16384 limit (1024 threads)
    ((3000, 5), 0),           # strided: the thread tier grows until 5 columns fit
    # lines longer than one workgroup holds: passes
    ((2, 20000), 1),          # 7-smooth: four-step 160 x 125, contiguous
    ((20000, 3), 0),          # four-step, strided
    ((1, 16807), 1),          # 7^5: four-step 7 x 2401
    ((2, 32768), 1),          # a power of two past the limit
    ((1, 8209), 1),           # prime above 8192: Bluestein of length 32768 in passes
    ((8209, 2), 0),           # the same, strided
    ((1, 20011), 1),          # prime: Bluestein of length 65536
]


@pytest.mark.parametrize("shape,axis", AXIS_CASES)
def test_fft_1d_matches_oracle(gpu, oracle, shape, axis):
    import torch

    from lucem_amd import fft

    rng = np.random.default_rng(sum(shape) + axis)
    re, im = rng.normal(size=shape), rng.normal(size=shape)
    for inverse in (False, True):
        ro, io = torch.empty(shape, device="cuda"), torch.empty(shape, device="cuda")
        fft.fft_1d_into(_dev(re), _dev(im), ro, io, axis=axis, inverse=inverse)
        torch.cuda.synchronize()
        wr, wi = oracle.fft_axis64(re.astype(np.float32), im.astype(np.float32), axes=[axis % len(shape)],
                                   inverse=inverse)
        e = _err(ro, io, wr, wi)
        print(f"fft {shape} axis {axis} inverse {inverse}: rel-L2 {e:.2e}")
        assert e < 2e-6


@pytest.mark.parametrize("shape", [(64, 80), (5, 12, 11), (3, 4, 5, 6), (17,)])
def test_fft_nd_matches_oracle_and_round_trips(gpu, oracle, shape):
    import torch

    from lucem_amd import fft

    rng = np.random.default_rng(len(shape) * 13 + shape[0])
    re, im = rng.normal(size=shape), rng.normal(size=shape)
    fr, fi = fft.fft_nd(_dev(re), _dev(im))
    wr, wi = oracle.fft_axis64(re.astype(np.float32), im.astype(np.float32))
    assert _err(fr, fi, wr, wi) < 2e-6
    br, bi = fft.fft_nd(fr, fi, inverse=True)
    torch.cuda.synchronize()
    assert _err(br, bi, re, im) < 2e-6
    if len(shape) >= 2:  # in place: outputs alias the inputs
        xr, xi = _dev(re), _dev(im)
        fft.fft_nd_into(xr, xi, xr, xi)
        torch.cuda.synchronize()
        assert _err(xr, xi, wr, wi) < 2e-6


@pytest.mark.parametrize("shape", [(3, 20000), (8209, 6), (4, 5, 17000)])
def test_fft_nd_long_axes_round_trip(gpu, oracle, shape):
    """N-D transforms with a long axis (four-step or Bluestein in passes), in place for the
    middle axes, against the oracle and back."""
    import torch

    from lucem_amd import fft

    rng = np.random.default_rng(sum(shape))
    re, im = rng.normal(size=shape), rng.normal(size=shape)
    fr, fi = fft.fft_nd(_dev(re), _dev(im))
    wr, wi = oracle.fft_axis64(re.astype(np.float32), im.astype(np.float32))
    e = _err(fr, fi, wr, wi)
    print(f"fft_nd {shape}: rel-L2 {e:.2e}")
    assert e < 2e-6
    br, bi = fft.fft_nd(fr, fi, inverse=True)
    assert _err(br, bi, re, im) < 2e-6


def test_fft_nd_scalar_and_empty(gpu):
    import torch

    from lucem_amd import fft

    r, i = torch.tensor(3.0, device="cuda"), torch.tensor(-2.0, device="cuda")
    ro, io = fft.fft_nd(r, i)
    assert float(ro) == 3.0 and float(io) == -2.0
    e = torch.empty((0, 5), device="cuda")
    fft.fft_nd_into(e, e.clone(), e.clone(), e.clone())


def test_complex_mul_matches_numpy(gpu):
    import torch

    from lucem_amd import fft

    rng = np.random.default_rng(3)
    a = rng.normal(size=(2, 3, 50)) + 1j * rng.normal(size=(2, 3, 50))
    b = rng.normal(size=(2, 3, 50)) + 1j * rng.normal(size=(2, 3, 50))
    for conj in (False, True):
        orr, oi = torch.empty((2, 3, 50), device="cuda"), torch.empty((2, 3, 50), device="cuda")
        fft.complex_mul_into(_dev(a.real), _dev(a.imag), _dev(b.real), _dev(b.imag), orr, oi, conj_b=conj)
        torch.cuda.synchronize()
        want = a * (np.conj(b) if conj else b)
        assert _err(orr, oi, want.real, want.imag) < 1e-6


def test_fft_errors(gpu):
    import torch

    import lucem_amd as LA
    from lucem_amd import fft

    x = torch.ones((4, 8), device="cuda")
    y = torch.ones((4, 8), device="cuda")
    with pytest.raises(LA.LucemError) as e:
        fft.fft_1d_into(x, y, torch.empty((4, 9), device="cuda"), torch.empty((4, 8), device="cuda"))
    assert e.value.kind == "SHAPE_MISMATCH"
    with pytest.raises(LA.LucemError) as e:
        fft.fft_1d_into(x, y, x, torch.empty((4, 8), device="cuda"))
    assert e.value.kind == "ALIASED_BUFFERS"
    with pytest.raises(LA.LucemError) as e:
        fft.fft_1d_into(x, y, torch.empty_like(x), torch.empty_like(x), axis=2)
    assert e.value.kind == "AXIS_OUT_OF_BOUNDS"
    v = torch.ones(8, device="cuda")
    with pytest.raises(LA.LucemError) as e:  # the 1-D N-D form goes through fft_1d_into's check
        fft.fft_nd_into(v, v.clone(), v, torch.empty_like(v))
    assert e.value.kind == "ALIASED_BUFFERS"

