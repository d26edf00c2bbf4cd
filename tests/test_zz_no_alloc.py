"""The allocation-free entry point (lucem_richardson_lucy_into_noalloc_f32) against the plain
2-D entry point it varies (lucem_richardson_lucy_into_f32), called through the C ABI.

Two calls in a row (the second on a smaller input, so the scratch keeps the first call's larger
contents), the direct form past the FFT length limit, the size limit, and that the call leaves
torch's allocator untouched.
"""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _settings(num_iter, clip=False, mode=0):
    from lucem_amd import _native as N

    s = N.RlSettings()
    N.lib().lucem_rl_settings_default(ctypes.byref(s))
    s.num_iter, s.clip, s.mode = num_iter, int(clip), mode
    return s


def _noalloc(client, x, p, o, s):
    from lucem_amd import _native as N

    return N.lib().lucem_richardson_lucy_into_noalloc_f32(
        client.handle, N._vp(x.data_ptr()), N.dims(x.shape), x.dim(), N._vp(p.data_ptr()), N.dims(p.shape),
        p.dim(), N._vp(o.data_ptr()), N.dims(o.shape), o.dim(), ctypes.byref(s))


def _plain(client, x, p, o, s):
    from lucem_amd import _native as N

    return N.lib().lucem_richardson_lucy_into_f32(
        client.handle, N._vp(x.data_ptr()), N.dims(x.shape), x.dim(), N._vp(p.data_ptr()), N.dims(p.shape),
        p.dim(), N._vp(o.data_ptr()), N.dims(o.shape), o.dim(), ctypes.byref(s), None)


def _case(shape, kshape, seed):
    import torch

    rng = np.random.default_rng(seed)
    y = torch.from_numpy(rng.uniform(20, 200, shape).astype(np.float32)).cuda()
    psf = torch.from_numpy(rng.uniform(0.1, 1.0, kshape).astype(np.float32)).cuda()
    return y, psf


def _rel(a, b):
    return float(((a - b).norm() / b.norm()).item())


@pytest.mark.parametrize("first,second", [
    (((300, 260), (11, 9)), ((96, 70), (5, 7))),
    (((1024, 1024), (25, 25)), ((1, 500), (1, 9))),
])
def test_two_calls_match_the_plain_entry_point(gpu, first, second):
    import torch

    from lucem_amd import _native as N

    print(f"no-alloc: {first} then {second}", flush=True)
    for k, ((shape, kshape), mode) in enumerate(((first, 0), (second, 3))):
        y, psf = _case(shape, kshape, 10 + k)
        s = _settings(8, clip=k == 1, mode=mode)
        ref = torch.empty_like(y)
        out = torch.full_like(y, float("nan"))
        N.check(_plain(gpu, y, psf, ref, s))
        N.check(_noalloc(gpu, y, psf, out, s))
        torch.cuda.synchronize()
        assert bool(torch.isfinite(out).all())
        err = _rel(out, ref)
        print(f"  call {k}: {shape} psf {kshape}: rel-L2 {err:.2e}, bit-equal {bool(torch.equal(out, ref))}")
        # the same launches with the same shapes and twiddles: equal bit for bit
        assert torch.equal(out, ref), err


def test_direct_form_past_the_fft_limit(gpu):
    import torch

    from lucem_amd import _native as N

    print("no-alloc: direct form 3x16400", flush=True)
    y, psf = _case((3, 16400), (3, 5), 3)
    s = _settings(4)
    ref, out = torch.empty_like(y), torch.full_like(y, float("nan"))
    N.check(_plain(gpu, y, psf, ref, s))
    N.check(_noalloc(gpu, y, psf, out, s))
    torch.cuda.synchronize()
    assert torch.equal(out, ref)


def test_rejects_a_problem_beyond_the_scratch(gpu):
    import torch

    from lucem_amd import _native as N

    L = N.lib()
    assert L.lucem_noalloc_scratch_bytes() == 96 << 20
    y = torch.empty((4096, 4096), device="cuda")
    psf = torch.ones((9, 9), device="cuda")
    out = torch.empty_like(y)
    st = _noalloc(gpu, y, psf, out, _settings(2))
    assert st == 9  # LUCEM_UNSUPPORTED_SHAPE
    assert b"bytes of scratch" in L.lucem_last_error()
    # the argument checks come first, in the reference's order
    assert _noalloc(gpu, y, psf, y, _settings(2)) == 2
    assert _noalloc(gpu, y, psf, out, _settings(0)) == 5


def test_allocates_nothing_through_torch(gpu):
    import torch

    from lucem_amd import _native as N

    y, psf = _case((200, 180), (7, 7), 4)
    out = torch.empty_like(y)
    s = _settings(5)
    N.check(_noalloc(gpu, y, psf, out, s))
    torch.cuda.synchronize()
    before = torch.cuda.memory_stats()["allocation.all.allocated"]
    for _ in range(3):
        N.check(_noalloc(gpu, y, psf, out, s))
    torch.cuda.synchronize()
    assert torch.cuda.memory_stats()["allocation.all.allocated"] == before
