"""The C ABI without a device: the library loads, exports every declared symbol, and its host
logic (PadMode::parse, defaults, argument checks in the reference's order) behaves as the
reference's (crates/lucem-ops/src/core/pad.rs:48-84, restore/deconvolve.rs:32-71)."""

from __future__ import annotations

import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


@pytest.fixture(scope="module")
def L():
    from lucem_amd import _native as N

    if not N.LIB_PATH.exists():
        from lucem_amd import build

        build.build()
    return N.lib()


def declared_symbols() -> set[str]:
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(lucem_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol(L):
    from lucem_amd import _native as N

    names = declared_symbols()
    assert len(names) >= 15
    for name in names:
        assert hasattr(L, name), f"{name} declared in include/ but not exported"
    assert set(N.EXPORTED) <= names


def test_version(L):
    assert b"gfx950" in L.lucem_version()


def test_settings_default_matches_reference(L):
    from lucem_amd import _native as N

    s = N.RlSettings()
    L.lucem_rl_settings_default(ctypes.byref(s))
    assert (s.num_iter, s.clip, s.has_filter_epsilon, s.mode) == (30, 1, 0, 0)


@pytest.mark.parametrize("name,want", [
    ("reflect", 0), ("symmetric", 0), ("grid-mirror", 0), ("constant", 1), ("grid-constant", 1), ("zeros", 1),
    ("nearest", 2), ("edge", 2), ("replicate", 2), ("mirror", 3), ("wrap", 4), ("grid-wrap", 4),
    ("periodic", 4), ("  Reflect ", 0), ("WRAP", 4),
])
def test_pad_mode_parse(L, name, want):
    out = ctypes.c_uint32(99)
    assert L.lucem_pad_mode_parse(name.encode(), ctypes.byref(out)) == 0
    assert out.value == want


def test_pad_mode_parse_rejects(L):
    out = ctypes.c_uint32(99)
    assert L.lucem_pad_mode_parse(b"invalid", ctypes.byref(out)) == 8
    assert b"not supported" in L.lucem_last_error()


def test_python_padmode_parse():
    from lucem_amd import PadMode

    assert PadMode.parse("grid-mirror") is PadMode.Reflect
    assert PadMode.parse("periodic") is PadMode.Wrap
    with pytest.raises(Exception):
        PadMode.parse("bogus")


def _call(L, in_dims, psf_dims, out_dims, num_iter=3, mode=0, inp=0x10000, out=0x20000000, nd=False):
    from lucem_amd import _native as N

    s = N.RlSettings()
    L.lucem_rl_settings_default(ctypes.byref(s))
    s.num_iter = num_iter
    s.mode = mode
    fn = L.lucem_richardson_lucy_nd_into_f32 if nd else L.lucem_richardson_lucy_into_f32
    # NULL client: the argument checks run first and never touch the (fake) device pointers
    return fn(None, N._vp(inp), N.dims(in_dims), len(in_dims), N._vp(0x30000000), N.dims(psf_dims), len(psf_dims),
              N._vp(out), N.dims(out_dims), len(out_dims), ctypes.byref(s), None)


def test_check_order_without_device(L):
    assert _call(L, (4, 5), (3, 3), (4, 6)) == 1  # same_shape
    assert b"does not match" in L.lucem_last_error()
    assert _call(L, (4, 5), (3, 3), (4, 5), inp=0x1000, out=0x1000 + 40) == 2  # overlapping buffers
    assert _call(L, (2, 4, 5), (3, 3), (2, 4, 5)) == 3  # input ndim
    assert b"expects 2-D" in L.lucem_last_error()
    assert _call(L, (4, 5), (3,), (4, 5)) == 3  # psf ndim
    assert _call(L, (4, 5), (0, 3), (4, 5)) == 4  # empty psf
    assert _call(L, (4, 5), (3, 3), (4, 5), num_iter=0) == 5
    assert _call(L, (4, 5), (3, 3), (4, 5), mode=7) == 8
    assert _call(L, (0, 5), (3, 3), (0, 5)) == 0  # empty input: Ok, nothing done
    assert _call(L, (4, 5), (3, 3), (4, 5)) == 11  # valid arguments, no client
    # a shape mismatch wins over every later check
    assert _call(L, (2, 4, 5), (0,), (2, 4), num_iter=0) == 1


def test_check_order_nd_without_device(L):
    assert _call(L, (2, 4, 5), (3, 3, 3), (2, 4, 5), nd=True) == 11
    assert _call(L, (2, 4, 5), (3, 3), (2, 4, 5), nd=True) == 3  # psf rank must match in N-D
    assert _call(L, (2, 2, 4, 5), (3, 3, 3, 3), (2, 2, 4, 5), nd=True) == 3
    assert _call(L, (65536, 65536, 2), (1, 1, 1), (65536, 65536, 2), nd=True, out=1 << 45) == 6  # > u32::MAX
    # a 2-D axis beyond the 16384 FFT limit takes the direct form (valid arguments, no client);
    # a 3-D one is refused
    assert _call(L, (8, 16400), (3, 3), (8, 16400)) == 11
    assert _call(L, (8, 16400), (3, 3), (8, 16400), nd=True) == 11
    assert _call(L, (2, 8, 16400), (3, 3, 3), (2, 8, 16400), nd=True) == 9


def test_oracle_is_not_imported_by_the_product():
    """The product path never imports the oracle (it is test infrastructure only)."""
    for p in [*(ROOT / "lucem_amd").rglob("*.py"), *(ROOT / "lucem").rglob("*.py")]:
        text = p.read_text()
        assert "oracle" not in re.sub(r"#.*", "", text).replace("oracle/", ""), p


def test_check_order_f64_without_device(L):
    from lucem_amd import _native as N

    s = N.RlSettings()
    L.lucem_rl_settings_default(ctypes.byref(s))

    def call(in_dims, psf_dims, out_dims, out=0x20000000):
        return L.lucem_richardson_lucy_into_f64(None, N._vp(0x10000), N.dims(in_dims), len(in_dims),
                                                N._vp(0x30000000), N.dims(psf_dims), len(psf_dims), N._vp(out),
                                                N.dims(out_dims), len(out_dims), ctypes.byref(s), None)

    assert call((4, 5), (3, 3), (4, 6)) == 1
    # 8-byte elements: 4 x 5 doubles starting at 0x10000 overlap an output 80 bytes on
    assert call((4, 5), (3, 3), (4, 5), out=0x10000 + 80) == 2
    assert call((4, 5), (3, 3), (4, 5), out=0x10000 + 160) == 11
    assert call((2, 4, 5), (3, 3, 3), (2, 4, 5)) == 3
    assert call((4, 5), (0, 3), (4, 5)) == 4
    assert call((0, 5), (3, 3), (0, 5)) == 0
    # the direct form has no FFT length limit
    assert call((8, 16400), (3, 3), (8, 16400)) == 11


def test_correlate_check_order_without_device(L):
    """correlate_2d_into's order (filter/convolve.rs:133-141): input ndim, weights ndim, same
    shape, aliasing, origin per axis (core/check.rs:59-72), then empty input."""
    from lucem_amd import _native as N

    s = N.CorrelateSettings()
    L.lucem_correlate_settings_default(ctypes.byref(s))
    assert (s.mode, tuple(s.origin), s.cval) == (0, (0, 0, 0), 0.0)

    def call(fn, in_dims, w_dims, out_dims, origin=(0, 0, 0), out=0x20000000, mode=0, f64=False):
        for i, o in enumerate(origin):
            s.origin[i] = o
        s.mode = mode
        args = [None, N._vp(0x10000), N.dims(in_dims), len(in_dims), N._vp(0x30000000), N.dims(w_dims), len(w_dims),
                N._vp(out), N.dims(out_dims), len(out_dims), ctypes.byref(s)]
        return fn(*args) if f64 else fn(*args, None)

    f32, nd, f64 = L.lucem_correlate_2d_into_f32, L.lucem_correlate_nd_into_f32, L.lucem_correlate_2d_into_f64
    assert call(f32, (2, 8, 8), (3, 3), (2, 8, 9)) == 3
    assert b"correlate_2d expects 2-D array, got 3-D shape" in L.lucem_last_error()
    assert call(f32, (8, 8), (3,), (8, 9)) == 3
    assert b"weights expects 2-D array" in L.lucem_last_error()
    assert call(f32, (8, 8), (3, 3), (8, 9)) == 1
    assert call(f32, (8, 8), (3, 3), (8, 8), out=0x10000 + 100) == 2
    assert call(f32, (8, 8), (0, 3), (8, 8)) == 4
    assert b"window: size must be at least 1" in L.lucem_last_error()
    assert call(f32, (8, 8), (3, 4), (8, 8), origin=(2, 0, 0)) == 14
    assert b"origin 2 is out of bounds for window of size 3 (must be in -1..=1)" in L.lucem_last_error()
    assert call(f32, (8, 8), (3, 4), (8, 8), origin=(0, -3, 0)) == 14
    assert call(f32, (8, 8), (3, 4), (8, 8), origin=(-1, -2, 0)) == 11  # in range: reaches the client check
    assert call(f32, (8, 8), (3, 4), (8, 8), origin=(1, 1, 0)) == 11
    assert call(f32, (8, 8), (3, 3), (8, 8), mode=7) == 8
    assert call(f32, (0, 8), (3, 3), (0, 8)) == 0
    assert call(f32, (8, 16400), (3, 3), (8, 16400)) == 11   # the direct sum beyond the FFT limit
    assert call(nd, (2, 8, 16400), (3, 3, 3), (2, 8, 16400)) == 9
    # the N-D form: 3-D arrays with weights of the same rank; origin per axis
    assert call(nd, (2, 8, 8), (3, 3, 3), (2, 8, 8)) == 11
    assert call(nd, (2, 8, 8), (3, 3), (2, 8, 8)) == 3
    assert call(nd, (2, 2, 8, 8), (3, 3, 3, 3), (2, 2, 8, 8)) == 3
    assert call(nd, (2, 8, 8), (3, 3, 3), (2, 8, 8), origin=(0, 0, 2)) == 14
    # the FP64 direct form: 8-byte elements, no FFT length limit
    assert call(f64, (4, 5), (3, 3), (4, 5), out=0x10000 + 80, f64=True) == 2
    assert call(f64, (4, 5), (3, 3), (4, 5), out=0x10000 + 160, f64=True) == 11
    assert call(f64, (8, 16400), (3, 3), (8, 16400), f64=True) == 11
    assert call(f64, (2, 4, 5), (3, 3, 3), (2, 4, 5), f64=True) == 3


def test_render_check_order_without_device(L):
    """The phantom renderers: rank, empty output, element and grid limits, then the handles."""
    from lucem_amd import _native as N

    def call(fn, dims, count=4, objects=0x40000):
        return fn(None, N._vp(0x10000), N.dims(dims), len(dims), N._vp(objects or None), count, 100.0)

    ell, sph = L.lucem_render_ellipses_f32, L.lucem_render_spheres_f32
    assert call(ell, (2, 8, 8)) == 3
    assert b"render_ellipses expects 2-D array, got 3-D shape" in L.lucem_last_error()
    assert call(sph, (8, 8)) == 3
    assert call(ell, (0, 8)) == 0
    assert call(sph, (4, 0, 8)) == 0
    assert call(ell, (1 << 17, 1 << 16)) == 6
    assert call(ell, (1048561, 1)) == 9
    assert call(ell, (1048560, 1)) == 11
    assert call(sph, (262141, 1, 1)) == 9
    assert call(sph, (1, 524281, 1)) == 9
    assert call(sph, (4, 8, 8)) == 11


def test_phase_correlation_check_order_without_device(L):
    """phase.rs:58-68: same shape, non-empty, upsample_factor >= 1; then rank and limits."""
    from lucem_amd import _native as N

    s = N.PhaseCorrelationSettings()
    L.lucem_phase_correlation_settings_default(ctypes.byref(s))
    assert (s.upsample_factor, s.normalize_phase) == (1, 1)
    out = N.PhaseCorrelationResult()

    def call(a, b, up=1):
        s.upsample_factor = up
        return L.lucem_phase_correlation_f32(None, N._vp(0x10000), N.dims(a), len(a), N._vp(0x20000000), N.dims(b),
                                             len(b), ctypes.byref(s), ctypes.byref(out), None)

    assert call((8, 8), (8, 9), up=0) == 1
    assert b"reference shape [8, 8] does not match moving shape [8, 9]" in L.lucem_last_error()
    assert call((0, 8), (0, 8), up=0) == 13
    assert b"phase_correlation cannot be computed on an empty array" in L.lucem_last_error()
    assert call((8, 8), (8, 8), up=0) == 5
    assert call((2, 2, 2, 2), (2, 2, 2, 2)) == 9
    assert call((8, 8), (8, 8)) == 11
    assert call((5,), (5,), up=3) == 11


def test_spectral_check_order_without_device(L):
    """spectral.rs: same_shape(re, im), then 1 or ndim values, then empty; butterworth checks its
    cutoff first."""
    from lucem_amd import _native as N

    vals = (ctypes.c_double * 3)(1.0, 2.0, 3.0)

    def call(fn, a, b, nv):
        return fn(None, N._vp(0x10000), N.dims(a), len(a), N._vp(0x20000000), N.dims(b), len(b), vals, nv)

    for fn in (L.lucem_fourier_gaussian_into_f32, L.lucem_fourier_box_into_f32, L.lucem_fourier_shift_into_f32):
        assert call(fn, (4, 5), (4, 6), 3) == 1
        assert call(fn, (4, 5), (4, 5), 3) == 5
        assert b"expected 1 or 2 values, got 3" in L.lucem_last_error()
        assert call(fn, (0, 5), (0, 5), 1) == 0
        assert call(fn, (4, 5), (4, 5), 2) == 11
        assert call(fn, (4, 5), (4, 5), 1) == 11

    def bw(a, b, c):
        return L.lucem_butterworth_f32(None, N._vp(0x10000), N.dims(a), len(a), N._vp(0x20000000), N.dims(b), len(b),
                                       c, 0, 2.0, None)

    assert bw((4, 5), (4, 6), 0.0) == 5
    assert b"cutoff_ratio must be in (0.0, 0.5], got 0" in L.lucem_last_error()
    assert bw((4, 5), (4, 6), 0.2) == 1
    assert bw((0, 5), (0, 5), 0.2) == 0
    assert bw((2, 2, 2, 2, 2), (2, 2, 2, 2, 2), 0.2) == 9
    assert bw((4, 5), (4, 5), 0.5) == 11


def test_correlate_1d_check_order_without_device(L):
    """correlate_1d_into (filter/convolve.rs:47-53): same shape, aliasing, origin, axis, empty."""
    from lucem_amd import _native as N

    s = N.Correlate1dSettings()
    L.lucem_correlate1d_settings_default(ctypes.byref(s))
    assert (s.axis, s.mode, s.origin, s.cval) == (-1, 0, 0, 0.0)

    def call(a, b, nw=3, axis=-1, origin=0, out=0x20000000):
        s.axis, s.origin = axis, origin
        return L.lucem_correlate_1d_into_f32(None, N._vp(0x10000), N.dims(a), len(a), N._vp(0x30000000), nw,
                                             N._vp(out), N.dims(b), len(b), ctypes.byref(s))

    assert call((4, 5), (4, 6), nw=0) == 1
    assert call((4, 5), (4, 5), nw=0, out=0x10000 + 40) == 2
    assert call((4, 5), (4, 5), nw=0) == 4
    assert call((4, 5), (4, 5), nw=3, origin=2) == 14
    assert call((4, 5), (4, 5), nw=3, axis=2) == 12
    assert call((4, 5), (4, 5), nw=3, axis=-3) == 12
    assert call((0, 5), (0, 5), nw=3) == 0
    assert call((4, 5), (4, 5), nw=4, origin=-2, axis=0) == 11


def test_wiener_check_order_without_device(L):
    from lucem_amd import _native as N

    s = N.WienerSettings()
    L.lucem_wiener_settings_default(ctypes.byref(s))
    assert (s.balance, s.clip) == (0.1, 1)

    def call(in_dims, psf_dims, out_dims, balance=0.1, out=0x20000000):
        s.balance = balance
        return L.lucem_wiener_deconvolve_into_f32(None, N._vp(0x10000), N.dims(in_dims), len(in_dims),
                                                  N._vp(0x30000000), N.dims(psf_dims), len(psf_dims), N._vp(out),
                                                  N.dims(out_dims), len(out_dims), ctypes.byref(s), None)

    assert call((8, 8), (3, 3), (8, 9)) == 1
    assert call((8, 8), (3, 3), (8, 8), out=0x10000 + 100) == 2
    assert call((2, 8, 8), (3, 3), (2, 8, 8)) == 3
    assert b"wiener_deconvolve input expects 2-D" in L.lucem_last_error()
    assert call((8, 8), (3,), (8, 8)) == 3
    assert call((8, 8), (0, 3), (8, 8)) == 4
    assert call((8, 8), (3, 3), (8, 8), balance=-0.5) == 5
    assert call((8, 8), (3, 3), (8, 8), balance=float("nan")) == 5
    assert call((0, 8), (30, 30), (0, 8)) == 0  # empty input wins over the PSF size check
    assert call((8, 8), (9, 3), (8, 8)) == 4
    assert b"must not exceed image" in L.lucem_last_error()
    assert call((8, 11), (3, 3), (8, 11)) == 11  # 11 is not 7-smooth: the general FFT route
    assert call((8, 8209), (3, 3), (8, 8209)) == 11  # a prime above 8192: Bluestein lines in passes
    assert call((8, 8), (3, 3), (8, 8)) == 11  # valid arguments, no client


def test_unsupervised_check_order_without_device(L):
    from lucem_amd import _native as N

    s = N.UnsupervisedWienerSettings()
    L.lucem_unsupervised_wiener_settings_default(ctypes.byref(s))
    assert (s.num_iter, s.burnin, s.clip, s.seed) == (30, 10, 1, 0)
    chain = (ctypes.c_double * 64)()
    got = ctypes.c_uint64(99)

    def call(dims, psf_dims, num_iter=12, burnin=4, cap=64, out=0x20000000):
        s.num_iter, s.burnin = num_iter, burnin
        return L.lucem_unsupervised_wiener_f32(None, N._vp(0x10000), N.dims(dims), len(dims), N._vp(0x30000000),
                                               N.dims(psf_dims), len(psf_dims), N._vp(out), ctypes.byref(s), chain,
                                               cap, ctypes.byref(got), None)

    assert call((2, 8, 8), (3, 3)) == 3
    assert call((8, 8), (3, 3, 1)) == 3
    assert call((8, 8), (3, 0)) == 4
    assert call((8, 8), (3, 3), num_iter=0, burnin=0) == 5
    assert call((8, 8), (3, 3), num_iter=4, burnin=4) == 5
    assert b"must be > burnin" in L.lucem_last_error()
    assert call((0, 8), (3, 3)) == 0 and got.value == 0
    assert call((8, 8), (3, 3), out=0x10000 + 4) == 2
    assert call((8, 8), (3, 3), cap=11) == 5
    assert call((8, 8), (9, 3)) == 4
    assert call((8, 8), (3, 3)) == 11


def test_fft_check_order_without_device(L):
    from lucem_amd import _native as N

    def arr(ptr, dims):
        return [N._vp(ptr), N.dims(dims), len(dims)]

    def call(dims=(4, 8), im=(4, 8), ro=(4, 8), io=(4, 8), axis=-1, out=0x20000000):
        return L.lucem_fft_1d_into_f32(None, *arr(0x10000, dims), *arr(0x40000, im), *arr(out, ro),
                                       *arr(0x60000000, io), axis, 0)

    assert call(im=(4, 9)) == 1
    assert call(ro=(8, 4)) == 1
    assert call(io=(4,)) == 1
    assert call(out=0x10000 + 64) == 2
    assert call(axis=2) == 12 and b"out of bounds" in L.lucem_last_error()
    assert call(axis=-3) == 12
    assert call((0, 8), (0, 8), (0, 8), (0, 8)) == 0
    assert call((0, 8), (0, 8), (0, 8), (0, 8), axis=5) == 12  # the axis is checked before emptiness
    assert call(axis=-2) == 11  # valid arguments, no client
    nd = L.lucem_fft_nd_into_f32(None, *arr(0x10000, (4, 8)), *arr(0x40000, (4, 8)), *arr(0x10000, (4, 8)),
                                 *arr(0x40000, (4, 8)), 0, None)
    assert nd == 11  # in place is allowed for ndim >= 2
    nd1 = L.lucem_fft_nd_into_f32(None, *arr(0x10000, (8,)), *arr(0x40000, (8,)), *arr(0x10000, (8,)),
                                  *arr(0x40000, (8,)), 0, None)
    assert nd1 == 2
    cm = L.lucem_complex_mul_into_f32(None, *arr(0x10000, (3, 4)), N._vp(0x1000), *arr(0x2000, (3, 5)),
                                      N._vp(0x3000), *arr(0x4000, (3, 4)), N._vp(0x5000), 0)
    assert cm == 1


def test_launch_shapes_fit_their_tiers(L):
    """Host-side planning invariants over every 7-smooth FFT length up to 16384 (tests/cpp/
    test_shapes.cpp, built by build()): row and line shapes fit the thread tier they name, and
    the radix plans multiply out to the length."""
    import subprocess

    exe = ROOT / "tests" / "cpp" / "test_shapes"
    if not exe.exists():
        from lucem_amd import build

        build.build()
        pytest.skip("tests/cpp/test_shapes is built by __graft_entry__.build()")
    res = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0 and res.stdout.strip() == "ok", res.stderr


def test_exit_hook_releases_plans_then_workspaces_then_clients(monkeypatch):
    """The interpreter-exit hook closes every live plan, then every workspace, then every client,
    whatever order they were made in (the C ABI also keeps a client alive while a workspace or
    plan made on it lives, so any order is safe; this one also frees each plan's graph while its
    stream is still there)."""
    import weakref

    from lucem_amd import _native as N

    order = []

    class Stub:
        def __init__(self, kind, name):
            self.kind, self.name = kind, name

        def close(self):
            order.append(self.kind)

    owned = {"plans": weakref.WeakSet(), "workspaces": weakref.WeakSet(), "clients": weakref.WeakSet()}
    monkeypatch.setattr(N, "_owned", owned)
    keep = [Stub("clients", "c0"), Stub("workspaces", "w0"), Stub("plans", "p0"), Stub("clients", "c1"),
            Stub("plans", "p1"), Stub("workspaces", "w1")]
    for s in keep:
        owned[s.kind].add(s)
    N._release_all()
    assert order == ["plans", "plans", "workspaces", "workspaces", "clients", "clients"]


def test_destroy_functions_accept_null(L):
    """Destroying a NULL handle is a no-op (as freeing a null pointer)."""
    assert L.lucem_client_destroy(None) == 0
    assert L.lucem_workspace_destroy(None) == 0


def test_direct_form_rejects_axes_beyond_int32(L):
    """An axis whose padded extent n + k - 1 exceeds 2^31 - 1 (possible within the 2^32 - 1 element
    limit when the other axis is 1) is refused before any device work: the direct kernels index
    within an axis with 32-bit integers. One element less fits (valid arguments, no client)."""
    from lucem_amd import _native as N

    big = (1, 2147483647)
    assert _call(L, big, (1, 3), big, out=1 << 45) == 9
    assert b"2147483647" in L.lucem_last_error()
    assert _call(L, (1, 2147483645), (1, 3), (1, 2147483645), out=1 << 45) == 11
    s = N.RlSettings()
    L.lucem_rl_settings_default(ctypes.byref(s))
    assert L.lucem_richardson_lucy_into_f64(None, N._vp(0x10000), N.dims(big), 2, N._vp(0x30000000), N.dims((1, 3)),
                                            2, N._vp(1 << 45), N.dims(big), 2, ctypes.byref(s), None) == 9
    cs = N.CorrelateSettings()
    L.lucem_correlate_settings_default(ctypes.byref(cs))
    assert L.lucem_correlate_2d_into_f64(None, N._vp(0x10000), N.dims(big), 2, N._vp(0x30000000), N.dims((1, 3)), 2,
                                         N._vp(1 << 45), N.dims(big), 2, ctypes.byref(cs)) == 9
