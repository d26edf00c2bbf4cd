"""GPU parity of the HIP Richardson-Lucy path against the CPU oracle, through the C ABI.

Bar: relative L2 <= 1e-4 against the FP32 restatement of the reference
(crates/lucem-ops/src/restore/deconvolve.rs:45-149), with the error against FP64 printed; the
reference KATs (deconvolve.rs:517-566) at their own tolerances; size-independent properties at
the full benchmark sizes.
"""

from __future__ import annotations

import subprocess
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]
RTOL_L2 = 1e-4


def _run(y, psf, num_iter, mode=0, clip=False, nd=None, eps=None):
    import torch

    import lucem_amd as LA

    yt = torch.from_numpy(np.ascontiguousarray(y, np.float32)).cuda()
    pt = torch.from_numpy(np.ascontiguousarray(psf, np.float32)).cuda()
    out = torch.empty_like(yt)
    s = LA.RichardsonLucy(num_iter=num_iter, clip=clip, mode=mode, filter_epsilon=eps)
    if (nd if nd is not None else y.ndim == 3):
        LA.richardson_lucy_nd_into(yt, pt, out, s)
    else:
        LA.richardson_lucy_into(yt, pt, out, s)
    torch.cuda.synchronize()
    return out.cpu().numpy()


@pytest.mark.parametrize("L", [1, 2, 3, 4, 5, 6, 7, 8, 9, 12, 15, 16, 25, 30, 49, 64, 105, 128, 160, 343, 525, 540,
                               1000, 1050, 2048, 4096, 6048, 8192, 9072, 16128, 16384])
def test_fft_lines_match_numpy(gpu, L):
    import ctypes

    import torch

    from lucem_amd import _native as N

    rng = np.random.default_rng(L)
    n = 5
    x = (rng.standard_normal((n, L)) + 1j * rng.standard_normal((n, L))).astype(np.complex64)
    t = torch.from_numpy(x.view(np.float32).copy()).cuda()
    f = torch.empty_like(t)
    b = torch.empty_like(t)
    N.check(N.lib().lucem_fft_c2c_lines_f32(gpu.handle, ctypes.c_void_p(t.data_ptr()), ctypes.c_void_p(f.data_ptr()),
                                            L, n, 0))
    N.check(N.lib().lucem_fft_c2c_lines_f32(gpu.handle, ctypes.c_void_p(f.data_ptr()), ctypes.c_void_p(b.data_ptr()),
                                            L, n, 1))
    torch.cuda.synchronize()
    got = f.cpu().numpy().view(np.complex64)
    ref = np.fft.fft(x.astype(np.complex128), axis=1)
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 2e-6
    back = b.cpu().numpy().view(np.complex64) / L
    assert np.linalg.norm(back - x) / np.linalg.norm(x) < 2e-6


CASES_2D = [
    ((8, 8), (3, 3)), ((37, 53), (5, 7)), ((64, 64), (4, 6)), ((40, 33), (9, 9)), ((33, 40), (7, 3)),
    ((50, 21), (5, 5)), ((1, 1), (3, 3)), ((1, 17), (3, 5)), ((19, 1), (5, 1)), ((5, 6), (11, 13)),
    ((129, 97), (17, 17)), ((256, 256), (25, 25)),
]


@pytest.mark.parametrize("mode", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("shape,kshape", CASES_2D)
def test_rl_2d_matches_oracle(gpu, oracle, mode, shape, kshape):
    rng = np.random.default_rng(hash((shape, kshape, mode)) % 2**32)
    y = rng.uniform(50, 150, shape).astype(np.float32)
    psf = rng.uniform(0.1, 1.0, kshape).astype(np.float32)
    got = _run(y, psf, 10, mode)
    st, ref = oracle.richardson_lucy(y, psf, 10, clip=False, mode=mode)
    assert st == 0
    _, ref64 = oracle.richardson_lucy(y.astype(np.float64), psf.astype(np.float64), 10, clip=False, mode=mode)
    e32, e64 = oracle.rel_l2(got, ref), oracle.rel_l2(got, ref64)
    print(f"{shape} {kshape} mode {mode}: rel-L2 vs FP32 {e32:.2e}, vs FP64 {e64:.2e}")
    assert e32 < RTOL_L2


CASES_3D = [((6, 9, 11), (3, 3, 5)), ((20, 30, 25), (5, 7, 3)), ((16, 32, 32), (7, 5, 5)), ((3, 4, 5), (5, 5, 5)),
            ((1, 12, 13), (1, 3, 3)), ((24, 40, 36), (31, 15, 15))]


@pytest.mark.parametrize("mode", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("shape,kshape", CASES_3D)
def test_rl_3d_matches_oracle(gpu, oracle, mode, shape, kshape):
    rng = np.random.default_rng(hash((shape, kshape, mode)) % 2**32)
    y = rng.uniform(50, 150, shape).astype(np.float32)
    psf = rng.uniform(0.1, 1.0, kshape).astype(np.float32)
    got = _run(y, psf, 6, mode, nd=True)
    st, ref = oracle.richardson_lucy(y, psf, 6, clip=False, mode=mode, nd=True)
    assert st == 0
    e = oracle.rel_l2(got, ref)
    print(f"{shape} {kshape} mode {mode}: rel-L2 vs FP32 {e:.2e}")
    assert e < RTOL_L2


def test_2d_through_nd_entry_matches(gpu, oracle):
    rng = np.random.default_rng(1)
    y = rng.uniform(50, 150, (45, 38)).astype(np.float32)
    psf = rng.uniform(0.1, 1.0, (5, 6)).astype(np.float32)
    a = _run(y, psf, 7, nd=False)
    b = _run(y, psf, 7, nd=True)
    np.testing.assert_array_equal(a, b)


def test_identity_kat(gpu):
    """deconvolve.rs:519-539 within 1e-4 (assert_close semantics)."""
    img = ((np.arange(64, dtype=np.float32) / 64.0) * 0.8 + 0.1).reshape(8, 8)
    delta = np.zeros((3, 3), np.float32)
    delta[1, 1] = 1.0
    out = _run(img, delta, 15)
    scale = np.abs(img).max()
    assert np.all(np.abs(out - img) <= 1e-4 * np.abs(img) + 1e-4 * scale)


def test_box_sharpening_kat(gpu, oracle):
    """deconvolve.rs:551-566: clip on, 25 iterations, peak gain > 0.15."""
    spot = np.full((8, 8), 0.1, np.float32)
    spot[4, 4] = 0.9
    box = np.full((3, 3), 1.0 / 9.0, np.float32)
    blurred = oracle.correlate(spot, box, 0, 0.0)
    out = _run(blurred, box, 25, clip=True)
    assert out[4, 4] > blurred[4, 4] + 0.15
    st, ref = oracle.richardson_lucy(blurred, box, 25, clip=True)
    assert oracle.rel_l2(out, ref) < RTOL_L2


def test_clip_only_on_last_iteration(gpu, oracle):
    rng = np.random.default_rng(2)
    y = rng.uniform(0.2, 0.9, (30, 31)).astype(np.float32)
    y[10:14, 10:14] = 3.0
    psf = rng.uniform(0.1, 1.0, (5, 5)).astype(np.float32)
    got = _run(y, psf, 6, clip=True)
    st, ref = oracle.richardson_lucy(y, psf, 6, clip=True)
    assert got.max() <= 1.0 and oracle.rel_l2(got, ref) < RTOL_L2


def test_clip_last_iteration_closed_form(gpu):
    """A 1 x 2 image under Wrap with PSF [[0.75, 0.25]]: two iterations written out by hand give
    (0.6092, 4.1990), clamped to (0.6092, 1.0) with clip; clamping every iteration would give
    0.8682 for the first pixel."""
    y = np.array([[3.0, 0.2]], np.float32)
    psf = np.array([[0.75, 0.25]], np.float32)
    a, b = 0.75, 0.25
    x = [0.5, 0.5]
    for _ in range(2):
        u = (b * x[1] + a * x[0], b * x[0] + a * x[1])
        r = (3.0 / u[0], 0.2 / u[1])
        x = [x[0] * (a * r[1] + b * r[0]), x[1] * (a * r[0] + b * r[1])]
    np.testing.assert_allclose(_run(y, psf, 2, mode=4), [x], rtol=1e-5)
    np.testing.assert_allclose(_run(y, psf, 2, clip=True, mode=4), [np.clip(x, -1, 1)], rtol=1e-5)


def test_shifted_delta_closed_form(gpu, oracle):
    rng = np.random.default_rng(3)
    y = rng.uniform(1.0, 2.0, (9, 7)).astype(np.float32)
    h = np.zeros((3, 3), np.float32)
    h[0, 1] = 1.0
    out = _run(y, h, 3)
    want = np.empty_like(y)
    want[1:] = y[:-1]
    want[0] = y[0]
    np.testing.assert_allclose(out, want, rtol=1e-5)


def test_zero_regions_and_filter_epsilon(gpu, oracle):
    """Zero data under Constant padding: the eps replacement keeps ratios finite."""
    y = np.zeros((40, 40), np.float32)
    y[15:25, 15:25] = 100.0
    psf = np.ones((5, 5), np.float32)
    for eps in (None, 1e-3):
        got = _run(y, psf, 5, mode=1, eps=eps)
        st, ref = oracle.richardson_lucy(y, psf, 5, clip=False, mode=1, filter_epsilon=eps)
        assert np.isfinite(got).all() and oracle.rel_l2(got, ref) < RTOL_L2


def test_eps_rule_uses_the_exact_correlation(gpu, oracle):
    """|u| < eps replaces u by eps (restore/kernels/deconvolve.rs:17-21). With data near eps the
    rule fires on some pixels and not others, so a correlation off by any constant factor (which
    RL alone would not show) changes the result."""
    rng = np.random.default_rng(8)
    y = rng.uniform(0.5e-3, 2.0e-3, (30, 34)).astype(np.float32)
    psf = rng.uniform(0.1, 1.0, (5, 5)).astype(np.float32)
    for eps in (1.0e-3, 1.5e-3):
        got = _run(y, psf, 3, eps=eps)
        st, ref = oracle.richardson_lucy(y, psf, 3, clip=False, filter_epsilon=eps)
        assert oracle.rel_l2(got, ref) < RTOL_L2
    # the case discriminates: a correlation twice too large is equivalent to eps / 2
    _, half = oracle.richardson_lucy(y, psf, 3, clip=False, filter_epsilon=0.75e-3)
    assert oracle.rel_l2(half, ref) > 1e-2


def test_plain2d_full_size(gpu, oracle):
    """1024x1024, 25x25 Gaussian sigma 2, 30 iterations (the 2-D benchmark workload)."""
    from lucem_amd import synth

    p = synth.plain2d()
    got = _run(p.image, p.psf, p.num_iter)
    st, ref = oracle.richardson_lucy(p.image, p.psf, p.num_iter, clip=False)
    e = oracle.rel_l2(got, ref)
    e64 = oracle.rel_l2(got, oracle.rl_fft64(p.image, p.psf, p.num_iter))
    print(f"plain2d: rel-L2 vs FP32 {e:.2e}, vs FP64 {e64:.2e}")
    assert e < RTOL_L2


def test_plain3d_reduced_matches_oracle(gpu, oracle):
    """The 3-D benchmark PSF (31x15x15, sigma (4,1.5,1.5)) on a 32x128x128 crop, 10 iterations."""
    from lucem_amd import synth

    p = synth.plain3d(shape=(32, 128, 128), count=60, num_iter=10)
    got = _run(p.image, p.psf, p.num_iter)
    st, ref = oracle.richardson_lucy(p.image, p.psf, p.num_iter, clip=False, nd=True)
    e = oracle.rel_l2(got, ref)
    print(f"plain3d crop: rel-L2 vs FP32 {e:.2e}")
    assert e < RTOL_L2


@pytest.mark.parametrize("depth,B,mode", [(128, 8, 0), (128, 12, 3), (130, 12, 1), (34, 32, 0), (66, 21, 4),
                                          (98, 16, 2), (162, 10, 0), (226, 8, 0), (130, 16, 0), (128, 20, 3),
                                          (128, 24, 1)])
def test_plain3d_compile_time_z_pass(gpu, oracle, monkeypatch, depth, B, mode):
    """The compile-time builds of the fused z pass (lines of depth + 30 = 64 to 256 points with a
    31-slice PSF, B columns per workgroup; two LDS buffers, or one for 16 to 24 columns of 160
    points) against the FP32 oracle and against the run-time build of the same pass, which they
    must match to rounding."""
    from lucem_amd import synth

    rng = np.random.default_rng(40 + depth + B + mode)
    y = rng.uniform(50, 150, (depth, 24, 20)).astype(np.float32)
    psf = synth.gaussian_psf((4.0, 1.5, 1.5), (31, 5, 5)).astype(np.float32)
    monkeypatch.setenv("LUCEM_AUTOTUNE", "0")
    monkeypatch.setenv("LUCEM_TUNE", f"col_B={B}")
    monkeypatch.setenv("LUCEM_COL_STATIC", "0")
    print(f"z pass depth={depth} B={B} mode={mode}: run-time build", flush=True)
    generic = _run(y, psf, 4, mode=mode)
    monkeypatch.setenv("LUCEM_COL_STATIC", "1")
    print(f"z pass depth={depth} B={B} mode={mode}: compile-time build", flush=True)
    static = _run(y, psf, 4, mode=mode)
    st, ref = oracle.richardson_lucy(y, psf, 4, mode=mode, clip=False, nd=True)
    e, d = oracle.rel_l2(static, ref), oracle.rel_l2(static, generic)
    print(f"z pass depth={depth} B={B} mode={mode}: rel-L2 vs FP32 oracle {e:.2e}, vs the run-time build {d:.2e}")
    assert e < RTOL_L2 and d < 1e-6


@pytest.mark.parametrize("P,mode,clip", [(2, 0, False), (3, 1, False), (3, 3, True), (2, 4, False), (3, 2, True)])
def test_rows_compile_time_pass(gpu, oracle, monkeypatch, P, mode, clip):
    """The compile-time ratio and update row passes (540-point rows: 526 columns and a 15-column
    PSF, P row pairs per workgroup, an odd row count) in 3-D and in the 2-D whole-array path,
    against the FP32 oracle and against the run-time build of the same passes."""
    rng = np.random.default_rng(60 + P + mode)
    lo, hi = (0.2, 1.6) if clip else (50, 150)
    cases = [(rng.uniform(lo, hi, (3, 7, 526)).astype(np.float32), rng.uniform(0.1, 1.0, (3, 3, 15)).astype(np.float32)),
             (rng.uniform(lo, hi, (21, 526)).astype(np.float32), rng.uniform(0.1, 1.0, (5, 15)).astype(np.float32))]
    monkeypatch.setenv("LUCEM_RL_TILED", "0")
    monkeypatch.setenv("LUCEM_AUTOTUNE", "0")
    monkeypatch.setenv("LUCEM_TUNE", f"row_P={P}")
    for y, psf in cases:
        monkeypatch.setenv("LUCEM_ROW_STATIC", "0")
        print(f"rows P={P} mode={mode} clip={clip} {y.shape}: run-time build", flush=True)
        generic = _run(y, psf, 4, mode=mode, clip=clip)
        monkeypatch.setenv("LUCEM_ROW_STATIC", "1")
        print(f"rows P={P} mode={mode} clip={clip} {y.shape}: compile-time build", flush=True)
        static = _run(y, psf, 4, mode=mode, clip=clip)
        st, ref = oracle.richardson_lucy(y, psf, 4, mode=mode, clip=clip, nd=y.ndim == 3)
        e, dg = oracle.rel_l2(static, ref), oracle.rel_l2(static, generic)
        print(f"rows P={P} mode={mode} clip={clip} {y.shape}: rel-L2 vs FP32 oracle {e:.2e}, vs the run-time build {dg:.2e}")
        assert e < RTOL_L2 and dg < 1e-6


@pytest.mark.parametrize("shape,kshape,P", [((3, 5, 256), (3, 3, 15), 6), ((4, 7, 256), (1, 3, 15), 4),
                                            ((9, 266), (3, 15), 6), ((2, 3, 1066), (3, 3, 15), 1)])
def test_rows_compile_time_pass_lengths(gpu, oracle, monkeypatch, shape, kshape, P):
    """The other compile-time row builds (270-, 280- and 1080-point rows) against the FP32 oracle
    and the run-time build."""
    rng = np.random.default_rng(70 + P + len(shape))
    y = rng.uniform(50, 150, shape).astype(np.float32)
    psf = rng.uniform(0.1, 1.0, kshape).astype(np.float32)
    monkeypatch.setenv("LUCEM_RL_TILED", "0")
    monkeypatch.setenv("LUCEM_AUTOTUNE", "0")
    monkeypatch.setenv("LUCEM_TUNE", f"row_P={P}")
    monkeypatch.setenv("LUCEM_ROW_STATIC", "0")
    generic = _run(y, psf, 4, mode=0)
    monkeypatch.setenv("LUCEM_ROW_STATIC", "1")
    print(f"rows {shape} P={P}: compile-time build", flush=True)
    static = _run(y, psf, 4, mode=0)
    st, ref = oracle.richardson_lucy(y, psf, 4, mode=0, clip=False, nd=y.ndim == 3)
    e, dg = oracle.rel_l2(static, ref), oracle.rel_l2(static, generic)
    print(f"rows {shape} P={P}: rel-L2 vs FP32 oracle {e:.2e}, vs the run-time build {dg:.2e}")
    assert e < RTOL_L2 and dg < 1e-6


@pytest.mark.parametrize("B,mode", [(14, 0), (12, 1), (14, 4), (12, 3)])
def test_plain3d_compile_time_y_passes(gpu, oracle, monkeypatch, B, mode):
    """The compile-time forward and inverse y passes of 3-D arrays (540-point lines: 526 rows and a
    15-row PSF, B columns per 512-thread workgroup) against the FP32 oracle and against the
    run-time builds."""
    rng = np.random.default_rng(80 + B + mode)
    y = rng.uniform(50, 150, (3, 526, 21)).astype(np.float32)
    psf = rng.uniform(0.1, 1.0, (3, 15, 3)).astype(np.float32)
    monkeypatch.setenv("LUCEM_AUTOTUNE", "0")
    monkeypatch.setenv("LUCEM_TUNE", f"col_B={B},col_wide=1")
    monkeypatch.setenv("LUCEM_COL_STATIC", "0")
    generic = _run(y, psf, 4, mode=mode)
    monkeypatch.setenv("LUCEM_COL_STATIC", "1")
    print(f"y passes B={B} mode={mode}: compile-time build", flush=True)
    static = _run(y, psf, 4, mode=mode)
    st, ref = oracle.richardson_lucy(y, psf, 4, mode=mode, clip=False, nd=True)
    e, dg = oracle.rel_l2(static, ref), oracle.rel_l2(static, generic)
    print(f"y passes B={B} mode={mode}: rel-L2 vs FP32 oracle {e:.2e}, vs the run-time build {dg:.2e}")
    assert e < RTOL_L2 and dg < 1e-6


def test_plain2d_whole_array_compile_time_y_pass(gpu, oracle, monkeypatch):
    """The 2-D whole-array passes with the compile-time fused y pass (130 rows and a 31-row PSF:
    160-point lines, 12 columns per workgroup) against the FP32 oracle."""
    rng = np.random.default_rng(48)
    y = rng.uniform(50, 150, (130, 90)).astype(np.float32)
    psf = rng.uniform(0.1, 1.0, (31, 7)).astype(np.float32)
    monkeypatch.setenv("LUCEM_RL_TILED", "0")
    monkeypatch.setenv("LUCEM_AUTOTUNE", "0")
    monkeypatch.setenv("LUCEM_TUNE", "col_B=12")
    monkeypatch.setenv("LUCEM_COL_STATIC", "1")
    got = _run(y, psf, 5)
    st, ref = oracle.richardson_lucy(y, psf, 5, clip=False)
    e = oracle.rel_l2(got, ref)
    print(f"2-D whole-array, compile-time y pass: rel-L2 vs FP32 oracle {e:.2e}")
    assert e < RTOL_L2


def test_plain3d_full_size_properties(gpu):
    """128x512x512 with the 31x15x15 PSF: a flat image is a fixed point of RL (Reflect), and
    the estimate of positive data stays positive and finite over 50 iterations."""
    import torch

    from lucem_amd import synth

    psf = synth.gaussian_psf((4.0, 1.5, 1.5), (31, 15, 15)).astype(np.float32)
    flat = np.full((128, 512, 512), 250.0, np.float32)
    out = _run(flat, psf, 5)
    assert np.abs(out - 250.0).max() / 250.0 < 1e-4
    rng = np.random.default_rng(4)
    y = rng.uniform(50, 150, (128, 512, 512)).astype(np.float32)
    out = _run(y, psf, 50)
    assert np.isfinite(out).all() and out.min() > 0
    del torch


def test_workspace_reuse_and_growth(gpu, oracle):
    import torch

    import lucem_amd as LA
    from lucem_amd import _native as N

    ws = N.Workspace(gpu)
    rng = np.random.default_rng(5)
    s = LA.RichardsonLucy(num_iter=4, clip=False)
    caps = []
    for shape in [(64, 64), (20, 30), (128, 100), (64, 64)]:
        y = rng.uniform(50, 150, shape).astype(np.float32)
        psf = rng.uniform(0.1, 1.0, (5, 5)).astype(np.float32)
        out = torch.empty(shape, device="cuda")
        LA.richardson_lucy_into(torch.from_numpy(y).cuda(), torch.from_numpy(psf).cuda(), out, s, ws, gpu)
        gpu.sync()
        caps.append(ws.capacity_bytes())
        st, ref = oracle.richardson_lucy(y, psf, 4, clip=False)
        assert oracle.rel_l2(out.cpu().numpy(), ref) < RTOL_L2
    assert caps == sorted(caps) and caps[-1] == caps[-2]


def test_errors_with_device(gpu):
    import torch

    import lucem_amd as LA

    x = torch.ones(8, 8, device="cuda")
    psf = torch.ones(3, 3, device="cuda")
    with pytest.raises(LA.LucemError) as e:
        LA.richardson_lucy_into(x, psf, x)
    assert e.value.kind == "ALIASED_BUFFERS"
    with pytest.raises(LA.LucemError) as e:
        LA.richardson_lucy_into(x, psf, torch.empty(8, 9, device="cuda"))
    assert e.value.kind == "SHAPE_MISMATCH"
    with pytest.raises(LA.LucemError) as e:
        LA.richardson_lucy_into(torch.ones(2, 8, 8, device="cuda"), psf, torch.empty(2, 8, 8, device="cuda"))
    assert e.value.kind == "INVALID_DIMENSION"
    with pytest.raises(LA.LucemError) as e:
        LA.richardson_lucy_into(x, psf, torch.empty_like(x), LA.RichardsonLucy(num_iter=0))
    assert e.value.kind == "INVALID_PARAMETER"


def test_cpp_caller(gpu):
    exe = ROOT / "tests" / "cpp" / "test_restore"
    assert exe.exists(), "build() compiles tests/cpp/test_restore"
    res = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0 and res.stdout.strip() == "ok", res.stderr


# Lines longer than 4096 after padding take the 512- and 1024-thread kernels (up to 16384).
LONG_CASES = [
    ((3, 6000), (3, 5)),     # rows of L = 6048: 512 threads
    ((2, 9000), (1, 7)),     # rows of L = 9072: 1024 threads, 71 KiB of LDS
    ((9000, 3), (5, 3)),     # columns of L = 9072: 1024-thread line pass
    ((5000, 3), (5, 3)),     # columns of L = 5040: 512-thread line pass (one line per workgroup)
    ((4500, 40), (9, 5)),    # columns of L = 4508 -> 4536: 512 threads
    ((5000, 300), (5, 3)),   # columns of L = 5040 with enough work to be autotuned
    ((2, 16000), (1, 3)),    # rows of L = 16128: 126 KiB of LDS
]


@pytest.mark.parametrize("shape,kshape", LONG_CASES)
def test_long_lines_match_oracle(gpu, oracle, shape, kshape):
    rng = np.random.default_rng(shape[0] * 7 + shape[1])
    y = rng.uniform(50, 150, shape).astype(np.float32)
    p = rng.uniform(0.1, 1.0, kshape).astype(np.float32)
    got = _run(y, p, 4)
    st, ref = oracle.richardson_lucy(y, p, 4, clip=False)
    assert st == 0
    e = oracle.rel_l2(got, ref)
    print(f"{shape} psf {kshape}: rel-L2 vs FP32 {e:.2e}")
    assert e < RTOL_L2


# A 2-D axis whose padded domain n + k - 1 is longer than 16384 takes the direct form: the
# reference's own correlations and order in FP32, so it matches the FP32 oracle to rounding
# (FMA contraction). 3-D arrays with such an axis are refused.
LONG_DIRECT_CASES = [
    ((2, 16400), (1, 3), 0),
    ((3, 16390), (3, 5), 3),
    ((4, 16380), (3, 9), 1),    # n fits, n + k - 1 does not
    ((16390, 2), (3, 1), 2),
    ((1100000, 2), (3, 1), 0),  # more than 65535 rows of 16-row blocks: the flat block grid
]


@pytest.mark.parametrize("shape,kshape,mode", LONG_DIRECT_CASES)
def test_longer_than_16384_takes_the_direct_form(gpu, oracle, shape, kshape, mode):
    rng = np.random.default_rng(shape[1] + kshape[1])
    y = rng.uniform(0.1, 1.0, shape).astype(np.float32)
    psf = rng.uniform(0.1, 1.0, kshape).astype(np.float32)
    print(f"direct f32 {shape} psf {kshape} mode {mode}", flush=True)
    got = _run(y, psf, 4, mode)
    st, ref = oracle.richardson_lucy(y, psf, num_iter=4, clip=False, mode=mode)
    assert st == 0
    e = oracle.rel_l2(got, ref)
    print(f"  rel-L2 vs FP32 {e:.2e}")
    assert e < 1e-6


def test_3d_longer_than_16384_is_refused(gpu):
    import torch

    import lucem_amd as LA

    y = torch.ones((2, 2, 16400), device="cuda")
    with pytest.raises(Exception, match="16384"):
        LA.richardson_lucy_nd_into(y, torch.ones((1, 1, 3), device="cuda"), torch.empty_like(y),
                                   LA.RichardsonLucy(num_iter=2))


# richardson_lucy_into::<f64>: FP64 direct correlations in the reference's accumulation order, so
# the result agrees with the FP64 oracle to rounding (FMA contraction and the PSF sum order aside).
RTOL_F64 = 1e-12
F64_CASES = [
    ((37, 53), (5, 7), 0, False),
    ((64, 64), (3, 3), 1, True),
    ((29, 41), (4, 6), 2, False),
    ((50, 33), (9, 2), 3, True),
    ((45, 45), (7, 7), 4, False),
    ((1, 30), (3, 3), 0, False),
    ((100, 90), (81, 81), 0, False),  # the input tile with its halo exceeds 64 KiB of LDS
]


@pytest.mark.parametrize("shape,kshape,mode,clip", F64_CASES)
def test_rl_f64_matches_fp64_oracle(gpu, oracle, shape, kshape, mode, clip):
    import torch

    import lucem_amd as LA

    rng = np.random.default_rng(shape[0] * 31 + kshape[1] + mode)
    y = rng.uniform(0.05, 0.95, shape) if clip else rng.uniform(50, 150, shape)
    p = rng.uniform(0.1, 1.0, kshape)
    out = torch.empty(shape, dtype=torch.float64, device="cuda")
    LA.richardson_lucy_into(torch.from_numpy(y).cuda(), torch.from_numpy(p).cuda(), out,
                            LA.RichardsonLucy(num_iter=6, clip=clip, mode=mode))
    torch.cuda.synchronize()
    st, ref = oracle.richardson_lucy(y, p, 6, clip=clip, mode=mode)
    assert st == 0 and ref.dtype == np.float64
    e = oracle.rel_l2(out.cpu().numpy(), ref)
    print(f"f64 {shape} psf {kshape} mode {mode}: rel-L2 vs FP64 {e:.2e}")
    assert e < RTOL_F64


def test_rl_f64_filter_epsilon_and_zeros(gpu, oracle):
    import torch

    import lucem_amd as LA

    rng = np.random.default_rng(71)
    y = rng.uniform(0.0, 2.0, (40, 40))
    y[10:30, 5:25] = 0.0
    p = rng.uniform(0.0, 1.0, (5, 5))
    p[2, :] = 0.0
    out = torch.empty_like(torch.from_numpy(y)).cuda()
    LA.richardson_lucy_into(torch.from_numpy(y).cuda(), torch.from_numpy(p).cuda(), out,
                            LA.RichardsonLucy(num_iter=8, clip=False, mode=LA.PadMode.Constant, filter_epsilon=1e-3))
    torch.cuda.synchronize()
    st, ref = oracle.richardson_lucy(y, p, 8, clip=False, filter_epsilon=1e-3, mode=1)
    assert st == 0
    assert oracle.rel_l2(out.cpu().numpy(), ref) < RTOL_F64


def test_rl_f64_errors(gpu):
    import torch

    import lucem_amd as LA

    y = torch.ones((8, 8), dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError, match="float64"):
        LA.richardson_lucy_into(y, torch.ones((3, 3), device="cuda"), torch.empty_like(y))
    with pytest.raises(LA.LucemError) as e:
        LA.richardson_lucy_into(y, torch.ones((3, 3), dtype=torch.float64, device="cuda"), y)
    assert e.value.kind == "ALIASED_BUFFERS"
    v = torch.ones((2, 8, 8), dtype=torch.float64, device="cuda")
    with pytest.raises(LA.LucemError) as e:
        LA.richardson_lucy_nd_into(v, torch.ones((3, 3, 3), dtype=torch.float64, device="cuda"), torch.empty_like(v))
    assert e.value.kind == "INVALID_DIMENSION"


@pytest.mark.parametrize("shape,kshape", [((5000, 2, 3), (3, 1, 3)), ((2, 4800, 3), (1, 5, 3)), ((3, 2, 9000), (3, 1, 5))])
def test_long_axes_3d_match_oracle(gpu, oracle, shape, kshape):
    """3-D volumes whose z, y or x line is longer than 4096 (512- and 1024-thread passes)."""
    import torch

    import lucem_amd as LA

    rng = np.random.default_rng(sum(shape))
    y = rng.uniform(50, 150, shape).astype(np.float32)
    p = rng.uniform(0.1, 1.0, kshape).astype(np.float32)
    out = torch.empty(shape, device="cuda")
    LA.richardson_lucy_nd_into(torch.from_numpy(y).cuda(), torch.from_numpy(p).cuda(), out,
                               LA.RichardsonLucy(num_iter=3, clip=False))
    torch.cuda.synchronize()
    st, ref = oracle.richardson_lucy(y, p, 3, clip=False, nd=True)
    assert st == 0
    e = oracle.rel_l2(out.cpu().numpy(), ref)
    print(f"{shape} psf {kshape}: rel-L2 vs FP32 {e:.2e}")
    assert e < RTOL_L2
