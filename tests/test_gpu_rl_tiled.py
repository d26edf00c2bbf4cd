"""The one-launch tiled iteration of 2-D arrays (rl_tiled.hip) against the FP32 oracle of the
reference (crates/lucem-ops/src/restore/deconvolve.rs:45-149), through the plain 2-D entry point.

Each output tile is correlated on its own small domain (overlap-save), both correlations, the
ratio and the update in one launch; the host admits a tile geometry only where every boundary fold
of rel stays inside the tile. The cases force the tiled form (LUCEM_RL_TILED=1) and check from the
profile that it ran, or that it declined (Wrap on an axis of several tiles) and the whole-array
passes ran instead; both against the oracle at the same 1e-4 relative L2 bar."""

from __future__ import annotations

import os
import zlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RTOL_L2 = 1e-4


@pytest.fixture
def tiled_mode():
    old = os.environ.get("LUCEM_RL_TILED")

    def set_mode(v):
        if v is None:
            os.environ.pop("LUCEM_RL_TILED", None)
        else:
            os.environ["LUCEM_RL_TILED"] = v

    yield set_mode
    set_mode(old)


def _run(client, y, psf, num_iter, mode=0, clip=False, eps=None):
    """Output and the kernel names the call launched."""
    import torch

    import lucem_amd as LA

    yt = torch.from_numpy(np.ascontiguousarray(y, np.float32)).cuda()
    pt = torch.from_numpy(np.ascontiguousarray(psf, np.float32)).cuda()
    out = torch.empty_like(yt)
    s = LA.RichardsonLucy(num_iter=num_iter, clip=clip, mode=mode, filter_epsilon=eps)
    client.profile(True)
    LA.richardson_lucy_into(yt, pt, out, s, None, client)
    stats = client.profile_read()
    client.profile(False)
    torch.cuda.synchronize()
    return out.cpu().numpy(), stats


CASES = [
    ((200, 300), (7, 9)), ((256, 256), (25, 25)), ((130, 70), (4, 6)), ((97, 129), (17, 17)), ((600, 40), (3, 5)),
    ((5, 6), (11, 13)), ((1, 300), (1, 9)), ((300, 1), (9, 1)), ((64, 64), (1, 1)), ((180, 190), (31, 30)),
    ((2, 2), (3, 3)), ((33, 500), (2, 3)),
]


@pytest.mark.parametrize("mode", [0, 1, 2, 3])
@pytest.mark.parametrize("shape,kshape", CASES)
def test_tiled_matches_oracle(gpu, oracle, tiled_mode, mode, shape, kshape):
    tiled_mode("1")
    rng = np.random.default_rng(zlib.crc32(repr((shape, kshape, mode)).encode()))
    y = rng.uniform(50, 150, shape).astype(np.float32)
    psf = rng.uniform(0.1, 1.0, kshape).astype(np.float32)
    got, stats = _run(gpu, y, psf, 8, mode)
    assert "tile_iter" in stats and stats["tile_iter"]["launches"] == 8, stats.keys()
    st, ref = oracle.richardson_lucy(y, psf, 8, clip=False, mode=mode)
    assert st == 0
    e = oracle.rel_l2(got, ref)
    print(f"{shape} {kshape} mode {mode}: rel-L2 vs FP32 {e:.2e}")
    assert e < RTOL_L2


def test_wrap_declines_on_several_tiles(gpu, oracle, tiled_mode):
    """Wrap folds a left-edge window onto the right edge, outside the tile: the whole-array passes
    run instead."""
    tiled_mode("1")
    rng = np.random.default_rng(5)
    y = rng.uniform(50, 150, (300, 280)).astype(np.float32)
    psf = rng.uniform(0.1, 1.0, (9, 9)).astype(np.float32)
    got, stats = _run(gpu, y, psf, 5, mode=4)
    assert "tile_iter" not in stats and "line_fused_y" in stats
    _, ref = oracle.richardson_lucy(y, psf, 5, clip=False, mode=4)
    assert oracle.rel_l2(got, ref) < RTOL_L2


def test_wrap_on_one_tile(gpu, oracle, tiled_mode):
    """An array that one tile covers keeps every Wrap target inside the tile."""
    tiled_mode("1")
    rng = np.random.default_rng(6)
    y = rng.uniform(50, 150, (40, 50)).astype(np.float32)
    psf = rng.uniform(0.1, 1.0, (5, 5)).astype(np.float32)
    got, stats = _run(gpu, y, psf, 6, mode=4)
    assert stats["tile_iter"]["launches"] == 6
    _, ref = oracle.richardson_lucy(y, psf, 6, clip=False, mode=4)
    assert oracle.rel_l2(got, ref) < RTOL_L2


def test_tiled_clip_eps_and_one_iteration(gpu, oracle, tiled_mode):
    tiled_mode("1")
    rng = np.random.default_rng(7)
    y = rng.uniform(0.2, 0.9, (90, 131)).astype(np.float32)
    y[40:50, 60:70] = 3.0
    psf = rng.uniform(0.1, 1.0, (5, 5)).astype(np.float32)
    got, stats = _run(gpu, y, psf, 6, clip=True)
    _, ref = oracle.richardson_lucy(y, psf, 6, clip=True)
    assert "tile_iter" in stats and got.max() <= 1.0 and oracle.rel_l2(got, ref) < RTOL_L2
    got, _ = _run(gpu, y, psf, 1)
    _, ref = oracle.richardson_lucy(y, psf, 1, clip=False)
    assert oracle.rel_l2(got, ref) < RTOL_L2
    # data near eps: the rule fires on some pixels and not on others
    z = rng.uniform(0.5e-3, 2.0e-3, (70, 74)).astype(np.float32)
    got, stats = _run(gpu, z, psf, 3, eps=1.0e-3)
    _, ref = oracle.richardson_lucy(z, psf, 3, clip=False, filter_epsilon=1.0e-3)
    assert "tile_iter" in stats and oracle.rel_l2(got, ref) < RTOL_L2


def test_tiled_against_whole_array_at_full_size(gpu, tiled_mode):
    """The 1024 x 1024 workload (25 x 25 Gaussian PSF, 30 iterations): the tiled and the
    whole-array transforms agree to rounding."""
    from lucem_amd import synth

    prob = synth.plain2d()
    tiled_mode("1")
    a, sa = _run(gpu, prob.image, prob.psf, prob.num_iter)
    tiled_mode("0")
    b, sb = _run(gpu, prob.image, prob.psf, prob.num_iter)
    assert "tile_iter" in sa and "tile_iter" not in sb
    e = float(np.linalg.norm(a - b) / np.linalg.norm(b))
    print(f"tiled vs whole-array: rel-L2 {e:.2e}")
    assert np.isfinite(a).all() and e < 2e-6


@pytest.mark.parametrize("mode", [0, 1, 2, 3])
@pytest.mark.parametrize("shape,kshape,grid", [((57, 85), (3, 4), "32,32,28,26"), ((129, 65), (5, 2), "48,48,40,46"),
                                              ((97, 97), (2, 7), "64,64,62,52")])
def test_edge_tiles_of_one_row(gpu, oracle, tiled_mode, monkeypatch, mode, shape, kshape, grid):
    """Grids whose last tiles are 1 to a few rows or columns high, with odd and even PSF extents:
    the row transforms run only over the row pairs each step reads (the rel window's, the output's),
    whose counts are odd or even here, on the compile-time builds."""
    tiled_mode("1")
    monkeypatch.setenv("LUCEM_TILE_GRID", grid)
    rng = np.random.default_rng(zlib.crc32(repr((shape, kshape, mode)).encode()))
    y = rng.uniform(50, 150, shape).astype(np.float32)
    psf = rng.uniform(0.1, 1.0, kshape).astype(np.float32)
    got, stats = _run(gpu, y, psf, 5, mode)
    assert "tile_iter" in stats
    _, ref = oracle.richardson_lucy(y, psf, 5, clip=False, mode=mode)
    e = oracle.rel_l2(got, ref)
    print(f"{shape} {kshape} grid {grid} mode {mode}: rel-L2 vs FP32 {e:.2e}")
    assert e < RTOL_L2


@pytest.mark.parametrize("shape,kshape,grid", [((1024, 1024), (25, 25), None), ((300, 200), (9, 9), "64,64,48,48"),
                                              ((500, 420), (17, 13), "96,96,64,72"), ((90, 70), (3, 3), "128,128,124,124")])
def test_constant_length_build_equals_run_time_build(gpu, tiled_mode, monkeypatch, shape, kshape, grid):
    """The kernel compiled with the tile lengths as constants runs the same radix sequence as the
    run-time-length build; the compiler fuses their multiply-adds differently, so they agree to
    rounding."""
    tiled_mode("1")
    if grid:
        monkeypatch.setenv("LUCEM_TILE_GRID", grid)
    rng = np.random.default_rng(zlib.crc32(repr(shape).encode()))
    y = rng.uniform(50, 150, shape).astype(np.float32)
    psf = rng.uniform(0.1, 1.0, kshape).astype(np.float32)
    monkeypatch.setenv("LUCEM_TILE_GENERIC", "1")
    a, sa = _run(gpu, y, psf, 4)
    monkeypatch.setenv("LUCEM_TILE_GENERIC", "0")
    b, sb = _run(gpu, y, psf, 4)
    assert "tile_iter" in sa and "tile_iter" in sb
    e = float(np.linalg.norm(a - b) / np.linalg.norm(b))
    print(f"{shape} {kshape}: constant vs run-time lengths rel-L2 {e:.2e}")
    assert e < 1e-6


def _tiled_batch_launches(num_iter, npsf):
    """Launches of a tiled batch: PSF preparation, two transfer-function builds of two kernels per
    PSF, one launch per iteration (the per-pass batch has four per iteration)."""
    return 1 + 4 * npsf + num_iter


@pytest.mark.parametrize("mode", [0, 1, 2, 3])
def test_tiled_batch_matches_oracle(gpu, oracle, tiled_mode, mode):
    """Uneven 2-D tiles with three PSFs in one tiled batch (one launch per iteration over every
    output tile of every array) against the FP32 oracle, and against the plain call of each tile."""
    import torch

    import lucem_amd as LA
    from lucem_amd.batched import RLBatch

    tiled_mode("1")
    rng = np.random.default_rng(40 + mode)
    shapes = [(64, 64), (17, 93), (128, 40), (5, 7), (150, 130), (200, 150), (33, 1), (1, 29)]
    psfs_np = [rng.uniform(0.1, 1.0, k).astype(np.float32) for k in [(5, 5), (7, 3), (11, 11)]]
    index = [0, 1, 2, 0, 2, 1, 0, 1]
    ys_np = [rng.uniform(50, 150, s).astype(np.float32) for s in shapes]
    ys = [torch.from_numpy(y).cuda() for y in ys_np]
    psfs = [torch.from_numpy(p).cuda() for p in psfs_np]
    s = LA.RichardsonLucy(num_iter=6, clip=False, mode=mode)
    b = RLBatch(shapes, [p.shape for p in psfs_np], index, s)
    assert b.launches == _tiled_batch_launches(6, 3)
    outs = [torch.empty_like(y) for y in ys]
    b.execute(ys, outs, psfs)
    torch.cuda.synchronize()
    for y, o, i in zip(ys_np, outs, index):
        got = o.cpu().numpy()
        _, ref = oracle.richardson_lucy(y, psfs_np[i], 6, clip=False, mode=mode)
        assert oracle.rel_l2(got, ref) < RTOL_L2
        plain, _ = _run(gpu, y, psfs_np[i], 6, mode)
        assert float(np.linalg.norm(got - plain) / np.linalg.norm(plain)) < 2e-6
    b.close()


def test_tiled_batch_layout_windows(gpu, tiled_mode):
    """Tiles given as strided views into one image, each storing a window into one result, through
    the tiled batch: the plain per-tile results inside the windows, the rest of the result untouched."""
    import torch

    import lucem_amd as LA
    from lucem_amd.batched import RLBatch, TileLayout

    tiled_mode("1")
    rng = np.random.default_rng(44)
    H, W = 160, 190
    img = torch.from_numpy(rng.uniform(50, 150, (H, W)).astype(np.float32)).cuda()
    res = torch.full((H, W), -7.0, device="cuda")
    # (read origin, read size, window origin in the tile, window size): y x
    tiles = [((0, 0), (90, 110), (4, 6), (70, 90)), ((30, 100), (130, 90), (10, 5), (100, 60)),
             ((100, 0), (60, 95), (0, 0), (0, 0))]
    psf = torch.from_numpy(rng.uniform(0.1, 1.0, (7, 5)).astype(np.float32)).cuda()
    s = LA.RichardsonLucy(num_iter=5, clip=False)
    plan = RLBatch([size for _, size, _, _ in tiles], [tuple(psf.shape)], [0] * len(tiles), s)
    assert plan.launches == _tiled_batch_launches(5, 1)
    ins, outs, lays = [], [], []
    for ro, size, wo, ws in tiles:
        ins.append(img.data_ptr() + 4 * (ro[0] * W + ro[1]))
        outs.append(res.data_ptr() + 4 * ((ro[0] + wo[0]) * W + ro[1] + wo[1]))
        lays.append(TileLayout(W, 0, W, 0, (0,) + wo, (0,) + ws))
    plan.execute_layout(ins, outs, lays, [psf])
    torch.cuda.synchronize()
    covered = torch.zeros((H, W), dtype=torch.bool, device="cuda")
    for ro, size, wo, ws in tiles:
        view = img[ro[0]:ro[0] + size[0], ro[1]:ro[1] + size[1]].contiguous()
        ref, _ = _run(gpu, view.cpu().numpy(), psf.cpu().numpy(), 5)
        wsz = [w if w else size[a] - wo[a] for a, w in enumerate(ws)]
        sl = (slice(ro[0] + wo[0], ro[0] + wo[0] + wsz[0]), slice(ro[1] + wo[1], ro[1] + wo[1] + wsz[1]))
        want = ref[wo[0]:wo[0] + wsz[0], wo[1]:wo[1] + wsz[1]]
        got = res[sl].cpu().numpy()
        assert float(np.linalg.norm(got - want) / np.linalg.norm(want)) < 2e-6
        covered[sl] = True
    assert bool((res[~covered] == -7.0).all())
    plan.close()


def test_batch_form_choice(gpu, tiled_mode):
    """Wrap across several tiles admits no tile grid: the per-pass batch is built. LUCEM_RL_TILED=0
    forces the per-pass batch, =1 the tiled one where it is admissible."""
    from lucem_amd.batched import RLBatch
    from lucem_amd.restore import RichardsonLucy

    tiled_mode(None)
    b = RLBatch([(300, 280)] * 4, [(9, 9)], [0] * 4, RichardsonLucy(num_iter=3, clip=False, mode=4))
    assert b.launches != _tiled_batch_launches(3, 1)
    b.close()
    s = RichardsonLucy(num_iter=3, clip=False)
    tiled_mode("0")
    b = RLBatch([(512, 512)] * 60, [(15, 15)], [0] * 60, s)
    assert b.launches != _tiled_batch_launches(3, 1)
    b.close()
    tiled_mode("1")
    b = RLBatch([(512, 512)] * 60, [(15, 15)], [0] * 60, s)
    assert b.launches == _tiled_batch_launches(3, 1)
    b.close()
