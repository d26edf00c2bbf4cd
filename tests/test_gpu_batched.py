"""GPU tests of the batched form: many uneven tiles in one graph-captured launch sequence must
give the plain per-tile result (same kernels, same geometry) and match the oracle."""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RTOL_L2 = 1e-4


def _plain(y, psf, settings):
    import torch

    import lucem_amd as LA

    out = torch.empty_like(y)
    LA.richardson_lucy_nd_into(y, psf, out, settings)
    return out


def test_batched_2d_uneven_tiles_match_plain_and_oracle(gpu, oracle, whole_array_rl):
    import torch

    import lucem_amd as LA
    from lucem_amd.batched import RLBatch

    rng = np.random.default_rng(20)
    shapes = [(64, 64), (17, 93), (128, 40), (5, 7), (64, 64), (200, 150), (33, 1), (1, 29)]
    psfs_np = [rng.uniform(0.1, 1.0, k).astype(np.float32) for k in [(5, 5), (7, 3), (11, 11)]]
    index = [0, 1, 2, 0, 2, 1, 0, 2]
    ys = [torch.from_numpy(rng.uniform(50, 150, s).astype(np.float32)).cuda() for s in shapes]
    psfs = [torch.from_numpy(p).cuda() for p in psfs_np]
    s = LA.RichardsonLucy(num_iter=7, clip=False)
    b = RLBatch(shapes, [p.shape for p in psfs_np], index, s)
    outs = [torch.empty_like(y) for y in ys]
    b.execute(ys, outs, psfs)
    torch.cuda.synchronize()
    assert b.launches > 7 * 4
    for y, o, i in zip(ys, outs, index):
        ref = _plain(y, psfs[i], s)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(o.cpu().numpy(), ref.cpu().numpy())
        st, want = oracle.richardson_lucy(y.cpu().numpy(), psfs_np[i], 7, clip=False)
        assert oracle.rel_l2(o.cpu().numpy(), want) < RTOL_L2
    # replay on other buffers (pointers change between calls)
    ys2 = [torch.from_numpy(rng.uniform(10, 20, s).astype(np.float32)).cuda() for s in shapes]
    outs2 = [torch.empty_like(y) for y in ys2]
    b.execute(ys2, outs2, psfs)
    b.execute(ys, outs, psfs)
    torch.cuda.synchronize()
    for y, o, i in zip(ys2, outs2, index):
        st, want = oracle.richardson_lucy(y.cpu().numpy(), psfs_np[i], 7, clip=False)
        assert oracle.rel_l2(o.cpu().numpy(), want) < RTOL_L2
    b.close()


def test_batched_3d_uneven_depths(gpu, oracle):
    import torch

    import lucem_amd as LA
    from lucem_amd.batched import richardson_lucy_batched

    rng = np.random.default_rng(21)
    shapes = [(8, 32, 32), (13, 32, 32), (5, 20, 24), (21, 16, 16)]
    psf = rng.uniform(0.1, 1.0, (5, 3, 3)).astype(np.float32)
    ys = [torch.from_numpy(rng.uniform(50, 150, s).astype(np.float32)).cuda() for s in shapes]
    s = LA.RichardsonLucy(num_iter=5, clip=False, mode=LA.PadMode.Mirror)
    outs = richardson_lucy_batched(ys, [torch.from_numpy(psf).cuda()], [0] * len(shapes), s)
    for y, o in zip(ys, outs):
        st, want = oracle.richardson_lucy(y.cpu().numpy(), psf, 5, clip=False, mode=3, nd=True)
        assert oracle.rel_l2(o.cpu().numpy(), want) < RTOL_L2


def test_batched_clip_and_modes(gpu, oracle):
    import torch

    import lucem_amd as LA
    from lucem_amd.batched import richardson_lucy_batched

    rng = np.random.default_rng(22)
    shapes = [(30, 31), (12, 50)]
    psf = rng.uniform(0.1, 1.0, (4, 6)).astype(np.float32)
    for mode in range(5):
        ys = [torch.from_numpy(rng.uniform(0.2, 0.9, s).astype(np.float32)).cuda() for s in shapes]
        s = LA.RichardsonLucy(num_iter=4, clip=True, mode=mode)
        outs = richardson_lucy_batched(ys, [torch.from_numpy(psf).cuda()], [0, 0], s)
        for y, o in zip(ys, outs):
            st, want = oracle.richardson_lucy(y.cpu().numpy(), psf, 4, clip=True, mode=mode)
            assert oracle.rel_l2(o.cpu().numpy(), want) < RTOL_L2


def test_batched_errors(gpu):
    import torch

    import lucem_amd as LA
    from lucem_amd.batched import RLBatch

    with pytest.raises(LA.LucemError) as e:
        RLBatch([(8, 8)], [(3, 3)], [1])
    assert e.value.kind == "INVALID_PARAMETER"
    with pytest.raises(LA.LucemError) as e:
        RLBatch([(8, 8)], [(3, 3)], [0], LA.RichardsonLucy(num_iter=0))
    assert e.value.kind == "INVALID_PARAMETER"
    b = RLBatch([(8, 8)], [(3, 3)], [0])
    x = torch.ones(8, 8, device="cuda")
    with pytest.raises(LA.LucemError) as e:
        b.execute([x], [x], [torch.ones(3, 3, device="cuda")])
    assert e.value.kind == "ALIASED_BUFFERS"
    with pytest.raises(LA.LucemError) as e:
        b.execute([torch.ones(8, 9, device="cuda")], [torch.ones(8, 9, device="cuda")], [torch.ones(3, 3, device="cuda")])
    assert e.value.kind == "SHAPE_MISMATCH"
    b.close()


def test_tiles_workload_matches_oracle(gpu, oracle):
    """The slide workload at reduced tile size (10 tiles, 2 PSFs): batched vs FP32 oracle."""
    import torch

    import lucem_amd as LA
    from lucem_amd import synth
    from lucem_amd.batched import richardson_lucy_batched

    ts = synth.tiles(channels=2, full=3, edge=2, scale=0.25, num_iter=20)
    ys = [torch.from_numpy(t).cuda() for t in ts.tiles]
    psfs = [torch.from_numpy(p).cuda() for p in ts.psfs]
    s = LA.RichardsonLucy(num_iter=ts.num_iter, clip=False)
    outs = richardson_lucy_batched(ys, psfs, ts.psf_index, s)
    for t, o, i in zip(ts.tiles, outs, ts.psf_index):
        st, want = oracle.richardson_lucy(t, ts.psfs[i], ts.num_iter, clip=False)
        assert oracle.rel_l2(o.cpu().numpy(), want) < RTOL_L2


def test_batched_long_and_short_tiles(gpu, oracle):
    """A 2 x 9000 tile (1024-thread row kernel) and a 9000 x 2 tile (1024-thread line kernel) in one
    batch with ordinary tiles: every launch runs at the widest tier its tiles need."""
    import torch

    import lucem_amd as LA
    from lucem_amd.batched import richardson_lucy_batched

    rng = np.random.default_rng(91)
    shapes = [(2, 9000), (40, 50), (9000, 2), (17, 300), (6000, 4), (5000, 30)]
    ys = [rng.uniform(50, 150, s).astype(np.float32) for s in shapes]
    psf = rng.uniform(0.1, 1.0, (3, 5)).astype(np.float32)
    s = LA.RichardsonLucy(num_iter=3, clip=False)
    outs = richardson_lucy_batched([torch.from_numpy(y).cuda() for y in ys], [torch.from_numpy(psf).cuda()],
                                   [0] * len(ys), s)
    for y, o in zip(ys, outs):
        st, ref = oracle.richardson_lucy(y, psf, 3, clip=False)
        assert st == 0
        assert oracle.rel_l2(o.cpu().numpy(), ref) < 1e-4


def test_batched_with_empty_tiles(gpu, oracle):
    """Empty tiles (a zero extent, as the reference's empty input) are skipped; the others are
    deconvolved as usual."""
    import torch

    import lucem_amd as LA
    from lucem_amd.batched import richardson_lucy_batched

    rng = np.random.default_rng(93)
    shapes = [(0, 16), (20, 24), (5, 0), (13, 9)]
    ys = [torch.from_numpy(rng.uniform(50, 150, s).astype(np.float32)).cuda() for s in shapes]
    psf = rng.uniform(0.1, 1.0, (3, 3)).astype(np.float32)
    outs = richardson_lucy_batched(ys, [torch.from_numpy(psf).cuda()], [0] * 4, LA.RichardsonLucy(num_iter=4, clip=False))
    for y, o in zip(ys, outs):
        assert tuple(o.shape) == tuple(y.shape)
        if y.numel():
            st, want = oracle.richardson_lucy(y.cpu().numpy(), psf, 4, clip=False)
            assert oracle.rel_l2(o.cpu().numpy(), want) < RTOL_L2


def test_batched_graph_replays_match_every_time(gpu):
    """The captured graph replayed five times, with eager work and synchronisations between the
    replays and the inputs changed and restored: every replay equals the first result bit for
    bit (no state left over from an earlier replay)."""
    import torch

    import lucem_amd as LA
    from lucem_amd.batched import RLBatch

    rng = np.random.default_rng(21)
    shapes = [(96, 80), (40, 130), (7, 9), (150, 150)]
    psf = torch.from_numpy(rng.uniform(0.1, 1.0, (9, 9)).astype(np.float32)).cuda()
    ys = [torch.from_numpy(rng.uniform(50, 150, s).astype(np.float32)).cuda() for s in shapes]
    b = RLBatch(shapes, [(9, 9)], [0] * len(shapes), LA.RichardsonLucy(num_iter=6, clip=True))
    outs = [torch.empty_like(y) for y in ys]
    b.execute(ys, outs, [psf])
    torch.cuda.synchronize()
    first = [o.cpu().numpy().copy() for o in outs]
    for rep in range(5):
        junk = torch.randn(1 << 20, device="cuda").sum()  # eager work between replays
        if rep % 2:
            for o in outs:
                o.fill_(-1.0)
        b.execute(ys, outs, [psf])
        torch.cuda.synchronize()
        assert bool(torch.isfinite(junk))
        for o, f in zip(outs, first):
            np.testing.assert_array_equal(o.cpu().numpy(), f)
    b.close()
