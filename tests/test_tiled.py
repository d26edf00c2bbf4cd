"""The tiled driver and the RL pipeline operator (lucem_amd/tiled.py).

CPU: the halo rule, parameter validation, and that the block walk owns every pixel exactly once.
GPU: with the exact halo, the tiled result equals the whole-array result (up to FFT rounding) and
the FP64 oracle on the whole array, in 2-D with several channels (odd and even PSF extents) and in
3-D; a short halo still returns a finite result of the shape.
"""

from __future__ import annotations

import numpy as np
import pytest

from lucem_amd.batched import schedule
from lucem_amd.tiled import RichardsonLucyOp


def test_halo_rule():
    assert RichardsonLucyOp(num_iter=20).halo((21, 21)) == 400
    assert RichardsonLucyOp(num_iter=3).halo((31, 15, 15)) == 90
    assert RichardsonLucyOp(num_iter=1).halo((1, 1)) == 0
    assert RichardsonLucyOp(num_iter=2).halo((4, 6)) == 12  # even k reaches k on one side


def test_validate():
    RichardsonLucyOp(num_iter=1).validate()
    with pytest.raises(ValueError, match="num_iter"):
        RichardsonLucyOp(num_iter=0).validate()
    with pytest.raises(ValueError, match="filter_epsilon"):
        RichardsonLucyOp(filter_epsilon=float("nan")).validate()


@pytest.mark.parametrize("dims,block,halo", [((1, 700, 900), (1, 256, 256), (0, 70, 70)),
                                             ((40, 120, 100), (16, 64, 64), (16, 16, 16)),
                                             ((1, 5, 7), (1, 512, 512), (0, 3, 3))])
def test_blocks_cover_every_pixel_once(dims, block, halo):
    owned = np.zeros(dims, np.int32)
    for b in schedule(dims, block, halo):
        sl = tuple(slice(o, o + s) for o, s in zip(b.origin, b.size))
        owned[sl] += 1
        for a in range(3):
            assert b.read_origin[a] <= b.origin[a]
            assert b.origin[a] + b.size[a] <= b.read_origin[a] + b.read_size[a] <= dims[a]
    assert (owned == 1).all()


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.gpu
def test_tiled_2d_channels_equal_whole_image(gpu):
    import torch

    from lucem_amd import synth
    from lucem_amd.tiled import deconvolve_blocks

    rng = np.random.default_rng(5)
    img = np.stack([rng.uniform(50, 150, (700, 900)), rng.uniform(20, 80, (700, 900))]).astype(np.float32)
    psfs = [synth.gaussian_psf(1.5, (11, 11)).astype(np.float32), rng.uniform(0.1, 1.0, (8, 6)).astype(np.float32)]
    op = RichardsonLucyOp(num_iter=5, clip=False)
    x = torch.from_numpy(img).cuda()
    p = [torch.from_numpy(q).cuda() for q in psfs]
    tiled = deconvolve_blocks(x, p, op, block=(256, 256)).cpu().numpy()
    for c in range(2):
        whole = op.run(x[c].contiguous(), p[c]).cpu().numpy()
        e = _rel(tiled[c], whole)
        print(f"channel {c}: tiled vs whole rel-L2 {e:.2e}")
        assert e < 1e-5


@pytest.mark.gpu
def test_tiled_3d_equal_whole_volume(gpu):
    import torch

    from lucem_amd import synth
    from lucem_amd.tiled import deconvolve_blocks

    rng = np.random.default_rng(6)
    vol = rng.uniform(50, 150, (40, 120, 100)).astype(np.float32)
    psf = synth.gaussian_psf((1.5, 1.0, 1.0), (5, 5, 5)).astype(np.float32)
    op = RichardsonLucyOp(num_iter=4, clip=False)
    x = torch.from_numpy(vol).cuda()
    p = torch.from_numpy(psf).cuda()
    tiled = deconvolve_blocks(x, p, op, block=(16, 64, 64)).cpu().numpy()
    whole = op.run(x, p).cpu().numpy()
    e = _rel(tiled, whole)
    print(f"3-D tiled vs whole rel-L2 {e:.2e}")
    assert e < 1e-5


@pytest.mark.gpu
def test_tiled_matches_the_oracle(gpu, oracle):
    """The block walk (every block of every channel in one batched plan, owned windows stored in
    place) against the FP64 oracle on the whole array: 2-D with two channels (odd and even PSF
    extents) and a 3-D volume."""
    import torch

    from lucem_amd import synth
    from lucem_amd.tiled import deconvolve_blocks

    rng = np.random.default_rng(15)
    img = np.stack([rng.uniform(50, 150, (700, 900)), rng.uniform(20, 80, (700, 900))]).astype(np.float32)
    psfs = [synth.gaussian_psf(1.5, (11, 11)).astype(np.float32), rng.uniform(0.1, 1.0, (8, 6)).astype(np.float32)]
    op = RichardsonLucyOp(num_iter=5, clip=False)
    print("tiled vs oracle: 2 x 700 x 900, blocks 256", flush=True)
    tiled = deconvolve_blocks(torch.from_numpy(img).cuda(), [torch.from_numpy(q).cuda() for q in psfs], op,
                              block=(256, 256)).cpu().numpy()
    for c in range(2):
        e = _rel(tiled[c], oracle.rl_fft64(img[c], psfs[c], 5))
        print(f"channel {c}: tiled vs FP64 oracle rel-L2 {e:.2e}")
        assert e < 1e-4
    vol = rng.uniform(50, 150, (40, 120, 100)).astype(np.float32)
    psf = synth.gaussian_psf((1.5, 1.0, 1.0), (5, 5, 5)).astype(np.float32)
    print("tiled vs oracle: 40 x 120 x 100, blocks 16 x 64 x 64", flush=True)
    tiled = deconvolve_blocks(torch.from_numpy(vol).cuda(), torch.from_numpy(psf).cuda(),
                              RichardsonLucyOp(num_iter=4, clip=False), block=(16, 64, 64)).cpu().numpy()
    e = _rel(tiled, oracle.rl_fft64(vol, psf, 4))
    print(f"3-D tiled vs FP64 oracle rel-L2 {e:.2e}")
    assert e < 1e-4


@pytest.mark.gpu
def test_tiled_short_halo_runs(gpu):
    import torch

    from lucem_amd.tiled import deconvolve_blocks

    rng = np.random.default_rng(7)
    x = torch.from_numpy(rng.uniform(50, 150, (300, 333)).astype(np.float32)).cuda()
    p = torch.from_numpy(rng.uniform(0.1, 1, (7, 7)).astype(np.float32)).cuda()
    out = deconvolve_blocks(x, p, RichardsonLucyOp(num_iter=3, clip=False), block=(128, 128), halo=8)
    assert out.shape == x.shape and bool(torch.isfinite(out).all())


@pytest.mark.gpu
@pytest.mark.parametrize("nd", [2, 3])
def test_layout_views_equal_plain_tiles(gpu, nd, whole_array_rl):
    """Tiles given as strided views into one image, each storing a window into one result, give
    the plain per-tile results inside their windows, and leave the rest of the result alone."""
    import numpy as np
    import torch

    import lucem_amd as LA
    from lucem_amd.batched import RLBatch, TileLayout

    rng = np.random.default_rng(30 + nd)
    full = (9, 40, 52) if nd == 3 else (1, 60, 77)
    img = torch.from_numpy(rng.uniform(50, 150, full).astype(np.float32)).cuda()
    res = torch.full(full, -7.0, device="cuda")
    # (read origin, read size, window origin in the tile, window size): z y x
    tiles = [((0, 0, 0), (full[0], 23, 31), (0, 2, 3), (full[0], 15, 20)),
             ((0, 10, 30), (full[0], 30, 22), (0, 5, 1), (full[0], 21, 20)),
             ((0, 35, 0), (full[0], full[1] - 35, 17), (0, 0, 0), (0, 0, 0))]
    if nd == 3:
        tiles.append(((2, 3, 40), (5, 20, 12), (1, 0, 2), (3, 4, 9)))  # windows do not overlap
    psf = torch.from_numpy(rng.uniform(0.1, 1.0, (3, 5, 3) if nd == 3 else (5, 3)).astype(np.float32)).cuda()
    s = LA.RichardsonLucy(num_iter=4, clip=False)
    shapes = [tuple(size[3 - nd:]) for _, size, _, _ in tiles]
    plan = RLBatch(shapes, [tuple(psf.shape)], [0] * len(tiles), s)
    H, W = full[1], full[2]
    esz = 4
    ins, outs, lays = [], [], []
    for (ro, size, wo, ws) in tiles:
        ins.append(img.data_ptr() + esz * (ro[0] * H * W + ro[1] * W + ro[2]))
        dst = tuple(r + w for r, w in zip(ro, wo))
        outs.append(res.data_ptr() + esz * (dst[0] * H * W + dst[1] * W + dst[2]))
        lays.append(TileLayout(W, H * W, W, H * W, wo, ws))
    plan.execute_layout(ins, outs, lays, [psf])
    torch.cuda.synchronize()
    covered = torch.zeros(full, dtype=torch.bool, device="cuda")
    for (ro, size, wo, ws) in tiles:
        view = img[ro[0]:ro[0] + size[0], ro[1]:ro[1] + size[1], ro[2]:ro[2] + size[2]].contiguous()
        if nd == 2:
            view = view[0]
        ref = torch.empty_like(view)
        LA.richardson_lucy_nd_into(view, psf, ref, s)
        torch.cuda.synchronize()
        ref3 = ref if nd == 3 else ref.unsqueeze(0)
        wsz = [w if w else size[a] - wo[a] for a, w in enumerate(ws)]
        sl = tuple(slice(ro[a] + wo[a], ro[a] + wo[a] + wsz[a]) for a in range(3))
        want = ref3[tuple(slice(wo[a], wo[a] + wsz[a]) for a in range(3))]
        np.testing.assert_array_equal(res[sl].cpu().numpy(), want.cpu().numpy())
        covered[sl] = True
    assert bool((res[~covered] == -7.0).all())
    plan.close()


@pytest.mark.gpu
def test_layout_errors(gpu):
    import torch

    import lucem_amd as LA
    from lucem_amd.batched import RLBatch, TileLayout

    img = torch.ones((30, 40), device="cuda")
    res = torch.zeros((30, 40), device="cuda")
    psf = torch.ones((3, 3), device="cuda")
    plan = RLBatch([(10, 12)], [(3, 3)], [0])
    with pytest.raises(LA.LucemError) as e:  # the window leaves the tile
        plan.execute_layout([img.data_ptr()], [res.data_ptr()], [TileLayout(40, 0, 40, 0, (0, 5, 0), (0, 6, 12))], [psf])
    assert e.value.kind == "INVALID_PARAMETER"
    with pytest.raises(LA.LucemError) as e:  # a row stride shorter than the tile's rows
        plan.execute_layout([img.data_ptr()], [res.data_ptr()], [TileLayout(8, 0, 40, 0)], [psf])
    assert e.value.kind == "INVALID_PARAMETER"
    with pytest.raises(LA.LucemError) as e:  # the stored window writes into the image being read
        plan.execute_layout([img.data_ptr()], [img.data_ptr() + 4 * 45], [TileLayout(40, 0, 40, 0)], [psf])
    assert e.value.kind == "ALIASED_BUFFERS"
    plan.close()

