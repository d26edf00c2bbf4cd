"""The whole-slide driver on the device: an OME-Zarr image on disk, read in blocks with their halo,
deconvolved in batched groups, written block by block to a new store, compared with the FP64 oracle
on the whole image (crates/lucem-ops/src/restore/deconvolve.rs:45-149; the walk of
engine/lucem-pipeline/src/custom.rs:233-251) and with the in-memory block walk."""

from __future__ import annotations

import threading

import numpy as np
import pytest

from lucem_amd.batched import schedule
from lucem_amd.slide import Pyramid, RunCancelled, SlideStats, WriterOptions, ZarrWriter, deconvolve_slide
from lucem_amd.tiled import RichardsonLucyOp

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def _store(root, img: np.ndarray, chunks, px=0.5):
    """``img`` as ``[c, z, y, x]`` float32 written to a new store."""
    shape = img.shape[1:]
    w = ZarrWriter.create(root, WriterOptions("slide", [f"c{c}" for c in range(img.shape[0])], shape, chunks, px, px))
    for b in schedule(shape, chunks, (0, 0, 0)):
        o, s = b.origin, b.size
        for c in range(img.shape[0]):
            w.write(c, o, s, img[c, o[0]:o[0] + s[0], o[1]:o[1] + s[1], o[2]:o[2] + s[2]])
    return w.finish()


def test_slide_2d_channels_match_the_oracle(gpu, oracle, tmp_path):
    from lucem_amd import synth

    rng = np.random.default_rng(31)
    img = np.stack([rng.uniform(50, 150, (1, 520, 610)), rng.uniform(20, 80, (1, 520, 610))]).astype(np.float32)
    psfs = [synth.gaussian_psf(1.5, (11, 11)).astype(np.float32), rng.uniform(0.1, 1.0, (8, 6)).astype(np.float32)]
    src = _store(tmp_path / "in.ome.zarr", img, (1, 128, 128))
    op = RichardsonLucyOp(num_iter=5, clip=False)
    seen = []
    st = SlideStats()
    print("slide: 2 x 520 x 610, blocks 128, groups of about 100k voxels", flush=True)
    out = deconvolve_slide(src, tmp_path / "out.ome.zarr", psfs, op, block=128, group_voxels=100_000,
                           progress=lambda d, t: seen.append((d, t)), stats=st)
    py = Pyramid.open(out)
    assert py.channels == ["c0", "c1"] and (py.height(), py.width()) == (520, 610)
    assert py.levels[0].tile_h == 128 and len(py.levels) > 1
    assert st.blocks == 2 * 5 * 5 and st.groups > 1 and seen[-1] == (st.blocks, st.blocks)
    assert all(a[0] < b[0] for a, b in zip(seen, seen[1:]))
    for c in range(2):
        got = py.read_block(0, c, (0, 0, 0), (1, 520, 610))[0]
        e = _rel(got, oracle.rl_fft64(img[c, 0], psfs[c], 5))
        print(f"channel {c}: slide vs FP64 oracle rel-L2 {e:.2e}")
        assert e < 1e-4


def test_slide_equals_the_in_memory_walk(gpu, tmp_path):
    import torch

    from lucem_amd import synth
    from lucem_amd.tiled import deconvolve_blocks

    rng = np.random.default_rng(32)
    img = rng.uniform(50, 150, (3, 1, 300, 400)).astype(np.float32)
    psf = synth.gaussian_psf(2.0, (9, 9)).astype(np.float32)
    src = _store(tmp_path / "in", img, (1, 100, 100))
    op = RichardsonLucyOp(num_iter=4, clip=False)
    # one group: the same tiles in one plan as the in-memory walk
    out = deconvolve_slide(src, tmp_path / "out", psf, op, block=100, channels=[2, 0])
    py = Pyramid.open(out)
    assert py.channels == ["c2", "c0"]
    mem = deconvolve_blocks(torch.from_numpy(img[[2, 0], 0]).cuda(), torch.from_numpy(psf).cuda(), op,
                            block=(100, 100)).cpu().numpy()
    for i in range(2):
        got = py.read_block(0, i, (0, 0, 0), (1, 300, 400))[0]
        e = _rel(got, mem[i])
        print(f"channel {i}: slide vs in-memory walk rel-L2 {e:.2e}")
        assert e < 1e-6


def test_slide_volume_matches_the_oracle(gpu, oracle, tmp_path):
    from lucem_amd import synth

    rng = np.random.default_rng(33)
    vol = rng.uniform(50, 150, (1, 40, 120, 100)).astype(np.float32)
    psf = synth.gaussian_psf((1.5, 1.0, 1.0), (5, 5, 5)).astype(np.float32)
    src = _store(tmp_path / "v", vol, (16, 64, 64))
    print("slide: volume 40 x 120 x 100, blocks 16 x 64 x 64", flush=True)
    out = deconvolve_slide(src, tmp_path / "vo", psf, RichardsonLucyOp(num_iter=4, clip=False), block=64, block_z=16,
                           group_voxels=200_000)
    got = Pyramid.open(out).read_block(0, 0, (0, 0, 0), (40, 120, 100))
    e = _rel(got, oracle.rl_fft64(vol[0], psf, 4))
    print(f"3-D slide vs FP64 oracle rel-L2 {e:.2e}")
    assert e < 1e-4


def test_slide_cancel_and_errors(gpu, tmp_path):
    from lucem_amd.slide import RunError

    img = np.full((1, 1, 256, 256), 100.0, np.float32)
    src = _store(tmp_path / "in", img, (1, 64, 64))
    psf = np.ones((5, 5), np.float32)
    ev = threading.Event()
    ev.set()
    with pytest.raises(RunCancelled):
        deconvolve_slide(src, tmp_path / "c", psf, RichardsonLucyOp(num_iter=2), block=64, cancel=ev)
    with pytest.raises(RunError, match="2 PSFs for 1 channels"):
        deconvolve_slide(src, tmp_path / "e1", [psf, psf], RichardsonLucyOp(num_iter=2))
    with pytest.raises(RunError, match="needs a volume"):
        deconvolve_slide(src, tmp_path / "e2", np.ones((3, 3, 3), np.float32), RichardsonLucyOp(num_iter=2))
    with pytest.raises(ValueError, match="num_iter"):
        deconvolve_slide(src, tmp_path / "e3", psf, RichardsonLucyOp(num_iter=0))
    # a flat image under a normalised PSF stays flat (Reflect): the fixed point, through the store
    out = deconvolve_slide(src, tmp_path / "flat", psf, RichardsonLucyOp(num_iter=3, clip=False), block=64)
    got = Pyramid.open(out).read_block(0, 0, (0, 0, 0), (1, 256, 256))
    np.testing.assert_allclose(got, 100.0, rtol=1e-5)


def test_slide_coarser_level_and_channel_subset(gpu, oracle, tmp_path):
    """A walk over level 1 of a store (the writer's 2x2 means) for two of three channels in reverse
    order, each with its own PSF, against the FP64 oracle on that level's image."""
    rng = np.random.default_rng(34)
    img = rng.uniform(50, 150, (3, 1, 400, 330)).astype(np.float32)
    src = _store(tmp_path / "in", img, (1, 128, 128))
    py = Pyramid.open(src)
    assert (py.levels[1].height, py.levels[1].width) == (200, 165)
    psfs = [rng.uniform(0.1, 1.0, (7, 5)).astype(np.float32), rng.uniform(0.1, 1.0, (4, 6)).astype(np.float32)]
    out = deconvolve_slide(py, tmp_path / "out", psfs, RichardsonLucyOp(num_iter=4, clip=False), level=1,
                           channels=[2, 0], block=64, group_voxels=30_000)
    po = Pyramid.open(out)
    assert po.channels == ["c2", "c0"] and (po.height(), po.width()) == (200, 165)
    assert po.pixel_size_um == 1.0  # level 1's pixels are twice level 0's 0.5 um
    for i, c in enumerate((2, 0)):
        lvl = py.read_block(1, c, (0, 0, 0), (1, 200, 165))[0]
        got = po.read_block(0, i, (0, 0, 0), (1, 200, 165))[0]
        e = _rel(got, oracle.rl_fft64(lvl, psfs[i], 4))
        print(f"level 1, channel {c}: slide vs FP64 oracle rel-L2 {e:.2e}")
        assert e < 1e-4
