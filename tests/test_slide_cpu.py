"""The on-disk side of the slide driver, without a device: the OME-Zarr writer and reader, the block
loader and the walk's host logic.

The writer's cases follow the reference's own tests (crates/lucem-io/src/zarr_writer.rs:352-442:
2-D and 3-D stores read back exactly, coarser levels built, an existing path never replaced, blocks
kept on the chunk grid) and the loader's (crates/lucem-io/src/loader.rs:127-155: blocks in walk
order whatever thread read them, an early close stops the threads). The coarser levels are pinned
by their closed form: each pixel the float32 mean of the 2 x 2 (x 2) pixels it covers, summed in
z, y, x order, fewer at odd edges.
"""

from __future__ import annotations

import gzip
import json
import threading
import zlib

import numpy as np
import pytest

from lucem_amd.batched import schedule
from lucem_amd.slide import (BlockLoader, IoError, Pyramid, RunError, RunOptions, WriterOptions, ZarrWriter)


def _write_store(root, shape, chunks, channels=2, px=0.5, dz=0.5, dtype=np.float32):
    n = shape[0] * shape[1] * shape[2]
    values = np.arange(n, dtype=np.float64).reshape(shape)
    opts = WriterOptions("t", [f"ch{c}" for c in range(channels)], shape, chunks, px, dz)
    w = ZarrWriter.create(root, opts, dtype)
    for b in schedule(shape, chunks, (0, 0, 0)):
        for c in range(channels):
            o, s = b.origin, b.size
            part = values[o[0]:o[0] + s[0], o[1]:o[1] + s[1], o[2]:o[2] + s[2]] + c
            w.write(c, o, s, part.astype(dtype))
    return w.finish(), values


@pytest.mark.parametrize("shape,chunks", [((1, 150, 100), (1, 64, 64)), ((20, 90, 70), (8, 32, 32))])
def test_store_reads_back_exactly(tmp_path, shape, chunks):
    path, values = _write_store(tmp_path / "s.ome.zarr", shape, chunks)
    py = Pyramid.open(path)
    assert (py.depth(), py.height(), py.width()) == shape
    assert py.channels == ["ch0", "ch1"]
    assert py.pixel_size_um == 0.5
    assert len(py.levels) > 1, "coarser levels are built"
    got = py.read_block(0, 1, (0, 0, 0), shape)
    np.testing.assert_array_equal(got, (values + 1).astype(np.float32))
    # a box across chunk borders, from a path inside the store
    py2 = Pyramid.open(path / "0")
    o, s = (shape[0] // 3, 17, 29), (max(shape[0] // 2, 1), 70, 40)
    np.testing.assert_array_equal(py2.read_block(0, 0, o, s),
                                  values[o[0]:o[0] + s[0], o[1]:o[1] + s[1], o[2]:o[2] + s[2]].astype(np.float32))


def _mean_reduce(src: np.ndarray, zf: int) -> np.ndarray:
    """Closed form of a coarser level: float32 sum over the covered pixels in (z, y, x) order over
    their count."""
    d, h, w = src.shape
    out = np.zeros((-(-d // zf), -(-h // 2), -(-w // 2)), np.float32)
    for z in range(out.shape[0]):
        for y in range(out.shape[1]):
            for x in range(out.shape[2]):
                acc, n = np.float32(0), 0
                for dz in range(zf):
                    for dy in range(2):
                        for dx in range(2):
                            zz, yy, xx = z * zf + dz, y * 2 + dy, x * 2 + dx
                            if zz < d and yy < h and xx < w:
                                acc = np.float32(acc + src[zz, yy, xx])
                                n += 1
                out[z, y, x] = np.float32(acc / np.float32(n))
    return out


def test_coarser_levels_are_the_mean_of_what_they_cover(tmp_path):
    rng = np.random.default_rng(1)
    shape, chunks = (1, 37, 29), (1, 8, 8)
    img = rng.uniform(0, 100, shape).astype(np.float32)
    w = ZarrWriter.create(tmp_path / "m", WriterOptions("m", ["a"], shape, chunks))
    for b in schedule(shape, chunks, (0, 0, 0)):
        o, s = b.origin, b.size
        w.write(0, o, s, img[:, o[1]:o[1] + s[1], o[2]:o[2] + s[2]])
    py = Pyramid.open(w.finish())
    # 37 x 29 -> 19 x 15 -> 10 x 8: stops once the larger side fits one chunk side
    assert [(lv.height, lv.width) for lv in py.levels] == [(37, 29), (19, 15), (10, 8), (5, 4)]
    prev = img
    for lv in range(1, len(py.levels)):
        want = _mean_reduce(prev, 1)
        got = py.read_block(lv, 0, (0, 0, 0), want.shape)
        np.testing.assert_array_equal(got, want)
        prev = got
    assert py.levels[1].downsample == 29 / 15


def test_volume_levels_halve_z_while_isotropic_enough(tmp_path):
    shape, chunks = (9, 40, 40), (4, 16, 16)
    rng = np.random.default_rng(2)
    vol = rng.uniform(0, 10, shape).astype(np.float32)
    opts = WriterOptions("v", ["a"], shape, chunks, pixel_size_um=1.0, z_spacing_um=1.0)
    w = ZarrWriter.create(tmp_path / "v", opts)
    for b in schedule(shape, chunks, (0, 0, 0)):
        o, s = b.origin, b.size
        w.write(0, o, s, vol[o[0]:o[0] + s[0], o[1]:o[1] + s[1], o[2]:o[2] + s[2]])
    py = Pyramid.open(w.finish())
    assert [(lv.depth, lv.height) for lv in py.levels] == [(9, 40), (5, 20), (3, 10)]
    np.testing.assert_array_equal(py.read_block(1, 0, (0, 0, 0), (5, 20, 20)), _mean_reduce(vol, 2))
    # z kept when its spacing is far coarser than the pixels
    opts = WriterOptions("v", ["a"], shape, chunks, pixel_size_um=0.25, z_spacing_um=2.0)
    w = ZarrWriter.create(tmp_path / "v2", opts)
    for b in schedule(shape, chunks, (0, 0, 0)):
        o, s = b.origin, b.size
        w.write(0, o, s, vol[o[0]:o[0] + s[0], o[1]:o[1] + s[1], o[2]:o[2] + s[2]])
    py = Pyramid.open(w.finish())
    assert [lv.depth for lv in py.levels] == [9, 9, 9]
    assert py.z_spacing_um == 2.0


def test_metadata_is_ome_zarr_v04(tmp_path):
    path, _ = _write_store(tmp_path / "s", (1, 150, 100), (1, 64, 64))
    attrs = json.loads((path / ".zattrs").read_text())
    ms = attrs["multiscales"][0]
    assert ms["version"] == "0.4" and ms["name"] == "t"
    assert [a["name"] for a in ms["axes"]] == ["c", "y", "x"]
    assert ms["axes"][1]["unit"] == "micrometer"
    assert [d["path"] for d in ms["datasets"]] == ["0", "1", "2"]
    assert ms["datasets"][1]["coordinateTransformations"][0]["scale"] == [1.0, 1.0, 1.0]
    assert attrs["omero"]["channels"] == [{"label": "ch0", "active": True}, {"label": "ch1", "active": True}]
    za = json.loads((path / "0" / ".zarray").read_text())
    assert za == {"zarr_format": 2, "shape": [2, 150, 100], "chunks": [1, 64, 64], "dtype": "<f4",
                  "compressor": None, "fill_value": 0, "order": "C", "filters": None, "dimension_separator": "/"}
    # edge chunks are whole chunks, zero-padded
    assert (path / "0" / "1" / "2" / "1").stat().st_size == 64 * 64 * 4


def test_existing_path_is_never_replaced_and_blocks_stay_on_the_grid(tmp_path):
    opts = WriterOptions("t", ["a"], (1, 10, 10), (1, 4, 4))
    w = ZarrWriter.create(tmp_path / "e", opts, np.uint32)
    with pytest.raises(IoError, match="already exists"):
        ZarrWriter.create(tmp_path / "e", opts, np.uint32)
    with pytest.raises(IoError, match="chunk grid"):
        w.write(0, (0, 1, 0), (1, 4, 4), np.zeros(16, np.uint32))
    with pytest.raises(IoError, match="whole chunk"):
        w.write(0, (0, 0, 0), (1, 4, 4), np.zeros(3, np.uint32))
    with pytest.raises(IoError, match="whole chunk"):
        w.write(0, (0, 8, 8), (1, 4, 4), np.zeros(16, np.uint32))
    with pytest.raises(IoError, match="element type"):
        w.write(0, (0, 0, 0), (1, 4, 4), np.zeros(16, np.float32))
    with pytest.raises(IoError, match="no channel"):
        w.write(1, (0, 0, 0), (1, 4, 4), np.zeros(16, np.uint32))
    with pytest.raises(IoError, match="non-empty"):
        ZarrWriter.create(tmp_path / "z", WriterOptions("t", [], (1, 4, 4), (1, 4, 4)))


def test_labels_are_sampled_not_averaged(tmp_path):
    shape, chunks = (1, 6, 6), (1, 4, 4)
    lab = np.arange(36, dtype=np.uint32).reshape(shape)
    w = ZarrWriter.create(tmp_path / "l", WriterOptions("l", ["a"], shape, chunks), np.uint32)
    for b in schedule(shape, chunks, (0, 0, 0)):
        o, s = b.origin, b.size
        w.write(0, o, s, lab[:, o[1]:o[1] + s[1], o[2]:o[2] + s[2]])
    py = Pyramid.open(w.finish())
    np.testing.assert_array_equal(py.read_block(1, 0, (0, 0, 0), (1, 3, 3)), lab[:, ::2, ::2].astype(np.float32))


def _blosc_frame(data: bytes, typesize: int, blocksize: int, shuffle: bool = True, code: int = 1) -> bytes:
    """A Blosc format-1 frame as c-blosc lays one out: header, block offsets, per block the byte
    planes of the shuffled block as separate streams (one stream for the last, partial block),
    each stream compressed by the real codec (pyarrow's LZ4 block, zstd, snappy; zlib), stored raw
    when that does not make it smaller."""
    import struct

    import pyarrow as pa

    codec = {1: "lz4_raw", 2: "snappy", 3: "zlib", 4: "zstd"}[code]

    def comp(x: bytes) -> bytes:
        return zlib.compress(x) if codec == "zlib" else pa.compress(x, codec=codec, asbytes=True)

    nblocks = -(-len(data) // blocksize) if data else 0
    body, starts = b"", []
    head = 16 + 4 * nblocks
    for b in range(nblocks):
        blk = np.frombuffer(data[b * blocksize:(b + 1) * blocksize], np.uint8)
        n = len(blk)
        if shuffle and typesize > 1:
            ne = n // typesize
            blk = np.concatenate([blk[:ne * typesize].reshape(ne, typesize).T.reshape(-1), blk[ne * typesize:]])
        split = typesize > 1 and n == blocksize and n % typesize == 0
        streams = np.split(blk, typesize) if split else [blk]
        starts.append(head + len(body))
        for st in streams:
            c = comp(st.tobytes())
            if len(c) >= len(st):
                c = st.tobytes()
            body += struct.pack("<I", len(c)) + c
    flags = (0x01 if shuffle else 0) | (code << 5)
    hdr = bytes([2, 1, flags, typesize]) + struct.pack("<III", len(data), blocksize, head + len(body))
    return hdr + struct.pack(f"<{nblocks}I", *starts) + body


def _encode(comp: dict | None, raw: bytes, typesize: int) -> bytes:
    import pyarrow as pa

    cid = None if comp is None else comp["id"]
    if cid is None:
        return raw
    if cid == "zlib":
        return zlib.compress(raw)
    if cid == "gzip":
        return gzip.compress(raw)
    if cid == "zstd":
        return pa.compress(raw, codec="zstd", asbytes=True)
    if cid == "lz4":
        return len(raw).to_bytes(4, "little") + pa.compress(raw, codec="lz4_raw", asbytes=True)
    code = {"lz4": 1, "snappy": 2, "zlib": 3, "zstd": 4}.get(comp.get("cname", "lz4"))
    if cid != "blosc" or code is None or comp.get("shuffle", 1) == 2:
        return raw  # a store the reader refuses when it opens it
    return _blosc_frame(raw, typesize, comp.get("_blocksize", 1024), bool(comp.get("shuffle", 1)), code)


def _hand_store(root, arr, chunks, sep, compressor, fill=0, axes=("c", "y", "x"), drop=()):
    root.mkdir()
    (root / ".zgroup").write_text('{"zarr_format": 2}')
    (root / ".zattrs").write_text(json.dumps({"multiscales": [{"version": "0.4", "axes": list(axes),
                                                               "datasets": [{"path": "0"}]}]}))
    d = root / "0"
    d.mkdir()
    comp = {"id": compressor} if isinstance(compressor, str) else compressor
    meta = {"zarr_format": 2, "shape": list(arr.shape), "chunks": list(chunks), "dtype": arr.dtype.str,
            "compressor": None if comp is None else {k: v for k, v in comp.items() if not k.startswith("_")},
            "fill_value": fill, "order": "C", "filters": None}
    if sep is not None:
        meta["dimension_separator"] = sep
    (d / ".zarray").write_text(json.dumps(meta))
    grid = [-(-s // c) for s, c in zip(arr.shape, chunks)]
    for idx in np.ndindex(*grid):
        if idx in drop:
            continue
        full = np.zeros(chunks, arr.dtype)
        sl = tuple(slice(i * c, min((i + 1) * c, s)) for i, c, s in zip(idx, chunks, arr.shape))
        part = arr[sl]
        full[tuple(slice(0, p) for p in part.shape)] = part
        raw = _encode(comp, full.tobytes(), arr.dtype.itemsize)
        p = d / (sep or ".").join(str(i) for i in idx)
        p.parent.mkdir(parents=True, exist_ok=True)
        p.write_bytes(raw)


@pytest.mark.parametrize("sep,compressor,dtype", [(".", "zlib", ">u2"), ("/", "gzip", "<i4"), (None, None, "|u1"),
                                                  ("/", None, "<f8")])
def test_reader_layouts(tmp_path, sep, compressor, dtype):
    rng = np.random.default_rng(3)
    arr = rng.integers(0, 200, (2, 45, 38)).astype(dtype)
    _hand_store(tmp_path / "h", arr, (1, 16, 16), sep, compressor, fill=7, drop={(1, 1, 1)})
    py = Pyramid.open(tmp_path / "h")
    assert py.channels == ["Channel 0", "Channel 1"]
    want = arr.astype(np.float32)
    want[1, 16:32, 16:32] = 7  # the missing chunk holds the fill value
    np.testing.assert_array_equal(py.read_block(0, 1, (0, 0, 0), (1, 45, 38))[0], want[1])
    np.testing.assert_array_equal(py.read_block(0, 0, (0, 5, 9), (1, 30, 20))[0], want[0, 5:35, 9:29])


@pytest.mark.parametrize("comp", [{"id": "zstd"}, {"id": "lz4"},
                                  {"id": "blosc", "cname": "lz4", "shuffle": 1, "_blocksize": 1000},
                                  {"id": "blosc", "cname": "zstd", "shuffle": 1, "_blocksize": 4096},
                                  {"id": "blosc", "cname": "zlib", "shuffle": 0, "_blocksize": 512},
                                  {"id": "blosc", "cname": "snappy", "shuffle": 1, "_blocksize": 333}])
@pytest.mark.parametrize("dtype", ["<f4", "<u2", "|u1"])
def test_reader_codecs(tmp_path, comp, dtype):
    """Stores whose chunks carry real LZ4, zstd, snappy and zlib payloads (pyarrow's and zlib's
    encoders), in Blosc frames with shuffled, split streams and partial last blocks, read back
    exactly."""
    rng = np.random.default_rng(4)
    arr = (rng.integers(0, 60, (2, 45, 38)) * (rng.random((2, 45, 38)) < 0.7)).astype(dtype)
    _hand_store(tmp_path / "h", arr, (1, 16, 16), "/", comp, drop={(0, 2, 2)})
    py = Pyramid.open(tmp_path / "h")
    want = arr.astype(np.float32)
    want[0, 32:45, 32:38] = 0
    for c in range(2):
        np.testing.assert_array_equal(py.read_block(0, c, (0, 0, 0), (1, 45, 38))[0], want[c])


def test_blosc_known_frame():
    """A frame laid out by hand: 16 bytes as 4 elements of 4 bytes, shuffled into 4 byte planes,
    each plane a raw 4-byte stream (its size equals its length), LZ4 code and shuffle flag."""
    import struct

    from lucem_amd.slide import blosc_decompress

    planes = [bytes([0, 4, 8, 12]), bytes([1, 5, 9, 13]), bytes([2, 6, 10, 14]), bytes([3, 7, 11, 15])]
    body = b"".join(struct.pack("<I", 4) + p for p in planes)
    frame = bytes([2, 1, 0x01 | (1 << 5), 4]) + struct.pack("<III", 16, 16, 20 + len(body)) + \
        struct.pack("<I", 20) + body
    assert blosc_decompress(frame, 16) == bytes(range(16))
    # the same bytes stored as they are (flag 0x02)
    stored = bytes([2, 1, 0x02, 4]) + struct.pack("<III", 16, 16, 32) + bytes(range(16))
    assert blosc_decompress(stored, 16) == bytes(range(16))
    # a block of 10 bytes of 4-byte elements keeps its 2 leftover bytes in place
    data = bytes(range(10))
    f = _blosc_frame(data, 4, 16)
    assert blosc_decompress(f, 10) == data


@pytest.mark.parametrize("typesize", [1, 2, 4, 8])
@pytest.mark.parametrize("code", [1, 2, 3, 4])
def test_blosc_round_trip(typesize, code):
    from lucem_amd.slide import blosc_decompress

    rng = np.random.default_rng(typesize * 10 + code)
    data = (rng.integers(0, 5, 5000) * 37).astype(np.uint8).tobytes() + rng.bytes(123)
    for blocksize in (64, 1000, 4096, 8192):
        bs = blocksize - blocksize % typesize
        assert blosc_decompress(_blosc_frame(data, typesize, bs, True, code), len(data)) == data
        assert blosc_decompress(_blosc_frame(data, typesize, bs, False, code), len(data)) == data


def test_blosc_refusals():
    import struct

    from lucem_amd.slide import blosc_decompress

    good = _blosc_frame(bytes(range(200)) * 10, 4, 512)
    with pytest.raises(IoError, match="header"):
        blosc_decompress(good[:10], 2000)
    with pytest.raises(IoError, match="needs 2001"):
        blosc_decompress(good, 2001)
    with pytest.raises(IoError, match="frame size"):
        blosc_decompress(good[:-1], 2000)
    bit = bytes([2, 1, good[2] | 0x04]) + good[3:]
    with pytest.raises(IoError, match="bit shuffle"):
        blosc_decompress(bit, 2000)
    blz = bytes([2, 1, good[2] & 0x1F]) + good[3:]
    with pytest.raises(IoError, match="BloscLZ"):
        blosc_decompress(blz, 2000)
    trunc = bytes([2, 1, good[2], 4]) + struct.pack("<III", 2000, 512, 18) + b"\0\0"
    with pytest.raises(IoError, match="table"):
        blosc_decompress(trunc, 2000)


def test_reader_refusals(tmp_path):
    arr = np.zeros((1, 8, 8), np.float32)
    _hand_store(tmp_path / "b", arr, (1, 8, 8), "/", {"id": "bz2"})
    with pytest.raises(IoError, match="bz2"):
        Pyramid.open(tmp_path / "b")
    _hand_store(tmp_path / "bl", arr, (1, 8, 8), "/", {"id": "blosc", "cname": "blosclz"})
    with pytest.raises(IoError, match="blosclz"):
        Pyramid.open(tmp_path / "bl")
    _hand_store(tmp_path / "bs", arr, (1, 8, 8), "/", {"id": "blosc", "cname": "lz4", "shuffle": 2})
    with pytest.raises(IoError, match="bit shuffle"):
        Pyramid.open(tmp_path / "bs")
    _hand_store(tmp_path / "ok", arr, (1, 8, 8), "/", None)
    py = Pyramid.open(tmp_path / "ok")
    with pytest.raises(IoError, match="outside level"):
        py.read_block(0, 0, (0, 4, 4), (1, 8, 8))
    with pytest.raises(IoError, match="no channel"):
        py.read_block(0, 1, (0, 0, 0), (1, 8, 8))
    with pytest.raises(IoError, match="no level"):
        py.read_block(1, 0, (0, 0, 0), (1, 8, 8))
    with pytest.raises(IoError, match="not an OME-Zarr"):
        Pyramid.open(tmp_path)
    (tmp_path / "ok" / ".zattrs").write_text(json.dumps({"multiscales": [{"datasets": [{"path": "../x"}]}]}))
    with pytest.raises(IoError, match="leaves the container"):
        Pyramid.open(tmp_path / "ok")


def test_blocks_arrive_in_order_whatever_thread_read_them(tmp_path):
    path, _ = _write_store(tmp_path / "s", (4, 64, 64), (2, 16, 16))
    py = Pyramid.open(path)
    blocks = schedule((4, 64, 64), (2, 16, 16), (0, 2, 2))
    with BlockLoader(py, 0, [0, 1], blocks, read_ahead=3, threads=4) as loader:
        assert loader.total == len(blocks) * 2
        n = 0
        for n, got in enumerate(loader):
            assert got.block == blocks[n // 2] and got.channel == n % 2
            np.testing.assert_array_equal(got.data, py.read_block(0, got.channel, got.block.read_origin,
                                                                  got.block.read_size))
        assert n + 1 == loader.total


def test_read_ahead_is_bounded_and_early_close_stops_threads(tmp_path):
    path, _ = _write_store(tmp_path / "s", (1, 64, 64), (1, 8, 8), channels=1)
    py = Pyramid.open(path)
    live, peak = [0], [0]
    lock = threading.Lock()
    orig = py.read_block

    def counted(*a):
        with lock:
            live[0] += 1
            peak[0] = max(peak[0], live[0])
        try:
            return orig(*a)
        finally:
            with lock:
                live[0] -= 1

    py.read_block = counted
    blocks = schedule((1, 64, 64), (1, 8, 8), (0, 0, 0))
    loader = BlockLoader(py, 0, [0], blocks, read_ahead=2, threads=8)
    first = next(loader)
    assert first.block == blocks[0]
    workers = list(loader._workers)
    assert len(workers) == 8
    loader.close()
    assert not any(w.is_alive() for w in workers)
    assert peak[0] <= 3  # reads in flight never pass the read-ahead window (plus the one handed out)
    assert loader._next_job <= 1 + 2


def test_loader_hands_read_errors_to_the_consumer(tmp_path):
    path, _ = _write_store(tmp_path / "s", (1, 32, 32), (1, 16, 16), channels=1)
    py = Pyramid.open(path)
    bad = schedule((1, 64, 64), (1, 16, 16), (0, 0, 0))  # blocks of a larger image: outside the level
    with BlockLoader(py, 0, [0], bad, read_ahead=2, threads=2) as loader:
        assert next(loader).block == bad[0]
        next(loader)
        with pytest.raises(IoError, match="outside level"):
            next(loader)


def test_walk_options(tmp_path):
    path, _ = _write_store(tmp_path / "s", (1, 600, 700), (1, 128, 128), channels=2)
    py = Pyramid.open(path)
    grid, blocks = RunOptions(halo=10).blocks(py)
    assert grid == (1, 256, 256)  # the chunk side, at least 256
    assert len(blocks) == 3 * 3 and blocks[0].read_size == (1, 266, 266)
    grid, blocks = RunOptions(block=300, halo=5).blocks(py)
    assert grid == (1, 300, 300) and len(blocks) == 2 * 3
    with pytest.raises(RunError, match="positive"):
        RunOptions(block=0).blocks(py)
    with pytest.raises(RunError, match="no level 9"):
        RunOptions(level=9).blocks(py)
    assert RunOptions().channel_list(py) == [0, 1]
    with pytest.raises(RunError, match="no channel 5"):
        RunOptions(channels=[5]).channel_list(py)
    with pytest.raises(RunError, match="no channels"):
        RunOptions(channels=[]).channel_list(py)
    wo = RunOptions(block=1000).writer_options(py, ["a"], (1, 1000, 1000), (1, 600, 700))
    assert wo.chunks == (1, 600, 700) and wo.pixel_size_um == 0.5
    v_path, _ = _write_store(tmp_path / "v", (40, 64, 64), (8, 32, 32), channels=1)
    grid, blocks = RunOptions(walk="volume", halo=3).blocks(Pyramid.open(v_path))
    assert grid == (32, 256, 256) and blocks[0].read_size == (35, 64, 64)
