"""The on-disk side of the whole-slide Richardson-Lucy driver.

An image on disk goes through RL block by block, as lucem-pipeline's ``custom::transform`` walks it
(engine/lucem-pipeline/src/custom.rs:64-251): blocks with a halo of context are read ahead of the
work on worker threads (``BlockLoader``, crates/lucem-io/src/loader.rs:37-125), deconvolved on the
device, and only the part each block owns is written to a new OME-Zarr store, one chunk at a time
(``ZarrWriter``, crates/lucem-io/src/zarr_writer.rs:87-345), whose coarser levels are built from the
chunks on disk when the walk ends.

The walk differs from the reference's in how the blocks reach the device. The reference uploads
each block, calls the function on it and downloads the owned part before the next block. Here the
blocks that the loader has ready are gathered into groups of a few tens of millions of voxels; a
group goes up in one copy, runs as one batched plan (every block of the group in one captured
graph, the owned window stored by the last iteration's update, so no crop pass), comes back in one
copy, and its chunks are written on worker threads while the next group is read and deconvolved.
Plans are kept for group compositions that repeat (the interior blocks of a regular grid).

The reader (``Pyramid``) covers the stores this driver and the reference's writer make and the
common Zarr v2 layouts: ``.zarray`` arrays of unsigned, signed or floating samples of either
byte order, ``.`` or ``/`` chunk keys, fill values for missing chunks, axes ``[c,] [z,] y, x``, and
chunks stored raw, with zlib, gzip, zstd or lz4, or in Blosc frames of LZ4, zstd, zlib or snappy
streams with or without byte shuffle (crates/lucem-io/src/pyramid.rs:1001-1086,
crates/lucem-io/src/ome_zarr.rs:146-343). LZ4, zstd and snappy are decoded by pyarrow's bundled
codecs. BloscLZ streams and bit shuffle raise a clear error.
"""

from __future__ import annotations

import gzip
import json
import math
import threading
import zlib
from collections import OrderedDict
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field
from pathlib import Path
from typing import Callable, Iterator, Sequence

import numpy as np

from .batched import Block, RLBatch, TileLayout, schedule

# Limits a store's metadata may declare (crates/lucem-io/src/pyramid.rs:36-43).
MAX_ARRAY_ELEMENTS = 1 << 48
MAX_CHUNK_BYTES = 1 << 31
MAX_LEVELS = 32


class IoError(RuntimeError):
    """A store that cannot be read or written as asked (crates/lucem-io/src/error.rs)."""


class RunError(RuntimeError):
    """A walk that cannot run as asked (custom.rs:36-47)."""


class RunCancelled(RunError):
    def __init__(self) -> None:
        super().__init__("cancelled")


def _safe_relative(path: str) -> bool:
    """crates/lucem-io/src/ome_zarr.rs:139-141: a non-empty path made of plain names only."""
    if not path:
        return False
    parts = Path(path).parts
    return not Path(path).is_absolute() and all(p not in ("", ".", "..") for p in parts)


def _to_micrometres(v: float, unit: str | None) -> float:
    scale = {None: 1.0, "micrometer": 1.0, "nanometer": 1e-3, "millimeter": 1e3, "centimeter": 1e4,
             "meter": 1e6, "angstrom": 1e-4}
    return v * scale.get(unit, 1.0)


# --------------------------------------------------------------------------------------------
# Chunk codecs


def _arrow_decompress(codec: str, data: bytes, size: int) -> bytes:
    """``size`` bytes from an LZ4 block (``lz4_raw``), a zstd frame or a raw snappy block, through
    pyarrow's bundled codecs."""
    try:
        import pyarrow as pa
    except ImportError as e:  # pragma: no cover - the image carries pyarrow
        raise IoError(f"{codec} chunks need pyarrow's decoders") from e
    try:
        out = pa.decompress(data, decompressed_size=size, codec=codec, asbytes=True)
    except Exception as e:
        raise IoError(f"{codec}: {e}") from e
    if len(out) != size:
        raise IoError(f"{codec}: {len(out)} bytes decoded, {size} expected")
    return out


# Blosc's compressor codes (bits 5-7 of its flags) and the stream decoders used for them.
_BLOSC_STREAMS = {1: "lz4_raw", 2: "snappy", 3: "zlib", 4: "zstd"}


def blosc_decompress(raw: bytes, expected: int) -> bytes:
    """A Blosc frame (format version 1, as the reference reads it: crates/lucem-io/src/ome_zarr.rs:
    237-343) of ``expected`` bytes: a 16-byte header (flags, type size, bytes, block size, frame
    size), then either the bytes as they are (flag 0x02) or a table of block offsets and per block
    one stream, or one stream per byte of the type when the block was split (found, as the
    reference does, from whether the first stream fills the block), each prefixed by its size and
    stored raw when that size is the stream's length; byte shuffle (flag 0x01) undone per block,
    the bytes past whole elements left in place. Streams of LZ4, zstd, zlib and snappy decode;
    BloscLZ streams and bit shuffle raise."""
    import struct

    if len(raw) < 16:
        raise IoError("Blosc header shorter than 16 bytes")
    flags, typesize = raw[2], max(raw[3], 1)
    nbytes, blocksize, cbytes = struct.unpack_from("<III", raw, 4)
    if cbytes > len(raw):
        raise IoError("Blosc frame size exceeds the chunk")
    if nbytes != expected:
        raise IoError(f"Blosc frame holds {nbytes} bytes, its array needs {expected}")
    if flags & 0x02:
        if len(raw) < 16 + nbytes:
            raise IoError("Blosc stored frame truncated")
        return bytes(raw[16:16 + nbytes])
    if flags & 0x04:
        raise IoError("Blosc bit shuffle is not supported")
    codec = _BLOSC_STREAMS.get((flags >> 5) & 7)
    if codec is None:
        raise IoError(f"Blosc compressor code {(flags >> 5) & 7} (BloscLZ) cannot be decoded here")
    if nbytes == 0:
        return b""
    if blocksize == 0:
        raise IoError("Blosc block size 0")
    nblocks = -(-nbytes // blocksize)
    if len(raw) < 16 + 4 * nblocks:
        raise IoError("Blosc block table truncated")
    starts = struct.unpack_from(f"<{nblocks}I", raw, 16)
    out = np.empty(nbytes, np.uint8)
    for b in range(nblocks):
        off = starts[b]
        end = starts[b + 1] if b + 1 < nblocks else cbytes
        bsize = min(blocksize, nbytes - b * blocksize)
        if off + 4 > len(raw):
            raise IoError("Blosc stream size truncated")
        first = struct.unpack_from("<I", raw, off)[0]
        nstreams = 1 if typesize <= 1 or off + 4 + first == end else typesize
        if bsize % nstreams:
            raise IoError("Blosc block does not split into its streams")
        slen = bsize // nstreams
        parts = []
        for _ in range(nstreams):
            if off + 4 > len(raw):
                raise IoError("Blosc stream size truncated")
            csize = struct.unpack_from("<I", raw, off)[0]
            off += 4
            if off + csize > len(raw):
                raise IoError("Blosc stream overruns the chunk")
            if csize == 0:
                parts.append(bytes(slen))
            elif csize == slen:
                parts.append(raw[off:off + csize])
            elif codec == "zlib":
                data = zlib.decompress(raw[off:off + csize])
                if len(data) != slen:
                    raise IoError("Blosc zlib stream of the wrong size")
                parts.append(data)
            else:
                parts.append(_arrow_decompress(codec, raw[off:off + csize], slen))
            off += csize
        blk = np.frombuffer(b"".join(parts), np.uint8)
        dst = out[b * blocksize:b * blocksize + bsize]
        if flags & 0x01 and typesize > 1:
            ne = bsize // typesize
            dst[:ne * typesize] = blk[:ne * typesize].reshape(typesize, ne).T.reshape(-1)
            dst[ne * typesize:] = blk[ne * typesize:]
        else:
            dst[:] = blk
    return out.tobytes()


def _decode_chunk(comp: dict | None, raw: bytes, expected: int) -> bytes:
    """A chunk's bytes under its Zarr v2 compressor: none, zlib, gzip, zstd, lz4 (a 4-byte size,
    then an LZ4 block) or blosc."""
    cid = None if comp is None else comp.get("id")
    if cid is None:
        return raw
    if cid == "zlib":
        return zlib.decompress(raw)
    if cid == "gzip":
        return gzip.decompress(raw)
    if cid == "zstd":
        return _arrow_decompress("zstd", raw, expected)
    if cid == "lz4":
        if len(raw) < 4 or int.from_bytes(raw[:4], "little") != expected:
            raise IoError("lz4 chunk declares another size than its array needs")
        return _arrow_decompress("lz4_raw", raw[4:], expected)
    if cid == "blosc":
        return blosc_decompress(raw, expected)
    raise IoError(f"Zarr compressor {cid!r} is not supported")


# --------------------------------------------------------------------------------------------
# Reading


class _ChunkCache:
    """Decoded chunks by (level, chunk index), least recently used first out, bounded in bytes.
    Neighbouring blocks read the same chunks for their halos; the loader's threads share it."""

    def __init__(self, max_bytes: int = 256 << 20):
        self.max_bytes = max_bytes
        self.bytes = 0
        self.items: OrderedDict = OrderedDict()
        self.lock = threading.Lock()

    def get(self, key):
        with self.lock:
            v = self.items.get(key)
            if v is not None:
                self.items.move_to_end(key)
            return v

    def put(self, key, value: np.ndarray) -> None:
        with self.lock:
            if key in self.items:
                return
            self.items[key] = value
            self.bytes += value.nbytes
            while self.bytes > self.max_bytes and len(self.items) > 1:
                _, old = self.items.popitem(last=False)
                self.bytes -= old.nbytes


class _ZarrArray:
    """One Zarr v2 array of a level (pyramid.rs:1001-1050, finish_open)."""

    def __init__(self, root: Path, prefix: str, axis_names: Sequence[str]):
        self.dir = root / prefix
        zpath = self.dir / ".zarray"
        if not zpath.is_file():
            raise IoError(f"{prefix}: no .zarray found (Zarr v2 arrays are read)")
        meta = json.loads(zpath.read_text())
        if meta.get("order") == "F":
            raise IoError("Fortran-ordered Zarr arrays are not supported")
        try:
            self.dtype = np.dtype(meta["dtype"])
        except TypeError as e:
            raise IoError(f"unsupported Zarr dtype {meta.get('dtype')!r}") from e
        if self.dtype.kind not in "uifb" or self.dtype.itemsize > 8:
            raise IoError(f"unsupported Zarr dtype {meta['dtype']!r}")
        comp = meta.get("compressor")
        self.compressor = comp
        cid = None if comp is None else comp.get("id")
        if cid not in (None, "zlib", "gzip", "zstd", "lz4", "blosc"):
            raise IoError(f"Zarr compressor {cid!r} is not supported (none, zlib, gzip, zstd, lz4 and blosc are)")
        if cid == "blosc":
            if comp.get("cname", "lz4") not in ("lz4", "zstd", "zlib", "snappy"):
                raise IoError(f"Blosc with {comp.get('cname')} is not supported (lz4, zstd, zlib and snappy are)")
            if comp.get("shuffle", 1) == 2:
                raise IoError("Blosc bit shuffle is not supported")
        if meta.get("filters"):
            raise IoError("Zarr filters are not supported")
        sep = meta.get("dimension_separator") or "."
        if sep not in (".", "/"):
            raise IoError(f"unsupported dimension separator {sep!r}")
        self.sep = sep
        fv = meta.get("fill_value")
        try:
            self.fill = float(fv) if fv is not None else 0.0
        except (TypeError, ValueError):
            self.fill = 0.0
        self.shape = [int(v) for v in meta["shape"]]
        self.chunks = [int(v) for v in meta["chunks"]]
        nd = len(self.shape)
        if not 2 <= nd <= 6 or len(self.chunks) != nd:
            raise IoError(f"unsupported Zarr shape {self.shape}")
        if 0 in self.shape or 0 in self.chunks:
            raise IoError(f"Zarr shape {self.shape} / chunks {self.chunks} has an empty axis")
        if math.prod(self.shape) > MAX_ARRAY_ELEMENTS:
            raise IoError(f"Zarr shape {self.shape} is too large")
        if math.prod(self.chunks) * self.dtype.itemsize > MAX_CHUNK_BYTES:
            raise IoError(f"Zarr chunks {self.chunks} are too large")
        names = list(axis_names)

        def find(name):
            return names.index(name) if name in names and names.index(name) < nd else None

        self.y = find("y") if find("y") is not None else nd - 2
        self.x = find("x") if find("x") is not None else nd - 1
        self.c = find("c") if find("c") is not None else (nd - 3 if nd >= 3 and not names else None)
        self.z = find("z")
        order = [a for a in (self.z, self.y, self.x) if a is not None]
        if order != sorted(order):
            raise IoError("axes must run z, y, x in the array's order")

    def _chunk(self, idx: tuple) -> np.ndarray | None:
        path = self.dir / self.sep.join(str(i) for i in idx)
        try:
            raw = path.read_bytes()
        except FileNotFoundError:
            return None
        n = math.prod(self.chunks)
        raw = _decode_chunk(self.compressor, raw, n * self.dtype.itemsize)
        if len(raw) != n * self.dtype.itemsize:
            raise IoError(f"{path} holds {len(raw)} bytes, its array needs {n * self.dtype.itemsize}")
        return np.frombuffer(raw, dtype=self.dtype).reshape(self.chunks)

    def read(self, level: int, cache: _ChunkCache, channel: int, origin, size) -> np.ndarray:
        """``size`` samples (z, y, x) from ``origin`` of ``channel``, as float32."""
        out = np.empty(tuple(size), dtype=np.float32)
        axes = [(self.z, 0), (self.y, 1), (self.x, 2)]
        ranges = []
        for a, k in axes:
            if a is None:
                ranges.append([(0, 0, size[k])])  # no z axis: the single plane
                continue
            c = self.chunks[a]
            spans = []
            for ci in range(origin[k] // c, (origin[k] + size[k] - 1) // c + 1):
                lo, hi = max(origin[k], ci * c), min(origin[k] + size[k], (ci + 1) * c)
                spans.append((ci, lo, hi - lo))
            ranges.append(spans)
        for zc, zl, zn in ranges[0]:
            for yc, yl, yn in ranges[1]:
                for xc, xl, xn in ranges[2]:
                    idx = [0] * len(self.shape)
                    sl: list = [0] * len(self.shape)
                    if self.c is not None:
                        idx[self.c] = channel // self.chunks[self.c]
                        sl[self.c] = channel % self.chunks[self.c]
                    for a, (ci, lo, n) in zip((self.z, self.y, self.x), ((zc, zl, zn), (yc, yl, yn), (xc, xl, xn))):
                        if a is None:
                            continue
                        idx[a] = ci
                        sl[a] = slice(lo - ci * self.chunks[a], lo - ci * self.chunks[a] + n)
                    key = (level, tuple(idx))
                    chunk = cache.get(key)
                    if chunk is None:
                        chunk = self._chunk(tuple(idx))
                        if chunk is not None:
                            cache.put(key, chunk)
                    dst = out[zl - origin[0]:zl - origin[0] + zn, yl - origin[1]:yl - origin[1] + yn,
                              xl - origin[2]:xl - origin[2] + xn]
                    if chunk is None:
                        dst[...] = self.fill
                    else:
                        dst[...] = chunk[tuple(sl)].reshape(dst.shape)
        return out


@dataclass
class LevelInfo:
    """A resolution level (pyramid.rs LevelInfo): sizes in its own pixels, chunk sizes, and level-0
    pixels per pixel."""

    path: str
    depth: int
    height: int
    width: int
    tile_d: int
    tile_h: int
    tile_w: int
    downsample: float = 1.0
    z_downsample: float = 1.0


class Pyramid:
    """An OME-Zarr image on disk, its metadata read and its chunks read on demand
    (pyramid.rs:507-650 ``open``/``open_store``, :879-895 ``read_block``)."""

    def __init__(self, root: Path, name: str, levels: list, arrays: list, channels: list,
                 pixel_size_um, z_spacing_um, cache_bytes: int = 256 << 20):
        self.path = root
        self.name = name
        self.levels = levels
        self._arrays = arrays
        self.channels = channels
        self.pixel_size_um = pixel_size_um
        self.z_spacing_um = z_spacing_um
        self._cache = _ChunkCache(cache_bytes)

    @classmethod
    def open(cls, path) -> "Pyramid":
        """The nearest OME-Zarr image at or above ``path`` (ome_zarr.rs:349-366), or a Zarr group
        at ``path``."""
        path = Path(path)
        root = None
        for anc in [path, *path.parents]:
            za = anc / ".zattrs"
            if anc.is_dir() and za.is_file() and (anc / ".zgroup").is_file() and "multiscales" in za.read_text():
                root = anc
                break
        if root is None:
            if path.is_dir() and (path / ".zgroup").is_file():
                root = path
            else:
                raise IoError(f"{path} is not an OME-Zarr image (no .zgroup with multiscales found)")
        attrs = json.loads((root / ".zattrs").read_text()) if (root / ".zattrs").is_file() else {}
        ms = (attrs.get("multiscales") or [{}])[0]
        axis_names = [a if isinstance(a, str) else (a.get("name") or "") if isinstance(a, dict) else ""
                      for a in ms.get("axes", [])]
        unit = next((a.get("unit") for a in reversed(ms.get("axes", [])) if isinstance(a, dict) and a.get("unit")), None)
        datasets = list(ms.get("datasets", []))
        if len(datasets) > MAX_LEVELS:
            raise IoError(f"{len(datasets)} resolution levels declared (at most {MAX_LEVELS})")
        bad = next((d for d in datasets if not _safe_relative(str(d.get("path", "")))), None)
        if bad is not None:
            raise IoError(f"level path {bad.get('path')!r} leaves the container")
        if not datasets:
            datasets = [{"path": c} for c in ("0", "1", "2", "3", "4", "5", "s0", "s1", "s2", "s3", "s4", "s5")
                        if (root / c / ".zarray").is_file()]
        if not datasets:
            raise IoError(f"no multiscale levels in {root.name}")

        def scale_of(ds):
            for t in ds.get("coordinateTransformations", []) or []:
                if t.get("type") == "scale":
                    return [float(v) for v in t.get("scale", [])]
            return None

        arrays, levels = [], []
        pixel_size_um = z_spacing_um = None
        for i, ds in enumerate(datasets):
            arr = _ZarrArray(root, str(ds["path"]).strip("/"), axis_names)
            d, td = (arr.shape[arr.z], arr.chunks[arr.z]) if arr.z is not None else (1, 1)
            if i == 0:
                s = scale_of(ds)
                if s and s[-1] > 0:
                    pixel_size_um = _to_micrometres(s[-1], unit)
                if s and arr.z is not None and arr.z < len(s) and s[arr.z] > 0:
                    z_spacing_um = _to_micrometres(s[arr.z], unit)
            levels.append(LevelInfo(str(ds["path"]), d, arr.shape[arr.y], arr.shape[arr.x], td, arr.chunks[arr.y],
                                    arr.chunks[arr.x]))
            arrays.append(arr)
        w0, d0 = max(levels[0].width, 1), max(levels[0].depth, 1)
        for lv in levels:
            lv.downsample = w0 / max(lv.width, 1)
            lv.z_downsample = d0 / max(lv.depth, 1)
        order = sorted(range(len(levels)), key=lambda i: -levels[i].width)
        levels = [levels[i] for i in order]
        arrays = [arrays[i] for i in order]
        a0 = arrays[0]
        nch = a0.shape[a0.c] if a0.c is not None else 1
        omero = (attrs.get("omero") or {}).get("channels", [])
        channels = [(omero[c].get("label") if c < len(omero) and omero[c].get("label") else f"Channel {c}")
                    for c in range(nch)]
        name = ms.get("name") or root.name
        return cls(root, name, levels, arrays, channels, pixel_size_um, z_spacing_um)

    def depth(self) -> int:
        return self.levels[0].depth

    def height(self) -> int:
        return self.levels[0].height

    def width(self) -> int:
        return self.levels[0].width

    def read_block(self, level: int, channel: int, origin, size) -> np.ndarray:
        """A box of one channel at ``level``: ``size`` values from ``origin`` (both ``[z, y, x]`` in
        that level's pixels) as float32 ``[z, y, x]``; it must lie inside the level."""
        if not 0 <= level < len(self.levels):
            raise IoError(f"no level {level}")
        lv = self.levels[level]
        dims = (lv.depth, lv.height, lv.width)
        if any(o < 0 or s < 0 or o + s > d for o, s, d in zip(origin, size, dims)):
            raise IoError(f"block {list(size)} at {list(origin)} is outside level {level} {list(dims)}")
        if not 0 <= channel < len(self.channels):
            raise IoError(f"no channel {channel} (the image has {len(self.channels)})")
        return self._arrays[level].read(level, self._cache, channel, tuple(origin), tuple(size))


# --------------------------------------------------------------------------------------------
# The loader


@dataclass
class LoadedBlock:
    """One block of one channel (loader.rs:17-24); ``data`` is the read part, float32 ``[z, y, x]``."""

    block: Block
    channel: int
    data: np.ndarray


class BlockLoader:
    """Blocks in walk order, each channel in turn for each block, read on ``threads`` threads while
    the caller works, at most ``read_ahead`` finished blocks waiting (loader.rs:37-125). Closing it
    (or leaving a ``with`` block) stops the threads."""

    def __init__(self, pyramid: Pyramid, level: int, channels: Sequence[int], blocks: Sequence[Block],
                 read_ahead: int = 4, threads: int = 4):
        self._jobs = [(b, c) for b in blocks for c in channels]
        self.total = len(self._jobs)
        self._read_ahead = max(int(read_ahead), 1)
        self._cv = threading.Condition()
        self._next_job = 0
        self._consumed = 0
        self._stopped = False
        self._done: dict = {}
        self._next = 0
        self._py, self._level = pyramid, level
        n = min(max(int(threads), 1), 32)
        self._workers = [threading.Thread(target=self._work, name=f"BlockLoader-{i}", daemon=True) for i in range(n)]
        for w in self._workers:
            w.start()

    def _work(self) -> None:
        while True:
            with self._cv:
                while not self._stopped and self._next_job < self.total and \
                        self._next_job >= self._consumed + self._read_ahead:
                    self._cv.wait()
                if self._stopped or self._next_job >= self.total:
                    return
                job = self._next_job
                self._next_job += 1
            block, channel = self._jobs[job]
            try:
                data = self._py.read_block(self._level, channel, block.read_origin, block.read_size)
                item = LoadedBlock(block, channel, data)
            except Exception as e:  # handed to the consumer in walk order
                item = e
            with self._cv:
                self._done[job] = item
                self._cv.notify_all()

    def __iter__(self) -> Iterator[LoadedBlock]:
        return self

    def __next__(self) -> LoadedBlock:
        if self._next >= self.total:
            raise StopIteration
        with self._cv:
            while self._next not in self._done:
                if self._stopped:
                    raise StopIteration
                self._cv.wait()
            item = self._done.pop(self._next)
            self._next += 1
            self._consumed = self._next
            self._cv.notify_all()
        if isinstance(item, Exception):
            raise item
        return item

    def close(self) -> None:
        with self._cv:
            self._stopped = True
            self._cv.notify_all()
        for w in self._workers:
            w.join()
        self._workers = []

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# --------------------------------------------------------------------------------------------
# Writing


@dataclass
class WriterOptions:
    """What to create (zarr_writer.rs:80-90): ``shape`` and ``chunks`` are ``[z, y, x]``; depth 1 is
    a 2-D image."""

    name: str
    channels: list
    shape: tuple
    chunks: tuple
    pixel_size_um: float | None = None
    z_spacing_um: float | None = None


_SAMPLES = {np.dtype(np.float32): "<f4", np.dtype(np.uint32): "<u4"}


class ZarrWriter:
    """An OME-Zarr v0.4 store (Zarr v2, raw chunks) written one level-0 chunk at a time; ``finish``
    builds the coarser levels from the chunks on disk and writes the metadata
    (zarr_writer.rs:87-345). Samples are float32 (coarser levels average) or uint32 labels (coarser
    levels sample). Axes are ``c, y, x`` for one plane and ``c, z, y, x`` for a volume."""

    def __init__(self, root: Path, options: WriterOptions, dtype):
        self.root = root
        self.options = options
        self.dtype = np.dtype(dtype)

    @classmethod
    def create(cls, root, options: WriterOptions, dtype=np.float32) -> "ZarrWriter":
        """Starts a store at ``root``, which must not exist yet (zarr_writer.rs:98-113)."""
        root = Path(root)
        dtype = np.dtype(dtype)
        if dtype not in _SAMPLES:
            raise IoError("a store holds float32 or uint32 samples")
        if root.exists():
            raise IoError(f"{root} already exists")
        shape, chunks = tuple(int(v) for v in options.shape), tuple(int(v) for v in options.chunks)
        if not options.channels or len(shape) != 3 or len(chunks) != 3 or 0 in shape or 0 in chunks or \
                min(shape + chunks) < 0:
            raise IoError("a store needs at least one channel and non-empty shape and chunks")
        if math.prod(shape) * math.prod(chunks) >= 1 << 128:
            raise IoError("the store's shape is too large")
        options = WriterOptions(options.name, list(options.channels), shape, chunks, options.pixel_size_um,
                                options.z_spacing_um)
        (root / "0").mkdir(parents=True)
        (root / ".zgroup").write_text('{\n  "zarr_format": 2\n}\n')
        return cls(root, options, dtype)

    def _chunk_path(self, level: int, channel: int, idx) -> Path:
        d = self.root / str(level) / str(channel)
        return d / str(idx[0]) / str(idx[1]) / str(idx[2]) if self.options.shape[0] > 1 else d / str(idx[1]) / str(idx[2])

    def _write_chunk(self, path: Path, data: np.ndarray, size, chunk) -> None:
        """``data`` of ``size`` as a whole chunk of ``chunk``, zero-padded at the image's edge."""
        if tuple(size) != tuple(chunk):
            full = np.zeros(chunk, dtype=self.dtype)
            full[:size[0], :size[1], :size[2]] = data.reshape(size)
            data = full
        path.parent.mkdir(parents=True, exist_ok=True)
        arr = np.ascontiguousarray(data, dtype=self.dtype.newbyteorder("<"))
        with open(path, "wb") as f:
            f.write(memoryview(arr.reshape(-1)).cast("B"))

    def write(self, channel: int, origin, size, data) -> None:
        """Stores the block of ``size`` at ``origin`` of ``channel`` as one chunk of level 0: ``origin``
        on the chunk grid and ``size`` a whole chunk, or what remains of one at the image's edge
        (zarr_writer.rs:129-144)."""
        data = np.asarray(data)
        if data.dtype != self.dtype:
            raise IoError("this store holds another element type")
        if not 0 <= channel < len(self.options.channels):
            raise IoError(f"no channel {channel}")
        c, s = self.options.chunks, self.options.shape
        origin, size = tuple(int(v) for v in origin), tuple(int(v) for v in size)
        if any(origin[a] % c[a] != 0 or origin[a] >= s[a] or origin[a] < 0 for a in range(3)):
            raise IoError(f"block at {list(origin)} is not on the chunk grid {list(c)} of shape {list(s)}")
        whole = all(size[a] == min(c[a], s[a] - origin[a]) for a in range(3))
        if not whole or data.size != math.prod(size):
            raise IoError(f"block of {list(size)} with {data.size} values at {list(origin)} is not a whole chunk "
                          f"of {list(c)}")
        idx = tuple(origin[a] // c[a] for a in range(3))
        self._write_chunk(self._chunk_path(0, channel, idx), data, size, c)

    def write_many(self, blocks, threads: int = 8) -> None:
        """Many ``(channel, origin, size, values)`` chunks at once, on ``threads`` threads (each chunk
        is its own file)."""
        with ThreadPoolExecutor(max_workers=max(int(threads), 1)) as ex:
            for f in [ex.submit(self.write, *b) for b in blocks]:
                f.result()

    def _read_box(self, level: int, channel: int, lshape, origin, size) -> np.ndarray:
        """A box of a level already on disk; missing chunks read as zeros (zarr_writer.rs:242-270)."""
        c = self.options.chunks
        out = np.zeros(size, dtype=self.dtype)
        for cz in range(origin[0] // c[0], (origin[0] + size[0] - 1) // c[0] + 1):
            for cy in range(origin[1] // c[1], (origin[1] + size[1] - 1) // c[1] + 1):
                for cx in range(origin[2] // c[2], (origin[2] + size[2] - 1) // c[2] + 1):
                    p = self._chunk_path(level, channel, (cz, cy, cx))
                    if not p.is_file():
                        continue
                    v = np.frombuffer(p.read_bytes(), dtype=self.dtype.newbyteorder("<")).reshape(c)
                    at = (cz * c[0], cy * c[1], cx * c[2])
                    lo = [max(origin[a], at[a]) for a in range(3)]
                    hi = [min(origin[a] + size[a], at[a] + c[a], lshape[a]) for a in range(3)]
                    if any(h <= l for l, h in zip(lo, hi)):
                        continue
                    out[lo[0] - origin[0]:hi[0] - origin[0], lo[1] - origin[1]:hi[1] - origin[1],
                        lo[2] - origin[2]:hi[2] - origin[2]] = \
                        v[lo[0] - at[0]:hi[0] - at[0], lo[1] - at[1]:hi[1] - at[1], lo[2] - at[2]:hi[2] - at[2]]
        return out

    def _reduce(self, src: np.ndarray, size, zf: int) -> np.ndarray:
        """``src`` reduced to ``size``: y and x halved, z divided by ``zf``. A float pixel is the
        float32 mean of the source pixels it covers, summed in (z, y, x) order; a label pixel is
        the first of them (zarr_writer.rs:385-409)."""
        sz, sy, sx = src.shape
        if self.dtype == np.uint32:
            return np.ascontiguousarray(src[::zf, ::2, ::2][:size[0], :size[1], :size[2]])
        total = np.zeros(size, dtype=np.float32)
        count = np.zeros(size, dtype=np.float32)
        for dz in range(zf):
            for dy in range(2):
                for dx in range(2):
                    part = src[dz::zf, dy::2, dx::2]
                    pz, py, px = part.shape
                    total[:pz, :py, :px] += part
                    count[:pz, :py, :px] += 1
        return total / count

    def _build_level(self, index: int, prev_shape, next_shape, halve_z: bool, threads: int) -> None:
        c = self.options.chunks
        zf = 2 if halve_z else 1

        def one(channel, idx):
            origin = tuple(idx[a] * c[a] for a in range(3))
            size = tuple(min(c[a], next_shape[a] - origin[a]) for a in range(3))
            so = (origin[0] * zf, origin[1] * 2, origin[2] * 2)
            ss = (min(size[0] * zf, prev_shape[0] - so[0]), min(size[1] * 2, prev_shape[1] - so[1]),
                  min(size[2] * 2, prev_shape[2] - so[2]))
            src = self._read_box(index - 1, channel, prev_shape, so, ss)
            self._write_chunk(self._chunk_path(index, channel, idx), self._reduce(src, size, zf), size, c)

        jobs = [(ch, (cz, cy, cx)) for ch in range(len(self.options.channels))
                for cz in range(-(-next_shape[0] // c[0])) for cy in range(-(-next_shape[1] // c[1]))
                for cx in range(-(-next_shape[2] // c[2]))]
        with ThreadPoolExecutor(max_workers=max(threads, 1)) as ex:
            for f in [ex.submit(one, *j) for j in jobs]:
                f.result()

    def finish(self, threads: int = 8) -> Path:
        """Builds the coarser levels (y and x halved until the larger side fits one chunk side, at
        most 12 levels; z halved too while its spacing stays within 1.5 times the pixel size),
        writes the metadata and returns the store's path (zarr_writer.rs:196-217)."""
        o = self.options
        levels = [(o.shape, (1.0, 1.0, 1.0))]
        px = o.pixel_size_um if o.pixel_size_um is not None else 1.0
        dz = o.z_spacing_um if o.z_spacing_um is not None else px
        largest = max(o.chunks[1], o.chunks[2])
        while max(levels[-1][0][1], levels[-1][0][2]) > largest and len(levels) < 12:
            pshape, pscale = levels[-1]
            halve_z = pshape[0] > 1 and dz * pscale[0] <= 1.5 * (px * pscale[1])
            nshape = (-(-pshape[0] // 2) if halve_z else pshape[0], -(-pshape[1] // 2), -(-pshape[2] // 2))
            nscale = (pscale[0] * (2.0 if halve_z else 1.0), pscale[1] * 2.0, pscale[2] * 2.0)
            self._build_level(len(levels), pshape, nshape, halve_z, threads)
            levels.append((nshape, nscale))
        n = len(o.channels)
        volume = o.shape[0] > 1
        for i, (shape, _) in enumerate(levels):
            meta = {"zarr_format": 2,
                    "shape": [n, *shape] if volume else [n, shape[1], shape[2]],
                    "chunks": [1, *o.chunks] if volume else [1, o.chunks[1], o.chunks[2]],
                    "dtype": _SAMPLES[self.dtype], "compressor": None, "fill_value": 0, "order": "C",
                    "filters": None, "dimension_separator": "/"}
            (self.root / str(i)).mkdir(parents=True, exist_ok=True)
            (self.root / str(i) / ".zarray").write_text(json.dumps(meta, indent=2))
        unit = "micrometer" if o.pixel_size_um is not None else None

        def space(name):
            return {"name": name, "type": "space", "unit": unit} if unit else {"name": name, "type": "space"}

        axes = [{"name": "c", "type": "channel"}] + ([space("z")] if volume else []) + [space("y"), space("x")]
        datasets = []
        for i, (_, sc) in enumerate(levels):
            scale = [1.0, dz * sc[0], px * sc[1], px * sc[2]] if volume else [1.0, px * sc[1], px * sc[2]]
            datasets.append({"path": str(i), "coordinateTransformations": [{"type": "scale", "scale": scale}]})
        attrs = {"multiscales": [{"version": "0.4", "name": o.name, "axes": axes, "datasets": datasets}],
                 "omero": {"channels": [{"label": c, "active": True} for c in o.channels]}}
        (self.root / ".zattrs").write_text(json.dumps(attrs, indent=2))
        return self.root


# --------------------------------------------------------------------------------------------
# The walk


@dataclass
class RunOptions:
    """How to walk an image (custom.rs:60-87): ``walk`` "planes" gives 2-D blocks ``[h, w]`` (each
    plane of a volume by itself), "volume" 3-D blocks ``[d, h, w]``. ``block``: side in y and x (the
    image's chunk side, at least 256, when None); ``block_z``: depth of a volume block (the chunk
    depth, at least 32); ``halo``: context per side in y and x; ``halo_z``: in z (``halo`` when
    None)."""

    level: int = 0
    channels: Sequence[int] | None = None
    walk: str = "planes"
    block: int | None = None
    block_z: int | None = None
    halo: int = 0
    halo_z: int | None = None
    read_ahead: int = 4
    threads: int = 4
    cancel: threading.Event | None = None
    progress: Callable[[int, int], None] | None = None

    def blocks(self, py: Pyramid):
        """The block grid and the blocks of the walk (custom.rs:111-126)."""
        if not 0 <= self.level < len(py.levels):
            raise RunError(f"no level {self.level} (the image has {len(py.levels)})")
        lv = py.levels[self.level]
        dims = (lv.depth, lv.height, lv.width)
        if self.block is not None:
            if self.block <= 0:
                raise RunError("the block side must be positive")
            by = bx = int(self.block)
        else:
            by, bx = max(lv.tile_h, 256), max(lv.tile_w, 256)
        if self.walk not in ("planes", "volume"):
            raise RunError(f"unknown walk {self.walk!r} (planes or volume)")
        if self.walk == "volume" and lv.depth > 1:
            bz = max(int(self.block_z) if self.block_z is not None else max(lv.tile_d, 32), 1)
            grid = (bz, by, bx)
            halo = (int(self.halo_z) if self.halo_z is not None else int(self.halo), int(self.halo), int(self.halo))
        else:
            grid, halo = (1, by, bx), (0, int(self.halo), int(self.halo))
        return grid, schedule(dims, grid, halo)

    def channel_list(self, py: Pyramid) -> list:
        """custom.rs:128-135."""
        lst = list(self.channels) if self.channels is not None else list(range(len(py.channels)))
        bad = next((c for c in lst if not 0 <= c < len(py.channels)), None)
        if bad is not None:
            raise RunError(f"no channel {bad} (the image has {len(py.channels)})")
        if not lst:
            raise RunError("no channels to process")
        return lst

    def writer_options(self, py: Pyramid, names, grid, shape) -> WriterOptions:
        """custom.rs:162-172: chunks are the block grid, cut to the image."""
        lv = py.levels[self.level]
        return WriterOptions(py.name, list(names), tuple(shape), tuple(min(g, s) for g, s in zip(grid, shape)),
                             None if py.pixel_size_um is None else py.pixel_size_um * lv.downsample,
                             None if py.z_spacing_um is None else py.z_spacing_um * lv.z_downsample)


@dataclass
class _Staging:
    """Grow-only pinned host and device buffers of one group in flight, and the chunk writes that
    still read its host output."""

    host_in: object = None
    host_out: object = None
    dev_in: object = None
    dev_out: object = None
    writes: list = field(default_factory=list)

    def need(self, n_in: int, n_out: int, device) -> None:
        import torch

        if self.host_in is None or self.host_in.numel() < n_in:
            self.host_in = torch.empty(n_in, dtype=torch.float32, pin_memory=True)
            self.dev_in = torch.empty(n_in, dtype=torch.float32, device=device)
        if self.host_out is None or self.host_out.numel() < n_out:
            self.host_out = torch.empty(n_out, dtype=torch.float32, pin_memory=True)
            self.dev_out = torch.empty(n_out, dtype=torch.float32, device=device)


@dataclass
class SlideStats:
    """What a slide run did: blocks, groups, plans built, voxels read and owned, and seconds spent
    waiting for the loader, packing and enqueueing groups and waiting for their results, waiting
    for the last writes, and building the coarser levels."""

    blocks: int = 0
    groups: int = 0
    plans_built: int = 0
    voxels_read: int = 0
    voxels_owned: int = 0
    wait_read_s: float = 0.0
    device_s: float = 0.0
    wait_write_s: float = 0.0
    finish_s: float = 0.0
    total_s: float = 0.0
    output: Path | None = None
    extra: dict = field(default_factory=dict)


def deconvolve_slide(source, output, psfs, op, *, level: int = 0, channels: Sequence[int] | None = None,
                     block: int | None = None, block_z: int | None = None, halo: int | None = None,
                     read_ahead: int = 8, threads: int = 4, group_voxels: int = 1 << 25,
                     cancel: threading.Event | None = None, progress: Callable[[int, int], None] | None = None,
                     client=None, stats: SlideStats | None = None) -> Path:
    """Richardson-Lucy over an OME-Zarr image, written to a new OME-Zarr store at ``output``; returns
    its path. This is ``custom::transform`` (custom.rs:233-251) with the RL operator as the
    function, on the device.

    ``source``: a path or a :class:`Pyramid`. ``psfs``: one PSF (NumPy or device array) or one per
    processed channel; 2-D PSFs walk planes, 3-D PSFs walk volume blocks. ``op``: a
    ``RichardsonLucyOp`` (its settings; its ``halo`` rule gives the default context, exact for
    every block). ``block``/``block_z``/``level``/``channels``/``read_ahead``/``threads``: as
    :class:`RunOptions`. ``group_voxels``: read voxels gathered into one batched run (a group ends at the end of a row
    of the block grid once it holds that many).
    ``cancel``: an event that stops the walk with :class:`RunCancelled`. ``progress(done, total)``:
    called after each group with the blocks done and the blocks in all. ``stats``: filled with
    what the run did."""
    import time

    import torch

    from . import _native as N

    t_start = time.perf_counter()
    op.validate()
    py = source if isinstance(source, Pyramid) else Pyramid.open(source)
    plist = list(psfs) if isinstance(psfs, (list, tuple)) else [psfs]
    plist = [N.as_tensor(p) if not isinstance(p, np.ndarray) else torch.from_numpy(np.ascontiguousarray(p, np.float32))
             for p in plist]
    nd = plist[0].dim()
    if nd not in (2, 3) or any(p.dim() != nd for p in plist):
        raise RunError("every PSF has the same rank, 2 or 3")
    device = torch.device("cuda", client.device if client is not None else 0)
    plist = [p.to(device=device, dtype=torch.float32).contiguous() for p in plist]
    h = op.halo(max((tuple(p.shape) for p in plist), key=max)) if halo is None else int(halo)
    ro = RunOptions(level=level, channels=channels, walk="volume" if nd == 3 else "planes", block=block,
                    block_z=block_z, halo=h, read_ahead=read_ahead, threads=threads, cancel=cancel,
                    progress=progress)
    chans = ro.channel_list(py)
    if len(plist) not in (1, len(chans)):
        raise RunError(f"{len(plist)} PSFs for {len(chans)} channels")
    lv = py.levels[level] if 0 <= level < len(py.levels) else None
    grid, blocks = ro.blocks(py)
    if nd == 3 and lv.depth == 1:
        raise RunError("a 3-D PSF needs a volume")
    psf_of = {c: (i if len(plist) == len(chans) else 0) for i, c in enumerate(chans)}
    order = {c: i for i, c in enumerate(chans)}
    writer = ZarrWriter.create(output, ro.writer_options(py, [py.channels[c] for c in chans], grid,
                                                         (lv.depth, lv.height, lv.width)))
    settings = op.settings()
    st = stats if stats is not None else SlideStats()
    cl = client if client is not None else N.default_client(device.index)
    # uploads, the plan and downloads all run in order on the client's stream
    tstream = torch.cuda.ExternalStream(cl.stream, device=device) if cl.stream else torch.cuda.default_stream(device)
    plans: OrderedDict = OrderedDict()
    stages = [_Staging(), _Staging()]
    pending: list = []
    done = 0
    loader = BlockLoader(py, level, chans, blocks, read_ahead, threads)
    pool = ThreadPoolExecutor(max_workers=max(int(threads), 1))

    def submit(items: list, stage: _Staging):
        """Packs a group into ``stage`` and enqueues its upload, plan and download; returns what
        ``land`` needs. The stage's previous group has landed and its writes are waited for here."""
        t0 = time.perf_counter()
        for f in stage.writes:
            f.result()
        stage.writes = []
        # a volume block keeps its z axis even when it is one plane deep (3-D PSFs take 3-D tiles)
        shapes = [tuple(it.block.read_size) if nd == 3 else tuple(it.block.read_size[1:]) for it in items]
        pidx = [psf_of[it.channel] for it in items]
        key = (tuple(shapes), tuple(pidx))
        plan = plans.pop(key, None)
        if plan is None:
            plan = RLBatch(shapes, [tuple(p.shape) for p in plist], pidx, settings, cl)
            st.plans_built += 1
        plans[key] = plan
        while len(plans) > 4:
            plans.popitem(last=False)[1].close()
        n_in = sum(it.data.size for it in items)
        n_out = sum(math.prod(it.block.size) for it in items)
        stage.need(n_in, n_out, device)
        hin = stage.host_in.numpy()
        in_ptrs, out_ptrs, layouts = [], [], []
        off_in = off_out = 0
        base_in, base_out = stage.dev_in.data_ptr(), stage.dev_out.data_ptr()
        for it in items:
            m = it.data.size
            hin[off_in:off_in + m] = it.data.reshape(-1)
            sz = it.block.size
            in_ptrs.append(base_in + 4 * off_in)
            out_ptrs.append(base_out + 4 * off_out)
            layouts.append(TileLayout(0, 0, sz[2], sz[1] * sz[2], it.block.inner_offset(), sz))
            off_in += m
            off_out += math.prod(sz)
        with torch.cuda.stream(tstream):
            stage.dev_in[:n_in].copy_(stage.host_in[:n_in], non_blocking=True)
        plan.execute_layout(in_ptrs, out_ptrs, layouts, plist)
        with torch.cuda.stream(tstream):
            stage.host_out[:n_out].copy_(stage.dev_out[:n_out], non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(tstream)
        st.device_s += time.perf_counter() - t0
        return items, stage, ev

    def land(flight) -> None:
        """Waits for a group's download and hands its owned blocks to the writers (views of the
        stage's pinned output, which the stage keeps until they are written)."""
        nonlocal done
        items, stage, ev = flight
        t0 = time.perf_counter()
        ev.synchronize()
        st.device_s += time.perf_counter() - t0
        hout = stage.host_out.numpy()
        off = 0
        for it in items:
            m = math.prod(it.block.size)
            f = pool.submit(writer.write, order[it.channel], it.block.origin, it.block.size,
                            hout[off:off + m].reshape(it.block.size))
            off += m
            stage.writes.append(f)
            pending.append(f)
            st.voxels_read += it.data.size
            st.voxels_owned += m
        st.groups += 1
        st.blocks += len(items)
        done += len(items)
        if progress is not None:
            progress(done, loader.total)

    flight = None
    submitted = 0
    try:
        group: list = []
        gvox = 0
        while True:
            if cancel is not None and cancel.is_set():
                raise RunCancelled()
            t0 = time.perf_counter()
            it = next(loader, None)
            st.wait_read_s += time.perf_counter() - t0
            last = it is None
            if not last:
                group.append(it)
                gvox += it.data.size
            # groups end with a row of the block grid (every channel of its last block), so the
            # interior rows of a regular grid give the same composition and reuse one plan
            row_end = not last and it.block.origin[2] + it.block.size[2] == lv.width and it.channel == chans[-1]
            if group and ((gvox >= group_voxels and row_end) or last):
                # the next group is packed and enqueued while the previous one still runs
                nxt = submit(group, stages[submitted % 2])
                submitted += 1
                if flight is not None:
                    land(flight)
                flight = nxt
                group, gvox = [], 0
            if last:
                break
            # chunks already written leave the queue; a failed write stops the walk
            while pending and pending[0].done():
                pending.pop(0).result()
        if flight is not None:
            land(flight)
            flight = None
        t0 = time.perf_counter()
        for f in pending:
            f.result()
        pending.clear()
        st.wait_write_s += time.perf_counter() - t0
    finally:
        loader.close()
        if flight is not None:
            flight[2].synchronize()
        for f in pending:
            f.cancel()
        pool.shutdown(wait=True)
        for p in plans.values():
            p.close()
        plans.clear()
    t0 = time.perf_counter()
    path = writer.finish(threads=max(int(threads), 1))
    st.finish_s = time.perf_counter() - t0
    st.output = path
    st.total_s = time.perf_counter() - t_start
    return path
