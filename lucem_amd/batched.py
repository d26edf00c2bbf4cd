"""Batched Richardson-Lucy over many tiles of uneven size, one graph-captured launch sequence.

The tiles are the read blocks of a block walk (crates/lucem-core/src/tiling.rs:13-95: ``Block``,
``schedule``; crates/lucem-core/src/compute/batch.rs:41-56: ``plan_tile_grid``), each an
independent RL problem with a PSF from a table. Results equal the plain entry point per tile.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Sequence

from . import _native as N
from ._native_ext import RlTile, RlTileLayout
from .restore import RichardsonLucy


@dataclass(frozen=True)
class TileSlice:
    """crates/lucem-core/src/compute/batch.rs:28-38: a tile's owned core and read region."""

    y0: int
    x0: int
    y1: int
    x1: int
    ry0: int
    rx0: int
    rh: int
    rw: int


def plan_tile_grid(width: int, height: int, core: int, halo: int) -> list[TileSlice]:
    """crates/lucem-core/src/compute/batch.rs:41-56 (same geometry, raster order)."""
    step = max(core, 1)
    ny, nx = -(-height // step), -(-width // step)
    out = []
    for ty in range(ny):
        for tx in range(nx):
            y0, x0 = ty * step, tx * step
            y1, x1 = min(y0 + step, height), min(x0 + step, width)
            ry0, rx0 = max(y0 - halo, 0), max(x0 - halo, 0)
            ry1, rx1 = min(y1 + halo, height), min(x1 + halo, width)
            out.append(TileSlice(y0, x0, y1, x1, ry0, rx0, ry1 - ry0, rx1 - rx0))
    return out


@dataclass(frozen=True)
class Block:
    """crates/lucem-core/src/tiling.rs:13-23: owned part and read part, ``[z, y, x]``."""

    index: int
    origin: tuple
    size: tuple
    read_origin: tuple
    read_size: tuple

    def inner_offset(self) -> tuple:
        return tuple(o - r for o, r in zip(self.origin, self.read_origin))

    def read_dims(self) -> list:
        return [self.read_size[1], self.read_size[2]] if self.read_size[0] == 1 else list(self.read_size)


def schedule(dims, block, halo) -> list[Block]:
    """crates/lucem-core/src/tiling.rs:54-78: blocks covering ``dims`` exactly once, raster order."""
    if 0 in dims:
        return []
    step = [dims[a] if block[a] == 0 else min(block[a], dims[a]) for a in range(3)]
    out = []

    def span(o, s, a):
        start = max(o - halo[a], 0)
        end = min(o + s + halo[a], dims[a])
        return start, end - start

    for z in range(0, dims[0], step[0]):
        for y in range(0, dims[1], step[1]):
            for x in range(0, dims[2], step[2]):
                size = (min(step[0], dims[0] - z), min(step[1], dims[1] - y), min(step[2], dims[2] - x))
                rz, dz = span(z, size[0], 0)
                ry, dy = span(y, size[1], 1)
                rx, dx = span(x, size[2], 2)
                out.append(Block(len(out), (z, y, x), size, (rz, ry, rx), (dz, dy, dx)))
    return out


@dataclass(frozen=True)
class TileLayout:
    """``lucem_rl_tile_layout``: element strides of a tile's input and output views (0: the tile's
    own contiguous layout) and the stored window (z, y, x; a 2-D tile uses y and x; size 0: to
    the tile's end)."""

    in_row_stride: int = 0
    in_plane_stride: int = 0
    out_row_stride: int = 0
    out_plane_stride: int = 0
    own_origin: tuple = (0, 0, 0)
    own_size: tuple = (0, 0, 0)

    def fill(self, c: RlTileLayout) -> None:
        c.in_row_stride, c.in_plane_stride = int(self.in_row_stride), int(self.in_plane_stride)
        c.out_row_stride, c.out_plane_stride = int(self.out_row_stride), int(self.out_plane_stride)
        for a in range(3):
            c.own_origin[a] = int(self.own_origin[a])
            c.own_size[a] = int(self.own_size[a])


def _tiles(inputs, outputs, psf_index) -> ctypes.Array:
    arr = (RlTile * max(len(inputs), 1))()
    for i, x in enumerate(inputs):
        t = arr[i]
        o = outputs[i] if outputs is not None else None
        t.input = N.device_ptr(x, f"input {i}") if (x.numel() and x.is_cuda) else None
        t.output = N.device_ptr(o, f"output {i}") if (o is not None and o.numel()) else None
        for a, d in enumerate(x.shape):
            t.dims[a] = int(d)
        t.ndim = x.dim()
        t.psf_index = int(psf_index[i])
    return arr


class RLBatch:
    """A planned batch: tile shapes, PSF shapes and settings fixed; buffers given per call."""

    def __init__(self, shapes: Sequence[Sequence[int]], psf_shapes: Sequence[Sequence[int]],
                 psf_index: Sequence[int], settings: RichardsonLucy | None = None, client: N.Client | None = None):
        import torch

        self.client = client or N.default_client()
        self.settings = settings or RichardsonLucy()
        self.shapes = [tuple(int(d) for d in s) for s in shapes]
        self.psf_index = [int(i) for i in psf_index]
        ranks = {len(p) for p in psf_shapes}
        if len(ranks) != 1:
            raise ValueError("every PSF of a batch has the same rank")
        self.psf_ndim = ranks.pop()
        self.psf_shapes = [tuple(int(d) for d in p) for p in psf_shapes]
        flat = [d for p in self.psf_shapes for d in p]
        self._psf_dims = (ctypes.c_uint64 * max(len(flat), 1))(*flat)
        fake = [torch.empty(s, device="meta") for s in self.shapes]
        arr = (RlTile * max(len(fake), 1))()
        for i, s in enumerate(self.shapes):
            for a, d in enumerate(s):
                arr[i].dims[a] = d
            arr[i].ndim = len(s)
            arr[i].psf_index = self.psf_index[i]
        h = N._vp()
        st = self.settings.to_c()
        N.check(N.lib().lucem_rl_batch_create(self.client.handle, arr, len(self.shapes), self._psf_dims,
                                              self.psf_ndim, len(self.psf_shapes), ctypes.byref(st), ctypes.byref(h)))
        self._h = h
        N.track(self, "plans")
        del fake

    @property
    def device_bytes(self) -> int:
        return int(N.lib().lucem_rl_batch_device_bytes(self._h))

    @property
    def launches(self) -> int:
        return int(N.lib().lucem_rl_batch_launches(self._h))

    def execute(self, inputs, outputs, psfs) -> None:
        """Enqueues the batch on these device tensors (shapes as planned)."""
        inputs = [N.as_tensor(x) for x in inputs]
        outputs = [N.as_tensor(x) for x in outputs]
        psfs = [N.as_tensor(p) for p in psfs]
        if len(inputs) != len(self.shapes) or len(outputs) != len(self.shapes) or len(psfs) != len(self.psf_shapes):
            raise ValueError("inputs, outputs and psfs must match the plan")
        arr = _tiles(inputs, outputs, self.psf_index)
        pp = (N._vp * max(len(psfs), 1))(*[N.device_ptr(p, f"psf {j}") for j, p in enumerate(psfs)])
        N.check(N.lib().lucem_rl_batch_execute(self._h, arr, pp))

    def execute_layout(self, in_ptrs: Sequence[int], out_ptrs: Sequence[int], layouts: Sequence["TileLayout"],
                       psfs) -> None:
        """Enqueues the batch on views into larger arrays: per tile a device address of its input's
        first element, of its stored window's first element in the output, and its layout."""
        n = len(self.shapes)
        if len(in_ptrs) != n or len(out_ptrs) != n or len(layouts) != n:
            raise ValueError("one input, output and layout per planned tile")
        psfs = [N.as_tensor(p) for p in psfs]
        if len(psfs) != len(self.psf_shapes):
            raise ValueError("psfs must match the plan")
        arr = (RlTile * max(n, 1))()
        lay = (RlTileLayout * max(n, 1))()
        for i, s in enumerate(self.shapes):
            t = arr[i]
            t.input = int(in_ptrs[i]) or None
            t.output = int(out_ptrs[i]) or None
            for a, d in enumerate(s):
                t.dims[a] = d
            t.ndim = len(s)
            t.psf_index = self.psf_index[i]
            layouts[i].fill(lay[i])
        pp = (N._vp * max(len(psfs), 1))(*[N.device_ptr(p, f"psf {j}") for j, p in enumerate(psfs)])
        N.check(N.lib().lucem_rl_batch_execute_layout(self._h, arr, lay, pp))

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            N.lib().lucem_rl_batch_destroy(self._h)
            self._h = N._vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def richardson_lucy_batched(inputs, psfs, psf_index, settings: RichardsonLucy | None = None,
                            client: N.Client | None = None):
    """One-shot batched RL; returns a list of outputs (same shapes as the inputs)."""
    import torch

    inputs = [N.as_tensor(x) for x in inputs]
    outputs = [torch.empty_like(x) for x in inputs]
    b = RLBatch([tuple(x.shape) for x in inputs], [tuple(N.as_tensor(p).shape) for p in psfs], psf_index, settings,
                client)
    b.execute(inputs, outputs, psfs)
    b.client.sync()
    b.close()
    return outputs
