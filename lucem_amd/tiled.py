"""Tiled whole-array Richardson-Lucy and the pipeline operator that registers it.

``deconvolve_blocks`` is the block walk of lucem-pipeline's ``custom::transform``
(engine/lucem-pipeline/src/custom.rs:171-251): each block is read with a halo
(crates/lucem-core/src/tiling.rs:54-78), deconvolved, and only its owned part is kept
(custom.rs:177-196). Here every block of every channel runs in one batched plan (one captured
graph) instead of one call per block, and nothing is copied: each tile reads its block straight
from the image and its last store writes only the owned part, straight into the result.

``RichardsonLucyOp`` is the operator a pipeline would register for RL, in the shape of the
reference's ``Op`` variants (engine/lucem-pipeline/src/ops.rs:97-203): its parameters, the
``halo`` rule (ops.rs:434-454), ``validate`` (ops.rs:695) and ``run`` (ops.rs:778). The halo is
exact: one RL iteration correlates twice with a PSF of extent k centred at k // 2 (once with the
flipped PSF, once with the PSF), so an output pixel depends on input pixels up to
``2 * (k // 2)`` away per iteration (k - 1 for odd k); with ``num_iter`` times that a tiled result
equals the whole-array result up to rounding. A smaller halo trades exactness at block seams
for less work.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Sequence

from . import _native as N
from .batched import RLBatch, TileLayout, schedule
from .restore import PadMode, RichardsonLucy, richardson_lucy_into, richardson_lucy_nd_into


@dataclass
class RichardsonLucyOp:
    """Richardson-Lucy deconvolution as a pipeline operator (tag ``richardson_lucy``): the image on
    the first port, the PSF on the second, whose extent per axis is at most ``psf_size`` (the
    bound the halo is computed from, as the reference's ``Op::halo`` takes no input,
    ops.rs:434)."""

    num_iter: int = 30
    clip: bool = True
    filter_epsilon: float | None = None
    mode: PadMode = PadMode.Reflect
    psf_size: int = 15
    tag: str = field(default="richardson_lucy", init=False)

    def halo(self, psf_shape: Sequence[int] | None = None) -> int:
        """Pixels of context per side for an exact result: ``num_iter * 2 * (k // 2)`` for the
        largest PSF extent k (``psf_size`` when no PSF shape is given)."""
        k = max(int(d) for d in psf_shape) if psf_shape is not None else int(self.psf_size)
        return int(self.num_iter) * 2 * (k // 2)

    def validate(self) -> None:
        """Parameter checks of the operator (the reference's RL check: num_iter >= 1)."""
        if int(self.num_iter) < 1:
            raise ValueError("invalid parameter `num_iter`: num_iter must be >= 1")
        if self.filter_epsilon is not None and not math.isfinite(float(self.filter_epsilon)):
            raise ValueError("invalid parameter `filter_epsilon`: not a finite number")
        if int(self.psf_size) < 1:
            raise ValueError("invalid parameter `psf_size`: psf_size must be >= 1")

    def params(self) -> dict:
        """The operator's parameters as the pipeline's JSON carries them (``op`` tag excluded)."""
        return {"num_iter": int(self.num_iter), "clip": bool(self.clip),
                "filter_epsilon": None if self.filter_epsilon is None else float(self.filter_epsilon),
                "mode": PadMode(self.mode).label, "psf_size": int(self.psf_size)}

    @classmethod
    def from_params(cls, params: dict) -> "RichardsonLucyOp":
        """The operator from its JSON parameters: every field given, no other (the reference's
        serde rules for ``Op``, ops.rs:97-99); the mode by its lowercase name (core/pad.rs:13-14)."""
        want = {"num_iter", "clip", "filter_epsilon", "mode", "psf_size"}
        unknown, missing = set(params) - want, want - set(params)
        if unknown:
            raise ValueError(f"unknown field(s) {sorted(unknown)} for richardson_lucy")
        if missing:
            raise ValueError(f"missing field(s) {sorted(missing)} for richardson_lucy")
        modes = {m.label: m for m in PadMode}
        if params["mode"] not in modes:
            raise ValueError(f"unknown variant `{params['mode']}`, expected one of {sorted(modes)}")
        for name in ("num_iter", "psf_size"):
            if isinstance(params[name], bool) or not isinstance(params[name], int) or params[name] < 0:
                raise ValueError(f"invalid value for `{name}`: expected a non-negative integer")
        if not isinstance(params["clip"], bool):
            raise ValueError("invalid value for `clip`: expected a boolean")
        eps = params["filter_epsilon"]
        if eps is not None and (isinstance(eps, bool) or not isinstance(eps, (int, float))):
            raise ValueError("invalid value for `filter_epsilon`: expected a number or null")
        return cls(num_iter=params["num_iter"], clip=params["clip"], filter_epsilon=eps,
                   mode=modes[params["mode"]], psf_size=params["psf_size"])

    def settings(self) -> RichardsonLucy:
        return RichardsonLucy(num_iter=int(self.num_iter), clip=bool(self.clip), mode=self.mode,
                              filter_epsilon=self.filter_epsilon)

    def run(self, image, psf):
        """The operator on a whole array (2-D, or 3-D with a 3-D PSF): a new output tensor."""
        import torch

        self.validate()
        image = N.as_tensor(image)
        if max(int(d) for d in N.as_tensor(psf).shape) > int(self.psf_size):
            raise ValueError(f"psf of shape {tuple(N.as_tensor(psf).shape)} is larger than psf_size {self.psf_size}")
        out = torch.empty_like(image)
        if image.dim() == 2:
            richardson_lucy_into(image, psf, out, self.settings())
        else:
            richardson_lucy_nd_into(image, psf, out, self.settings())
        return out


def _blocks(dims3, block, halo):
    return schedule(tuple(dims3), tuple(block), tuple(halo))


def deconvolve_blocks(image, psfs, op: RichardsonLucyOp, block: Sequence[int] = (512, 512), halo: int | None = None,
                      channel_axis: bool | None = None, client: N.Client | None = None):
    """Deconvolves ``image`` block by block, all blocks in one batched plan; returns the output.

    ``image``: float32 device tensor ``[H, W]``, ``[C, H, W]`` (channels, 2-D PSFs), ``[D, H, W]``
    (one volume, 3-D PSF) or ``[C, D, H, W]``. ``psfs``: one PSF, or one per channel. ``block``:
    the owned block size per spatial axis (2 or 3 values). ``halo``: context per side; default
    ``op.halo`` of the largest PSF (exact). ``channel_axis``: whether axis 0 is channels (default:
    inferred from the PSF rank).
    """
    import torch

    op.validate()
    image = N.as_tensor(image)
    psf_list = [N.as_tensor(p) for p in (psfs if isinstance(psfs, (list, tuple)) else [psfs])]
    nd = psf_list[0].dim()
    if any(p.dim() != nd for p in psf_list) or nd not in (2, 3):
        raise ValueError("every PSF has the same rank, 2 or 3")
    if channel_axis is None:
        channel_axis = image.dim() == nd + 1
    chans = image if channel_axis else image.unsqueeze(0)
    if chans.dim() != nd + 1:
        raise ValueError(f"image of rank {image.dim()} does not fit {nd}-D PSFs")
    C = chans.shape[0]
    if len(psf_list) not in (1, C):
        raise ValueError(f"{len(psf_list)} PSFs for {C} channels")
    pidx = [c if len(psf_list) == C else 0 for c in range(C)]
    spatial = tuple(int(d) for d in chans.shape[1:])
    dims3 = (1,) * (3 - nd) + spatial
    blk = (1,) * (3 - nd) + tuple(int(b) for b in block)
    if len(blk) != 3:
        raise ValueError(f"block needs {nd} sizes")
    h = op.halo(max((tuple(p.shape) for p in psf_list), key=max)) if halo is None else int(halo)
    hal = (0,) * (3 - nd) + (h,) * nd
    blocks = _blocks(dims3, blk, hal)

    if not chans.is_cuda or chans.dtype != torch.float32:
        raise ValueError("image must be a float32 device tensor")
    chans = chans.contiguous()
    result = torch.empty_like(chans)
    if not blocks:
        return result if channel_axis else result[0]
    # Zero-copy block walk: every tile reads its block straight from the image (row and plane
    # strides of the channel) and stores only its owned part straight into the result.
    esz = chans.element_size()
    row = spatial[-1]
    plane = spatial[-1] * spatial[-2]
    shapes, in_ptrs, out_ptrs, layouts, index = [], [], [], [], []
    for c in range(C):
        base_in, base_out = chans[c].data_ptr(), result[c].data_ptr()
        for b in blocks:
            ro, o = b.read_origin, b.origin
            shapes.append(tuple(b.read_size[3 - nd:]))
            in_ptrs.append(base_in + esz * (ro[0] * plane + ro[1] * row + ro[2]))
            out_ptrs.append(base_out + esz * (o[0] * plane + o[1] * row + o[2]))
            layouts.append(TileLayout(row, plane, row, plane, b.inner_offset(), b.size))
            index.append(pidx[c])
    plan = RLBatch(shapes, [tuple(p.shape) for p in psf_list], index, op.settings(), client)
    plan.execute_layout(in_ptrs, out_ptrs, layouts, psf_list)
    if client is not None:
        client.sync()  # the caller's next work may run on torch's current stream
    plan.close()
    return result if channel_axis else result[0]
