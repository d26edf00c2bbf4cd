"""``lucem._native`` for the Richardson-Lucy path on MI355X.

The reference builds one extension module under this name (bindings/python/pyproject.toml:34,
bindings/python/lucem-py/src/lib.rs:8-15). Its operator catalog functions
(bindings/python/lucem-py-core/src/ops.rs) are what the reference's ``lucem.ops`` calls: here
``op_catalog``, ``op_halo`` and ``run_op`` serve the ``richardson_lucy`` operator (the variant a
pipeline would add to engine/lucem-pipeline/src/ops.rs:97-203, its patch in INTEGRATION.md),
running on ``liblucem_hip.so`` through :mod:`lucem_amd`. Arrays cross as NumPy (uploaded, run,
downloaded, as the reference does) or as device tensors (torch or any DLPack exporter, left on the
device). Bad specifications and parameters raise ``ValueError``, failures inside the library
``RuntimeError`` (lucem-py-core/src/lib.rs:22-29).

The module also carries the device-level forms: ``richardson_lucy_into`` / ``richardson_lucy`` (and
their N-D twins), the batched plan ``RLBatch`` with ``richardson_lucy_batched``, and the
differentiable form ``RichardsonLucyNLL`` (a ``torch.autograd.Function``) with ``rl_nll`` and
``gaussian_psf``.
"""

from __future__ import annotations

import json

import numpy as np

from lucem_amd import _native as _N
from lucem_amd.autograd import RichardsonLucyNLL, RLGradPlan, gaussian_psf, rl_nll  # noqa: F401
from lucem_amd.batched import RLBatch  # noqa: F401
from lucem_amd.restore import (  # noqa: F401
    PadMode,
    RichardsonLucy,
    richardson_lucy,
    richardson_lucy_into,
    richardson_lucy_nd,
    richardson_lucy_nd_into,
)
from lucem_amd.tiled import RichardsonLucyOp

__version__ = "0.0.6+mi355x"

_RL = RichardsonLucyOp(num_iter=30, clip=False)  # catalog defaults: unclipped photon counts


def _entry() -> dict:
    return {"tag": "richardson_lucy", "title": "Richardson-Lucy", "family": "Restore", "params": _RL.params(),
            "halo": _RL.halo(), "inputs": [{"type": "image", "name": "image"}, {"type": "image", "name": "psf"}],
            "output": "image"}


def op_catalog() -> str:
    """Every operator this module runs, as JSON (tag, title, family, parameters with defaults,
    halo, inputs, output), in the shape of the reference's ``op_catalog``."""
    return json.dumps([_entry()])


def _parse(spec: str) -> RichardsonLucyOp:
    try:
        d = json.loads(spec)
    except (TypeError, json.JSONDecodeError) as e:
        raise ValueError(f"operator spec is not JSON: {e}") from None
    if not isinstance(d, dict) or "op" not in d:
        raise ValueError("missing field `op`")
    tag = d.pop("op")
    if tag != "richardson_lucy":
        raise ValueError(f"unknown variant `{tag}`: this build runs `richardson_lucy`")
    op = RichardsonLucyOp.from_params(d)
    op.validate()
    return op


def op_halo(spec: str) -> int:
    """Pixels of context the operator (its JSON) reads around a pixel."""
    return _parse(spec).halo()


def run_op(spec: str, inputs):
    """Runs one operator on its inputs, in port order (for ``richardson_lucy``: image, PSF).

    NumPy inputs are uploaded and the result comes back as a float32 NumPy array; device tensors
    stay on the device and the result is a device tensor."""
    import torch

    op = _parse(spec)
    if len(inputs) != 2:
        raise ValueError(f"Richardson-Lucy takes 2 inputs (image, psf), {len(inputs)} given")
    host = not any(isinstance(a, torch.Tensor) for a in inputs)
    if host:
        image, psf = (np.ascontiguousarray(a, dtype=np.float32) for a in inputs)
        if image.ndim not in (2, 3) or psf.ndim != image.ndim:
            raise ValueError(f"Richardson-Lucy: image and psf must both be 2-D or 3-D, got {image.ndim}-D "
                             f"and {psf.ndim}-D")
        out = op.run(torch.from_numpy(image).cuda(), torch.from_numpy(psf).cuda())
        return out.cpu().numpy()
    image, psf = (_N.as_tensor(a) for a in inputs)
    return op.run(image, psf)


def run_chain(specs, input):
    """The reference runs one-input operators back to back on the device; Richardson-Lucy takes a
    second input (the PSF), so it cannot stand in a chain: a chain of it raises ``ValueError``."""
    for s in specs:
        _parse(s)
    raise ValueError("richardson_lucy takes two inputs (image, psf) and cannot be a stage of a chain; use run_op")


def richardson_lucy_batched(images, outputs, psfs, psf_index, settings: RichardsonLucy | None = None):
    """Runs many tiles of uneven size (device tensors) in one batched call; ``psf_index[i]`` names
    tile i's PSF in ``psfs``. Plans once per call; keep an ``RLBatch`` to reuse a plan."""
    plan = RLBatch([tuple(t.shape) for t in images], [tuple(p.shape) for p in psfs], psf_index, settings)
    try:
        plan.execute(images, outputs, psfs)
    finally:
        plan.close()
