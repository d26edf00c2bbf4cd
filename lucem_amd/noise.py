"""Host-side mirror of lucem-ops' ``add_noise`` (crates/lucem-ops/src/intensity/noise.rs:17-143).

``NoiseModel`` mirrors the reference enum (Gaussian, Speckle, Salt, Pepper, SaltAndPepper,
Poisson); :func:`add_noise` returns a new float32 device tensor, :func:`add_noise_into` writes a
given one (which may be the input). The per-pixel counter generator and the seed folding are the
reference's, so a seed gives the reference's uniform draws.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass

from . import _native as N

GAUSSIAN, SPECKLE, SALT, PEPPER, SALT_AND_PEPPER, POISSON = range(6)


@dataclass(frozen=True)
class NoiseModel:
    """One noise model and its parameters (noise.rs:17-32)."""

    kind: int
    p0: float = 0.0
    p1: float = 0.0

    @classmethod
    def Gaussian(cls, mean: float, std: float) -> "NoiseModel":
        return cls(GAUSSIAN, mean, std)

    @classmethod
    def Speckle(cls, mean: float, std: float) -> "NoiseModel":
        return cls(SPECKLE, mean, std)

    @classmethod
    def Salt(cls, amount: float) -> "NoiseModel":
        return cls(SALT, amount)

    @classmethod
    def Pepper(cls, amount: float) -> "NoiseModel":
        return cls(PEPPER, amount)

    @classmethod
    def SaltAndPepper(cls, amount: float, salt_vs_pepper: float) -> "NoiseModel":
        return cls(SALT_AND_PEPPER, amount, salt_vs_pepper)

    @classmethod
    def Poisson(cls, scale: float) -> "NoiseModel":
        return cls(POISSON, scale)

    def to_c(self) -> N.NoiseModelC:
        m = N.NoiseModelC()
        m.kind, m.p0, m.p1 = int(self.kind), float(self.p0), float(self.p1)
        return m


def add_noise_into(input, output, model: NoiseModel, seed: int = 0, clip: bool = True,
                   client: N.Client | None = None) -> None:
    """Writes ``input`` with ``model`` noise into ``output`` (same shape; may be ``input``)."""
    x, o = N.as_tensor(input), N.as_tensor(output)
    if tuple(x.shape) != tuple(o.shape):
        raise ValueError(f"output shape {tuple(o.shape)} differs from input shape {tuple(x.shape)}")
    if client is None:
        client = N.default_client(x.device.index if x.is_cuda else None)
    xp = N.device_ptr(x, "input") if x.numel() else 0
    op = N.device_ptr(o, "output") if o.numel() else 0
    m = model.to_c()
    N.check(N.lib().lucem_add_noise_f32(client.handle, N._vp(xp or None), N.dims(x.shape), x.dim(),
                                        N._vp(op or None), ctypes.byref(m), int(seed) & 0xFFFFFFFFFFFFFFFF,
                                        1 if clip else 0))


def add_noise(input, model: NoiseModel, seed: int = 0, clip: bool = True, client: N.Client | None = None):
    """add_noise (noise.rs:35-143): a new tensor holding ``input`` with ``model`` noise."""
    import torch

    x = N.as_tensor(input)
    out = torch.empty_like(x)
    add_noise_into(x, out, model, seed, clip, client)
    return out
