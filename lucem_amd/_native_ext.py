"""Signatures of the batched and differentiable entry points of ``liblucem_hip.so``."""

from __future__ import annotations

import ctypes


def register(L: ctypes.CDLL) -> None:
    """Declares argument types of the optional entry points present in ``L``."""
    return None
