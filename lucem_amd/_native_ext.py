"""Signatures of the batched and differentiable entry points of ``liblucem_hip.so``."""

from __future__ import annotations

import ctypes

_vp = ctypes.c_void_p
_u64p = ctypes.POINTER(ctypes.c_uint64)


class RlTile(ctypes.Structure):
    """``lucem_rl_tile``: one tile of a batch (device pointers, dims, PSF index)."""

    _fields_ = [
        ("input", _vp),
        ("output", _vp),
        ("dims", ctypes.c_uint64 * 3),
        ("ndim", ctypes.c_uint32),
        ("psf_index", ctypes.c_uint32),
    ]


class RlTileLayout(ctypes.Structure):
    """``lucem_rl_tile_layout``: strides of a tile's input and output views and its stored window."""

    _fields_ = [
        ("in_row_stride", ctypes.c_uint64),
        ("in_plane_stride", ctypes.c_uint64),
        ("out_row_stride", ctypes.c_uint64),
        ("out_plane_stride", ctypes.c_uint64),
        ("own_origin", ctypes.c_uint64 * 3),
        ("own_size", ctypes.c_uint64 * 3),
    ]


def register(L: ctypes.CDLL) -> None:
    """Declares argument types of the optional entry points present in ``L``."""
    from ._native import RlSettings

    if hasattr(L, "lucem_rl_batch_create"):
        L.lucem_rl_batch_create.argtypes = [_vp, ctypes.POINTER(RlTile), ctypes.c_uint64, _u64p, ctypes.c_uint32,
                                            ctypes.c_uint64, ctypes.POINTER(RlSettings), ctypes.POINTER(_vp)]
        L.lucem_rl_batch_execute.argtypes = [_vp, ctypes.POINTER(RlTile), ctypes.POINTER(_vp)]
        L.lucem_rl_batch_execute_layout.argtypes = [_vp, ctypes.POINTER(RlTile), ctypes.POINTER(RlTileLayout),
                                                    ctypes.POINTER(_vp)]
        L.lucem_rl_batch_destroy.argtypes = [_vp]
        L.lucem_rl_batch_device_bytes.argtypes = [_vp]
        L.lucem_rl_batch_device_bytes.restype = ctypes.c_uint64
        L.lucem_rl_batch_launches.argtypes = [_vp]
        L.lucem_rl_batch_launches.restype = ctypes.c_uint64
        L.lucem_richardson_lucy_batched_f32.argtypes = [_vp, ctypes.POINTER(RlTile), ctypes.c_uint64,
                                                        ctypes.POINTER(_vp), _u64p, ctypes.c_uint32, ctypes.c_uint64,
                                                        ctypes.POINTER(RlSettings)]
