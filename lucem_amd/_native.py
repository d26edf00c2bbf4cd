"""ctypes binding of ``liblucem_hip.so`` (the C ABI declared in ``include/lucem_hip.h``).

This is the Python side of the drop-in boundary: arrays cross it as raw device pointers taken
from torch tensors (or anything exporting DLPack, converted with ``torch.from_dlpack``). The
library must be built (``python -m lucem_amd.build``); there is no fallback path, and a missing
library or device raises.
"""

from __future__ import annotations

import atexit
import ctypes
import os
import threading
import weakref
from pathlib import Path

_PKG = Path(__file__).resolve().parent
LIB_PATH = Path(os.environ.get("LUCEM_HIP_LIB", str(_PKG / "liblucem_hip.so")))

LUCEM_OK = 0
STATUS_NAMES = {
    0: "OK",
    1: "SHAPE_MISMATCH",
    2: "ALIASED_BUFFERS",
    3: "INVALID_DIMENSION",
    4: "INVALID_WEIGHTS",
    5: "INVALID_PARAMETER",
    6: "TOO_MANY_ELEMENTS",
    7: "OUT_OF_DEVICE_MEMORY",
    8: "UNKNOWN_PAD_MODE",
    9: "UNSUPPORTED_SHAPE",
    10: "HIP_ERROR",
    11: "INVALID_HANDLE",
    12: "AXIS_OUT_OF_BOUNDS",
    13: "EMPTY_INPUT",
    14: "INVALID_ORIGIN",
}

# Every symbol include/lucem_hip.h declares (checked by the CPU tests).
EXPORTED = [
    "lucem_last_error", "lucem_version", "lucem_rl_settings_default", "lucem_pad_mode_parse",
    "lucem_client_create", "lucem_client_destroy", "lucem_client_sync", "lucem_client_stream",
    "lucem_client_num_cus", "lucem_workspace_create", "lucem_workspace_destroy",
    "lucem_workspace_capacity_bytes", "lucem_richardson_lucy_into_f32", "lucem_richardson_lucy_f32",
    "lucem_richardson_lucy_nd_into_f32", "lucem_richardson_lucy_into_f64", "lucem_richardson_lucy_f64",
    "lucem_fft_c2c_lines_f32", "lucem_client_profile",
    "lucem_client_profile_read", "lucem_rl_batch_create", "lucem_rl_batch_execute", "lucem_rl_batch_execute_layout", "lucem_rl_batch_destroy",
    "lucem_rl_batch_device_bytes", "lucem_rl_batch_launches", "lucem_richardson_lucy_batched_f32",
    "lucem_wiener_settings_default", "lucem_wiener_deconvolve_into_f32", "lucem_wiener_deconvolve_f32",
    "lucem_unsupervised_wiener_settings_default", "lucem_unsupervised_wiener_f32",
    "lucem_fft_1d_into_f32", "lucem_fft_nd_into_f32", "lucem_complex_mul_into_f32", "lucem_add_noise_f32",
    "lucem_correlate_settings_default", "lucem_correlate_2d_into_f32", "lucem_correlate_nd_into_f32",
    "lucem_correlate_2d_into_f64", "lucem_render_ellipses_f32", "lucem_render_spheres_f32",
    "lucem_phase_correlation_settings_default", "lucem_phase_correlation_f32",
    "lucem_fourier_gaussian_into_f32", "lucem_fourier_box_into_f32", "lucem_fourier_shift_into_f32",
    "lucem_butterworth_f32", "lucem_gabor_settings_default", "lucem_gabor_2d_f32",
    "lucem_correlate1d_settings_default", "lucem_correlate_1d_into_f32", "lucem_correlate_1d_into_f64",
    "lucem_richardson_lucy_into_noalloc_f32", "lucem_noalloc_scratch_bytes", "lucem_rl_grad_forward_target",
    "lucem_parse_bytes", "lucem_set_device_memory_limit", "lucem_device_memory_limit", "lucem_device_bytes_in_use",
    "lucem_check_allocation",
]


class LucemError(RuntimeError):
    """A non-OK status of the C ABI (mirrors lucem_ops::OpsError variants)."""

    def __init__(self, status: int, message: str):
        self.status = status
        self.kind = STATUS_NAMES.get(status, str(status))
        super().__init__(f"{self.kind}: {message}")


class KernelStat(ctypes.Structure):
    _fields_ = [
        ("name", ctypes.c_char * 48),
        ("launches", ctypes.c_uint64),
        ("total_ms", ctypes.c_double),
        ("bytes", ctypes.c_double),
    ]


class RlSettings(ctypes.Structure):
    _fields_ = [
        ("num_iter", ctypes.c_uint64),
        ("clip", ctypes.c_int32),
        ("has_filter_epsilon", ctypes.c_int32),
        ("filter_epsilon", ctypes.c_double),
        ("mode", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
    ]


class WienerSettings(ctypes.Structure):
    _fields_ = [("balance", ctypes.c_double), ("clip", ctypes.c_int32), ("reserved", ctypes.c_uint32)]


class UnsupervisedWienerSettings(ctypes.Structure):
    _fields_ = [
        ("num_iter", ctypes.c_uint64),
        ("burnin", ctypes.c_uint64),
        ("clip", ctypes.c_int32),
        ("reserved", ctypes.c_uint32),
        ("seed", ctypes.c_uint64),
    ]


class CorrelateSettings(ctypes.Structure):
    _fields_ = [("mode", ctypes.c_uint32), ("origin", ctypes.c_int32 * 3), ("cval", ctypes.c_double)]


class PhaseCorrelationSettings(ctypes.Structure):
    _fields_ = [("upsample_factor", ctypes.c_uint64), ("normalize_phase", ctypes.c_int32),
                ("reserved", ctypes.c_uint32)]


class PhaseCorrelationResult(ctypes.Structure):
    _fields_ = [("shift", ctypes.c_double * 3), ("ndim", ctypes.c_uint32), ("reserved", ctypes.c_uint32),
                ("error", ctypes.c_double), ("phasediff", ctypes.c_double)]


class Ellipse(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("cy", "cx", "a", "b", "cos_theta", "sin_theta", "value")]


class Sphere(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("cz", "cy", "cx", "radius", "radius_sq", "value")]


class Correlate1dSettings(ctypes.Structure):
    _fields_ = [("axis", ctypes.c_int64), ("mode", ctypes.c_uint32), ("origin", ctypes.c_int32),
                ("cval", ctypes.c_double)]


class GaborSettings(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("frequency", "theta", "sigma_x", "sigma_y", "n_stds", "offset")] + [
        ("mode", ctypes.c_uint32), ("reserved", ctypes.c_uint32), ("cval", ctypes.c_double)]


class NoiseModelC(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_uint32), ("reserved", ctypes.c_uint32), ("p0", ctypes.c_double),
                ("p1", ctypes.c_double)]


_lib = None
_lock = threading.Lock()
_u64p = ctypes.POINTER(ctypes.c_uint64)
_vp = ctypes.c_void_p


def lib() -> ctypes.CDLL:
    """The loaded library. Raises if it has not been built."""
    global _lib
    with _lock:
        if _lib is None:
            if not LIB_PATH.exists():
                raise RuntimeError(f"{LIB_PATH} is not built: run `python -m lucem_amd.build` (hipcc, gfx950)")
            L = ctypes.CDLL(str(LIB_PATH))
            L.lucem_last_error.restype = ctypes.c_char_p
            L.lucem_version.restype = ctypes.c_char_p
            L.lucem_rl_settings_default.argtypes = [ctypes.POINTER(RlSettings)]
            L.lucem_pad_mode_parse.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_uint32)]
            L.lucem_parse_bytes.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_uint64)]
            L.lucem_parse_bytes.restype = ctypes.c_int32
            L.lucem_set_device_memory_limit.argtypes = [ctypes.c_int32, ctypes.c_uint64]
            L.lucem_set_device_memory_limit.restype = None
            L.lucem_device_memory_limit.argtypes = [ctypes.POINTER(ctypes.c_uint64)]
            L.lucem_device_memory_limit.restype = ctypes.c_int32
            L.lucem_device_bytes_in_use.argtypes = [ctypes.c_int]
            L.lucem_device_bytes_in_use.restype = ctypes.c_uint64
            L.lucem_check_allocation.argtypes = [ctypes.c_int, ctypes.c_char_p, ctypes.c_uint64]
            L.lucem_client_create.argtypes = [ctypes.c_int, _vp, ctypes.c_int32, ctypes.POINTER(_vp)]
            L.lucem_client_destroy.argtypes = [_vp]
            L.lucem_client_sync.argtypes = [_vp]
            L.lucem_client_stream.argtypes = [_vp]
            L.lucem_client_stream.restype = _vp
            L.lucem_client_num_cus.argtypes = [_vp]
            L.lucem_workspace_create.argtypes = [_vp, ctypes.POINTER(_vp)]
            L.lucem_workspace_destroy.argtypes = [_vp]
            L.lucem_workspace_capacity_bytes.argtypes = [_vp]
            L.lucem_workspace_capacity_bytes.restype = ctypes.c_uint64
            rl = [_vp, _vp, _u64p, ctypes.c_uint32, _vp, _u64p, ctypes.c_uint32, _vp, _u64p, ctypes.c_uint32,
                  ctypes.POINTER(RlSettings), _vp]
            L.lucem_richardson_lucy_into_f32.argtypes = rl
            L.lucem_richardson_lucy_nd_into_f32.argtypes = rl
            L.lucem_richardson_lucy_f32.argtypes = [_vp, _vp, _u64p, ctypes.c_uint32, _vp, _u64p, ctypes.c_uint32,
                                                    _vp, ctypes.POINTER(RlSettings)]
            L.lucem_richardson_lucy_into_f64.argtypes = rl
            L.lucem_richardson_lucy_into_noalloc_f32.argtypes = rl[:-1]
            L.lucem_noalloc_scratch_bytes.argtypes = []
            L.lucem_noalloc_scratch_bytes.restype = ctypes.c_uint64
            L.lucem_richardson_lucy_f64.argtypes = L.lucem_richardson_lucy_f32.argtypes
            L.lucem_fft_c2c_lines_f32.argtypes = [_vp, _vp, _vp, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int32]
            L.lucem_add_noise_f32.argtypes = [_vp, _vp, _u64p, ctypes.c_uint32, _vp, ctypes.POINTER(NoiseModelC),
                                              ctypes.c_uint64, ctypes.c_int32]
            arr = [_vp, _u64p, ctypes.c_uint32]
            L.lucem_fft_1d_into_f32.argtypes = [_vp] + arr * 4 + [ctypes.c_int64, ctypes.c_int32]
            L.lucem_fft_nd_into_f32.argtypes = [_vp] + arr * 4 + [ctypes.c_int32, _vp]
            L.lucem_complex_mul_into_f32.argtypes = [_vp] + arr + [_vp] + arr + [_vp] + arr + [_vp, ctypes.c_int32]
            L.lucem_wiener_settings_default.argtypes = [ctypes.POINTER(WienerSettings)]
            L.lucem_wiener_settings_default.restype = None
            L.lucem_wiener_deconvolve_into_f32.argtypes = rl[:-2] + [ctypes.POINTER(WienerSettings), _vp]
            L.lucem_wiener_deconvolve_f32.argtypes = [_vp, _vp, _u64p, ctypes.c_uint32, _vp, _u64p, ctypes.c_uint32,
                                                      _vp, ctypes.POINTER(WienerSettings)]
            L.lucem_unsupervised_wiener_settings_default.argtypes = [ctypes.POINTER(UnsupervisedWienerSettings)]
            L.lucem_unsupervised_wiener_settings_default.restype = None
            L.lucem_unsupervised_wiener_f32.argtypes = [
                _vp, _vp, _u64p, ctypes.c_uint32, _vp, _u64p, ctypes.c_uint32, _vp,
                ctypes.POINTER(UnsupervisedWienerSettings), ctypes.POINTER(ctypes.c_double), ctypes.c_uint64,
                ctypes.POINTER(ctypes.c_uint64), _vp]
            L.lucem_correlate_settings_default.argtypes = [ctypes.POINTER(CorrelateSettings)]
            L.lucem_correlate_settings_default.restype = None
            L.lucem_correlate_2d_into_f32.argtypes = rl[:-2] + [ctypes.POINTER(CorrelateSettings), _vp]
            L.lucem_correlate_nd_into_f32.argtypes = L.lucem_correlate_2d_into_f32.argtypes
            L.lucem_correlate_2d_into_f64.argtypes = rl[:-2] + [ctypes.POINTER(CorrelateSettings)]
            L.lucem_render_ellipses_f32.argtypes = [_vp, _vp, _u64p, ctypes.c_uint32, _vp, ctypes.c_uint64,
                                                    ctypes.c_double]
            L.lucem_render_spheres_f32.argtypes = L.lucem_render_ellipses_f32.argtypes
            L.lucem_phase_correlation_settings_default.argtypes = [ctypes.POINTER(PhaseCorrelationSettings)]
            L.lucem_phase_correlation_settings_default.restype = None
            L.lucem_phase_correlation_f32.argtypes = [_vp, _vp, _u64p, ctypes.c_uint32, _vp, _u64p, ctypes.c_uint32,
                                                      ctypes.POINTER(PhaseCorrelationSettings),
                                                      ctypes.POINTER(PhaseCorrelationResult), _vp]
            spec = [_vp, _vp, _u64p, ctypes.c_uint32, _vp, _u64p, ctypes.c_uint32, ctypes.POINTER(ctypes.c_double),
                    ctypes.c_uint32]
            L.lucem_fourier_gaussian_into_f32.argtypes = spec
            L.lucem_fourier_box_into_f32.argtypes = spec
            L.lucem_fourier_shift_into_f32.argtypes = spec
            L.lucem_butterworth_f32.argtypes = [_vp, _vp, _u64p, ctypes.c_uint32, _vp, _u64p, ctypes.c_uint32,
                                                ctypes.c_double, ctypes.c_int32, ctypes.c_double, _vp]
            L.lucem_correlate1d_settings_default.argtypes = [ctypes.POINTER(Correlate1dSettings)]
            L.lucem_correlate1d_settings_default.restype = None
            c1 = [_vp, _vp, _u64p, ctypes.c_uint32, _vp, ctypes.c_uint64, _vp, _u64p, ctypes.c_uint32,
                  ctypes.POINTER(Correlate1dSettings)]
            L.lucem_correlate_1d_into_f32.argtypes = c1
            L.lucem_correlate_1d_into_f64.argtypes = c1
            L.lucem_gabor_settings_default.argtypes = [ctypes.POINTER(GaborSettings)]
            L.lucem_gabor_settings_default.restype = None
            L.lucem_gabor_2d_f32.argtypes = [_vp, _vp, _u64p, ctypes.c_uint32, _vp, _vp, ctypes.POINTER(GaborSettings),
                                             _vp]
            L.lucem_client_profile.argtypes = [_vp, ctypes.c_int32]
            L.lucem_client_profile_read.argtypes = [_vp, ctypes.POINTER(KernelStat), ctypes.c_uint32,
                                                    ctypes.POINTER(ctypes.c_uint32)]
            _register_optional(L)
            _lib = L
        return _lib


def _register_optional(L: ctypes.CDLL) -> None:
    """Signatures of the batched and gradient entry points (bound in their modules)."""
    from . import _native_ext

    _native_ext.register(L)


def check(status: int) -> None:
    if status != LUCEM_OK:
        msg = lib().lucem_last_error().decode(errors="replace")
        raise LucemError(int(status), msg)


def dims(shape) -> ctypes.Array:
    shape = [int(s) for s in shape]
    return (ctypes.c_uint64 * max(len(shape), 1))(*shape)


def version() -> str:
    return lib().lucem_version().decode()


def parse_bytes(text: str) -> int | None:
    """Bytes of a size such as ``"7GiB"``, ``" 6.5 GB "`` or ``"7516192768"`` (``None`` if
    unparsable), as lucem-core's ``parse_bytes`` (memory/budget.rs:263-279)."""
    out = ctypes.c_uint64()
    return int(out.value) if lib().lucem_parse_bytes(text.encode(), ctypes.byref(out)) else None


def set_device_memory_limit(limit: int | None) -> None:
    """Sets (or with ``None`` clears) the device memory budget of this library's workspace slots
    and plans; overrides ``LUCEM_DEVICE_MEMORY`` / ``LUCEM_DEVICE_MEMORY_LIMIT`` (budget.rs:175-179)."""
    lib().lucem_set_device_memory_limit(0 if limit is None else 1, 0 if limit is None else int(limit))


def device_memory_limit() -> int | None:
    """The device memory limit in bytes, or ``None`` without one (budget.rs:182-187)."""
    out = ctypes.c_uint64()
    return int(out.value) if lib().lucem_device_memory_limit(ctypes.byref(out)) else None


def device_bytes_in_use(device: int = 0) -> int:
    """Bytes this library holds on ``device`` in workspace slots and plans."""
    return int(lib().lucem_device_bytes_in_use(int(device)))


def check_allocation(what: str, nbytes: int, device: int = 0) -> None:
    """Raises ``LucemError`` (``OUT_OF_DEVICE_MEMORY``) when ``nbytes`` more would not fit the
    budget (budget.rs:232-244); no-op without a limit or below 64 MiB."""
    check(lib().lucem_check_allocation(int(device), what.encode(), int(nbytes)))


def last_error() -> str:
    return lib().lucem_last_error().decode(errors="replace")


_owned = {"plans": weakref.WeakSet(), "workspaces": weakref.WeakSet(), "clients": weakref.WeakSet()}
_exit_hook = False


def track(obj, kind: str) -> None:
    """Registers a library object (kind "plans", "workspaces" or "clients") for release at exit.

    The exit hook releases plans first and clients last, while the HIP runtime and torch are
    still whole: it is registered after torch has been imported, so it runs before torch's own
    exit handlers (atexit runs last-in, first-out), instead of leaving the order to the garbage
    collection of interpreter shutdown.
    """
    global _exit_hook
    _owned[kind].add(obj)
    if not _exit_hook:
        atexit.register(_release_all)
        _exit_hook = True


def _release_all() -> None:
    for kind in ("plans", "workspaces", "clients"):
        for obj in list(_owned[kind]):
            try:
                obj.close()
            except Exception:
                pass


class Client:
    """A device and a HIP stream (mirrors lucem_core::compute::Client).

    By default the client runs on torch's current stream of ``device`` so that library work is
    ordered with the torch operations that produce its inputs and consume its outputs.
    """

    def __init__(self, device: int = 0, stream: int | None = None, own_stream: bool = False):
        import torch

        if not torch.cuda.is_available():
            raise RuntimeError("no HIP device is visible: lucem_amd runs on MI355X only")
        if stream is None and not own_stream:
            stream = torch.cuda.current_stream(device).cuda_stream
        h = _vp()
        check(lib().lucem_client_create(int(device), _vp(stream or None), 1 if own_stream else 0, ctypes.byref(h)))
        self._h = h
        self.device = int(device)
        track(self, "clients")

    @property
    def handle(self) -> ctypes.c_void_p:
        return self._h

    def sync(self) -> None:
        check(lib().lucem_client_sync(self._h))

    @property
    def stream(self) -> int:
        return int(lib().lucem_client_stream(self._h) or 0)

    @property
    def num_cus(self) -> int:
        return int(lib().lucem_client_num_cus(self._h))

    def profile(self, enable: bool = True) -> None:
        """Starts (clearing earlier records) or stops per-kernel event timing."""
        check(lib().lucem_client_profile(self._h, 1 if enable else 0))

    def profile_read(self) -> dict:
        """``{kernel: {"launches", "total_ms", "bytes"}}`` since profiling was enabled (clears)."""
        arr = (KernelStat * 64)()
        n = ctypes.c_uint32()
        check(lib().lucem_client_profile_read(self._h, arr, 64, ctypes.byref(n)))
        out = {}
        for i in range(min(n.value, 64)):
            s = arr[i]
            out[s.name.decode()] = {"launches": int(s.launches), "total_ms": float(s.total_ms),
                                    "bytes": float(s.bytes)}
        return out

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            lib().lucem_client_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Workspace:
    """Grow-only device buffers reused across calls (mirrors lucem_core::memory::Workspace)."""

    def __init__(self, client: Client):
        h = _vp()
        check(lib().lucem_workspace_create(client.handle, ctypes.byref(h)))
        self._h = h
        self._client = client
        track(self, "workspaces")

    @property
    def handle(self) -> ctypes.c_void_p:
        return self._h

    def capacity_bytes(self) -> int:
        return int(lib().lucem_workspace_capacity_bytes(self._h))

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            lib().lucem_workspace_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_clients: dict[tuple[int, int], Client] = {}


def default_client(device: int | None = None) -> Client:
    """One client per (device, torch current stream)."""
    import torch

    dev = torch.cuda.current_device() if device is None else int(device)
    key = (dev, int(torch.cuda.current_stream(dev).cuda_stream))
    c = _default_clients.get(key)
    if c is None:
        c = Client(dev, key[1])
        _default_clients[key] = c
    return c


def as_tensor(x):
    """A torch tensor view of ``x`` (torch tensor or any DLPack exporter)."""
    import torch

    if isinstance(x, torch.Tensor):
        return x
    if hasattr(x, "__dlpack__"):
        return torch.from_dlpack(x)
    raise TypeError(f"expected a torch tensor or a DLPack-exporting array, got {type(x).__name__}")


def device_ptr(t, what: str, dtype=None) -> int:
    """Raw device pointer of a contiguous HIP tensor of ``dtype`` (default float32)."""
    import torch

    dtype = torch.float32 if dtype is None else dtype
    if not t.is_cuda:
        raise ValueError(f"{what} must be a device (HIP) tensor")
    if t.dtype != dtype:
        raise ValueError(f"{what} must be {dtype}, got {t.dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{what} must be C-contiguous")
    return int(t.data_ptr())
