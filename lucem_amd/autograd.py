"""Differentiable Richardson-Lucy for PyTorch (ROCm): ``torch.autograd.Function`` over the HIP
forward-with-checkpoints and hand-written adjoint (``lucem_rl_grad_*`` in ``include/lucem_hip.h``).

The model is the reference RL iteration (crates/lucem-ops/src/restore/deconvolve.rs:109-147,
no clip) run for ``num_iter`` iterations from a given initial estimate ``x0``, followed by the
Poisson negative log-likelihood of the re-blurred estimate, per tile:

    loss_t = sum(u - y_t * log(u)),   u = A x_T,   A = correlate with the flipped normalised PSF

Gradients flow to ``x0`` and to the PSF (through the reference's normalisation psf / sum(psf),
deconvolve.rs:77-79), so a parametric PSF built in torch (e.g. :func:`gaussian_psf`) gets
gradients for its parameters by ordinary autograd.
"""

from __future__ import annotations

import ctypes
from collections import OrderedDict
from typing import Sequence

import torch

from . import _native as N
from .restore import PadMode, RichardsonLucy

_u64p = ctypes.POINTER(ctypes.c_uint64)


def _register(L) -> None:
    if getattr(L, "_lucem_grad_registered", False):
        return
    L.lucem_rl_grad_create.argtypes = [N._vp, _u64p, ctypes.c_uint32, ctypes.c_uint64, _u64p,
                                       ctypes.POINTER(N.RlSettings), ctypes.c_uint32, ctypes.POINTER(N._vp)]
    L.lucem_rl_grad_forward.argtypes = [N._vp, ctypes.POINTER(N._vp), ctypes.POINTER(N._vp), N._vp,
                                        ctypes.POINTER(N._vp), N._vp]
    L.lucem_rl_grad_forward_target.argtypes = [N._vp, ctypes.POINTER(N._vp), ctypes.POINTER(N._vp),
                                               ctypes.POINTER(N._vp), N._vp, ctypes.POINTER(N._vp), N._vp]
    L.lucem_rl_grad_backward.argtypes = [N._vp, N._vp, ctypes.POINTER(N._vp), N._vp]
    L.lucem_rl_grad_destroy.argtypes = [N._vp]
    L.lucem_rl_grad_device_bytes.argtypes = [N._vp]
    L.lucem_rl_grad_device_bytes.restype = ctypes.c_uint64
    L._lucem_grad_registered = True


class RLGradPlan:
    """Plan of the differentiable form for fixed tile shapes, PSF shape and settings."""

    def __init__(self, shapes: Sequence[Sequence[int]], psf_shape: Sequence[int], num_iter: int,
                 checkpoint_every: int = 5, mode=PadMode.Reflect, filter_epsilon: float | None = None,
                 client: N.Client | None = None):
        L = N.lib()
        _register(L)
        self.client = client or N.default_client()
        self.shapes = [tuple(int(d) for d in s) for s in shapes]
        ranks = {len(s) for s in self.shapes}
        if len(ranks) != 1 or len(psf_shape) != next(iter(ranks)):
            raise ValueError("tiles and PSF must share one rank (2 or 3)")
        self.ndim = len(psf_shape)
        self.psf_shape = tuple(int(d) for d in psf_shape)
        flat = [d for s in self.shapes for d in s]
        dims = (ctypes.c_uint64 * len(flat))(*flat)
        pd = (ctypes.c_uint64 * self.ndim)(*self.psf_shape)
        s = RichardsonLucy(num_iter=num_iter, clip=False, mode=mode, filter_epsilon=filter_epsilon).to_c()
        h = N._vp()
        N.check(L.lucem_rl_grad_create(self.client.handle, dims, self.ndim, len(self.shapes), pd, ctypes.byref(s),
                                       int(checkpoint_every), ctypes.byref(h)))
        self._h = h
        N.track(self, "plans")
        self.generation = 0  # bumped by every forward; a backward needs its own forward resident

    @property
    def device_bytes(self) -> int:
        return int(N.lib().lucem_rl_grad_device_bytes(self._h))

    def forward(self, x0s, ys, psf, x_outs=None, targets=None):
        """Enqueues the forward; returns the per-tile loss as a float64 device tensor. ``targets``
        (per tile, or None): the counts the loss compares the re-blurred estimate with, in place
        of ``ys`` (RL still runs on ``ys``)."""
        n = len(self.shapes)
        xp = (N._vp * n)(*[N.device_ptr(x, "x0") for x in x0s])
        yp = (N._vp * n)(*[N.device_ptr(y, "y") for y in ys])
        op = None
        if x_outs is not None:
            op = (N._vp * n)(*[N.device_ptr(o, "x_out") if o is not None else None for o in x_outs])
        tp = None
        if targets is not None:
            tp = (N._vp * n)(*[N.device_ptr(t, "target") if t is not None else None for t in targets])
        loss = torch.empty(n, dtype=torch.float64, device=psf.device)
        N.check(N.lib().lucem_rl_grad_forward_target(self._h, xp, yp, tp, N._vp(N.device_ptr(psf, "psf")), op,
                                                     N._vp(loss.data_ptr())))
        self.generation += 1
        return loss

    def backward(self, weight, gx0s, gpsf) -> None:
        n = len(self.shapes)
        gp = (N._vp * n)(*[N.device_ptr(g, "grad_x0") for g in gx0s])
        w = N._vp(N.device_ptr(weight, "loss weight")) if weight is not None else None
        N.check(N.lib().lucem_rl_grad_backward(self._h, w, gp, N._vp(N.device_ptr(gpsf, "grad_psf"))))

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            N.lib().lucem_rl_grad_destroy(self._h)
            self._h = N._vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_plans: "OrderedDict[tuple, RLGradPlan]" = OrderedDict()
_plan_cache_size = 4


def set_plan_cache_size(n: int) -> None:
    """Keeps at most ``n`` (>= 1) plans of :func:`rl_nll`, least recently used first out (default
    4). A plan is keyed by tile shapes, PSF shape and settings, and holds its checkpoints and
    spectra on the device (about 10 KB per pixel of its tiles), so a fit over changing tile sets
    would otherwise keep every plan it ever made."""
    global _plan_cache_size
    if int(n) < 1:
        raise ValueError("the plan cache holds at least one plan")
    _plan_cache_size = int(n)
    while len(_plans) > _plan_cache_size:
        _plans.popitem(last=False)


def clear_plan_cache() -> None:
    """Drops every cached plan of :func:`rl_nll`. A plan still needed by a pending backward stays
    alive through that graph and is freed with it."""
    _plans.clear()


def _plan(shapes, psf_shape, num_iter, checkpoint_every, mode, eps, device) -> RLGradPlan:
    key = (tuple(tuple(int(d) for d in s) for s in shapes), tuple(psf_shape), int(num_iter), int(checkpoint_every),
           int(mode), eps, device)
    p = _plans.get(key)
    if p is None:
        p = RLGradPlan(shapes, psf_shape, num_iter, checkpoint_every, mode, eps, N.default_client(device))
        _plans[key] = p
        while len(_plans) > _plan_cache_size:
            _plans.popitem(last=False)  # the autograd graphs that use it keep it until they are freed
    else:
        _plans.move_to_end(key)
    return p


class RichardsonLucyNLL(torch.autograd.Function):
    """loss[t] = Poisson NLL of A x_T for tile t after ``num_iter`` RL iterations from x0[t]."""

    @staticmethod
    def forward(ctx, psf, num_iter, checkpoint_every, mode, eps, n_tiles, *tensors):
        # tensors: n_tiles x0, n_tiles y, then n_tiles targets or none
        x0s = [t.detach().contiguous() for t in tensors[:n_tiles]]
        ys = [t.detach().contiguous() for t in tensors[n_tiles:2 * n_tiles]]
        tg = [t.detach().contiguous() for t in tensors[2 * n_tiles:]] or None
        psf_c = psf.detach().contiguous()
        plan = _plan([x.shape for x in x0s], psf_c.shape, num_iter, checkpoint_every, mode, eps, psf.device.index)
        loss = plan.forward(x0s, ys, psf_c, targets=tg)
        ctx.plan = plan
        ctx.generation = plan.generation
        ctx.n_tiles = n_tiles
        ctx.n_inputs = len(tensors)
        ctx.save_for_backward(psf_c, *x0s, *ys, *(tg or []))
        return loss

    @staticmethod
    def backward(ctx, grad_loss):
        plan = ctx.plan
        n = ctx.n_tiles
        saved = ctx.saved_tensors
        psf_c, x0s, ys, tg = saved[0], saved[1:1 + n], saved[1 + n:1 + 2 * n], saved[1 + 2 * n:]
        if plan.generation != ctx.generation:  # another forward used the plan since: redo ours
            plan.forward(list(x0s), list(ys), psf_c, targets=list(tg) or None)
            ctx.generation = plan.generation
        w = grad_loss.detach().to(torch.float32).contiguous()
        gx0 = [torch.empty_like(x) for x in x0s]
        gpsf = torch.empty_like(psf_c)
        plan.backward(w, gx0, gpsf)
        return (gpsf, None, None, None, None, None, *gx0, *([None] * (ctx.n_inputs - n)))


def rl_nll(x0, y, psf, num_iter: int, checkpoint_every: int = 5, mode=PadMode.Reflect,
           filter_epsilon: float | None = None, target=None):
    """Per-tile Poisson NLL after unrolled RL (see the module docstring).

    ``x0`` and ``y``: a float32 device tensor ``[N, h, w]`` / ``[N, d, h, w]`` (tiles of one shape)
    or lists of tensors (tiles of different shapes, same rank). ``psf``: float32 device tensor.
    ``target`` (same layout, optional): the counts the NLL is taken against in place of ``y``
    (RL still runs on ``y``). Returns a float64 tensor of N losses, differentiable w.r.t. ``x0``
    and ``psf``.
    """
    x0s = list(x0.unbind(0)) if isinstance(x0, torch.Tensor) else list(x0)
    ys = list(y.unbind(0)) if isinstance(y, torch.Tensor) else list(y)
    if len(x0s) != len(ys):
        raise ValueError("x0 and y must hold the same number of tiles")
    tg = []
    if target is not None:
        tg = list(target.unbind(0)) if isinstance(target, torch.Tensor) else list(target)
        if len(tg) != len(ys):
            raise ValueError("target and y must hold the same number of tiles")
    eps = None if filter_epsilon is None else float(filter_epsilon)
    return RichardsonLucyNLL.apply(psf, int(num_iter), int(checkpoint_every), int(PadMode(mode)), eps, len(x0s),
                                   *x0s, *ys, *tg)


def gaussian_psf(sigma: torch.Tensor, shape: Sequence[int]) -> torch.Tensor:
    """Sampled Gaussian on ``shape`` centred at ``k // 2`` per axis, normalised to sum 1, as a
    differentiable function of ``sigma`` (one value per axis, or a scalar)."""
    sigma = torch.as_tensor(sigma)
    nd = len(shape)
    sig = sigma.reshape(-1).expand(nd) if sigma.numel() == 1 else sigma.reshape(nd)
    grids = torch.meshgrid(*[torch.arange(k, dtype=sig.dtype, device=sig.device) - (k // 2) for k in shape],
                           indexing="ij")
    r2 = sum((g / s) ** 2 for g, s in zip(grids, sig))
    p = torch.exp(-0.5 * r2)
    return p / p.sum()
