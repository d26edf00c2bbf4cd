"""PSF fitting through the adjoint: Adam on the parameters of a sampled Gaussian PSF, each step
one batched differentiable RL forward + backward over all volumes (uneven depths allowed).

Two objectives, both through ``num_iter`` unrolled RL iterations from the reference's flat start
(0.5 everywhere, crates/lucem-ops/src/restore/deconvolve.rs:93), see :mod:`lucem_amd.autograd`:

- ``"nll"``: the Poisson NLL of the re-blurred estimate against the counts RL ran on. A delta PSF
  minimises it (one RL iteration then returns the data, and the data are the NLL's own minimum), so
  it does not identify the PSF.
- ``"heldout"``: binomial thinning splits every volume's counts y into two independent Poisson
  halves, y1 ~ Binomial(y, 1/2) and y2 = y - y1, each with half of y's mean. RL runs on y1 and the
  NLL is taken against y2, which RL never saw: an estimate that fits y1's noise predicts y2 badly,
  so the minimum moves to the PSF that explains the data, not to the delta. The minimum still
  depends on ``num_iter``: a short unroll leaves the estimate blurred and favours a narrow PSF, a
  long one fits y1's noise and favours a wide one.
- ``"known"``: the held-out likelihood with RL started from the known object (half its counts,
  like y1) instead of the flat start: the calibration case of beads or a phantom. At the true PSF
  the iterations stay near the object whatever ``num_iter``, so the minimum sits at the true PSF.
"""

from __future__ import annotations

import time
from dataclasses import dataclass, field

import torch

from .autograd import gaussian_psf, rl_nll


@dataclass
class FitResult:
    sigma: list
    history: list = field(default_factory=list)   # (step, loss, sigma) per step
    seconds: float = 0.0
    ms_per_step: float = 0.0


def thin(volumes, seed: int = 0):
    """Binomial thinning: (y1, y2) with y1 ~ Binomial(y, 1/2) and y2 = y - y1 per volume of counts."""
    g = torch.Generator(device=volumes[0].device).manual_seed(int(seed))
    y1 = [torch.binomial(v.round().clamp(min=0), torch.full_like(v, 0.5), generator=g) for v in volumes]
    y2 = [v.round().clamp(min=0) - a for v, a in zip(volumes, y1)]
    return y1, y2


def fit_gaussian_sigma(volumes, psf_shape, sigma0, num_iter: int = 10, steps: int = 100, lr: float = 0.05,
                       betas=(0.9, 0.999), checkpoint_every: int = 5, x0_value: float = 0.5,
                       log_sigma: bool = False, objective: str = "nll", seed: int = 0,
                       objects=None) -> FitResult:
    """Fits sigma (one per axis) of a Gaussian PSF on ``psf_shape`` to ``volumes`` (float32 device
    tensors of one rank, photon counts). ``log_sigma`` optimises log(sigma) instead of sigma.
    ``objective``: ``"nll"``, ``"heldout"`` or ``"known"`` (module docstring; ``"known"`` takes
    the true ``objects``, one float32 device tensor per volume); ``seed`` draws the thinning."""
    if objective not in ("nll", "heldout", "known"):
        raise ValueError(f"unknown objective {objective!r} (nll, heldout or known)")
    if objective == "known" and (objects is None or len(objects) != len(volumes)):
        raise ValueError("the known-object objective needs one object per volume")
    dev = volumes[0].device
    s0 = torch.as_tensor(sigma0, dtype=torch.float32, device=dev).clone()
    param = (s0.log() if log_sigma else s0).requires_grad_(True)
    opt = torch.optim.Adam([param], lr=lr, betas=betas)
    if objective == "known":
        x0 = [(0.5 * o).contiguous() for o in objects]
    else:
        x0 = [torch.full_like(v, x0_value) for v in volumes]
    rl_in, target = (volumes, None) if objective == "nll" else thin(volumes, seed)
    res = FitResult(sigma=s0.tolist())
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for step in range(steps):
        opt.zero_grad(set_to_none=True)
        sigma = param.exp() if log_sigma else param
        psf = gaussian_psf(sigma, psf_shape).float().contiguous()
        loss = rl_nll(x0, rl_in, psf, num_iter, checkpoint_every=checkpoint_every, target=target).sum()
        loss.backward()
        opt.step()
        with torch.no_grad():
            if not log_sigma:
                param.clamp_(min=0.05)
        res.history.append((step, float(loss.detach()), (param.exp() if log_sigma else param).detach().tolist()))
    torch.cuda.synchronize()
    res.seconds = time.perf_counter() - t0
    res.ms_per_step = res.seconds * 1e3 / max(steps, 1)
    res.sigma = (param.exp() if log_sigma else param).detach().tolist()
    return res


def main(argv=None) -> int:
    """The PSF-fitting workload: 16 volumes of 32..128 x 256 x 256 (random spheres blurred by a
    Gaussian of sigma (3.5, 1.4, 1.6) on 31x15x15, Poisson noise), sigma fitted from (2.5, 1.0,
    1.0) by Adam (lr 0.05, betas (0.9, 0.999)) through 10 unrolled iterations, 100 steps. Prints
    one JSON line with the trajectory."""
    import argparse
    import json

    from . import synth

    ap = argparse.ArgumentParser(description="fit a Gaussian PSF through the differentiable RL")
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--lr", type=float, default=0.05)
    ap.add_argument("--log-sigma", action="store_true")
    ap.add_argument("--objective", choices=["nll", "heldout", "known"], default="heldout")
    ap.add_argument("--num-iter", type=int, default=10)
    args = ap.parse_args(argv)
    vols, _, objs = synth.psffit3d(return_truth=True)
    vt = [torch.from_numpy(v).cuda() for v in vols]
    ot = [torch.from_numpy(o).cuda() for o in objs] if args.objective == "known" else None
    true_sigma = [3.5, 1.4, 1.6]
    res = fit_gaussian_sigma(vt, (31, 15, 15), (2.5, 1.0, 1.0), num_iter=args.num_iter, steps=args.steps,
                             lr=args.lr, log_sigma=args.log_sigma, objective=args.objective, objects=ot)
    err = [abs(a - b) / b for a, b in zip(res.sigma, true_sigma)]
    print(json.dumps({"volumes": [list(v.shape) for v in vols], "objective": args.objective,
                      "num_iter": args.num_iter, "steps": args.steps, "lr": args.lr,
                      "ms_per_step": round(res.ms_per_step, 3), "sigma_true": true_sigma,
                      "sigma_final": [round(s, 5) for s in res.sigma], "rel_error": [round(e, 4) for e in err],
                      "trajectory": [[h[0], h[1], [round(s, 5) for s in h[2]]] for h in res.history]}), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
