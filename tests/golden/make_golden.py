"""Writes the golden fixtures of this directory from the FP64 oracle (run from the repo root:
``python tests/golden/make_golden.py``).

``rl_cases.npz``: small 2-D and 3-D Richardson-Lucy cases in every PadMode, with clip on and
off and a custom filter epsilon; per case the float32 input and PSF, the FP64 direct-oracle
output (the reference algorithm, crates/lucem-ops/src/restore/deconvolve.rs:45-149, run in
double precision) and the FP32 direct-oracle output. ``grad_case.npz``: one 2-D and one 3-D
case of the unrolled-RL Poisson NLL with its FP64 loss and gradients. ``restore_cases.npz``: the
Wiener filter (restore/deconvolve.rs:213-315) and the unsupervised Wiener loop (:357-509) in FP64,
the complex FFT over all axes (core/fft.rs:53-100) in FP64, and add_noise
(intensity/noise.rs:35-143) impulse and Poisson draws in float32. ``filter_cases.npz``: the
correlation (filter/convolve.rs:126-166) in FP64, 2-D and 3-D, every PadMode, window origins and
Constant cval. The metadata of each case is in ``cases.json``. Arrays only (no pickled objects): load with ``numpy.load`` defaults.
"""

from __future__ import annotations

import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parents[1]))

from oracle import oracle as O  # noqa: E402

CASES = [
    # (shape, psf shape, mode, num_iter, clip, filter_epsilon)
    ((24, 31), (5, 4), 0, 6, False, None),
    ((24, 31), (5, 4), 1, 6, False, None),
    ((24, 31), (5, 4), 2, 6, False, None),
    ((24, 31), (5, 4), 3, 6, False, None),
    ((24, 31), (5, 4), 4, 6, False, None),
    ((17, 13), (7, 7), 0, 5, True, None),
    ((40, 9), (3, 9), 3, 4, False, 1e-3),
    ((1, 12), (1, 5), 0, 3, False, None),
    ((10, 12, 9), (3, 5, 3), 0, 4, False, None),
    ((7, 11, 13), (5, 3, 4), 4, 3, False, None),
]


def main() -> None:
    rng = np.random.default_rng(424242)
    arrays, meta = {}, []
    for i, (shape, kshape, mode, it, clip, eps) in enumerate(CASES):
        y = rng.uniform(10.0, 200.0, shape).astype(np.float32)
        if i == 5:
            y /= 400.0  # values below 1 so that the final clip to [-1, 1] is exercised both ways
            y[::3] *= 8.0
        p = rng.uniform(0.05, 1.0, kshape).astype(np.float32)
        nd = len(shape) == 3
        st64, o64 = O.richardson_lucy(y.astype(np.float64), p.astype(np.float64), it, clip=clip, filter_epsilon=eps,
                                      mode=mode, nd=nd)
        st32, o32 = O.richardson_lucy(y, p, it, clip=clip, filter_epsilon=eps, mode=mode, nd=nd)
        assert st64 == 0 and st32 == 0
        arrays[f"c{i}_image"] = y
        arrays[f"c{i}_psf"] = p
        arrays[f"c{i}_out64"] = o64
        arrays[f"c{i}_out32"] = o32
        meta.append({"case": i, "shape": list(shape), "psf": list(kshape), "mode": mode, "num_iter": it, "clip": clip,
                     "filter_epsilon": eps})
    np.savez_compressed(HERE / "rl_cases.npz", **arrays)
    grads, gmeta = {}, []
    for j, (shape, kshape, it) in enumerate([((12, 13), (3, 4), 3), ((5, 8, 7), (3, 3, 3), 2)]):
        y = rng.uniform(50.0, 150.0, shape).astype(np.float32)
        x0 = rng.uniform(0.5, 2.0, shape).astype(np.float32)
        p = rng.uniform(0.1, 1.0, kshape).astype(np.float32)
        loss, gx, gp = O.rl_nll_grad64(y, x0, p, it)
        grads.update({f"g{j}_y": y, f"g{j}_x0": x0, f"g{j}_psf": p, f"g{j}_loss": np.array([loss]),
                      f"g{j}_gx0": gx, f"g{j}_gpsf": gp})
        gmeta.append({"case": j, "shape": list(shape), "psf": list(kshape), "num_iter": it, "mode": 0})
    np.savez_compressed(HERE / "grad_case.npz", **grads)
    rest, rmeta = {}, {"wiener": [], "unsupervised": [], "fft": [], "noise": []}
    for j, (shape, kshape, balance, clip) in enumerate([((24, 30), (5, 5), 0.1, False), ((23, 26), (3, 4), 0.01, True),
                                                         ((2, 9), (1, 3), 0.5, False)]):
        y = rng.uniform(0.0, 1.0, shape).astype(np.float32)
        p = rng.uniform(0.1, 1.0, kshape).astype(np.float32)
        rest.update({f"w{j}_image": y, f"w{j}_psf": p, f"w{j}_out64": O.wiener64(y, p, balance, clip)})
        rmeta["wiener"].append({"case": j, "shape": list(shape), "psf": list(kshape), "balance": balance, "clip": clip})
    for j, (shape, kshape, it, burn, seed) in enumerate([((16, 20), (3, 3), 8, 3, 5), ((13, 16), (5, 3), 6, 2, 11)]):
        y = rng.uniform(0.0, 1.0, shape).astype(np.float32)
        p = rng.uniform(0.1, 1.0, kshape).astype(np.float32)
        img, chain = O.unsupervised_wiener64(y, p, it, burn, False, seed)
        rest.update({f"u{j}_image": y, f"u{j}_psf": p, f"u{j}_out64": img, f"u{j}_chain": np.array(chain)})
        rmeta["unsupervised"].append({"case": j, "shape": list(shape), "psf": list(kshape), "num_iter": it,
                                      "burnin": burn, "clip": False, "seed": seed})
    for j, (shape, inverse) in enumerate([((6, 10, 7), False), ((11, 13), True)]):
        re = rng.normal(size=shape).astype(np.float32)
        im = rng.normal(size=shape).astype(np.float32)
        fr, fi = O.fft_axis64(re, im, inverse=inverse)
        rest.update({f"f{j}_re": re, f"f{j}_im": im, f"f{j}_out_re": fr, f"f{j}_out_im": fi})
        rmeta["fft"].append({"case": j, "shape": list(shape), "inverse": inverse})
    for j, (kind, a, b, seed, clip) in enumerate([(4, 0.3, 0.4, (1 << 40) + 3, False), (5, 20.0, 0.0, 9, False)]):
        x = rng.uniform(0.0, 1.0, (32, 32)).astype(np.float32)
        rest.update({f"n{j}_image": x, f"n{j}_out": O.add_noise32(x, kind, a, b, seed=seed, clip=clip)})
        rmeta["noise"].append({"case": j, "kind": kind, "p0": a, "p1": b, "seed": seed, "clip": clip})
    np.savez_compressed(HERE / "restore_cases.npz", **rest)
    frng = np.random.default_rng(535353)
    filt, fmeta = {}, []
    for j, (shape, kshape, mode, cval, origin) in enumerate([
            ((20, 27), (5, 4), 0, 0.0, (0, 0)), ((19, 23), (4, 6), 1, 1.5, (-2, 2)), ((16, 9), (7, 3), 2, 0.0, (3, -1)),
            ((9, 11), (15, 21), 3, 0.0, (0, 0)), ((25, 17), (3, 3), 4, 0.0, (1, -1)),
            ((6, 9, 8), (3, 4, 5), 0, 0.0, (1, -2, 2)), ((5, 7, 6), (4, 3, 3), 1, -0.75, (-2, 1, 0))]):
        x = frng.uniform(-1.0, 1.0, shape).astype(np.float32)
        w = frng.standard_normal(kshape).astype(np.float32)
        out = O.correlate(x.astype(np.float64), w.astype(np.float64), mode, cval, origin=origin)
        filt.update({f"c{j}_input": x, f"c{j}_weights": w, f"c{j}_out64": out})
        fmeta.append({"case": j, "shape": list(shape), "weights": list(kshape), "mode": mode, "cval": cval,
                      "origin": list(origin)})
    np.savez_compressed(HERE / "filter_cases.npz", **filt)
    (HERE / "cases.json").write_text(json.dumps({"rl": meta, "grad": gmeta, **rmeta, "correlate": fmeta}, indent=1)
                                     + "\n")


if __name__ == "__main__":
    main()
