"""Writes the golden fixtures of this directory from the FP64 oracle (run from the repo root:
``python tests/golden/make_golden.py``).

``rl_cases.npz``: small 2-D and 3-D Richardson-Lucy cases in every PadMode, with clip on and
off and a custom filter epsilon; per case the float32 input and PSF, the FP64 direct-oracle
output (the reference algorithm, crates/lucem-ops/src/restore/deconvolve.rs:45-149, run in
double precision) and the FP32 direct-oracle output. ``grad_case.npz``: one 2-D and one 3-D
case of the unrolled-RL Poisson NLL with its FP64 loss and gradients. The metadata of each case
is in ``cases.json``. Arrays only (no pickled objects): load with ``numpy.load`` defaults.
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
    (HERE / "cases.json").write_text(json.dumps({"rl": meta, "grad": gmeta}, indent=1) + "\n")


if __name__ == "__main__":
    main()
