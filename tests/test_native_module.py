"""``lucem._native`` (lucem/_native.py): the reference's extension-module surface for the
Richardson-Lucy operator (bindings/python/lucem-py-core/src/ops.rs), as the reference's own
``lucem.ops`` consumes it: the catalog entry, the halo rule, the spec checks (CPU), and ``run_op``
on NumPy and device arrays against the oracle, plus the batched and differentiable forms (GPU)."""

from __future__ import annotations

import json

import numpy as np
import pytest


def _spec(**kw):
    from lucem import _native as M

    d = {"op": "richardson_lucy", **json.loads(M.op_catalog())[0]["params"], **kw}
    return json.dumps(d)


def test_catalog_entry():
    from lucem import _native as M

    cat = json.loads(M.op_catalog())
    assert [e["tag"] for e in cat] == ["richardson_lucy"]
    e = cat[0]
    assert e["params"] == {"num_iter": 30, "clip": False, "filter_epsilon": None, "mode": "reflect", "psf_size": 15}
    assert [p["type"] for p in e["inputs"]] == ["image", "image"] and e["output"] == "image"
    assert e["halo"] == 30 * 2 * 7
    assert M.__version__


def test_halo_and_spec_checks():
    from lucem import _native as M

    assert M.op_halo(_spec(num_iter=3, psf_size=31)) == 3 * 2 * 15
    assert M.op_halo(_spec(num_iter=2, psf_size=4)) == 2 * 2 * 2
    for bad, msg in [(_spec(num_iter=0), "num_iter"), (_spec(mode="edge"), "unknown variant"),
                     (_spec(filter_epsilon="x"), "filter_epsilon"), (_spec(psf_size=0), "psf_size"),
                     (json.dumps({"op": "richardson_lucy", "num_iter": 3}), "missing field"),
                     (_spec(sigma=2.0), "unknown field"), (json.dumps({"op": "gaussian_blur", "sigma": 2.0}),
                                                          "unknown variant"), ("{", "JSON")]:
        with pytest.raises(ValueError, match=msg):
            M.op_halo(bad)
    with pytest.raises(ValueError, match="chain"):
        M.run_chain([_spec()], np.zeros((4, 4), np.float32))


@pytest.mark.gpu
def test_run_op_against_the_oracle(gpu, oracle):
    import torch

    from lucem import _native as M

    rng = np.random.default_rng(12)
    y = rng.uniform(20, 200, (96, 80)).astype(np.float32)
    psf = rng.uniform(0.1, 1.0, (9, 7)).astype(np.float32)
    print("lucem._native.run_op: 96 x 80, psf 9 x 7, 12 iterations, mirror", flush=True)
    got = M.run_op(_spec(num_iter=12, mode="mirror"), [y, psf])
    assert isinstance(got, np.ndarray) and got.dtype == np.float32 and got.shape == y.shape
    st, want = oracle.richardson_lucy(y, psf, 12, clip=False, mode=3)
    assert st == 0 and oracle.rel_l2(got, want) < 1e-4
    dev = M.run_op(_spec(num_iter=12, mode="mirror"), [torch.from_numpy(y).cuda(), torch.from_numpy(psf).cuda()])
    assert dev.is_cuda and np.array_equal(dev.cpu().numpy(), got)
    vol = rng.uniform(20, 200, (12, 30, 26)).astype(np.float32)
    p3 = rng.uniform(0.1, 1.0, (3, 5, 5)).astype(np.float32)
    got3 = M.run_op(_spec(num_iter=5, clip=True), [vol, p3])
    st, want3 = oracle.richardson_lucy(vol, p3, 5, clip=True, nd=True)
    assert oracle.rel_l2(got3, want3) < 1e-4
    with pytest.raises(ValueError, match="psf_size"):
        M.run_op(_spec(psf_size=5), [y, psf])


@pytest.mark.gpu
def test_batched_and_differentiable_forms(gpu, oracle):
    import torch

    from lucem import _native as M

    rng = np.random.default_rng(13)
    tiles = [rng.uniform(20, 200, s).astype(np.float32) for s in ((40, 50), (33, 61))]
    psfs = [rng.uniform(0.1, 1.0, (5, 5)).astype(np.float32)]
    ys = [torch.from_numpy(t).cuda() for t in tiles]
    outs = [torch.empty_like(t) for t in ys]
    M.richardson_lucy_batched(ys, outs, [torch.from_numpy(psfs[0]).cuda()], [0, 0],
                              M.RichardsonLucy(num_iter=6, clip=False))
    torch.cuda.synchronize()
    for t, o in zip(tiles, outs):
        st, want = oracle.richardson_lucy(t, psfs[0], 6, clip=False)
        assert oracle.rel_l2(o.cpu().numpy(), want) < 1e-4
    x0 = torch.full_like(ys[0], 0.5, requires_grad=True)
    loss = M.rl_nll([x0], [ys[0]], torch.from_numpy(psfs[0]).cuda(), 4).sum()
    loss.backward()
    l64, g64, _ = oracle.rl_nll_grad64(tiles[0], np.full(tiles[0].shape, 0.5), psfs[0], 4)
    assert abs(float(loss) - l64) <= 1e-6 * abs(l64)
    assert oracle.rel_l2(x0.grad.cpu().numpy(), g64) < 1e-3
    assert issubclass(M.RichardsonLucyNLL, torch.autograd.Function)
