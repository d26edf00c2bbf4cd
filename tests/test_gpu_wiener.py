"""GPU tests of the Wiener forms (restore/deconvolve.rs:184-509) against the FP64 oracle: the
regularised filter on several shapes, PSF sizes and balances, the reference's identity KAT,
exact inversion of a circular blur, the unsupervised loop's image and balance chain, and the
reference's checks."""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RTOL_L2 = 1e-4


def _dev(a):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


def _wiener(y, psf, balance, clip):
    import lucem_amd as LA

    out = LA.wiener_deconvolve(_dev(y), _dev(psf), LA.WienerDeconvolve(balance=balance, clip=clip))
    return out.cpu().numpy()


WIENER_CASES = [
    ((64, 64), (5, 5), 0.1, False),
    ((48, 80), (7, 4), 0.01, True),
    ((30, 42), (30, 42), 0.1, False),   # the PSF as large as the image
    ((2, 16), (1, 3), 0.1, False),      # an axis shorter than 3: the regulariser is 1
    ((1, 64), (1, 5), 0.5, False),
    ((64, 1), (3, 1), 0.5, False),
    ((3, 3), (3, 3), 0.2, False),
    ((125, 98), (9, 6), 0.05, False),   # odd and 7-smooth sizes
    ((1024, 1024), (25, 25), 0.1, False),
    # sizes that are not 7-smooth: the general complex FFT with Bluestein lines
    ((37, 53), (5, 5), 0.1, False),
    ((100, 97), (9, 7), 0.02, True),
    ((11, 1), (3, 1), 0.5, False),
    ((509, 768), (15, 15), 0.1, False),
    ((4200, 16), (7, 3), 0.1, False),   # a y line of 4200 (7-smooth) takes the 512-thread pass
    ((16, 4200), (3, 7), 0.1, False),   # an x line of 4200 takes the 512-thread row pass
    ((8209, 16), (5, 3), 0.1, False),   # a prime above 8192: Bluestein lines in passes
    ((8, 20000), (3, 9), 0.05, False),  # 7-smooth above 16384: four-step lines
]


@pytest.mark.parametrize("shape,kshape,balance,clip", WIENER_CASES)
def test_wiener_matches_fp64_oracle(gpu, oracle, shape, kshape, balance, clip):
    rng = np.random.default_rng(shape[0] * 7 + shape[1] + kshape[0])
    y = rng.uniform(0.0, 1.0, shape)
    psf = rng.uniform(0.1, 1.0, kshape)
    got = _wiener(y, psf, balance, clip)
    want = oracle.wiener64(y.astype(np.float32), psf.astype(np.float32), balance, clip)
    e = oracle.rel_l2(got, want)
    print(f"wiener {shape} psf {kshape} balance {balance}: rel-L2 vs FP64 {e:.2e}")
    assert e < RTOL_L2


def test_wiener_identity_kat(gpu):
    """deconvolve.rs:542-550: a delta PSF at balance 1e-7 returns the image within 1e-3."""
    img = (np.arange(64, dtype=np.float32) / 64.0 * 0.8 + 0.1).reshape(8, 8)
    delta = np.zeros((3, 3), np.float32)
    delta[1, 1] = 1.0
    out = _wiener(img, delta, 1e-7, False)
    assert np.max(np.abs(out - img)) < 1e-3


def test_wiener_inverts_a_circular_blur(gpu):
    rng = np.random.default_rng(12)
    x = rng.uniform(0.0, 1.0, (96, 120))
    psf = rng.uniform(0.0, 0.01, (7, 5))
    psf[3, 2] += 1.0
    y = np.zeros_like(x)
    for ky in range(7):
        for kx in range(5):
            y += psf[ky, kx] * np.roll(x, (ky - 3, kx - 2), axis=(0, 1))
    out = _wiener(y, psf, 0.0, False)
    assert np.max(np.abs(out - x)) < 1e-4


def test_wiener_errors(gpu):
    import torch

    import lucem_amd as LA

    y = torch.ones((16, 16), device="cuda")
    with pytest.raises(LA.LucemError) as e:
        LA.wiener_deconvolve(y, torch.ones((3, 3), device="cuda"), LA.WienerDeconvolve(balance=-1.0))
    assert e.value.kind == "INVALID_PARAMETER"
    with pytest.raises(LA.LucemError) as e:
        LA.wiener_deconvolve(y, torch.ones((17, 3), device="cuda"))
    assert e.value.kind == "INVALID_WEIGHTS"
    with pytest.raises(LA.LucemError) as e:
        LA.wiener_deconvolve_into(y, torch.ones((3, 3), device="cuda"), y)
    assert e.value.kind == "ALIASED_BUFFERS"
    empty = torch.ones((0, 16), device="cuda")
    assert LA.wiener_deconvolve(empty, torch.ones((3, 3), device="cuda")).shape == (0, 16)


def _blurred_spot():
    """The reference test's input (deconvolve.rs:552-557): a 0.9 spot on 0.1, box-blurred."""
    spot = np.full((8, 8), 0.1)
    spot[4, 4] = 0.9
    box = np.full((3, 3), 1.0 / 9.0)
    pad = np.pad(spot, 1, mode="symmetric")
    blurred = sum(box[ky, kx] * pad[ky:ky + 8, kx:kx + 8] for ky in range(3) for kx in range(3))
    return blurred.astype(np.float32), box.astype(np.float32)


def test_unsupervised_reference_case(gpu, oracle):
    """deconvolve.rs:569-578: 12 iterations, burn-in 4, clip, seed 7 on the blurred spot."""
    import lucem_amd as LA

    blurred, box = _blurred_spot()
    img, chain = LA.unsupervised_wiener(_dev(blurred), _dev(box), LA.UnsupervisedWiener(num_iter=12, burnin=4,
                                                                                      clip=True, seed=7))
    assert len(chain) == 12
    assert tuple(img.shape) == (8, 8)
    want_img, want_chain = oracle.unsupervised_wiener64(blurred, box, 12, 4, True, 7)
    c = np.array(chain)
    w = np.array(want_chain)
    print("chain rel err", np.max(np.abs(c - w) / w), "image rel-L2", oracle.rel_l2(img.cpu().numpy(), want_img))
    assert np.max(np.abs(c - w) / w) < 1e-4
    assert oracle.rel_l2(img.cpu().numpy(), want_img) < RTOL_L2


@pytest.mark.parametrize("shape,kshape,seed", [((64, 48), (5, 5), 3), ((100, 128), (9, 7), 11), ((2, 30), (1, 3), 0),
                                                ((37, 41), (5, 3), 5), ((61, 64), (7, 7), 9)])
def test_unsupervised_matches_fp64_oracle(gpu, oracle, shape, kshape, seed):
    import lucem_amd as LA

    rng = np.random.default_rng(seed + 40)
    x = rng.uniform(0.0, 1.0, shape)
    psf = rng.uniform(0.1, 1.0, kshape)
    y = (x + rng.normal(0.0, 0.01, shape)).astype(np.float32)
    s = LA.UnsupervisedWiener(num_iter=20, burnin=5, clip=False, seed=seed)
    img, chain = LA.unsupervised_wiener(_dev(y), _dev(psf), s)
    want_img, want_chain = oracle.unsupervised_wiener64(y, psf.astype(np.float32), 20, 5, False, seed)
    c, w = np.array(chain), np.array(want_chain)
    e_chain = float(np.max(np.abs(c - w) / w))
    e_img = oracle.rel_l2(img.cpu().numpy(), want_img)
    print(f"unsupervised {shape}: chain rel err {e_chain:.2e}, image rel-L2 {e_img:.2e}")
    assert e_chain < 1e-4
    assert e_img < RTOL_L2


def test_unsupervised_errors_and_empty(gpu):
    import torch

    import lucem_amd as LA

    y = torch.ones((16, 16), device="cuda")
    with pytest.raises(LA.LucemError) as e:
        LA.unsupervised_wiener(y, torch.ones((3, 3), device="cuda"), LA.UnsupervisedWiener(num_iter=5, burnin=5))
    assert e.value.kind == "INVALID_PARAMETER"
    with pytest.raises(LA.LucemError) as e:
        LA.unsupervised_wiener(y, torch.ones((0, 3), device="cuda"))
    assert e.value.kind == "INVALID_WEIGHTS"
    img, chain = LA.unsupervised_wiener(torch.ones((0, 8), device="cuda"), torch.ones((3, 3), device="cuda"))
    assert chain == [] and tuple(img.shape) == (0, 8)
