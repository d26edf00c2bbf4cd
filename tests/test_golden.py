"""The committed golden fixtures (tests/golden/, written by make_golden.py from the FP64 oracle).

CPU: the oracle still reproduces every fixture (FP64 to 1e-12, FP32 to 1e-6 relative L2), so a
later change to the oracle cannot silently move the target. GPU: the HIP entry points reproduce
the FP64 outputs within the parity bar of 1e-4 relative L2 (the FP32 oracle's own distance to
FP64 is printed beside it), the differentiable form reproduces the FP64 loss and gradients
within 1e-3, and the correlation its FP64 fixtures within 1e-6 (FP32) and 1e-12 (FP64).
"""

from __future__ import annotations

import json
from pathlib import Path

import numpy as np
import pytest

G = Path(__file__).resolve().parent / "golden"
META = json.loads((G / "cases.json").read_text())
RL = np.load(G / "rl_cases.npz")
GR = np.load(G / "grad_case.npz")


def _rel(a, b):
    dt = np.complex128 if np.iscomplexobj(a) or np.iscomplexobj(b) else np.float64
    a = np.asarray(a, dt)
    b = np.asarray(b, dt)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


@pytest.mark.parametrize("case", [m["case"] for m in META["rl"]])
def test_oracle_reproduces_rl_fixture(oracle, case):
    m = META["rl"][case]
    y, p = RL[f"c{case}_image"], RL[f"c{case}_psf"]
    nd = len(m["shape"]) == 3
    kw = dict(clip=m["clip"], filter_epsilon=m["filter_epsilon"], mode=m["mode"], nd=nd)
    st, o64 = oracle.richardson_lucy(y.astype(np.float64), p.astype(np.float64), m["num_iter"], **kw)
    assert st == 0 and _rel(o64, RL[f"c{case}_out64"]) < 1e-12
    st, o32 = oracle.richardson_lucy(y, p, m["num_iter"], **kw)
    assert st == 0 and _rel(o32, RL[f"c{case}_out32"]) < 1e-6
    assert _rel(RL[f"c{case}_out32"], RL[f"c{case}_out64"]) < 1e-5


@pytest.mark.parametrize("case", [m["case"] for m in META["grad"]])
def test_oracle_reproduces_grad_fixture(oracle, case):
    m = META["grad"][case]
    loss, gx, gp = oracle.rl_nll_grad64(GR[f"g{case}_y"], GR[f"g{case}_x0"], GR[f"g{case}_psf"], m["num_iter"])
    assert abs(loss - GR[f"g{case}_loss"][0]) <= 1e-12 * abs(loss)
    assert _rel(gx, GR[f"g{case}_gx0"]) < 1e-12
    assert _rel(gp, GR[f"g{case}_gpsf"]) < 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("case", [m["case"] for m in META["rl"]])
def test_hip_matches_rl_fixture(gpu, case):
    import torch

    import lucem_amd as LA

    m = META["rl"][case]
    y = torch.from_numpy(RL[f"c{case}_image"]).cuda()
    p = torch.from_numpy(RL[f"c{case}_psf"]).cuda()
    out = torch.empty_like(y)
    s = LA.RichardsonLucy(num_iter=m["num_iter"], clip=m["clip"], mode=m["mode"], filter_epsilon=m["filter_epsilon"])
    if len(m["shape"]) == 2:
        LA.richardson_lucy_into(y, p, out, s)
    else:
        LA.richardson_lucy_nd_into(y, p, out, s)
    got = out.cpu().numpy()
    e64 = _rel(got, RL[f"c{case}_out64"])
    e32 = _rel(got, RL[f"c{case}_out32"])
    print(f"case {case} {m['shape']} mode {m['mode']}: HIP vs FP64 {e64:.2e}, vs FP32 oracle {e32:.2e}, "
          f"FP32 oracle vs FP64 {_rel(RL[f'c{case}_out32'], RL[f'c{case}_out64']):.2e}")
    assert e64 < 1e-4 and e32 < 1e-4


@pytest.mark.gpu
@pytest.mark.parametrize("case", [m["case"] for m in META["grad"]])
def test_hip_matches_grad_fixture(gpu, case):
    import torch

    from lucem_amd.autograd import rl_nll

    m = META["grad"][case]
    x0 = torch.from_numpy(GR[f"g{case}_x0"]).cuda().requires_grad_(True)
    y = torch.from_numpy(GR[f"g{case}_y"]).cuda()
    p = torch.from_numpy(GR[f"g{case}_psf"]).cuda().requires_grad_(True)
    loss = rl_nll([x0], [y], p, m["num_iter"], checkpoint_every=2)
    loss.sum().backward()
    L = float(loss[0])
    assert abs(L - GR[f"g{case}_loss"][0]) <= 1e-6 * abs(L)
    assert _rel(x0.grad.cpu().numpy(), GR[f"g{case}_gx0"]) < 1e-3
    assert _rel(p.grad.cpu().numpy(), GR[f"g{case}_gpsf"]) < 1e-3


RS = np.load(G / "restore_cases.npz")


@pytest.mark.parametrize("case", [m["case"] for m in META["wiener"]])
def test_oracle_reproduces_wiener_fixture(oracle, case):
    m = META["wiener"][case]
    out = oracle.wiener64(RS[f"w{case}_image"], RS[f"w{case}_psf"], m["balance"], m["clip"])
    assert _rel(out, RS[f"w{case}_out64"]) < 1e-12


@pytest.mark.parametrize("case", [m["case"] for m in META["unsupervised"]])
def test_oracle_reproduces_unsupervised_fixture(oracle, case):
    m = META["unsupervised"][case]
    img, chain = oracle.unsupervised_wiener64(RS[f"u{case}_image"], RS[f"u{case}_psf"], m["num_iter"], m["burnin"],
                                              m["clip"], m["seed"])
    assert _rel(img, RS[f"u{case}_out64"]) < 1e-12
    np.testing.assert_allclose(chain, RS[f"u{case}_chain"], rtol=1e-12)


@pytest.mark.parametrize("case", [m["case"] for m in META["fft"]])
def test_oracle_reproduces_fft_fixture(oracle, case):
    m = META["fft"][case]
    fr, fi = oracle.fft_axis64(RS[f"f{case}_re"], RS[f"f{case}_im"], inverse=m["inverse"])
    assert _rel(fr + 1j * fi, RS[f"f{case}_out_re"] + 1j * RS[f"f{case}_out_im"]) < 1e-12


@pytest.mark.parametrize("case", [m["case"] for m in META["noise"]])
def test_oracle_reproduces_noise_fixture(oracle, case):
    m = META["noise"][case]
    out = oracle.add_noise32(RS[f"n{case}_image"], m["kind"], m["p0"], m["p1"], seed=m["seed"], clip=m["clip"])
    np.testing.assert_array_equal(out, RS[f"n{case}_out"])


@pytest.mark.gpu
@pytest.mark.parametrize("case", [m["case"] for m in META["wiener"]])
def test_hip_matches_wiener_fixture(gpu, case):
    import torch

    import lucem_amd as LA

    m = META["wiener"][case]
    out = LA.wiener_deconvolve(torch.from_numpy(RS[f"w{case}_image"]).cuda(),
                               torch.from_numpy(RS[f"w{case}_psf"]).cuda(),
                               LA.WienerDeconvolve(balance=m["balance"], clip=m["clip"]))
    e = _rel(out.cpu().numpy(), RS[f"w{case}_out64"])
    print(f"wiener case {case} {m['shape']}: HIP vs FP64 {e:.2e}")
    assert e < 1e-4


@pytest.mark.gpu
@pytest.mark.parametrize("case", [m["case"] for m in META["unsupervised"]])
def test_hip_matches_unsupervised_fixture(gpu, case):
    import torch

    import lucem_amd as LA

    m = META["unsupervised"][case]
    img, chain = LA.unsupervised_wiener(torch.from_numpy(RS[f"u{case}_image"]).cuda(),
                                        torch.from_numpy(RS[f"u{case}_psf"]).cuda(),
                                        LA.UnsupervisedWiener(num_iter=m["num_iter"], burnin=m["burnin"],
                                                              clip=m["clip"], seed=m["seed"]))
    assert _rel(img.cpu().numpy(), RS[f"u{case}_out64"]) < 1e-4
    np.testing.assert_allclose(chain, RS[f"u{case}_chain"], rtol=1e-4)


@pytest.mark.gpu
@pytest.mark.parametrize("case", [m["case"] for m in META["fft"]])
def test_hip_matches_fft_fixture(gpu, case):
    import torch

    from lucem_amd import fft

    m = META["fft"][case]
    fr, fi = fft.fft_nd(torch.from_numpy(RS[f"f{case}_re"]).cuda(), torch.from_numpy(RS[f"f{case}_im"]).cuda(),
                        inverse=m["inverse"])
    got = fr.cpu().numpy() + 1j * fi.cpu().numpy()
    assert _rel(got, RS[f"f{case}_out_re"] + 1j * RS[f"f{case}_out_im"]) < 2e-6


@pytest.mark.gpu
@pytest.mark.parametrize("case", [m["case"] for m in META["noise"]])
def test_hip_matches_noise_fixture(gpu, case):
    import torch

    import lucem_amd as LA

    m = META["noise"][case]
    out = LA.add_noise(torch.from_numpy(RS[f"n{case}_image"]).cuda(), LA.NoiseModel(m["kind"], m["p0"], m["p1"]),
                       m["seed"], m["clip"])
    np.testing.assert_array_equal(out.cpu().numpy(), RS[f"n{case}_out"])


FC = np.load(G / "filter_cases.npz")


@pytest.mark.parametrize("case", [m["case"] for m in META["correlate"]])
def test_oracle_reproduces_correlate_fixture(oracle, case):
    m = META["correlate"][case]
    out = oracle.correlate(FC[f"c{case}_input"].astype(np.float64), FC[f"c{case}_weights"].astype(np.float64),
                           m["mode"], m["cval"], origin=m["origin"])
    assert _rel(out, FC[f"c{case}_out64"]) < 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("case", [m["case"] for m in META["correlate"]])
def test_hip_matches_correlate_fixture(gpu, case):
    """FP32 FFT form within 1e-6 relative L2; the FP64 direct form (2-D cases) within 1e-12."""
    import torch

    import lucem_amd as LA

    m = META["correlate"][case]
    s = LA.Correlate2d(mode=m["mode"], cval=m["cval"], origin=tuple(m["origin"]))
    x, w = FC[f"c{case}_input"], FC[f"c{case}_weights"]
    want = FC[f"c{case}_out64"]
    got = LA.correlate_nd(torch.from_numpy(x).cuda(), torch.from_numpy(w).cuda(), s).cpu().numpy()
    e = _rel(got, want)
    print(f"correlate case {case} {m['shape']}: HIP f32 vs FP64 {e:.2e}")
    assert e < 1e-6
    if x.ndim == 2:
        got64 = LA.correlate_2d(torch.from_numpy(x.astype(np.float64)).cuda(),
                                torch.from_numpy(w.astype(np.float64)).cuda(), s).cpu().numpy()
        assert _rel(got64, want) < 1e-12
