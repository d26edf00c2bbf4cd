"""Argument checks of the PSF fit (lucem_amd/psf_fit.py) that need no device."""

from __future__ import annotations

import pytest


def test_psf_fit_objective_checks():
    """The PSF fit names an unknown objective and asks for the objects of the known-object one
    before it touches a device."""
    import torch

    from lucem_amd.psf_fit import fit_gaussian_sigma

    v = [torch.zeros(4, 4)]
    with pytest.raises(ValueError, match="objective"):
        fit_gaussian_sigma(v, (3, 3), (1.0, 1.0), objective="other")
    with pytest.raises(ValueError, match="one object per volume"):
        fit_gaussian_sigma(v, (3, 3), (1.0, 1.0), objective="known")
