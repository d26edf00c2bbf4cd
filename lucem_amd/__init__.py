"""MI355X-native (gfx950) Richardson-Lucy deconvolution: a drop-in for lucem-ops'
``richardson_lucy_into`` behind a C ABI (``include/lucem_hip.h``), with HIP FFT kernels.

Public surface (mirrors the reference's names):

- :class:`PadMode`, :class:`RichardsonLucy`        crates/lucem-ops/src/core/pad.rs:13-57,
                                                   restore/deconvolve.rs:19-41
- :func:`richardson_lucy_into`, :func:`richardson_lucy`
                                                   restore/deconvolve.rs:45-162
- :class:`WienerDeconvolve`, :func:`wiener_deconvolve_into`, :func:`wiener_deconvolve`,
  :class:`UnsupervisedWiener`, :func:`unsupervised_wiener`
                                                   restore/deconvolve.rs:164-509
- :mod:`lucem_amd.fft`: ``fft_1d_into``, ``fft_nd_into``, ``complex_mul_into``
                                                   core/fft.rs:13-135
- :class:`NoiseModel`, :func:`add_noise`          intensity/noise.rs:17-143
- :class:`Correlate2d`, :func:`correlate_2d_into`, :func:`correlate_2d`, :class:`Correlate1d`,
  :func:`correlate_1d_into`, :func:`correlate_1d`, :func:`convolve_1d`
                                                   filter/convolve.rs:15-177
- :func:`fourier_gaussian_into`, :func:`fourier_box_into`, :func:`fourier_shift_into`,
  :func:`butterworth`, :func:`gabor_2d`            filter/spectral.rs:20-268
- :class:`GaussianBlur`, :func:`gaussian_blur`, :func:`gaussian_1d`
                                                   filter/gaussian.rs:18-252
- :class:`PhaseCorrelation`, :func:`phase_correlation`
                                                   register/phase.rs:18-250
- :class:`Client`, :class:`Workspace`              lucem-core Client / Workspace
"""

from .restore import (  # noqa: F401
    PadMode,
    RichardsonLucy,
    richardson_lucy,
    richardson_lucy_into,
    richardson_lucy_nd,
    richardson_lucy_nd_into,
)
from ._native import Client, LucemError, Workspace, version  # noqa: F401
from ._native import device_memory_limit, parse_bytes, set_device_memory_limit  # noqa: F401
from .wiener import (  # noqa: F401
    UnsupervisedWiener,
    WienerDeconvolve,
    unsupervised_wiener,
    wiener_deconvolve,
    wiener_deconvolve_into,
)
from . import fft  # noqa: F401,E402
from .noise import NoiseModel, add_noise, add_noise_into  # noqa: F401,E402
from . import tiled  # noqa: F401,E402
from .filter import (  # noqa: F401,E402
    Correlate1d,
    Correlate2d,
    convolve_1d,
    correlate_1d,
    correlate_1d_into,
    butterworth,
    correlate_2d,
    correlate_2d_into,
    correlate_nd,
    correlate_nd_into,
    fourier_box_into,
    fourier_gaussian_into,
    fourier_shift_into,
    gabor_2d,
    Gaussian1d,
    GaussianBlur,
    gaussian_1d,
    gaussian_1d_into,
    gaussian_blur,
    gaussian_blur_into,
    gaussian_kernel_1d,
)
from .register import PhaseCorrelation, PhaseCorrelationResult, phase_correlation  # noqa: F401,E402
