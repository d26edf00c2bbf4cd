"""``lucem`` as the MI355X build provides it: the ``lucem._native`` module for the Richardson-Lucy
path (see :mod:`lucem._native`). The reference's Python package (bindings/python/python/lucem)
builds its ``lucem.ops`` functions from ``_native.op_catalog()`` and runs them through
``_native.run_op``; with this module in its place, ``lucem.ops.richardson_lucy`` runs on the HIP
library."""

from lucem._native import __version__  # noqa: F401
