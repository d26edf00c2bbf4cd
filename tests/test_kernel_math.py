"""Host checks of integer arithmetic the HIP kernels rely on (no device needed).

- ``Divisor`` (lucem_amd/csrc/fft_lds.hpp): q = umulhi(u, m) with m = floor((2^32 - 1) / d) + 1
  must equal u // d for every index u < 2^17 and divisor d <= 16384 (line counts, half-spectrum
  widths up to 8193); checked exhaustively against integer division.
- The halo walk of the adjoint folds (fold_rows / fold_lines): positions [0, left) then
  [left + n, end), in increasing order, exactly the padded positions outside [left, left + n).
- ``pad_index`` (fft_lds.hpp), whose positions within one fold take closed forms without a
  division: equal to the oracle's restatement of core/kernels/pad.rs:11-34 at every position
  from five array lengths before the array to six after, lengths 1..70, all five modes.
"""

from __future__ import annotations

import numpy as np


def test_multiply_high_division_is_exact():
    u = np.arange(0, 1 << 17, dtype=np.uint64)
    for d in range(2, 16385):
        m = np.uint64(0xFFFFFFFF // d + 1)
        q = (u * m) >> np.uint64(32)
        assert np.array_equal(q, u // np.uint64(d)), d


def _halo_walk(left, n, end):
    out, right = [], left + n
    h = 0 if left > 0 else right
    while h < end:
        out.append(h)
        h = right if h + 1 == left else h + 1
    return out


def test_halo_walk_visits_exactly_the_halo_in_order():
    for left in range(0, 24):
        for n in range(1, 48):
            for end in range(0, 96):
                want = [h for h in range(end) if not (left <= h < left + n)]
                assert _halo_walk(left, n, end) == want


def test_kernel_pad_index_matches_the_oracle(oracle):
    import ctypes
    import subprocess
    from pathlib import Path

    root = Path(__file__).resolve().parent
    so = root / "cpp" / "libpad_probe.so"
    if not so.exists():
        from lucem_amd import build

        subprocess.run([build.hipcc(), "-O2", "-std=c++17", "-fPIC", "-shared", "-o", str(so),
                        str(root / "cpp" / "pad_probe.cpp")], check=True, capture_output=True)
    lib = ctypes.CDLL(str(so))
    ip = ctypes.POINTER(ctypes.c_int)
    lib.lucem_probe_pad_index.argtypes = [ip, ctypes.c_int, ctypes.c_int, ctypes.c_int, ip]
    for length in range(1, 71):
        p = np.arange(-5 * length - 40, 6 * length + 40, dtype=np.int32)
        for mode in range(5):
            got = np.empty_like(p)
            lib.lucem_probe_pad_index(p.ctypes.data_as(ip), len(p), length, mode, got.ctypes.data_as(ip))
            want = np.array([oracle.pad_index(int(q), length, mode) for q in p], dtype=np.int32)
            assert np.array_equal(got, want), (length, mode)
