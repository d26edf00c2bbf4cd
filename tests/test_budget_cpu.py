"""The device memory budget without a device (crates/lucem-core/src/memory/budget.rs): size
parsing (its own test, budget.rs:285-293), the limit from the environment and from the setter, and
check_allocation's threshold, arithmetic and message (budget.rs:232-244, lucem-core error.rs:27-29)."""

from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
MIB = 1 << 20


@pytest.fixture(scope="module")
def N():
    from lucem_amd import _native as N

    if not N.LIB_PATH.exists():
        from lucem_amd import build

        build.build()
    N.lib()
    return N


@pytest.fixture
def no_limit(N):
    before = N.device_memory_limit()
    N.set_device_memory_limit(None)
    yield N
    N.set_device_memory_limit(before)


# budget.rs:285-293, plus the unit table of :268-277 and the forms Rust's f64 parse accepts
@pytest.mark.parametrize("text,want", [
    ("1024", 1024), ("7GiB", 7 << 30), (" 6.5 GB ", 6_500_000_000), ("512mib", 512 << 20), ("lots", None),
    ("5 TB", None), ("10b", 10), ("3kb", 3000), ("2KiB", 2048), ("1.5MB", 1_500_000), ("1.", 1), (".5kib", 512),
    ("", None), (".", None), ("1.2.3", None), ("-5", None), ("7 GiB", 7 << 30), ("0", 0),
])
def test_parse_bytes(N, text, want):
    assert N.parse_bytes(text) == want


def test_no_limit_allows_anything(no_limit):
    N = no_limit
    assert N.device_memory_limit() is None
    N.check_allocation("huge", 1 << 50)


def test_small_allocations_are_not_checked(no_limit):
    N = no_limit
    N.set_device_memory_limit(1)
    N.check_allocation("small", 64 * MIB - 1)  # below CHECKED_ALLOCATION_BYTES (budget.rs:31)
    with pytest.raises(N.LucemError) as e:
        N.check_allocation("at threshold", 64 * MIB)
    assert e.value.status == 7


def test_out_of_device_memory_message(no_limit):
    N = no_limit
    limit = 100 * MIB
    N.set_device_memory_limit(limit)
    assert N.device_memory_limit() == limit
    used = N.device_bytes_in_use(0)
    N.check_allocation("fits", limit - used)
    with pytest.raises(N.LucemError) as e:
        N.check_allocation("rl_spec_a", limit - used + 1)
    assert e.value.status == 7  # LUCEM_OUT_OF_DEVICE_MEMORY
    assert str(e.value).endswith(
        f"rl_spec_a needs {limit - used + 1} bytes of device memory; the budget leaves {limit - used} (limit {limit})")
    N.set_device_memory_limit(None)
    N.check_allocation("rl_spec_a", limit + 1)


def _limit_in_fresh_process(env_extra: dict) -> str:
    env = {k: v for k, v in os.environ.items() if not k.startswith("LUCEM_DEVICE_MEMORY")}
    env.update(env_extra)
    code = "from lucem_amd import _native as N; print(N.device_memory_limit())"
    r = subprocess.run([sys.executable, "-c", code], cwd=str(ROOT), env=env, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return r.stdout.strip().splitlines()[-1]


def test_limit_from_environment(N):
    assert _limit_in_fresh_process({}) == "None"
    assert _limit_in_fresh_process({"LUCEM_DEVICE_MEMORY": "100MiB"}) == str(100 * MIB)
    assert _limit_in_fresh_process({"LUCEM_DEVICE_MEMORY_LIMIT": "7GiB"}) == str(7 << 30)
    # the short name wins (budget.rs:250-253); an unparsable value sets nothing
    assert _limit_in_fresh_process({"LUCEM_DEVICE_MEMORY": "2GB", "LUCEM_DEVICE_MEMORY_LIMIT": "7GiB"}) == "2000000000"
    assert _limit_in_fresh_process({"LUCEM_DEVICE_MEMORY": "lots"}) == "None"
