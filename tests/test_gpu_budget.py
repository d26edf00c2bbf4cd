"""The device memory budget on the device: a workspace slot or a batched plan of 64 MiB or more
that exceeds the limit fails with OUT_OF_DEVICE_MEMORY before anything is allocated
(crates/lucem-core/src/memory/budget.rs:232-244, workspace.rs:32-35); without the limit the same
calls run, and the bytes the library holds rise and fall with its workspaces and plans."""

from __future__ import annotations

import pytest

pytestmark = pytest.mark.gpu

MIB = 1 << 20


@pytest.fixture
def limit_guard(gpu):
    from lucem_amd import _native as N

    before = N.device_memory_limit()
    yield N
    N.set_device_memory_limit(before)


def test_workspace_growth_is_checked(gpu, limit_guard):
    import torch

    import lucem_amd as LA

    N = limit_guard
    y = torch.rand((4096, 4096), device="cuda") + 0.5
    psf = torch.ones((3, 3), device="cuda")
    out = torch.empty_like(y)
    s = LA.RichardsonLucy(num_iter=2, clip=False)
    ws = N.Workspace(gpu)
    base = N.device_bytes_in_use(0)
    N.set_device_memory_limit(base + MIB)
    with pytest.raises(N.LucemError) as e:
        LA.richardson_lucy_into(y, psf, out, s, ws, gpu)
    assert e.value.status == 7
    msg = str(e.value)
    assert "bytes of device memory; the budget leaves" in msg and f"(limit {base + MIB})" in msg, msg
    assert N.device_bytes_in_use(0) <= base + MIB

    N.set_device_memory_limit(None)
    LA.richardson_lucy_into(y, psf, out, s, ws, gpu)
    torch.cuda.synchronize()
    assert bool(torch.isfinite(out).all())
    held = N.device_bytes_in_use(0)
    assert held >= base + 2 * 4096 * 4096 * 4  # at least the estimate and a spectrum slot
    ws.close()
    assert N.device_bytes_in_use(0) <= base + MIB


def test_batched_plan_is_checked(gpu, limit_guard):
    import torch

    from lucem_amd.batched import RLBatch
    from lucem_amd.restore import RichardsonLucy

    N = limit_guard
    shapes = [(512, 512)] * 40
    s = RichardsonLucy(num_iter=2, clip=False)
    b = RLBatch(shapes, [(9, 9)], [0] * len(shapes), s, client=gpu)
    need = b.device_bytes
    b.close()
    assert need >= 64 * MIB
    base = N.device_bytes_in_use(0)
    N.set_device_memory_limit(base + need - 1)
    with pytest.raises(N.LucemError) as e:
        RLBatch(shapes, [(9, 9)], [0] * len(shapes), s, client=gpu)
    assert e.value.status == 7 and f"batch plan needs {need} bytes" in str(e.value)
    N.set_device_memory_limit(base + need)
    b = RLBatch(shapes, [(9, 9)], [0] * len(shapes), s, client=gpu)
    assert N.device_bytes_in_use(0) == base + need
    ys = [torch.rand(sh, device="cuda") + 0.5 for sh in shapes]
    outs = [torch.empty_like(t) for t in ys]
    b.execute(ys, outs, [torch.ones((9, 9), device="cuda")])
    torch.cuda.synchronize()
    b.close()
    assert N.device_bytes_in_use(0) == base
