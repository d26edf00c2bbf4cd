"""Release order of clients, workspaces and plans (C ABI lifetime rules).

A workspace or plan keeps the client it was made on alive (the client's handle and each dependent
hold one reference), so they may be destroyed in any order: a plan or workspace destroyed after
its client no longer reads a freed client to find its stream. Each case runs in a fresh
interpreter, whose exit status must be 0:

- the client destroyed first, a plan still executed afterwards, then the plan, workspace and
  gradient plan destroyed;
- nothing destroyed explicitly and the package's exit hook removed, so interpreter shutdown
  finalises the objects in whatever order it takes.
"""

from __future__ import annotations

import subprocess
import sys
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]

_PRELUDE = f"""
import sys
sys.path.insert(0, {str(ROOT)!r})
import ctypes
import numpy as np
import torch
import lucem_amd as LA
from lucem_amd import _native as N
from lucem_amd.batched import RLBatch
from lucem_amd.autograd import RLGradPlan
rng = np.random.default_rng(1)
y = torch.from_numpy(rng.uniform(10, 100, (40, 50)).astype(np.float32)).cuda()
psf = torch.from_numpy(rng.uniform(0.1, 1, (5, 5)).astype(np.float32)).cuda()
client = N.Client(0, own_stream=True)
ws = N.Workspace(client)
plan = RLBatch([(40, 50)], [(5, 5)], [0], LA.RichardsonLucy(num_iter=3, clip=False), client=client)
grad = RLGradPlan([(40, 50)], (5, 5), 4, 2, client=client)
out = torch.empty_like(y)
torch.cuda.synchronize()
LA.richardson_lucy_into(y, psf, out, LA.RichardsonLucy(num_iter=3, clip=False), ws, client)
client.sync()
"""


def _run(body: str) -> subprocess.CompletedProcess:
    return subprocess.run([sys.executable, "-c", _PRELUDE + body], capture_output=True, text=True, timeout=300)


def test_client_destroyed_before_its_dependents(gpu):
    print("teardown: client first, then plan, workspace and gradient plan", flush=True)
    r = _run("""
ref = out.clone()
client.close()                      # the caller's reference; the plans keep the client alive
out2 = torch.empty_like(y)
plan.execute([y], [out2], [psf])    # still runs on the client's own stream
torch.cuda.synchronize()
assert torch.allclose(out2, ref, rtol=1e-4), float((out2 - ref).abs().max())
plan.close(); ws.close(); grad.close()
print("ok", flush=True)
""")
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0 and "ok" in r.stdout


def test_interpreter_exit_without_the_hook(gpu):
    print("teardown: objects left to interpreter shutdown, exit hook removed", flush=True)
    r = _run("""
import atexit
atexit.unregister(N._release_all)
print("ok", flush=True)
""")
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0 and "ok" in r.stdout
