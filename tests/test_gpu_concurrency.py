"""Concurrency contract (crates/lucem-core/src/memory/workspace.rs:15-16: one workspace per thread
of work; a client is one device and one stream): two host threads, each with its own client (own
stream) and workspace, run Richardson-Lucy, the batched form and the correlation at the same time;
every result equals the one computed alone. The C ABI calls release the GIL (ctypes), so the
threads really overlap on the host and their streams on the device."""

from __future__ import annotations

import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _dev(a):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


def _work(seed, client, ws, reps):
    """RL 2-D and 3-D, a batched plan and a correlation on one client; returns host results."""
    import torch

    import lucem_amd as LA
    from lucem_amd.batched import RLBatch

    rng = np.random.default_rng(seed)
    y2 = _dev(rng.uniform(10, 100, (300, 260)))
    p2 = _dev(rng.uniform(0.1, 1.0, (9, 7)))
    y3 = _dev(rng.uniform(10, 100, (20, 64, 48)))
    p3 = _dev(rng.uniform(0.1, 1.0, (5, 5, 3)))
    tiles = [_dev(rng.uniform(10, 100, s)) for s in [(64, 80), (100, 37), (50, 50)]]
    s = LA.RichardsonLucy(num_iter=8, clip=False)
    with torch.cuda.stream(torch.cuda.ExternalStream(client.stream)):
        b = RLBatch([tuple(t.shape) for t in tiles], [(9, 7)], [0, 0, 0], s, client=client)
        o2, o3 = torch.empty_like(y2), torch.empty_like(y3)
        oc = torch.empty_like(y2)
        ob = [torch.empty_like(t) for t in tiles]
        for _ in range(reps):
            LA.richardson_lucy_into(y2, p2, o2, s, ws, client)
            LA.richardson_lucy_nd_into(y3, p3, o3, s, ws, client)
            b.execute(tiles, ob, [p2])
            LA.correlate_2d_into(y2, _dev(np.ones((21, 21), np.float32) / 441.0), oc, LA.Correlate2d(), ws, client)
        client.sync()
        return [o2.cpu().numpy(), o3.cpu().numpy(), oc.cpu().numpy()] + [o.cpu().numpy() for o in ob]


def test_two_clients_on_two_threads_match_sequential(gpu, monkeypatch):
    from lucem_amd import _native as N

    monkeypatch.setenv("LUCEM_AUTOTUNE", "0")  # the same launch shapes in every call, so bit-equal results

    clients = [N.Client(0, own_stream=True) for _ in range(2)]
    wss = [N.Workspace(c) for c in clients]
    alone = [_work(10 + i, clients[i], wss[i], 1) for i in range(2)]
    results = [None, None]
    errors = []

    def run(i):
        try:
            results[i] = _work(10 + i, clients[i], wss[i], 3)
        except Exception as e:  # noqa: BLE001 - reported below
            errors.append(e)

    threads = [threading.Thread(target=run, args=(i,)) for i in range(2)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert not errors, errors
    for i in range(2):
        assert results[i] is not None
        for got, want in zip(results[i], alone[i]):
            np.testing.assert_array_equal(got, want)


def test_default_stream_client_beside_a_capturing_thread(gpu):
    """A client on the null stream (torch's default) uploads twiddle tables and tuning inputs through
    that stream while another thread creates batched plans, whose creation captures a graph on a
    private non-blocking stream in thread-local capture mode: both sides complete, and the results
    equal the ones computed alone."""
    import torch

    import lucem_amd as LA
    from lucem_amd import _native as N
    from lucem_amd.batched import RLBatch

    rng = np.random.default_rng(77)
    shapes = [(90 + 7 * i, 70 + 5 * i) for i in range(6)]  # new geometries: new twiddles and tuning
    ys = [_dev(rng.uniform(10, 100, s)) for s in shapes]
    p = _dev(rng.uniform(0.1, 1.0, (7, 7)))
    s = LA.RichardsonLucy(num_iter=5, clip=False)
    torch.cuda.synchronize()
    default = N.Client(0)  # the null stream: not an own stream
    assert default.stream in (0, None)
    outs = [torch.empty_like(y) for y in ys]
    errors = []

    def plans():
        try:
            other = N.Client(0, own_stream=True)
            for i in range(4):
                t = [_dev(rng.uniform(10, 100, (40 + i, 50))), _dev(rng.uniform(10, 100, (33, 61 + i)))]
                b = RLBatch([tuple(x.shape) for x in t], [(7, 7)], [0, 0], s, client=other)
                b.execute(t, [torch.empty_like(x) for x in t], [p])
                b.close()
            other.sync()
        except Exception as e:  # noqa: BLE001 - reported below
            errors.append(e)

    th = threading.Thread(target=plans)
    th.start()
    for y, o in zip(ys, outs):
        LA.richardson_lucy_into(y, p, o, s, None, default)
    default.sync()
    th.join(timeout=120)
    assert not errors, errors
    for y, o in zip(ys, outs):
        ref = torch.empty_like(y)
        LA.richardson_lucy_into(y, p, ref, s, None, default)
        default.sync()
        np.testing.assert_array_equal(o.cpu().numpy(), ref.cpu().numpy())
