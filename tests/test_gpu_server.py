"""Serving loop: requests batched while others run give the same result as
stepping each scene alone, and every request completes with a latency."""
import threading
import time

import numpy as np
import pytest

import octopusengine_amd as oe
from scenes import max_err

pytestmark = pytest.mark.gpu

SPECS = [
    (oe.TETRA10, (6, 2, 2), 5e5, 0.49),
    (oe.HEXA27, (4, 2, 2), 1e6, 0.45),
    (oe.TETRA10, (3, 2, 1), 2e6, 0.3),
    (oe.HEXA27, (2, 2, 1), 3e5, 0.49),
    (oe.TETRA10, (8, 1, 1), 1e5, 0.4),
]


def _scene(engine, spec):
    el, cells, E, nu = spec
    return oe.make_scene(engine, el, cells, young=E, poisson=nu)


def test_served_equals_alone(engine):
    served = [_scene(engine, s) for s in SPECS]
    alone = [_scene(engine, s) for s in SPECS]
    srv = oe.Server(engine, max_batch=3, streams=2)
    slots = [srv.add_scene(s) for s in served]
    rng = np.random.default_rng(11)
    counts = np.zeros(len(SPECS), int)
    # a producer thread submits while the worker is running batches
    order = rng.integers(0, len(SPECS), 60)

    def produce():
        for k in order:
            srv.submit(slots[k], 0.005, 1)
            time.sleep(float(rng.uniform(0, 3e-4)))

    th = threading.Thread(target=produce)
    th.start()
    th.join()
    srv.drain()
    for k in order:
        counts[k] += 1
    for s, n in zip(alone, counts):
        for _ in range(n):
            s.step(0.005, 1)
    for a, b, spec, n in zip(served, alone, SPECS, counts):
        err = max_err(a.positions(), b.positions())
        print(f"served el={spec[0]} cells={spec[1]} steps={n} |served-alone|={err:.3e}", flush=True)
        assert err <= 1e-6
    st = srv.stats()
    print(st, flush=True)
    assert st["completed"] == len(order)
    assert st["vertex_updates"] == int(sum(s.n_vertices * 10 * n for s, n in zip(served, counts)))
    lat = srv.latencies()
    assert lat.size == len(order) and (lat > 0).all()
    assert st["batches"] >= len(order) / 3
    srv.close()


def test_server_rejects_bad_requests(engine):
    s = _scene(engine, SPECS[0])
    srv = oe.Server(engine, max_batch=2, streams=1)
    slot = srv.add_scene(s)
    with pytest.raises(oe.OEError):
        srv.submit(slot + 5, 0.005, 1)
    with pytest.raises(oe.OEError):
        srv.submit(slot, -1.0, 1)
    srv.submit(slot, 0.005, 2)
    srv.drain()
    assert srv.stats()["completed"] == 1
    srv.close()


def test_budget_admission(engine):
    """Scenes beyond the engine's memory budget are refused at creation."""
    small = oe.Engine(0, budget_bytes=8 << 20)
    ok = oe.make_scene(small, oe.TETRA10, (4, 2, 2))
    kept = []
    with pytest.raises(oe.OEError) as ei:
        for _ in range(64):
            kept.append(oe.make_scene(small, oe.TETRA10, (16, 4, 4)))
    assert ei.value.status == 3   # OE_ERR_BUDGET
    used, budget = small.memory()
    assert used <= budget
    assert len(kept) >= 1
    ok.step(0.005, 1)
    # freeing scenes returns their memory and admits new ones
    kept.clear()
    again = oe.make_scene(small, oe.TETRA10, (16, 4, 4))
    again.step(0.005, 1)
