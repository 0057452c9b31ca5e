"""The served workload's shape under test: a pool of resident P2 / Q2 beams
with log-normal vertex counts (median 20k; here 47 drawn scenes plus one Q2
beam of about 110k vertices), 4 server streams, step requests (dt = 0.005 s,
10 VBD iterations) submitted by a producer thread with Poisson gaps
(2000 requests/s) while the server's ticks run.

Each served scene must end within 1e-6 m of its twin stepped alone (the
same number of steps through oe_scene_step), per-ticket wait/poll must
report completion, and the two smallest scenes must stay within 1e-5 of the
bounding-box diagonal of the FP64 oracle.
"""
import threading
import time

import numpy as np
import pytest

import octopusengine_amd as oe
from conftest import beam_diag
from scenes import max_err

pytestmark = pytest.mark.gpu

STEPS_PER_SCENE = 3


def _pool_specs():
    import bench

    specs = bench.served_specs(47, seed=5)
    # one Q2 beam of aspect 4:1:1 with about 110k vertices: (2*57+1)*(2*15+1)^2 = 110,515
    specs.append((oe.HEXA27, (57, 15, 15), 2.0e6, 0.45))
    return specs


def _build(engine, specs):
    import bench

    return bench.build_scenes(engine, specs, 8)


def test_served_pool_equals_alone_and_oracle(engine):
    import pyoracle as O

    specs = _pool_specs()
    served = _build(engine, specs)
    nv = np.array([s.n_vertices for s in served])
    print(f"pool: {len(served)} scenes, {int(nv.sum())} vertices, median {int(np.median(nv))}, "
          f"largest {int(nv.max())}", flush=True)
    assert nv.max() >= 100_000
    srv = oe.Server(engine, max_batch=64, streams=4)
    slots = [srv.add_scene(s) for s in served]
    rng = np.random.default_rng(21)
    order = rng.permutation(np.repeat(np.arange(len(served)), STEPS_PER_SCENE))
    gaps = rng.exponential(1.0 / 2000.0, order.size)
    tickets = [None] * order.size

    def produce():
        t0 = time.perf_counter()
        due = 0.0
        for i, (k, gap) in enumerate(zip(order, gaps)):
            due += gap
            while time.perf_counter() - t0 < due:
                pass
            tickets[i] = srv.submit(slots[k], 0.005, 1)

    th = threading.Thread(target=produce)
    th.start()
    th.join()
    # per-ticket completion: wait on every scene's last request
    last = {}
    for i, k in enumerate(order):
        last[int(k)] = tickets[i]
    for k, t in last.items():
        srv.wait(t)
        assert srv.poll(t)
    srv.drain()
    st = srv.stats()
    print(f"served: {st['completed']} requests in {st['batches']} ticks, p50 {st['p50_ms']:.2f} ms, "
          f"p99 {st['p99_ms']:.2f} ms", flush=True)
    assert st["completed"] == order.size
    assert all(srv.poll(t) for t in tickets)
    srv.close()

    # twins stepped alone, one scene at a time
    worst = 0.0
    for spec, s in zip(specs, served):
        (twin,) = _build(engine, [spec])
        for _ in range(STEPS_PER_SCENE):
            twin.step(0.005, 1)
        err = max_err(s.positions(), twin.positions())
        worst = max(worst, err)
        assert err <= 1e-6, f"{spec}: served vs alone {err:.3e}"
        twin.close()
    print(f"served vs alone: max {worst:.3e} m over {len(served)} scenes", flush=True)

    # the two smallest scenes against the oracle
    diag = beam_diag()
    for i in np.argsort(nv)[:2]:
        el, cells, E, nu = specs[i]
        xyz, topo = oe.beam_mesh(el, cells, (1.0, 0.25, 0.25))
        o = O.Scene(el, xyz, topo, material=O.STABLE_NEOHOOKE, young=E, poisson=nu, iterations=10)
        o.fix(np.nonzero(xyz[:, 0] < 1e-6)[0])
        o.box((-10, -5, -10), (10, 5, 10))
        for _ in range(STEPS_PER_SCENE):
            o.step(0.005, 1)
        err = max_err(served[i].positions(), o.positions())
        print(f"served scene el={el} cells={cells} ({nv[i]} vertices) vs oracle: {err:.3e}", flush=True)
        assert err <= 1e-5 * diag
    for s in served:
        s.close()
