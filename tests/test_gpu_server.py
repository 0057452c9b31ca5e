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


def test_server_tile_sets_the_served_layout(engine):
    """A server's Q2 layout (incidences per block, latency threshold) applies
    to every tick; pencil blocks of 14 and 28 incidences give the same steps."""
    specs = [(oe.HEXA27, (6, 3, 2), 1e6, 0.45), (oe.HEXA27, (5, 2, 2), 3e5, 0.49)]
    results = []
    for inc in (14, 28):
        scenes = [_scene(engine, sp) for sp in specs]
        srv = oe.Server(engine, max_batch=2, streams=1)
        slots = [srv.add_scene(sc) for sc in scenes]
        t = srv.get_tile_ex(oe.HEXA27)
        assert t["mode"] == oe.PASS_TASKS and t["inc_per_block"] == 14 and t["latency_blocks"] == 512
        # latency_blocks 0: every colour in the throughput form, so the layout is used
        srv.set_tile_ex(oe.HEXA27, t["lanes"], t["qpl"], t["mode"], inc, 0)
        t = srv.get_tile_ex(oe.HEXA27)
        assert t["inc_per_block"] == inc and t["latency_blocks"] == 0
        for _ in range(3):
            for sl in slots:
                srv.submit(sl, 0.005, 1)
        srv.drain()
        results.append([sc.positions() for sc in scenes])
        with pytest.raises(oe.OEError):
            srv.set_tile_ex(oe.HEXA27, 64, 1, oe.PASS_TASKS, 36, 0)   # above the kernel's 35
        with pytest.raises(oe.OEError):
            srv.set_tile_ex(99, 64, 1, oe.PASS_TASKS, 14, 0)
        srv.close()
        for sc in scenes:
            sc.close()
    alone = [_scene(engine, sp) for sp in specs]
    for sc in alone:
        for _ in range(3):
            sc.step(0.005, 1)
    for k, (sp, sc) in enumerate(zip(specs, alone)):
        d_inc = max_err(results[0][k], results[1][k])
        d_alone = max_err(results[0][k], sc.positions())
        print(f"served Q2 cells={sp[1]}: |inc14-inc28|={d_inc:.3e} |served-alone|={d_alone:.3e}", flush=True)
        assert d_inc <= 1e-7
        assert d_alone <= 1e-6


def test_server_tile_refused_while_requests_are_in_flight(engine):
    """The layout changes only while nothing is pending or in flight: a step of
    5 substeps of this scene (about 1,450 dependent launches) takes
    milliseconds, and the call right after the submission lands inside it."""
    sc = _scene(engine, (oe.HEXA27, (16, 4, 4), 1e6, 0.45))
    srv = oe.Server(engine, max_batch=2, streams=1)
    slot = srv.add_scene(sc)
    srv.submit(slot, 0.005, 5)
    with pytest.raises(oe.OEError):
        srv.set_tile_ex(oe.HEXA27, 64, 1, oe.PASS_TASKS, 28, 0)
    srv.drain()
    srv.set_tile_ex(oe.HEXA27, 64, 1, oe.PASS_TASKS, 28, 0)
    assert srv.get_tile_ex(oe.HEXA27)["inc_per_block"] == 28
    srv.submit(slot, 0.005, 1)
    srv.drain()
    assert np.isfinite(sc.positions()).all()
    srv.close()
    sc.close()


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


def test_wait_ticket_then_read_equals_alone(engine):
    """A read after wait(ticket) of a scene's last request sees every step
    submitted for it; poll reports completion."""
    served = _scene(engine, SPECS[1])
    alone = _scene(engine, SPECS[1])
    srv = oe.Server(engine, max_batch=4, streams=2)
    slot = srv.add_scene(served)
    tickets = [srv.submit(slot, 0.005, 1) for _ in range(4)]
    srv.wait(tickets[-1])
    assert all(srv.poll(t) for t in tickets)
    for _ in range(4):
        alone.step(0.005, 1)
    err = max_err(served.positions(), alone.positions())
    print(f"after wait(ticket): |served - alone| = {err:.3e}", flush=True)
    assert err <= 1e-6
    with pytest.raises(oe.OEError):
        srv.wait(tickets[-1] + 100)
    srv.close()


def test_reads_of_a_served_scene_wait_for_its_requests(engine):
    """positions() on a served scene returns the state after every request
    submitted so far, without drain or wait."""
    served = _scene(engine, SPECS[0])
    alone = _scene(engine, SPECS[0])
    srv = oe.Server(engine, max_batch=2, streams=2)
    slot = srv.add_scene(served)
    for _ in range(3):
        srv.submit(slot, 0.005, 1)
    p = served.positions()
    for _ in range(3):
        alone.step(0.005, 1)
    assert max_err(p, alone.positions()) <= 1e-6
    srv.close()


def test_served_scene_refuses_direct_use(engine):
    """While a server holds a scene, direct steps, state writes and constraint
    changes are refused (OE_ERR_STATE); after the server is gone they work."""
    s = _scene(engine, SPECS[0])
    srv = oe.Server(engine, max_batch=2, streams=1)
    slot = srv.add_scene(s)
    srv.submit(slot, 0.005, 1)
    for call in (lambda: s.step(0.005, 1), lambda: s.set_state(p=s.positions()),
                 lambda: s.fix_plane((1e-6, 0, 0), (-1, 0, 0)), lambda: s.box_limit(),
                 lambda: s.set_tile(64, 1, oe.PASS_TASKS)):
        with pytest.raises(oe.OEError) as ei:
            call()
        assert ei.value.status == 4
    srv.drain()
    srv.close()
    s.step(0.005, 1)


def test_duplicate_and_shared_scenes_are_rejected(engine):
    a = _scene(engine, SPECS[0])
    b = _scene(engine, SPECS[2])
    with pytest.raises(oe.OEError) as ei:
        oe.Batch(engine, [a, b, a])
    assert ei.value.status == 1
    srv = oe.Server(engine, max_batch=2, streams=1)
    srv.add_scene(a)
    with pytest.raises(oe.OEError) as ei:
        srv.add_scene(a)
    assert ei.value.status == 1
    with pytest.raises(oe.OEError) as ei:     # served scenes cannot join a batch
        oe.Batch(engine, [a, b])
    assert ei.value.status == 4
    bat = oe.Batch(engine, [b])
    with pytest.raises(oe.OEError) as ei:     # batched scenes cannot be served
        srv.add_scene(b)
    assert ei.value.status == 4
    with pytest.raises(oe.OEError) as ei:     # the batch copied b's constraints
        b.box_limit()
    assert ei.value.status == 4
    bat.close()
    b.box_limit()
    srv.close()


def test_destroying_a_served_scene_releases_its_slot(engine):
    a = _scene(engine, SPECS[0])
    keep = _scene(engine, SPECS[2])
    srv = oe.Server(engine, max_batch=2, streams=2)
    sa = srv.add_scene(a)
    sk = srv.add_scene(keep)
    for _ in range(3):
        srv.submit(sa, 0.005, 1)
        srv.submit(sk, 0.005, 1)
    a.close()                                  # waits for its requests
    with pytest.raises(oe.OEError):
        srv.submit(sa, 0.005, 1)
    t = srv.submit(sk, 0.005, 1)
    srv.wait(t)
    srv.drain()
    assert srv.stats()["completed"] == 7
    srv.close()


def test_destroyed_scene_stops_its_batch(engine):
    a = _scene(engine, SPECS[0])
    b = _scene(engine, SPECS[2])
    bat = oe.Batch(engine, [a, b])
    bat.step(0.005, 1)
    a.close()
    with pytest.raises(oe.OEError) as ei:
        bat.step(0.005, 1)
    assert ei.value.status == 4
    bat.close()
    b.step(0.005, 1)


def test_requests_of_one_scene_run_in_submission_order(engine):
    """Steps of different dt do not commute.  With one stream busy on a large
    scene, the queue holds [B dt1, A dt2, A dt1]: the tick formed for dt1 must
    not take A's second request ahead of its first."""
    big = oe.make_scene(engine, oe.HEXA27, (24, 6, 6))
    a = _scene(engine, SPECS[0])
    b = _scene(engine, SPECS[2])
    a_alone = _scene(engine, SPECS[0])
    srv = oe.Server(engine, max_batch=8, streams=1)
    sg, sa, sb = srv.add_scene(big), srv.add_scene(a), srv.add_scene(b)
    dt1, dt2 = 0.005, 0.0025
    srv.submit(sg, dt1, 1)
    srv.submit(sb, dt1, 1)
    srv.submit(sa, dt2, 1)
    srv.submit(sa, dt1, 1)
    srv.drain()
    a_alone.step(dt2, 1)
    a_alone.step(dt1, 1)
    err = max_err(a.positions(), a_alone.positions())
    print(f"mixed dt, one scene: |served - alone in submission order| = {err:.3e}", flush=True)
    assert err <= 1e-6
    srv.close()


def test_served_pool_shaped_like_the_served_workload(engine):
    """48 scenes drawn like the served workload's pool (log-normal vertex
    counts around 20k, P2 or Q2, E log-uniform, nu uniform), among them the
    pool's largest Q2 scene (346,275 vertices), on 4 streams with Poisson
    submissions (2000 requests/s) from a producer thread while ticks run.
    Every scene ends within 1e-6 m of its twin stepped alone as often, and
    the two smallest within 1e-5 of the beam diagonal of the FP64 oracle."""
    import sys
    from pathlib import Path

    import pyoracle as O
    from scenes import fixed_ids

    sys.path.insert(0, str(Path(__file__).resolve().parent.parent))
    import bench

    pool = bench.served_specs(1024, 5)
    nv = [(2 * c[0] + 1) * (2 * c[1] + 1) * (2 * c[2] + 1) for _, c, _, _ in pool]
    largest_q2 = max((n, i) for i, (n, sp) in enumerate(zip(nv, pool)) if sp[0] == oe.HEXA27)[1]
    specs = pool[:47] + [pool[largest_q2]]
    assert nv[largest_q2] >= 100_000
    served = [_scene(engine, s) for s in specs]
    alone = [_scene(engine, s) for s in specs]
    srv = oe.Server(engine, max_batch=64, streams=4)
    slots = [srv.add_scene(s) for s in served]
    rng = np.random.default_rng(7)
    n_req = 300
    which = rng.integers(0, len(specs), n_req)
    which[-1] = len(specs) - 1               # the large Q2 scene at least once
    gaps = rng.exponential(1.0 / 2000.0, n_req)

    def produce():
        for k, g in zip(which, gaps):
            time.sleep(float(g))
            srv.submit(slots[k], 0.005, 1)

    th = threading.Thread(target=produce)
    th.start()
    th.join()
    srv.drain()
    counts = np.bincount(which, minlength=len(specs))
    for s, n in zip(alone, counts):
        for _ in range(n):
            s.step(0.005, 1)
    worst = 0.0
    for a, b in zip(served, alone):
        worst = max(worst, max_err(a.positions(), b.positions()))
    st = srv.stats()
    print(f"served pool: {len(specs)} scenes, {n_req} requests, ticks {st['batches']}, "
          f"worst |served - alone| = {worst:.3e}", flush=True)
    assert st["completed"] == n_req
    assert worst <= 1e-6
    order = np.argsort([s.n_vertices for s in served])[:2]
    for i in order:
        el, cells, E, nu = specs[i]
        xyz, topo = oe.beam_mesh(el, cells, (1.0, 0.25, 0.25))
        o = O.Scene(el, xyz, topo, material=O.STABLE_NEOHOOKE, young=E, poisson=nu, iterations=10)
        o.fix(fixed_ids(xyz))
        o.box((-10, -5, -10), (10, 5, 10))
        for _ in range(counts[i]):
            o.step(0.005, 1)
        err = max_err(served[i].positions(), o.positions())
        print(f"served {cells} el={el} steps={counts[i]}: |served - oracle| = {err:.3e}", flush=True)
        assert err <= 1e-5 * float(np.sqrt(1.0 + 2 * 0.25 ** 2))
    srv.close()
