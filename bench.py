"""Benchmark: VBD vertex-updates/s and substep latency on one MI355X.

Default workload (N=1): plain VBD, one scene - the 1.0 x 0.25 x 0.25 m
cantilever beam as 64x16x16 hexes split into quadratic P2 tetrahedra
(140,481 vertices, 98,304 elements), Stable Neo-Hookean E=5e5 Pa, nu=0.49,
x=0 face fixed, gravity, dt=0.005 s, 10 VBD iterations per step, FP32 state
with FP64 3x3 solves, DSAT colouring, ascending colour order.

A step is one call of step(dt, 1): predictor, 10 iterations of (one gather
launch per colour + fixation + box limit), velocity update.  One
vertex-update = one vertex's local 3x3 solve in one iteration; throughput =
N_vertices * iterations * steps / wall time of the timed region.

Prints one JSON line (see README of the harness contract).
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import subprocess
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
SIZE = (1.0, 0.25, 0.25)


def algorithmic_bytes_per_iteration(n_vertices: int, n_elements: int, S: int, Q: int) -> float:
    """Compulsory bytes of one VBD sweep over all colours: per incidence
    (element, node) the element's FP32 JX_inv + V per quadrature point
    (40 Q), its S node indices and positions (16 S), the owner record (8);
    per vertex p, last_p, y, m, mask read and p written (56)."""
    return n_elements * S * (40.0 * Q + 16.0 * S + 8.0) + 56.0 * n_vertices


def cpu_info() -> tuple[str, int]:
    model = platform.processor() or "unknown"
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return model, os.cpu_count() or 0


def run_cpu_baseline(xyz, topo, steps: int = 2) -> dict:
    """FP64 oracle VBD (single thread) on a bounded sample of the same workload."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import pyoracle as O

    o = O.Scene(O.TETRA10, xyz, topo, material=O.STABLE_NEOHOOKE, young=5e5, poisson=0.49, iterations=10)
    o.fix(np.nonzero(xyz[:, 0] < 1e-6)[0])
    o.box((-10, -5, -10), (10, 5, 10))
    t0 = time.perf_counter()
    for _ in range(steps):
        o.step(0.005, 1)
    dt = time.perf_counter() - t0
    vu = xyz.shape[0] * 10 * steps
    model, ncpu = cpu_info()
    return {
        "value": vu / dt,
        "unit": "vertex-updates/s",
        "cores": 1,
        "kind": "port",
        "sample": f"{steps} steps (x10 iterations) of the same P2 64x16x16 scene, FP64 oracle, 1 thread; "
                  f"host {model}, {ncpu} cpus visible",
        "seconds": dt,
    }, o.positions()


def run_cpu_xpbd_reference() -> dict:
    """Reference CPU XPBD path (float, single thread) on its own config:
    P1 16x4x4 beam (425 vertices, 1536 tets), NeoHooke E=1e6 nu=0.45,
    dt=0.01, 20 substeps, 100 steps."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import octopusengine_amd as oe
    import pyoracle as O

    xyz, topo = oe.beam_mesh(oe.TETRA, (16, 4, 4), SIZE)
    fixed = np.nonzero(xyz[:, 0] < 1e-6)[0]
    c = O.CpuXPBD(O.TETRA, xyz, topo, fixed, material=O.NEOHOOKE, young=1e6, poisson=0.45, iterations=1,
                  substeps=20)
    t0 = time.perf_counter()
    for _ in range(100):
        c.step(0.01)
    dt = time.perf_counter() - t0
    return {"steps_per_s": 100 / dt, "constraint_projections_per_s": 1536 * 2 * 20 * 100 / dt,
            "vertex_substeps_per_s": 425 * 20 * 100 / dt, "seconds": dt, "cores": 1,
            "tip_y": float(c.positions()[:, 1].min())}


def plain_vbd(args) -> dict:
    import octopusengine_amd as oe

    eng = oe.Engine(0)
    xyz, topo = oe.beam_mesh(oe.TETRA10, (64, 16, 16), SIZE)
    t_setup = time.perf_counter()
    sc = oe.Scene(eng, oe.TETRA10, xyz, topo, material=oe.STABLE_NEOHOOKE, young=5e5, poisson=0.49, density=1000.0,
                  distribution=oe.MASS_FEMSHAPE, iterations=10)
    n_fixed = sc.fix_plane((1e-6, 0.0, 0.0), (-1.0, 0.0, 0.0))
    sc.box_limit((-10, -5, -10), (10, 5, 10))
    if args.tile:
        sc.set_tile(*[int(v) for v in args.tile.split(",")])
    setup_s = time.perf_counter() - t_setup
    N, E, S, Q = sc.n_vertices, sc.n_elements, sc.S, sc.Q
    dt, iters = 0.005, 10

    for _ in range(args.warmup):
        sc.step(dt, 1)
    eng.synchronize()

    # timed region: eager launches with HIP events around every gather launch
    sc.set_timing(True)
    sc.get_timing()
    eng.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        sc.step(dt, 1)
    eng.synchronize()
    wall = time.perf_counter() - t0
    gather_ms, launches = sc.get_timing()
    sc.set_timing(False)

    # the same steps replayed from a captured HIP graph (launch overhead removed)
    sc.set_graph(True)
    sc.step(dt, 1)
    eng.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        sc.step(dt, 1)
    eng.synchronize()
    wall_graph = time.perf_counter() - t0

    # per-substep latency: step + synchronise, host clock
    lat = []
    for _ in range(min(args.steps, 200)):
        t = time.perf_counter()
        sc.step(dt, 1)
        eng.synchronize()
        lat.append((time.perf_counter() - t) * 1e3)
    sc.set_graph(False)
    p = sc.positions()
    finite = bool(np.isfinite(p).all())

    vu_per_step = N * iters
    value = vu_per_step * args.steps / wall
    bytes_iter = algorithmic_bytes_per_iteration(N, E, S, Q)
    bytes_per_launch = bytes_iter * iters * args.steps / max(launches, 1)
    avg_launch_s = gather_ms / 1e3 / max(launches, 1)
    achieved = bytes_per_launch / avg_launch_s / 1e9
    out = {
        "metric": "VBD vertex-updates/s (plain, one scene)",
        "value": value,
        "unit": "vertex-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic beam mesh generated in-process (no dataset)",
        "config": {"workload": "plain VBD, P2 beam 64x16x16 (140481 vertices), SNH E=5e5 nu=0.49, "
                               "dt=0.005, 10 iterations, 1 scene",
                   "vertices": N, "elements": E, "colors": sc.n_colors, "fixed": n_fixed,
                   "tile": list(sc.get_tile()), "launch": "eager + per-colour HIP events"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_vertex_gather (VBD colour pass)",
                     "bytes_per_launch": bytes_per_launch, "avg_launch_us": avg_launch_s * 1e6,
                     "launches": launches},
        "p50_substep_ms": float(np.percentile(lat, 50)),
        "p99_substep_ms": float(np.percentile(lat, 99)),
        "graph_value": vu_per_step * args.steps / wall_graph,
        "graph_ms_per_step": wall_graph / args.steps * 1e3,
        "setup_s": setup_s,
        "positions_finite": finite,
    }
    eng_objs = (sc, eng)
    return out, xyz, topo, eng_objs


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--tile", default="", help="lanes,qpl override of the gather tile")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline legs")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; scenes shard across GPUs only by design "
                                   "(see DESIGN.md), --gpus must be 1"}))
        sys.exit(2)
    out, xyz, topo, keep = plain_vbd(args)
    if not args.no_cpu:
        cpu, _ = run_cpu_baseline(xyz, topo)
        out["cpu_baseline"] = cpu
        out["cpu_reference_xpbd_config1"] = run_cpu_xpbd_reference()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
