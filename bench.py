"""Benchmark: VBD vertex-updates/s and substep latency on one MI355X,
batched vs plain.

Headline workload (N = 1 GPU): batched VBD - 256 independent cantilever beams
(1.0 x 0.25 x 0.25 m) whose grid sizes are drawn from a fixed seed
(nx ~ U{8..64}, ny, nz ~ U{4..16}), each a P2-tetrahedron or Q2-hexahedron
mesh with probability 1/2, Stable Neo-Hookean with E ~ logU[1e5, 1e7] Pa and
nu ~ U[0.3, 0.49], x = 0 face fixed, gravity, dt = 0.005 s, 10 VBD
iterations per step, FP32 state and FP64 3x3 solves.  One step of the batch
advances every scene by one step; all scenes of one global colour are solved
by one launch per element type.

Also measured in the same run:
  * the plain per-scene loop on the same 256 scenes (each scene stepped on its
    own, one launch per colour) - the "plain" form of the metric;
  * the plain single-scene P2 64x16x16 beam (140,481 vertices, E = 5e5,
    nu = 0.49);
  * the CPU baseline: the FP64 oracle (single thread) on a bounded sample of
    the single-scene workload, and the reference CPU XPBD path on its own
    425-vertex beam.

A vertex-update is one vertex's local 3x3 solve in one iteration; throughput =
sum over scenes of N * iterations * steps / wall time of the timed region.
Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent

# HIP maps a process's streams onto GPU_MAX_HW_QUEUES hardware queues (4 by
# default).  The served leg runs 4 server streams next to the engine's own
# stream and the null stream; with 4 queues two of them would share one and
# serialise, so the bench asks for 8 (raising a lower setting).  It must be set
# before the HIP runtime starts, i.e. before the engine library is used.
if int(os.environ.get("GPU_MAX_HW_QUEUES", "4") or 4) < 8:
    os.environ["GPU_MAX_HW_QUEUES"] = "8"
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
# incidences per block of the Q2 tiles that select the pencil form of the throughput pass
Q2_PENCIL_INC = (14, 21, 28)
SIZE = (1.0, 0.25, 0.25)
DT = 0.005
ITERS = 10


def algorithmic_bytes_per_iteration(n_vertices: int, n_elements: int, S: int, Q: int) -> float:
    """Compulsory bytes of one VBD sweep over all colours: per incidence
    (element, node) the element's FP32 JX_inv + V per quadrature point
    (40 Q), its S node indices and positions (16 S), the owner record (8);
    per vertex p, last_p, y, m, mask read and p written (56)."""
    return n_elements * S * (40.0 * Q + 16.0 * S + 8.0) + 56.0 * n_vertices


def progress(msg: str) -> None:
    """One line per finished leg on stderr (stdout carries only the JSON line)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


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


def batched_beam_specs(n: int = 256, seed: int = 4):
    """(element, cells, E, nu) of the batched workload, drawn from `seed`."""
    import octopusengine_amd as oe

    rng = np.random.default_rng(seed)
    specs = []
    for _ in range(n):
        nx = int(rng.integers(8, 65))
        ny = int(rng.integers(4, 17))
        nz = int(rng.integers(4, 17))
        el = oe.TETRA10 if rng.random() < 0.5 else oe.HEXA27
        E = float(10.0 ** rng.uniform(5.0, 7.0))
        nu = float(rng.uniform(0.3, 0.49))
        specs.append((el, (nx, ny, nz), E, nu))
    return specs


def build_scenes(eng, specs, workers: int):
    import octopusengine_amd as oe

    def one(sp):
        el, cells, E, nu = sp
        return oe.make_scene(eng, el, cells, SIZE, young=E, poisson=nu, iterations=ITERS)

    with ThreadPoolExecutor(max_workers=workers) as ex:
        return list(ex.map(one, specs))


def timed(fn, steps, sync):
    sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    sync()
    return time.perf_counter() - t0


def measured_traffic(kernel_prefix: str, workload: str):
    """HBM bytes per launch of a kernel on a workload, from the rocprof
    FETCH_SIZE / WRITE_SIZE runs summarised in profiles/traffic.json (written
    from the counter CSVs; each counter already scaled by its copy-kernel
    calibration).  None when no such measurement exists."""
    path = ROOT / "profiles" / "traffic.json"
    try:
        rec = json.loads(path.read_text())
    except (OSError, ValueError):
        return None
    for r in rec.get("kernels", []):
        if r.get("kernel", "").startswith(kernel_prefix) and r.get("workload") == workload:
            return {"bytes_per_launch": r["read_bytes_per_launch"] + r["write_bytes_per_launch"],
                    "read_bytes_per_launch": r["read_bytes_per_launch"],
                    "write_bytes_per_launch": r["write_bytes_per_launch"],
                    "dispatches": r["dispatches"], "source": "profiles/traffic.json"}
    return None


def kernel_roofline(name: str, alg_bytes_per_launch: float, avg_launch_s: float, launches: int, traffic) -> dict:
    """Roofline of one colour-pass kernel: `achieved` = the algorithmic bytes
    of the hot-path model per launch / the kernel's event-timed average launch
    duration; `traffic` = HBM bytes it really moves per launch (rocprof), and
    `traffic_frac` = that traffic / duration / peak."""
    achieved = alg_bytes_per_launch / max(avg_launch_s, 1e-12) / 1e9
    out = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": achieved / HBM_PEAK_GBS, "traffic": traffic["bytes_per_launch"] if traffic else None,
           "kernel": name, "bytes_per_launch": alg_bytes_per_launch, "avg_launch_us": avg_launch_s * 1e6,
           "launches": launches}
    if traffic:
        out["traffic_gbs"] = traffic["bytes_per_launch"] / max(avg_launch_s, 1e-12) / 1e9
        out["traffic_frac"] = out["traffic_gbs"] / HBM_PEAK_GBS
        out["traffic_detail"] = traffic
    return out


def batched_leg(args, eng) -> dict:
    import octopusengine_amd as oe

    specs = batched_beam_specs(args.scenes, args.seed)
    t0 = time.perf_counter()
    scenes = build_scenes(eng, specs, args.workers)
    setup_s = time.perf_counter() - t0
    nv = int(sum(s.n_vertices for s in scenes))
    n_p2 = sum(1 for s in scenes if s.element == oe.TETRA10)
    b = oe.Batch(eng, scenes)
    b.set_streams(args.batch_streams)
    info = b.info()
    vu_step = info["vertex_updates"]
    for _ in range(args.warmup):
        b.step(DT, 1)
    # the two smallest scenes (one per element type) after the warm-up steps,
    # for the accuracy check against the oracle
    early = {"steps": args.warmup, "specs": [], "positions": []}
    for el in (oe.TETRA10, oe.HEXA27):
        idx = [i for i, sp in enumerate(specs) if sp[0] == el]
        if idx:
            i = min(idx, key=lambda k: scenes[k].n_vertices)
            early["specs"].append(specs[i])
            early["positions"].append(scenes[i].positions())
    # timed region (value): eager launches, as Batch.step issues them
    wall = timed(lambda: b.step(DT, 1), args.steps, eng.synchronize)
    # roofline: as many steps again with HIP events around every colour-pass
    # launch, on the stream the kernels run on (kernel durations only; with
    # the events on, every pass runs on one stream, so each duration is the
    # kernel's own and not shared with a pass beside it)
    b.set_timing(True)
    b.get_timing()
    q2_pencil = b.get_tile_ex(oe.HEXA27)["inc_per_block"] in Q2_PENCIL_INC
    kinds = [(oe.TETRA10, "k_batch_fused_tasks<10,4> (P2 colour pass)", "k_batch_fused_tasks<10, 4, false>"),
             (oe.HEXA27, "k_batch_q2_pencil (Q2 colour pass, pencil tasks)", "k_batch_q2_pencil") if q2_pencil else
             (oe.HEXA27, "k_batch_q2_fused (Q2 colour pass)", "void k_batch_q2_fused<false>")]
    for el, _, _ in kinds:
        b.get_timing_element(el)
    wall_events = timed(lambda: b.step(DT, 1), args.steps, eng.synchronize)
    per_kernel = {}
    for el, name, prefix in kinds:
        ms, n = b.get_timing_element(el)
        ie = b.info_element(el)
        if n == 0 or ie["gather_launches"] == 0:
            continue
        per_kernel[el] = kernel_roofline(name, ie["gather_bytes"] / ie["gather_launches"], ms / 1e3 / n, n,
                                         measured_traffic(prefix, f"batched{args.scenes}_seed{args.seed}"))
        per_kernel[el]["ms_per_step"] = ms / args.steps
    gather_ms, launches = b.get_timing()
    b.set_timing(False)
    wall_graph, wall_plain, lat = wall, float("nan"), [float("nan")]
    plain_steps = max(1, args.plain_steps)
    if not args.eager_only:
        # same steps from a captured HIP graph
        b.set_graph(True)
        b.step(DT, 1)
        wall_graph = timed(lambda: b.step(DT, 1), args.steps, eng.synchronize)
        lat = []
        for _ in range(args.steps):
            t = time.perf_counter()
            b.step(DT, 1)
            eng.synchronize()
            lat.append((time.perf_counter() - t) * 1e3)
        b.set_graph(False)

        # plain per-scene loop on the same scenes (each scene steps on its own)
        def plain_step():
            for s in scenes:
                s.step(DT, 1)

        plain_step()
        wall_plain = timed(plain_step, plain_steps, eng.synchronize)
    finite = all(bool(np.isfinite(s.positions()).all()) for s in scenes)
    bytes_total = info["gather_bytes"] * args.steps
    avg_launch_s = gather_ms / 1e3 / max(launches, 1)
    achieved = bytes_total / max(launches, 1) / avg_launch_s / 1e9
    b.close()
    # the kernel with the largest share of the step is the headline roofline
    dominant = max(per_kernel.values(), key=lambda r: r["ms_per_step"])
    return {
        "value": vu_step * args.steps / wall,
        "graph_value": vu_step * args.steps / wall_graph,
        "ms_per_step": wall / args.steps * 1e3,
        "graph_ms_per_step": wall_graph / args.steps * 1e3,
        "events_ms_per_step": wall_events / args.steps * 1e3,
        "p50_substep_ms": float(np.percentile(lat, 50)),
        "p99_substep_ms": float(np.percentile(lat, 99)),
        "plain_value": vu_step * plain_steps / wall_plain,
        "plain_ms_per_step": wall_plain / plain_steps * 1e3,
        # both eager: the headline timed region against the plain loop's eager launches
        "speedup_vs_plain": (vu_step * args.steps / wall) / (vu_step * plain_steps / wall_plain),
        "scenes": len(scenes), "p2_scenes": n_p2, "vertices": nv, "vertex_updates_per_step": vu_step,
        "gather_launches_per_step": info["gather_launches"], "setup_s": setup_s, "positions_finite": finite,
        "roofline": dominant,
        "roofline_by_kernel": list(per_kernel.values()),
        "colour_pass_model_frac_all_kernels": achieved / HBM_PEAK_GBS,
        "colour_pass_avg_launch_us_all_kernels": avg_launch_s * 1e6,
        "_scenes": scenes,
        "_early": early,
    }


def xpbd_explicit_leg(eng) -> dict:
    """The high-order explicit and XPBD workload on the GPU: Q2 and P3 beams of
    32 x 8 x 8 cells, Neo-Hookean E = 1e6 Pa, nu = 0.45, +-1e-3 m initial noise,
    x = 0 face fixed; XPBD frames of dt = 0.01 s with 50 substeps (V2 kernels)
    and explicit substeps of 1e-5 s (timed as 10 calls of 200 substeps)."""
    import octopusengine_amd as oe

    out = {}
    for el, name in ((oe.HEXA27, "q2"), (oe.TETRA20, "p3")):
        xyz, topo = oe.beam_mesh(el, (32, 8, 8), (1.0, 0.25, 0.25))
        rng = np.random.default_rng(11)
        p0 = (xyz + rng.uniform(-1e-3, 1e-3, xyz.shape)).astype(np.float32)
        for solver, kind in ((oe.SOLVER_XPBD, "xpbd"), (oe.SOLVER_EXPLICIT, "explicit")):
            sc = oe.Scene(eng, el, xyz, topo, material=oe.NEOHOOKE, young=1e6, poisson=0.45, density=1000.0,
                          distribution=oe.MASS_FEMSHAPE, solver=solver, xpbd_version=2)
            sc.fix_plane((1e-6, 0.0, 0.0), (-1.0, 0.0, 0.0))
            sc.box_limit((-10.0, -5.0, -10.0), (10.0, 5.0, 10.0))
            sc.set_state(p=p0, last_p=p0)
            if kind == "xpbd":
                dt, sub, calls = 0.01, 50, 5
            else:
                dt, sub, calls = 200 * 1e-5, 200, 10
            sc.step(dt, sub)
            wall = timed(lambda: sc.step(dt, sub), calls, eng.synchronize)
            n = sc.n_vertices
            rec = {"vertices": n, "substeps_per_call": sub, "calls": calls,
                   "us_per_substep": wall / (calls * sub) * 1e6,
                   "vertex_substeps_per_s": n * sub * calls / wall,
                   "positions_finite": bool(np.isfinite(sc.positions()).all())}
            if kind == "xpbd":
                rec["ms_per_frame"] = wall / calls * 1e3
            out[f"{kind}_{name}_32x8x8"] = rec
            sc.close()
    return out


def plain_single_leg(args, eng) -> dict:
    import octopusengine_amd as oe

    sc = oe.make_scene(eng, oe.TETRA10, (64, 16, 16), SIZE, young=5e5, poisson=0.49, iterations=ITERS)
    N, E, S, Q = sc.n_vertices, sc.n_elements, sc.S, sc.Q
    for _ in range(args.warmup):
        sc.step(DT, 1)
    wall = timed(lambda: sc.step(DT, 1), args.steps, eng.synchronize)
    sc.set_timing(True)
    sc.get_timing()
    wall_events = timed(lambda: sc.step(DT, 1), args.steps, eng.synchronize)
    gather_ms, launches = sc.get_timing()
    sc.set_timing(False)
    sc.set_graph(True)
    sc.step(DT, 1)
    wall_graph = timed(lambda: sc.step(DT, 1), args.steps, eng.synchronize)
    lat = []
    for _ in range(args.steps):
        t = time.perf_counter()
        sc.step(DT, 1)
        eng.synchronize()
        lat.append((time.perf_counter() - t) * 1e3)
    sc.set_graph(False)
    bytes_per_launch = algorithmic_bytes_per_iteration(N, E, S, Q) * ITERS * args.steps / max(launches, 1)
    avg_launch_s = gather_ms / 1e3 / max(launches, 1)
    achieved = bytes_per_launch / avg_launch_s / 1e9
    out = {
        "value": N * ITERS * args.steps / wall,
        "graph_value": N * ITERS * args.steps / wall_graph,
        "ms_per_step": wall / args.steps * 1e3,
        "graph_ms_per_step": wall_graph / args.steps * 1e3,
        "events_ms_per_step": wall_events / args.steps * 1e3,
        "p50_substep_ms": float(np.percentile(lat, 50)),
        "p99_substep_ms": float(np.percentile(lat, 99)),
        "vertices": N, "elements": E, "colors": sc.n_colors, "tile": list(sc.get_tile()),
        "gather_frac_of_hbm_peak": achieved / HBM_PEAK_GBS, "gather_achieved_gbs": achieved,
        "gather_avg_launch_us": avg_launch_s * 1e6,
    }
    sc.close()
    return out


def cpu_baseline_leg(steps: int) -> dict:
    """The reference's CPU VBD path (VertexBlockDescent + VBD_FEM: sequential
    Gauss-Seidel over the vertices in id order, float, one thread), restated
    in C, on a bounded sample of the single-scene workload."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import octopusengine_amd as oe
    import pyoracle as O

    xyz, topo = oe.beam_mesh(oe.TETRA10, (64, 16, 16), SIZE)
    c = O.CpuVBD(O.TETRA10, xyz, topo, np.nonzero(xyz[:, 0] < 1e-6)[0], material=O.STABLE_NEOHOOKE, young=5e5,
                 poisson=0.49, iterations=ITERS)
    t0 = time.perf_counter()
    for _ in range(steps):
        c.step(DT)
    dt = time.perf_counter() - t0
    model, ncpu = cpu_info()
    return {"value": xyz.shape[0] * ITERS * steps / dt, "unit": "vertex-updates/s", "cores": 1, "kind": "port",
            "sample": f"{steps} steps x {ITERS} iterations of the P2 64x16x16 beam (140481 vertices) on the "
                      f"reference's CPU VBD path restated in C (float, sequential Gauss-Seidel), 1 thread, host "
                      f"'{model}' ({ncpu} cpus visible)",
            "seconds": dt}


def _oracle_run(spec, steps):
    """FP64 oracle positions of a beam scene (element, cells, E, nu) after
    `steps` VBD steps (fixed x = 0 face, box limit, dt = 0.005, 10 iterations)."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import octopusengine_amd as oe
    import pyoracle as O

    el, cells, E, nu = spec
    xyz, topo = oe.beam_mesh(el, cells, SIZE)
    o = O.Scene(el, xyz, topo, material=O.STABLE_NEOHOOKE, young=E, poisson=nu, iterations=ITERS)
    o.fix(np.nonzero(xyz[:, 0] < 1e-6)[0])
    o.box((-10, -5, -10), (10, 5, 10))
    t0 = time.perf_counter()
    for _ in range(steps):
        o.step(DT, 1)
    return o.positions(), time.perf_counter() - t0


def accuracy_leg(args, early) -> dict:
    """Max vertex-position error of GPU results against FP64 oracle runs of
    the same scenes, as a fraction of the bounding-box diagonal (targets:
    1e-5 after 1 step, 1e-3 after 100):
      * the two smallest scenes (one P2, one Q2) of the batched workload,
        read from the timed batch after its warm-up steps;
      * the same two scenes stepped 100 times in a batch of their own (the
        batched kernels);
      * the single-scene P2 64x16x16 beam after `cpu_steps` steps.
    The oracle runs go to a thread pool (the C oracle releases the GIL)."""
    import octopusengine_amd as oe

    diag = float(np.sqrt(sum(v * v for v in SIZE)))
    long_steps = 100
    eng = oe.Engine(0)
    small = early["specs"]
    twins = [oe.make_scene(eng, el, cells, SIZE, young=E, poisson=nu, iterations=ITERS) for el, cells, E, nu in small]
    tb = oe.Batch(eng, twins)
    for _ in range(long_steps):
        tb.step(DT, 1)
    long_gpu = [t.positions() for t in twins]
    tb.close()
    for t in twins:
        t.close()
    single = oe.make_scene(eng, oe.TETRA10, (64, 16, 16), SIZE, young=5e5, poisson=0.49, iterations=ITERS)
    for _ in range(args.cpu_steps):
        single.step(DT, 1)
    single_gpu = single.positions()
    single.close()
    eng.close()
    jobs = [(sp, early["steps"]) for sp in small] + [(sp, long_steps) for sp in small] + \
           [((oe.TETRA10, (64, 16, 16), 5e5, 0.49), args.cpu_steps)]
    with ThreadPoolExecutor(max_workers=len(jobs)) as ex:
        res = list(ex.map(lambda j: _oracle_run(*j), jobs))
    out = []
    names = ["P2", "Q2"]
    for i, sp in enumerate(small):
        el, cells, E, nu = sp
        tag = f"{names[i]} {cells[0]}x{cells[1]}x{cells[2]} (E={E:.3g}, nu={nu:.3f})"
        e1 = float(np.abs(early["positions"][i].astype(np.float64) - res[i][0]).max())
        e2 = float(np.abs(long_gpu[i].astype(np.float64) - res[len(small) + i][0]).max())
        out.append({"scene": tag + ", the timed 256-scene batch", "steps": early["steps"], "max_err_m": e1,
                    "frac_of_diag": e1 / diag})
        out.append({"scene": tag + ", batch of the two smallest scenes", "steps": long_steps, "max_err_m": e2,
                    "frac_of_diag": e2 / diag})
    e3 = float(np.abs(single_gpu.astype(np.float64) - res[-1][0]).max())
    out.append({"scene": "P2 64x16x16 single scene (E=5e5, nu=0.49)", "steps": args.cpu_steps, "max_err_m": e3,
                "frac_of_diag": e3 / diag})
    return {"vs": "FP64 oracle of the same scene (oracle/oracle.c)", "diag_m": diag, "results": out,
            "max_frac_of_diag_after_100_steps": max(r["frac_of_diag"] for r in out if r["steps"] >= 100),
            "oracle_seconds": sum(r[1] for r in res)}


def cpu_xpbd_reference_leg() -> dict:
    """Reference CPU XPBD path (float, single thread) on its own configuration:
    P1 16x4x4 beam (425 vertices, 1536 tets), NeoHooke E = 1e6, nu = 0.45,
    dt = 0.01, 20 substeps, 100 steps."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import octopusengine_amd as oe
    import pyoracle as O

    xyz, topo = oe.beam_mesh(oe.TETRA, (16, 4, 4), SIZE)
    fixed = np.nonzero(xyz[:, 0] < 1e-6)[0]
    c = O.CpuXPBD(O.TETRA, xyz, topo, fixed, material=O.NEOHOOKE, young=1e6, poisson=0.45, iterations=1, substeps=20)
    t0 = time.perf_counter()
    for _ in range(100):
        c.step(0.01)
    dt = time.perf_counter() - t0
    return {"steps_per_s": 100 / dt, "constraint_projections_per_s": 1536 * 2 * 20 * 100 / dt,
            "vertex_substeps_per_s": 425 * 20 * 100 / dt, "seconds": dt, "cores": 1}


def copy_calibration(nbytes: int) -> dict:
    """Device-to-device copy of nbytes with the HIP runtime (its copy kernel):
    the reference point for scaling rocprof FETCH_SIZE / WRITE_SIZE readings."""
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
    hip.hipDeviceSynchronize()
    t0 = time.perf_counter()
    assert hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) == 0   # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    dt = time.perf_counter() - t0
    hip.hipFree(a)
    hip.hipFree(b)
    return {"bytes": nbytes, "seconds": dt, "gbs_read_plus_write": 2 * nbytes / dt / 1e9}


def served_specs(n: int, seed: int):
    """(element, cells, E, nu) of the served pool: vertex counts drawn
    log-normal (median 20k, sigma 1, capped at 500k), element P2 or Q2 with
    probability 1/2, beams of aspect 4:1:1 sized to the drawn count:
    N = (2 nx + 1)(2 k + 1)^2 with ny = nz = k."""
    import octopusengine_amd as oe

    rng = np.random.default_rng(seed)
    specs = []
    for _ in range(n):
        target = min(float(np.exp(np.log(20000.0) + rng.normal())), 500000.0)
        k = max(1, int(round((target / 32.0) ** (1.0 / 3.0))))
        nx = max(1, int(round((target / (2 * k + 1) ** 2 - 1.0) / 2.0)))
        el = oe.TETRA10 if rng.random() < 0.5 else oe.HEXA27
        E = float(10.0 ** rng.uniform(5.0, 7.0))
        nu = float(rng.uniform(0.3, 0.49))
        specs.append((el, (nx, k, k), E, nu))
    return specs


def served_leg(args) -> dict:
    """Poisson step requests against a pool of resident scenes through the
    continuous-batching server; latency = completion - submission."""
    import octopusengine_amd as oe

    eng = oe.Engine(0, budget_bytes=int(args.serve_budget_gb * 2**30))
    specs = served_specs(args.serve_pool, args.seed + 1)
    t0 = time.perf_counter()
    scenes = build_scenes(eng, specs, args.workers)
    setup_s = time.perf_counter() - t0
    nv = np.array([s.n_vertices for s in scenes])
    srv = oe.Server(eng, max_batch=args.serve_max_batch, streams=args.serve_streams)
    slots = [srv.add_scene(s) for s in scenes]
    if args.serve_q2_inc or args.serve_q2_lat_blocks >= 0:
        q2 = srv.get_tile_ex(oe.HEXA27)
        srv.set_tile_ex(oe.HEXA27, q2["lanes"], q2["qpl"], q2["mode"],
                        args.serve_q2_inc or q2["inc_per_block"],
                        args.serve_q2_lat_blocks if args.serve_q2_lat_blocks >= 0 else q2["latency_blocks"])
    for sl in slots:                          # warm-up: every scene once
        srv.submit(sl, DT, 1)
    srv.drain()
    base = srv.stats()
    n_lat0 = base["completed"]
    rng = np.random.default_rng(args.seed + 2)
    n_req = int(args.serve_rate * args.serve_seconds)
    gaps = rng.exponential(1.0 / args.serve_rate, n_req)
    which = rng.integers(0, len(slots), n_req)
    arrivals = np.cumsum(gaps)
    lag = 0.0
    t_start = time.perf_counter()
    for a, k in zip(arrivals, which):
        while True:
            now = time.perf_counter() - t_start
            if now >= a:
                break
            if a - now > 1e-3:
                time.sleep(a - now - 5e-4)
        lag = max(lag, time.perf_counter() - t_start - a)
        srv.submit(slots[k], DT, 1)
    srv.drain()
    wall = time.perf_counter() - t_start
    st = srv.stats()
    lat = srv.latencies()[n_lat0:]
    lat_slot = srv.latency_slots()[n_lat0:]
    vu = st["vertex_updates"] - base["vertex_updates"]
    # the tail by scene: element type and size class of each request's scene
    els = np.array([sp[0] for sp in specs])[lat_slot]
    nvs = nv[lat_slot]
    by_class = []
    for el, name in ((oe.TETRA10, "P2"), (oe.HEXA27, "Q2")):
        for lo, hi in ((0, 25_000), (25_000, 100_000), (100_000, 10**9)):
            sel = (els == el) & (nvs >= lo) & (nvs < hi)
            if sel.any():
                by_class.append({"element": name, "vertices": [lo, min(hi, int(nv.max()))], "requests": int(sel.sum()),
                                 "p50_ms": float(np.percentile(lat[sel], 50)),
                                 "p99_ms": float(np.percentile(lat[sel], 99))})
    slowest = lat >= np.percentile(lat, 99)
    out = {
        "pool_scenes": len(scenes), "pool_vertices": int(nv.sum()), "median_scene_vertices": float(np.median(nv)),
        "max_scene_vertices": int(nv.max()), "offered_rate_req_s": args.serve_rate, "requests": n_req,
        "seconds": args.serve_seconds, "wall_s": wall, "max_submit_lag_ms": lag * 1e3,
        "offered_vertex_updates_per_s": float(nv[which].sum() * ITERS / max(arrivals[-1], 1e-9)),
        "value": vu / wall, "unit": "vertex-updates/s",
        "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
        "max_ms": float(lat.max()), "ticks": st["batches"] - base["batches"], "mean_batch": st["mean_batch"],
        "by_scene_class": by_class,
        "slowest_1pct_share_q2": float((els[slowest] == oe.HEXA27).mean()),
        "slowest_1pct_median_vertices": float(np.median(nvs[slowest])),
        "max_batch": args.serve_max_batch, "streams": args.serve_streams,
        "q2_tile": srv.get_tile_ex(oe.HEXA27),
        "hw_queues": int(os.environ.get("GPU_MAX_HW_QUEUES", "4")), "setup_s": setup_s,
        "device_bytes": int(eng.memory()[0]), "budget_bytes": int(eng.memory()[1]),
    }
    srv.close()
    for s in scenes:
        s.close()
    return out


def _finite_or_null(obj):
    """JSON has no NaN: legs that were skipped report null."""
    if isinstance(obj, float) and not np.isfinite(obj):
        return None
    if isinstance(obj, dict):
        return {k: _finite_or_null(v) for k, v in obj.items()}
    if isinstance(obj, list):
        return [_finite_or_null(v) for v in obj]
    return obj


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--scenes", type=int, default=256)
    ap.add_argument("--seed", type=int, default=4)
    ap.add_argument("--plain-steps", type=int, default=3)
    ap.add_argument("--workers", type=int, default=min(16, os.cpu_count() or 4))
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline legs")
    ap.add_argument("--no-single", action="store_true", help="skip the single-scene leg")
    ap.add_argument("--no-serve", action="store_true", help="skip the served leg")
    ap.add_argument("--serve-pool", type=int, default=1024, help="resident scenes in the served pool")
    ap.add_argument("--serve-seconds", type=float, default=60.0, help="length of the Poisson request stream")
    ap.add_argument("--serve-rate", type=float, default=2000.0, help="step requests per second")
    ap.add_argument("--serve-max-batch", type=int, default=64)
    ap.add_argument("--serve-streams", type=int, default=4)
    ap.add_argument("--serve-budget-gb", type=float, default=200.0)
    ap.add_argument("--serve-q2-inc", type=int, default=0,
                    help="incidences per block of the served Q2 passes (0: the server's default)")
    ap.add_argument("--serve-q2-lat-blocks", type=int, default=-1,
                    help="throughput-form blocks below which a served Q2 colour runs the latency form (-1: default)")
    ap.add_argument("--eager-only", action="store_true",
                    help="batched leg: only the eager timed region (no graph, latency or plain loop)")
    ap.add_argument("--batch-streams", type=int, default=2, choices=(1, 2),
                    help="streams of a batch step: 2 runs the Q2 colour passes beside the P2 ones")
    ap.add_argument("--serve-only", action="store_true", help="run only the served leg")
    ap.add_argument("--xpbd-explicit-only", action="store_true",
                    help="run only the explicit / XPBD leg (Q2 and P3 32x8x8) and print its figures (profiling)")
    ap.add_argument("--single-only", action="store_true",
                    help="run only the single-scene leg (P2 64x16x16) and print its figures (profiling)")
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="also run a device copy of this many bytes (rocprof counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "one GPU per process: scenes shard across GPUs by scene (design only); "
                                   "run with --gpus 1"}))
        sys.exit(2)
    import octopusengine_amd as oe

    eng = oe.Engine(0)
    if args.single_only:
        print(json.dumps(_finite_or_null({"plain_single_scene_p2_64x16x16": plain_single_leg(args, eng)})), flush=True)
        eng.close()
        return
    if args.xpbd_explicit_only:
        print(json.dumps(_finite_or_null({"xpbd_explicit_32x8x8": xpbd_explicit_leg(eng)})), flush=True)
        eng.close()
        return
    if args.serve_only:
        eng.close()
        print(json.dumps(_finite_or_null({"served": served_leg(args)})), flush=True)
        return
    bat = batched_leg(args, eng)
    progress(f"batched leg: {bat['ms_per_step']:.2f} ms per step")
    scenes = bat.pop("_scenes")
    early = bat.pop("_early")
    roof = bat.pop("roofline")
    roof_by_kernel = bat.pop("roofline_by_kernel")
    out = {
        "metric": "VBD vertex-updates/s, batched (256 mixed P2/Q2 scenes) vs plain per-scene loop",
        "value": bat["value"],
        "unit": "vertex-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": bat["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic beam meshes generated in-process from a fixed seed (no dataset)",
        "config": {"workload": f"batched VBD: {bat['scenes']} beams ({bat['p2_scenes']} P2 / "
                               f"{bat['scenes'] - bat['p2_scenes']} Q2), nx~U[8,64], ny,nz~U[4,16], "
                               f"E~logU[1e5,1e7], nu~U[0.3,0.49], SNH, dt=0.005, 10 iterations, seed {args.seed}",
                   "vertices": bat["vertices"], "vertex_updates_per_step": bat["vertex_updates_per_step"],
                   "gather_launches_per_step": bat["gather_launches_per_step"],
                   "batch_streams": args.batch_streams,
                   "launch": "eager (value; with 2 batch streams the Q2 colour passes run beside the P2 ones); "
                             "roofline from a second run of as many steps with HIP events around every colour-pass "
                             "launch, all on one stream, for the kernel with the largest share of a step "
                             "(every colour-pass kernel in roofline_by_kernel)"},
        "roofline": roof,
        "roofline_by_kernel": roof_by_kernel,
        "p50_substep_ms": bat["p50_substep_ms"],
        "p99_substep_ms": bat["p99_substep_ms"],
        "latency_source": "batched step of all 256 scenes (graph replay)",
        "batched": {k: bat[k] for k in ("graph_value", "graph_ms_per_step", "events_ms_per_step", "plain_value",
                                        "plain_ms_per_step",
                                        "speedup_vs_plain", "p50_substep_ms", "p99_substep_ms", "setup_s",
                                        "positions_finite", "colour_pass_model_frac_all_kernels",
                                        "colour_pass_avg_launch_us_all_kernels")},
    }
    for sc in scenes:
        sc.close()
    del scenes
    # one engine (one stream) at a time: the server adds its own streams, and a
    # process gets 4 hardware queues, beyond which streams share a queue and
    # serialise
    eng.close()
    if not args.no_serve:
        out["served"] = served_leg(args)
        progress(f"served leg: p50 {out['served']['p50_ms']:.2f} ms, p99 {out['served']['p99_ms']:.2f} ms")
        # the substep-latency target is a serving figure: request (one step)
        # latency under the Poisson load
        out["p50_substep_ms"] = out["served"]["p50_ms"]
        out["p99_substep_ms"] = out["served"]["p99_ms"]
        out["latency_source"] = "served: Poisson step requests, latency = completion - submission"
    if not args.no_single:
        eng = oe.Engine(0)
        out["plain_single_scene_p2_64x16x16"] = plain_single_leg(args, eng)
        progress(f"single-scene leg: {out['plain_single_scene_p2_64x16x16']['ms_per_step']:.3f} ms per step")
        out["xpbd_explicit_32x8x8"] = xpbd_explicit_leg(eng)
        progress("config-3 leg: " + ", ".join(f"{k} {v['us_per_substep']:.1f} us/substep"
                                              for k, v in out["xpbd_explicit_32x8x8"].items()))
        eng.close()
    if args.copy_bytes > 0:
        out["copy_calibration"] = copy_calibration(args.copy_bytes)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline_leg(args.cpu_steps)
        progress(f"cpu baseline: {out['cpu_baseline']['value']:.3e} vertex-updates/s")
        out["cpu_reference_xpbd_p1_16x4x4"] = cpu_xpbd_reference_leg()
        if early["steps"] > 0:
            out["accuracy"] = accuracy_leg(args, early)
            progress(f"accuracy leg: max {out['accuracy']['max_frac_of_diag_after_100_steps']:.2e} of the diagonal")
    print(json.dumps(_finite_or_null(out)), flush=True)


if __name__ == "__main__":
    main()
