"""MI355X soft-body FEM stepping engine (gfx950, HIP) - Python front end.

The compute path is liboe.so (C ABI in include/octopus_engine.h); this module
only converts numpy arrays to pointers.  The object names follow the
reference's script components so that a scene reads like the reference's:

    Cuda_VBD_FEM_Dynamic(density, distrib, young, poisson, material,
                         iteration, sub_iteration, damping, rho, version, target_color)
        -> vbd_scene(...)          (include/Script/Dynamic/Cuda_VBD_FEM_Dynamic.h:10-15)
    Cuda_XPBD_FEM_Dynamic(density, distrib, young, poisson, material,
                          sub_it, damping, coupled, version)
        -> xpbd_scene(...)         (include/Script/Dynamic/Cuda_XPBD_FEM_Dynamic.h:13-19)
    Cuda_FEM_Explicit_Dynamic(sub_it, density, distrib, young, poisson,
                              material, damping, integrator)
        -> explicit_scene(...)     (include/Script/Dynamic/Cuda_FEM_Explicit_Dynamic.h:5-21)
    Cuda_Constraint_Rigid_Controller(Plane(o, n))  -> Scene.fix_plane(o, n)
    update()                        -> Scene.update() = step(dt, sub_iterations)
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from ._lib import (HEXA, HEXA27, MASS_FEMSHAPE, MASS_UNIFORM, NEOHOOKE, NODES, SOLVER_EXPLICIT, SOLVER_VBD,
                   SOLVER_XPBD, STABLE_NEOHOOKE, TETRA, TETRA10, TETRA20, VBD_BASE, VBD_BETTER_COLORATION,
                   VBD_REDUCTION_SYMMETRY, VBD_THREADED_QUADRATURE, XPBD_OPTIGROUP, XPBD_OPTISHARED,
                   XPBD_QUADRATURE, XPBD_XPBD, OEError)

__all__ = [
    "Engine", "Scene", "Batch", "Server", "make_scene", "beam_mesh", "vbd_scene", "xpbd_scene", "explicit_scene", "OEError",
    "host_shape", "host_color_primal", "host_color_dual", "host_fem_constants", "host_masses",
    "TETRA", "HEXA", "TETRA10", "TETRA20", "HEXA27", "NEOHOOKE", "STABLE_NEOHOOKE",
    "MASS_UNIFORM", "MASS_FEMSHAPE", "SOLVER_VBD", "SOLVER_XPBD", "SOLVER_EXPLICIT",
    "VBD_BASE", "VBD_THREADED_QUADRATURE", "VBD_REDUCTION_SYMMETRY", "VBD_BETTER_COLORATION",
    "XPBD_XPBD", "XPBD_QUADRATURE", "XPBD_OPTISHARED", "XPBD_OPTIGROUP", "NODES",
    "VtkWriter", "Recorder", "load_mesh", "flexion_error", "rotation_error",
    "PASS_GATHER", "PASS_CACHE", "PASS_TASKS",
]


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


def _i32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int32)


# ------------------------------------------------------------------ host side
def beam_mesh(element: int, cells=(16, 4, 4), size=(1.0, 0.25, 0.25)):
    """Structured beam (reference BeamMeshGenerator + converters): (xyz (N,3), topology (E,S))."""
    nv, ne = C.c_int32(), C.c_int32()
    L.check(L.lib.oe_host_beam_mesh(element, *cells, *[float(s) for s in size], C.byref(nv), C.byref(ne), None, None))
    xyz = np.empty((nv.value, 3), np.float32)
    topo = np.empty((ne.value, NODES[element]), np.int32)
    L.check(L.lib.oe_host_beam_mesh(element, *cells, *[float(s) for s in size], None, None, L.fptr(xyz), L.iptr(topo)))
    return xyz, topo


def host_shape(element: int):
    S, Q = C.c_int32(), C.c_int32()
    L.check(L.lib.oe_host_shape(element, C.byref(S), C.byref(Q), None, None, None))
    dN = np.empty((Q.value, S.value, 3), np.float64)
    w = np.empty(Q.value, np.float64)
    qp = np.empty((Q.value, 3), np.float64)
    L.check(L.lib.oe_host_shape(element, None, None, L.dptr(dN), L.dptr(w), L.dptr(qp)))
    return dN, w, qp


def host_color_primal(element: int, topo, n_vertices: int, target_color: int = -1):
    topo = _i32(topo)
    colors = np.empty(n_vertices, np.int32)
    nc, nconf = C.c_int32(), C.c_int64()
    L.check(L.lib.oe_host_color_primal(element, L.iptr(topo), topo.shape[0], n_vertices, target_color,
                                       L.iptr(colors), C.byref(nc), C.byref(nconf)))
    return colors, nc.value, nconf.value


def host_color_dual(element: int, topo, n_vertices: int):
    topo = _i32(topo)
    colors = np.empty(topo.shape[0], np.int32)
    nc, nconf = C.c_int32(), C.c_int64()
    L.check(L.lib.oe_host_color_dual(element, L.iptr(topo), topo.shape[0], n_vertices, L.iptr(colors),
                                     C.byref(nc), C.byref(nconf)))
    return colors, nc.value, nconf.value


def host_fem_constants(element: int, xyz, topo):
    xyz, topo = _f32(xyz), _i32(topo)
    Q = host_shape(element)[1].shape[0]
    out = np.empty((topo.shape[0], Q, 10), np.float32)
    L.check(L.lib.oe_host_fem_constants(element, L.fptr(xyz), xyz.shape[0], L.iptr(topo), topo.shape[0], L.fptr(out)))
    return out


def host_affine_records(element: int, xyz, topo):
    """(records (E, 10), weights (Q,)) of an affine mesh, or None when some
    element is curved."""
    xyz, topo = _f32(xyz), _i32(topo)
    Q = host_shape(element)[1].shape[0]
    rec = np.empty((topo.shape[0], 10), np.float32)
    w = np.empty(Q, np.float32)
    flag = C.c_int32()
    L.check(L.lib.oe_host_affine_records(element, L.fptr(xyz), xyz.shape[0], L.iptr(topo), topo.shape[0],
                                         L.fptr(rec), L.fptr(w), C.byref(flag)))
    return (rec, w) if flag.value else None


def host_masses(element: int, xyz, topo, density: float, distribution: int):
    xyz, topo = _f32(xyz), _i32(topo)
    out = np.empty(xyz.shape[0], np.float32)
    L.check(L.lib.oe_host_masses(element, L.fptr(xyz), xyz.shape[0], L.iptr(topo), topo.shape[0], density,
                                 distribution, L.fptr(out)))
    return out


def load_mesh(path):
    """(xyz (N, 3) float32, {element: topology (E, S) int32}) from a .vtk or
    .msh file (the reference's VTK_Loader / Msh_Loader formats)."""
    h = C.c_void_p()
    L.check(L.lib.oe_mesh_load(str(path).encode(), C.byref(h)))
    try:
        nv = C.c_int32()
        counts = np.zeros(10, np.int32)
        L.check(L.lib.oe_mesh_info(h, C.byref(nv), L.iptr(counts)))
        xyz = np.empty((nv.value, 3), np.float32)
        L.check(L.lib.oe_mesh_vertices(h, L.fptr(xyz)))
        topos = {}
        for el in (TETRA, HEXA, TETRA10, TETRA20, HEXA27):
            if counts[el]:
                t = np.empty((int(counts[el]), NODES[el]), np.int32)
                L.check(L.lib.oe_mesh_topology(h, el, L.iptr(t)))
                topos[el] = t
        return xyz, topos
    finally:
        L.lib.oe_mesh_destroy(h)


class VtkWriter:
    """Legacy-VTK 5.1 ASCII writer with the reference's file layout
    (VTK_Formater: open, save_mesh, start_point_data / start_cell_data,
    add_scalar_data / add_vector_data, close)."""

    def __init__(self, path):
        h = C.c_void_p()
        L.check(L.lib.oe_vtk_open(str(path).encode(), C.byref(h)))
        self._h = h

    def mesh(self, element: int, xyz, topo):
        xyz, topo = _f32(xyz), _i32(topo)
        L.check(L.lib.oe_vtk_mesh(self._h, element, L.fptr(xyz), xyz.shape[0], L.iptr(topo), topo.shape[0]))
        return self

    def point_data(self):
        L.check(L.lib.oe_vtk_point_data(self._h))
        return self

    def cell_data(self):
        L.check(L.lib.oe_vtk_cell_data(self._h))
        return self

    def scalars(self, name: str, data):
        d = _f32(data).reshape(-1)
        L.check(L.lib.oe_vtk_scalars(self._h, name.encode(), L.fptr(d), d.size))
        return self

    def vectors(self, name: str, data):
        d = _f32(data).reshape(-1, 3)
        L.check(L.lib.oe_vtk_vectors(self._h, name.encode(), L.fptr(d), d.shape[0]))
        return self

    def close(self):
        if getattr(self, "_h", None):
            h, self._h = self._h, None
            L.check(L.lib.oe_vtk_close(h))


class Recorder:
    """The reference's DataRecorder JSON file ({"name" : value,...}, entries in
    the order added, each written as its recorder writes it).  `save()` writes
    the file and frees the handle; `text()` returns the JSON text."""

    def __init__(self, path):
        h = C.c_void_p()
        L.check(L.lib.oe_record_open(str(path).encode(), C.byref(h)))
        self._h = h

    def time(self, frame: int, time: float, dt: float, fixed_time: float, fixed_dt: float):
        L.check(L.lib.oe_record_time(self._h, frame, time, dt, fixed_time, fixed_dt))
        return self

    def mesh(self, counts: dict, n_vertices: int):
        """counts: {element id: number of elements}"""
        c = np.zeros(10, np.int32)
        for e, n in counts.items():
            c[int(e)] = int(n)
        L.check(L.lib.oe_record_mesh(self._h, L.iptr(c), int(n_vertices)))
        return self

    def material(self, material: int, poisson: float, young: float, sub_iterations: int, density: float,
                 iterations: int | None = None):
        """FEM_Dynamic_Recorder, or XPBD_FEM_Dynamic_Recorder when iterations is given."""
        xpbd = iterations is not None
        L.check(L.lib.oe_record_material(self._h, int(xpbd), material, poisson, young,
                                         int(iterations) if xpbd else 0, sub_iterations, density))
        return self

    def string(self, name: str, value: str):
        L.check(L.lib.oe_record_string(self._h, name.encode(), value.encode()))
        return self

    def flexion(self, error: float, p, target):
        p, target = _f32(p).reshape(3), _f32(target).reshape(3)
        L.check(L.lib.oe_record_flexion(self._h, error, L.fptr(p), L.fptr(target)))
        return self

    def rotation(self, rot_max: float, beam_length: float, dist, angles):
        d, a = _f32(dist).reshape(-1), _f32(angles).reshape(-1)
        if d.size != a.size:
            raise ValueError("dist and angles differ in length")
        L.check(L.lib.oe_record_rotation(self._h, rot_max, beam_length, L.fptr(d), L.fptr(a), d.size))
        return self

    def text(self) -> str:
        n = C.c_int64()
        L.check(L.lib.oe_record_text(self._h, None, 0, C.byref(n)))
        buf = C.create_string_buffer(n.value + 1)
        L.check(L.lib.oe_record_text(self._h, buf, n.value + 1, C.byref(n)))
        return buf.raw[:n.value].decode()

    def save(self):
        if getattr(self, "_h", None):
            h, self._h = self._h, None
            L.check(L.lib.oe_record_close(h))


def flexion_error(rest, pos, follow, target):
    """(error, p, vertex) of FEM_Flexion_error_recorder; vertex -1 when no rest
    position is within 1e-3 of follow."""
    rest, pos = _f32(rest).reshape(-1, 3), _f32(pos).reshape(-1, 3)
    f, t = _f32(follow).reshape(3), _f32(target).reshape(3)
    err, p, v = C.c_float(), np.zeros(3, np.float32), C.c_int32()
    L.check(L.lib.oe_flexion_error(L.fptr(rest), L.fptr(pos), rest.shape[0], L.fptr(f), L.fptr(t), C.byref(err),
                                   L.fptr(p), C.byref(v)))
    return float(err.value), p, int(v.value)


def rotation_error(element: int, topo, rest, pos, samples_per_edge: int, beam_length: float):
    """(dist, angles) of FEM_Torsion_error_recorder.compute_errors."""
    topo = _i32(topo)
    rest, pos = _f32(rest).reshape(-1, 3), _f32(pos).reshape(-1, 3)
    cap = max(1, topo.shape[0] * max(2, int(samples_per_edge)) * 4)
    d, a, n = np.zeros(cap, np.float32), np.zeros(cap, np.float32), C.c_int32(cap)
    L.check(L.lib.oe_rotation_error(element, L.iptr(topo), topo.shape[0], L.fptr(rest), L.fptr(pos), rest.shape[0],
                                    int(samples_per_edge), beam_length, L.fptr(d), L.fptr(a), C.byref(n)))
    return d[:n.value].copy(), a[:n.value].copy()


# VBD colour-pass modes (oe_scene_set_tile)
PASS_GATHER, PASS_CACHE, PASS_TASKS = 0, 1, 2


# ------------------------------------------------------------------ device side
class Engine:
    """One GPU, one stream, one device-memory budget (bytes, 0 = none)."""

    def __init__(self, device: int = 0, budget_bytes: int = 0):
        h = C.c_void_p()
        L.check(L.lib.oe_engine_create(device, budget_bytes, C.byref(h)))
        self._h = h

    @property
    def handle(self):
        return self._h

    def set_autotune(self, enable: bool):
        """Tune the layout of every VBD / explicit scene created on this engine."""
        L.check(L.lib.oe_engine_set_autotune(self._h, int(enable)))

    def synchronize(self):
        L.check(L.lib.oe_engine_synchronize(self._h))

    def memory(self):
        used, budget = C.c_uint64(), C.c_uint64()
        L.check(L.lib.oe_engine_memory(self._h, C.byref(used), C.byref(budget)))
        return used.value, budget.value

    def close(self):
        if getattr(self, "_h", None):
            L.lib.oe_engine_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Scene:
    """Mesh + particle system + integrator (+ FEM dynamic + constraints)."""

    def __init__(self, engine: Engine, element: int, xyz, topo, *, material=STABLE_NEOHOOKE, young=1e6,
                 poisson=0.45, density=1000.0, distribution=MASS_FEMSHAPE, solver=SOLVER_VBD, iterations=10,
                 rho=0.0, damping=0.0, global_damping=0.0, vbd_version=VBD_REDUCTION_SYMMETRY, target_color=-1,
                 xpbd_version=XPBD_OPTISHARED, gravity=(0.0, -9.81, 0.0), shuffle_seed=0, sub_iterations=1):
        self.engine = engine
        self._xyz = _f32(xyz).reshape(-1, 3)
        self._topo = _i32(topo).reshape(-1, NODES[element])
        d = L.SceneDesc()
        L.lib.oe_scene_desc_default(C.byref(d))
        d.element = element
        d.n_vertices = self._xyz.shape[0]
        d.xyz = L.fptr(self._xyz)
        d.n_elements = self._topo.shape[0]
        d.topology = L.iptr(self._topo)
        d.material = material
        d.young, d.poisson, d.density = young, poisson, density
        d.mass_distribution = distribution
        d.solver = solver
        d.iterations = iterations
        d.rho, d.damping, d.global_damping = rho, damping, global_damping
        d.vbd_version, d.target_color, d.xpbd_version = vbd_version, target_color, xpbd_version
        d.gravity[:] = list(gravity)
        d.shuffle_seed = shuffle_seed
        h = C.c_void_p()
        L.check(L.lib.oe_scene_create(engine.handle, C.byref(d), C.byref(h)))
        self._h = h
        self.element = element
        self.sub_iterations = sub_iterations
        nv, ne, S, Q, nc, nb = (C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32(), C.c_uint64())
        L.check(L.lib.oe_scene_info(h, C.byref(nv), C.byref(ne), C.byref(S), C.byref(Q), C.byref(nc), C.byref(nb)))
        self.n_vertices, self.n_elements = nv.value, ne.value
        self.S, self.Q, self.n_colors, self.device_bytes = S.value, Q.value, nc.value, nb.value

    @property
    def handle(self):
        return self._h

    # constraints
    def fix_plane(self, origin, normal) -> int:
        n = C.c_int32()
        L.check(L.lib.oe_scene_fix_plane(self._h, L.fptr(_f32(origin)), L.fptr(_f32(normal)), C.byref(n)))
        return n.value

    def fix_vertices(self, ids):
        ids = _i32(ids)
        L.check(L.lib.oe_scene_fix_vertices(self._h, L.iptr(ids), ids.size))

    def box_limit(self, pmin=(-10.0, -5.0, -10.0), pmax=(10.0, 5.0, 10.0)):
        L.check(L.lib.oe_scene_box_limit(self._h, L.fptr(_f32(pmin)), L.fptr(_f32(pmax))))

    # stepping
    def step(self, dt: float, substeps: int = 1):
        L.check(L.lib.oe_scene_step(self._h, dt, substeps))

    def update(self, dt: float):
        self.step(dt, self.sub_iterations)

    def set_graph(self, enable: bool):
        L.check(L.lib.oe_scene_set_graph(self._h, int(enable)))

    def set_tile(self, lanes: int, qpl: int, mode: int = PASS_GATHER):
        """VBD colour-pass layout: lanes per vertex, quadrature points per lane
        and mode (PASS_GATHER; PASS_CACHE: cached gradients with qpl = 1;
        PASS_TASKS: element tasks, which ignore lanes and qpl)."""
        L.check(L.lib.oe_scene_set_tile(self._h, lanes, qpl, int(mode)))

    def set_tile_ex(self, lanes=64, qpl=1, mode=PASS_GATHER, inc_per_block=0, latency_blocks=512):
        t = L.Tile(lanes, qpl, int(mode), inc_per_block, latency_blocks)
        L.check(L.lib.oe_scene_set_tile_ex(self._h, C.byref(t)))

    def get_tile_ex(self) -> dict:
        t = L.Tile()
        L.check(L.lib.oe_scene_get_tile_ex(self._h, C.byref(t)))
        return t.as_dict()

    def autotune(self, force: bool = False) -> dict:
        """Tune (or take the cached) colour-pass layout; returns it with the
        time per sweep and the number of layouts timed (0 = from the cache)."""
        t, ms, n = L.Tile(), C.c_double(), C.c_int32()
        L.check(L.lib.oe_scene_autotune(self._h, int(force), C.byref(t), C.byref(ms), C.byref(n)))
        out = t.as_dict()
        out.update(sweep_ms=ms.value, candidates=n.value)
        return out

    def get_tile(self):
        """(lanes, quadrature points per lane, mode)."""
        a, b, c = C.c_int32(), C.c_int32(), C.c_int32()
        L.check(L.lib.oe_scene_get_tile(self._h, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def element_diagnostics(self):
        """dict of per-element stress (von Mises x V), volume, volume_diff, inverted."""
        out = {k: np.empty(self.n_elements, np.float32) for k in ("stress", "volume", "volume_diff", "inverted")}
        L.check(L.lib.oe_scene_element_diagnostics(self._h, *(L.fptr(out[k]) for k in
                                                              ("stress", "volume", "volume_diff", "inverted"))))
        return out

    def residual(self, dt: float):
        """Residual force of every free vertex's local VBD problem, (N, 3)."""
        r = np.empty((self.n_vertices, 3), np.float32)
        L.check(L.lib.oe_scene_residual(self._h, dt, L.fptr(r)))
        return r

    def residual_scratch(self, dt: float):
        """The same residual through the allocation-free entry point (static device scratch)."""
        r = np.empty((self.n_vertices, 3), np.float32)
        L.check(L.lib.oe_scene_residual_scratch(self._h, dt, L.fptr(r)))
        return r

    def save_vtk(self, path):
        """Rest geometry + point data "weights" (masses) and "u" (displacement)."""
        L.check(L.lib.oe_scene_save_vtk(self._h, str(path).encode()))

    def affine(self) -> bool:
        """True when the element records are stored once per element."""
        a = C.c_int32()
        L.check(L.lib.oe_scene_get_layout(self._h, C.byref(a)))
        return bool(a.value)

    def set_timing(self, enable: bool):
        L.check(L.lib.oe_scene_set_timing(self._h, int(enable)))

    def get_timing(self):
        ms, n = C.c_double(), C.c_int64()
        L.check(L.lib.oe_scene_get_timing(self._h, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    # read-back
    def _get3(self, fn):
        out = np.empty((self.n_vertices, 3), np.float32)
        L.check(fn(self._h, L.fptr(out)))
        return out

    def positions(self):
        return self._get3(L.lib.oe_scene_get_positions)

    def last_positions(self):
        return self._get3(L.lib.oe_scene_get_last_positions)

    def init_positions(self):
        return self._get3(L.lib.oe_scene_get_init_positions)

    def velocities(self):
        return self._get3(L.lib.oe_scene_get_velocities)

    def forces(self):
        return self._get3(L.lib.oe_scene_get_forces)

    def masses(self):
        out = np.empty(self.n_vertices, np.float32)
        L.check(L.lib.oe_scene_get_masses(self._h, L.fptr(out)))
        return out

    def masks(self):
        out = np.empty(self.n_vertices, np.int32)
        L.check(L.lib.oe_scene_get_masks(self._h, L.iptr(out)))
        return out

    def colors(self):
        n = self.n_elements if self._is_xpbd else self.n_vertices
        out = np.empty(n, np.int32)
        L.check(L.lib.oe_scene_get_colors(self._h, L.iptr(out)))
        return out

    @property
    def _is_xpbd(self):
        return getattr(self, "_solver", None) == SOLVER_XPBD

    def set_state(self, p=None, last_p=None, v=None):
        args = [L.fptr(_f32(a)) if a is not None else None for a in (p, last_p, v)]
        L.check(L.lib.oe_scene_set_state(self._h, *args))

    def close(self):
        if getattr(self, "_h", None):
            L.lib.oe_scene_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Many VBD scenes stepped together (one launch per global colour and
    element type).  The scenes must stay alive while the batch exists."""

    def __init__(self, engine: Engine, scenes):
        self.engine = engine
        self.scenes = list(scenes)
        arr = (C.c_void_p * len(self.scenes))(*[s.handle for s in self.scenes])
        h = C.c_void_p()
        L.check(L.lib.oe_batch_create(engine.handle, arr, len(self.scenes), C.byref(h)))
        self._h = h

    def step(self, dt: float, substeps: int = 1):
        L.check(L.lib.oe_batch_step(self._h, dt, substeps))

    def set_graph(self, enable: bool):
        L.check(L.lib.oe_batch_set_graph(self._h, int(enable)))

    def set_streams(self, n: int):
        """1: every colour pass on the engine's stream; 2 (default): Q2 passes on a second stream."""
        L.check(L.lib.oe_batch_set_streams(self._h, n))

    def set_tile(self, element: int, lanes: int, qpl: int, mode: int = PASS_GATHER):
        L.check(L.lib.oe_batch_set_tile(self._h, element, lanes, qpl, int(mode)))

    def set_timing(self, enable: bool):
        L.check(L.lib.oe_batch_set_timing(self._h, int(enable)))

    def get_timing(self):
        ms, n = C.c_double(), C.c_int64()
        L.check(L.lib.oe_batch_get_timing(self._h, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def info(self):
        nl, vu, gb = C.c_int32(), C.c_int64(), C.c_double()
        L.check(L.lib.oe_batch_info(self._h, C.byref(nl), C.byref(vu), C.byref(gb)))
        return {"gather_launches": nl.value, "vertex_updates": vu.value, "gather_bytes": gb.value}

    def set_tile_ex(self, element: int, lanes=64, qpl=1, mode=PASS_GATHER, inc_per_block=0, latency_blocks=512):
        t = L.Tile(lanes, qpl, int(mode), inc_per_block, latency_blocks)
        L.check(L.lib.oe_batch_set_tile_ex(self._h, element, C.byref(t)))

    def get_tile_ex(self, element: int) -> dict:
        t = L.Tile()
        L.check(L.lib.oe_batch_get_tile_ex(self._h, element, C.byref(t)))
        return t.as_dict()

    def autotune(self, force: bool = False):
        L.check(L.lib.oe_batch_autotune(self._h, int(force)))

    def get_timing_element(self, element: int):
        """Event-timed colour-pass milliseconds and launches of one element type since its last read."""
        ms, n = C.c_double(), C.c_int64()
        L.check(L.lib.oe_batch_get_timing_element(self._h, element, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def info_element(self, element: int):
        nl, vu, gb = C.c_int32(), C.c_int64(), C.c_double()
        L.check(L.lib.oe_batch_info_element(self._h, element, C.byref(nl), C.byref(vu), C.byref(gb)))
        return {"gather_launches": nl.value, "vertex_updates": vu.value, "gather_bytes": gb.value}

    def close(self):
        if getattr(self, "_h", None):
            L.lib.oe_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching request loop over resident VBD scenes."""

    def __init__(self, engine: Engine, max_batch: int = 64, streams: int = 2):
        self.engine = engine
        h = C.c_void_p()
        L.check(L.lib.oe_server_create(engine.handle, max_batch, streams, C.byref(h)))
        self._h = h
        self._scenes = []

    def add_scene(self, scene: Scene) -> int:
        slot = C.c_int32()
        L.check(L.lib.oe_server_add_scene(self._h, scene.handle, C.byref(slot)))
        self._scenes.append(scene)
        return slot.value

    def submit(self, slot: int, dt: float, substeps: int = 1) -> int:
        t = C.c_uint64()
        L.check(L.lib.oe_server_submit(self._h, slot, dt, substeps, C.byref(t)))
        return t.value

    def set_tile_ex(self, element: int, lanes=64, qpl=1, mode=PASS_TASKS, inc_per_block=0, latency_blocks=512):
        """Layout of one element type's passes in every tick (refused while requests are in flight)."""
        t = L.Tile(lanes, qpl, int(mode), inc_per_block, latency_blocks)
        L.check(L.lib.oe_server_set_tile_ex(self._h, element, C.byref(t)))

    def get_tile_ex(self, element: int) -> dict:
        t = L.Tile()
        L.check(L.lib.oe_server_get_tile_ex(self._h, element, C.byref(t)))
        return t.as_dict()

    def drain(self):
        L.check(L.lib.oe_server_drain(self._h))

    def wait(self, ticket: int):
        """Block until the request with this ticket is complete."""
        L.check(L.lib.oe_server_wait(self._h, ticket))

    def poll(self, ticket: int) -> bool:
        done = C.c_int32()
        L.check(L.lib.oe_server_poll(self._h, ticket, C.byref(done)))
        return bool(done.value)

    def stats(self):
        c, vu, nb = C.c_int64(), C.c_int64(), C.c_int64()
        p50, p99, mx, mb = C.c_double(), C.c_double(), C.c_double(), C.c_double()
        L.check(L.lib.oe_server_stats(self._h, C.byref(c), C.byref(p50), C.byref(p99), C.byref(mx), C.byref(vu),
                                      C.byref(nb), C.byref(mb)))
        return {"completed": c.value, "p50_ms": p50.value, "p99_ms": p99.value, "max_ms": mx.value,
                "vertex_updates": vu.value, "batches": nb.value, "mean_batch": mb.value}

    def latencies(self):
        n = C.c_int64()
        L.check(L.lib.oe_server_latencies(self._h, None, 0, C.byref(n)))
        out = np.empty(n.value, np.float64)
        L.check(L.lib.oe_server_latencies(self._h, L.dptr(out), n.value, C.byref(n)))
        return out

    def latency_slots(self):
        """Slot of each request, in the order of latencies()."""
        n = C.c_int64()
        L.check(L.lib.oe_server_latency_slots(self._h, None, 0, C.byref(n)))
        out = np.empty(n.value, np.int32)
        L.check(L.lib.oe_server_latency_slots(self._h, L.iptr(out), n.value, C.byref(n)))
        return out

    def close(self):
        if getattr(self, "_h", None):
            L.lib.oe_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def make_scene(engine, element, cells, size=(1.0, 0.25, 0.25), *, young=5e5, poisson=0.49,
               material=STABLE_NEOHOOKE, iterations=10, fix_x0=True, box=True, **kw):
    """A beam scene as the benchmarks use it: x = 0 face fixed (Plane at x = 1e-6
    facing -x), box limit (-10,-5,-10)..(10,5,10)."""
    xyz, topo = beam_mesh(element, cells, size)
    s = Scene(engine, element, xyz, topo, material=material, young=young, poisson=poisson, iterations=iterations,
              **kw)
    s._solver = kw.get("solver", SOLVER_VBD)
    if fix_x0:
        s.fix_plane((1e-6, 0.0, 0.0), (-1.0, 0.0, 0.0))
    if box:
        s.box_limit((-10.0, -5.0, -10.0), (10.0, 5.0, 10.0))
    return s


def vbd_scene(engine, element, xyz, topo, density, distrib, young, poisson, material, iteration=30,
              sub_iteration=1, damping=0.0, rho=0.0, version=VBD_REDUCTION_SYMMETRY, target_color=-1, **kw):
    """Cuda_VBD_FEM_Dynamic equivalent.  The reference's default version
    (Better_Coloration with target_color=-1) is rejected there too; the default
    here is Reduction_Symmetry."""
    s = Scene(engine, element, xyz, topo, material=material, young=young, poisson=poisson, density=density,
              distribution=distrib, solver=SOLVER_VBD, iterations=iteration, rho=rho, damping=damping,
              vbd_version=version, target_color=target_color, sub_iterations=sub_iteration, **kw)
    s._solver = SOLVER_VBD
    return s


def xpbd_scene(engine, element, xyz, topo, density, distrib, young, poisson, material, sub_it=30, damping=0.0,
               version=XPBD_OPTISHARED, **kw):
    """Cuda_XPBD_FEM_Dynamic equivalent (GPU_PBD(global damping) + GPU_PBD_FEM)."""
    s = Scene(engine, element, xyz, topo, material=material, young=young, poisson=poisson, density=density,
              distribution=distrib, solver=SOLVER_XPBD, global_damping=damping, xpbd_version=version,
              sub_iterations=sub_it, **kw)
    s._solver = SOLVER_XPBD
    return s


def explicit_scene(engine, element, xyz, topo, sub_it, density, distrib, young, poisson, material, damping=0.0,
                   **kw):
    """Cuda_FEM_Explicit_Dynamic equivalent with the semi-implicit Euler integrator."""
    s = Scene(engine, element, xyz, topo, material=material, young=young, poisson=poisson, density=density,
              distribution=distrib, solver=SOLVER_EXPLICIT, damping=damping, sub_iterations=sub_it, **kw)
    s._solver = SOLVER_EXPLICIT
    return s


def autotune_cache():
    """The autotuner's cache: one dict per (device, element, size bucket, form)."""
    n = C.c_int32(0)
    L.check(L.lib.oe_autotune_cache(C.byref(n), None, None, None, None, None, None, None))
    cap = n.value
    dev, el, bk, fm, cand = (np.zeros(cap, np.int32) for _ in range(5))
    ms = np.zeros(cap, np.float64)
    tiles = (L.Tile * max(cap, 1))()
    n = C.c_int32(cap)
    L.check(L.lib.oe_autotune_cache(C.byref(n), L.iptr(dev), L.iptr(el), L.iptr(bk), L.iptr(fm), tiles,
                                    ms.ctypes.data_as(C.POINTER(C.c_double)), L.iptr(cand)))
    out = []
    for i in range(min(cap, n.value)):
        d = tiles[i].as_dict()
        d.update(device=int(dev[i]), element=int(el[i]), bucket=int(bk[i]), form="batch" if fm[i] else "scene",
                 ms=float(ms[i]), candidates=int(cand[i]))
        out.append(d)
    return out


def autotune_clear():
    L.check(L.lib.oe_autotune_clear())
