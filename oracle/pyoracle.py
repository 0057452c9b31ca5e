"""ctypes binding of the FP64 CPU oracle (oracle/build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker.  The engine never imports it.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "build" / "liboracle.so"


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def _load():
    if not LIB_PATH.exists():
        build()
    return C.CDLL(str(LIB_PATH))


lib = _load()

TETRA, HEXA, TETRA10, TETRA20, HEXA27 = 3, 6, 7, 8, 9
NEOHOOKE, STABLE_NEOHOOKE = 2, 3
VBD, XPBD, EXPLICIT = 0, 1, 2
NODES = {TETRA: 4, HEXA: 8, TETRA10: 10, TETRA20: 20, HEXA27: 27}

f32p = C.POINTER(C.c_float)
f64p = C.POINTER(C.c_double)
i32p = C.POINTER(C.c_int32)
P = C.c_void_p

lib.orc_nodes.argtypes = [C.c_int]
lib.orc_quadrature.argtypes = [C.c_int]
lib.orc_shape_eval.argtypes = [C.c_int, f64p, f64p, f64p]
lib.orc_quadrature_rule.argtypes = [C.c_int, f64p, f64p]
lib.orc_beam.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float, i32p, i32p, f32p, i32p]
lib.orc_color_dsat.argtypes = [C.c_int, i32p, C.c_int, C.c_int, C.c_int, i32p]
lib.orc_color_slf.argtypes = [C.c_int, i32p, C.c_int, C.c_int, i32p]
lib.orc_primal_conflicts.argtypes = [C.c_int, i32p, C.c_int, C.c_int, i32p]
lib.orc_primal_conflicts.restype = C.c_int64
lib.orc_fem_constants.argtypes = [C.c_int, f32p, C.c_int, i32p, C.c_int, f64p, f64p]
lib.orc_masses.argtypes = [C.c_int, f32p, C.c_int, i32p, C.c_int, C.c_double, C.c_int, f64p]
lib.orc_vbd_material.argtypes = [C.c_int, C.c_double, C.c_double, f64p, f64p, C.c_double, f64p, f64p]
lib.orc_xpbd_material.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, f64p, f64p, f64p]
lib.orc_scene_create.restype = P
lib.orc_scene_create.argtypes = [C.c_int, C.c_int, f32p, C.c_int, i32p, C.c_int, C.c_double, C.c_double, C.c_double,
                                 C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, f64p]
lib.orc_scene_destroy.argtypes = [P]
lib.orc_scene_fix.argtypes = [P, i32p, C.c_int]
lib.orc_scene_box.argtypes = [P, f64p, f64p]
lib.orc_scene_color_order.argtypes = [P, i32p]
lib.orc_scene_vbd_options.argtypes = [P, C.c_double, C.c_int, C.c_int]
lib.orc_scene_diagnostics.argtypes = [P, f64p, f64p, f64p, f64p]
lib.orc_scene_residual.argtypes = [P, C.c_double, f64p]
lib.orc_scene_step.argtypes = [P, C.c_double, C.c_int]
lib.orc_scene_get.argtypes = [P, C.c_int, f64p]
lib.orc_scene_set.argtypes = [P, C.c_int, f64p]
lib.orc_scene_ncolors.argtypes = [P]
lib.orc_scene_colors.argtypes = [P, i32p]
for sfx in ("_f", "_d"):
    cr = getattr(lib, "orc_cpu_xpbd_create" + sfx)
    cr.restype = P
    cr.argtypes = [C.c_int, C.c_int, f32p, C.c_int, i32p, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int,
                   C.c_int, C.c_int, C.c_double, i32p, C.c_int, f64p]
    getattr(lib, "orc_cpu_xpbd_step" + sfx).argtypes = [P, C.c_double]
    getattr(lib, "orc_cpu_xpbd_positions" + sfx).argtypes = [P, f64p]
    getattr(lib, "orc_cpu_xpbd_destroy" + sfx).argtypes = [P]
    cr = getattr(lib, "orc_cpu_vbd_create" + sfx)
    cr.restype = P
    cr.argtypes = [C.c_int, C.c_int, f32p, C.c_int, i32p, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int,
                   C.c_int, C.c_int, C.c_double, C.c_double, i32p, C.c_int, f64p]
    getattr(lib, "orc_cpu_vbd_step" + sfx).argtypes = [P, C.c_double]
    getattr(lib, "orc_cpu_vbd_positions" + sfx).argtypes = [P, f64p]
    getattr(lib, "orc_cpu_vbd_destroy" + sfx).argtypes = [P]


def _f32(a):
    return np.ascontiguousarray(a, np.float32)


def _i32(a):
    return np.ascontiguousarray(a, np.int32)


def _f64(a):
    return np.ascontiguousarray(a, np.float64)


def fp(a):
    return a.ctypes.data_as(f32p)


def dp(a):
    return a.ctypes.data_as(f64p)


def ip(a):
    return a.ctypes.data_as(i32p)


def shape_eval(element, x):
    S = NODES[element]
    N = np.empty(S)
    dN = np.empty((S, 3))
    lib.orc_shape_eval(element, dp(_f64(x)), dp(N), dp(dN))
    return N, dN


def quadrature(element):
    Q = lib.orc_quadrature(element)
    qp = np.empty((Q, 3))
    w = np.empty(Q)
    lib.orc_quadrature_rule(element, dp(qp), dp(w))
    return qp, w


def dN_table(element):
    qp, _ = quadrature(element)
    return np.stack([shape_eval(element, x)[1] for x in qp])


def beam(element, cells=(16, 4, 4), size=(1.0, 0.25, 0.25)):
    nv, ne = C.c_int(), C.c_int()
    lib.orc_beam(element, *cells, *size, C.byref(nv), C.byref(ne), None, None)
    xyz = np.empty((nv.value, 3), np.float32)
    topo = np.empty((ne.value, NODES[element]), np.int32)
    lib.orc_beam(element, *cells, *size, None, None, fp(xyz), ip(topo))
    return xyz, topo


def color_dsat(element, topo, nv, target=-1):
    topo = _i32(topo)
    c = np.empty(nv, np.int32)
    nc = lib.orc_color_dsat(element, ip(topo), topo.shape[0], nv, target, ip(c))
    return c, nc


def color_slf(element, topo, nv):
    topo = _i32(topo)
    c = np.empty(topo.shape[0], np.int32)
    nc = lib.orc_color_slf(element, ip(topo), topo.shape[0], nv, ip(c))
    return c, nc


def primal_conflicts(element, topo, nv, colors):
    topo = _i32(topo)
    return lib.orc_primal_conflicts(element, ip(topo), topo.shape[0], nv, ip(_i32(colors)))


def fem_constants(element, xyz, topo):
    xyz, topo = _f32(xyz), _i32(topo)
    Q = lib.orc_quadrature(element)
    ji = np.empty((topo.shape[0], Q, 9))
    V = np.empty((topo.shape[0], Q))
    lib.orc_fem_constants(element, fp(xyz), xyz.shape[0], ip(topo), topo.shape[0], dp(ji), dp(V))
    return ji, V


def masses(element, xyz, topo, density, femshape=True):
    xyz, topo = _f32(xyz), _i32(topo)
    m = np.empty(xyz.shape[0])
    lib.orc_masses(element, fp(xyz), xyz.shape[0], ip(topo), topo.shape[0], density, int(femshape), dp(m))
    return m


def vbd_material(material, lam, mu, F, g, V):
    """P (3x3, row/col as math) and the per-vertex Hessian block for dF/dx = g."""
    Fc = _f64(np.asarray(F, np.float64).T.reshape(-1))   # column-major
    P_ = np.empty(9)
    H = np.empty(9)
    lib.orc_vbd_material(material, lam, mu, dp(Fc), dp(_f64(g)), V, dp(P_), dp(H))
    return P_.reshape(3, 3).T, H.reshape(3, 3).T


def xpbd_material(material, m, lam, mu, F):
    Fc = _f64(np.asarray(F, np.float64).T.reshape(-1))
    P_ = np.empty(9)
    e = C.c_double()
    lib.orc_xpbd_material(material, m, lam, mu, dp(Fc), dp(P_), C.byref(e))
    return P_.reshape(3, 3).T, e.value


class Scene:
    """FP64 oracle scene with the same parameters as the engine's Scene."""

    def __init__(self, element, xyz, topo, *, material=STABLE_NEOHOOKE, young=1e6, poisson=0.45, density=1000.0,
                 femshape=True, solver=VBD, iterations=10, damping=0.0, global_damping=0.0, xpbd_version=2,
                 gravity=(0.0, -9.81, 0.0)):
        self._xyz, self._topo = _f32(xyz), _i32(topo)
        self.N = self._xyz.shape[0]
        self.E = self._topo.shape[0]
        self.solver = solver
        g = _f64(gravity)
        self._h = lib.orc_scene_create(element, self.N, fp(self._xyz), self.E, ip(self._topo), material, young,
                                       poisson, density, int(femshape), solver, iterations, damping, global_damping,
                                       xpbd_version, dp(g))
        if not self._h:
            raise ValueError("oracle scene creation failed")

    def fix(self, ids):
        ids = _i32(ids)
        lib.orc_scene_fix(self._h, ip(ids), ids.size)

    def box(self, pmin=(-10, -5, -10), pmax=(10, 5, 10)):
        lib.orc_scene_box(self._h, dp(_f64(pmin)), dp(_f64(pmax)))

    def vbd_options(self, rho=0.0, jacobi=False, target_color=-1):
        """Chebyshev radius, Jacobi (v4) colour sweep, colour reduction target."""
        n = lib.orc_scene_vbd_options(self._h, rho, int(jacobi), target_color)
        if n < 0:
            raise ValueError("VBD options on a non-VBD scene")
        return n

    def element_diagnostics(self):
        out = {k: np.empty(self.E) for k in ("stress", "volume", "volume_diff", "inverted")}
        lib.orc_scene_diagnostics(self._h, *(dp(out[k]) for k in ("stress", "volume", "volume_diff", "inverted")))
        return out

    def residual(self, dt):
        r = np.empty((self.N, 3))
        lib.orc_scene_residual(self._h, dt, dp(r))
        return r

    def color_order(self, order):
        if order is None:
            lib.orc_scene_color_order(self._h, None)
        else:
            o = _i32(order)
            lib.orc_scene_color_order(self._h, ip(o))

    def step(self, dt, substeps=1):
        lib.orc_scene_step(self._h, dt, substeps)

    def get(self, what):
        idx = {"p": 0, "last_p": 1, "v": 2, "f": 3, "y": 4, "init_p": 5, "m": 6, "mask": 7}[what]
        out = np.empty(self.N if idx >= 6 else (self.N, 3))
        lib.orc_scene_get(self._h, idx, dp(out))
        return out

    def set(self, what, arr):
        idx = {"p": 0, "last_p": 1, "v": 2, "f": 3, "y": 4, "init_p": 5}[what]
        lib.orc_scene_set(self._h, idx, dp(_f64(arr)))

    def positions(self):
        return self.get("p")

    @property
    def ncolors(self):
        return lib.orc_scene_ncolors(self._h)

    def colors(self):
        out = np.empty(self.E if self.solver == XPBD else self.N, np.int32)
        lib.orc_scene_colors(self._h, ip(out))
        return out

    def __del__(self):
        try:
            lib.orc_scene_destroy(self._h)
        except Exception:
            pass


class CpuXPBD:
    """Reference CPU XPBD path (float = timed baseline, double = its FP64 twin)."""

    def __init__(self, element, xyz, topo, fixed, *, material=NEOHOOKE, young=1e6, poisson=0.45, density=1000.0,
                 femshape=True, iterations=1, substeps=20, global_damping=0.0, gravity=(0.0, -9.81, 0.0),
                 precision="f"):
        self._sfx = "_f" if precision == "f" else "_d"
        self._xyz, self._topo, self._fix = _f32(xyz), _i32(topo), _i32(fixed)
        self.N = self._xyz.shape[0]
        create = getattr(lib, "orc_cpu_xpbd_create" + self._sfx)
        self._h = create(element, self.N, fp(self._xyz), self._topo.shape[0], ip(self._topo), material, young, poisson,
                         density, int(femshape), iterations, substeps, global_damping, ip(self._fix), self._fix.size,
                         dp(_f64(gravity)))

    def step(self, dt):
        getattr(lib, "orc_cpu_xpbd_step" + self._sfx)(self._h, dt)

    def positions(self):
        out = np.empty((self.N, 3))
        getattr(lib, "orc_cpu_xpbd_positions" + self._sfx)(self._h, dp(out))
        return out

    def __del__(self):
        try:
            getattr(lib, "orc_cpu_xpbd_destroy" + self._sfx)(self._h)
        except Exception:
            pass


class CpuVBD:
    """Reference CPU VBD path: sequential Gauss-Seidel over the vertices in id
    order (float = timed baseline, double = its FP64 twin)."""

    def __init__(self, element, xyz, topo, fixed, *, material=STABLE_NEOHOOKE, young=5e5, poisson=0.49,
                 density=1000.0, femshape=True, iterations=10, substeps=1, damping=0.0, rho=0.0,
                 gravity=(0.0, -9.81, 0.0), precision="f"):
        self._sfx = "_f" if precision == "f" else "_d"
        self._xyz, self._topo, self._fix = _f32(xyz), _i32(topo), _i32(fixed)
        self.N = self._xyz.shape[0]
        create = getattr(lib, "orc_cpu_vbd_create" + self._sfx)
        self._h = create(element, self.N, fp(self._xyz), self._topo.shape[0], ip(self._topo), material, young, poisson,
                         density, int(femshape), iterations, substeps, damping, rho, ip(self._fix), self._fix.size,
                         dp(_f64(gravity)))
        if not self._h:
            raise ValueError("CPU VBD supports Stable Neo-Hookean only")

    def step(self, dt):
        getattr(lib, "orc_cpu_vbd_step" + self._sfx)(self._h, dt)

    def positions(self):
        out = np.empty((self.N, 3))
        getattr(lib, "orc_cpu_vbd_positions" + self._sfx)(self._h, dp(out))
        return out

    def __del__(self):
        try:
            getattr(lib, "orc_cpu_vbd_destroy" + self._sfx)(self._h)
        except Exception:
            pass
