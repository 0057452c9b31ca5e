"""Pins of the oracle's integrator frames, damping branches and the restated
reference CPU paths (CPU only).  Each expected value comes from a closed form
or from finite differences of the energies, never from the oracle's own
formulas, so that a dropped term, a wrong sign or a swapped order fails.

Reference (octopus/):
  XPBD velocity update with global damping   src/GPU/PBD/GPU_PBD.cu:8-18
  semi-implicit frame (XPBD, explicit)       src/GPU/GPU_Integrator.cu:27-34
  VBD damping branch                         src/GPU/VBD/GPU_VBD_FEM.cu:107-112
  explicit damping branch                    src/GPU/Explicit/GPU_FEM_Explicit.cu:49-53
  XPBD V0 / V2 projections                   src/GPU/PBD/GPU_PBD_FEM.cu:13-93, :206-263
  CPU XPBD path                              src/Dynamic/PBD/PositionBasedDynamic.cpp:7-32,76-109,
                                             src/Dynamic/PBD/XPBD_Constraint.cpp:4-28,
                                             src/Dynamic/PBD/XPBD_FEM_Generic.cpp:9-45,
                                             src/Dynamic/Base/Constraint.cpp:20-38
  CPU VBD path                               src/Dynamic/VBD/VertexBlockDescent.cpp:12-51,
                                             src/Dynamic/VBD/VBD_FEM.cpp:13-111,241-254
"""
import numpy as np
import pytest

import pyoracle as O

G = np.array([0.0, -9.81, 0.0])


# --------------------------------------------------------------- helpers
def _tet_energy(kind, lam, mu, m, F):
    """XPBD constraint energies per unit volume (GPU_PBD_FEM_Materials.cu:97-118;
    PBD_ContinuousMaterial.cpp:56-74,176-196), written from their definitions."""
    J = np.linalg.det(F)
    n2 = float(np.sum(F * F))
    if kind == O.STABLE_NEOHOOKE:
        return (J - 1.0) ** 2 if m == 0 else n2 - 3.0 - 2.0 * (J - 1.0)
    alpha = 1.0 + mu / lam
    return (J - alpha) ** 2 if m == 0 else n2 - 3.0


def _p1_energy(kind, lam, mu, m, X, x):
    """E_m = V e_m(F) of one linear tet, F = Ds Dm^-1, V = |det Dm| / 6."""
    Dm = (X[1:] - X[0]).T
    Ds = (x[1:] - x[0]).T
    F = Ds @ np.linalg.inv(Dm)
    return abs(np.linalg.det(Dm)) / 6.0 * _tet_energy(kind, lam, mu, m, F)


def _fd_grad(fun, x, h=1e-5):
    """Fourth-order central differences."""
    g = np.zeros_like(x)
    for idx in np.ndindex(*x.shape):
        vals = []
        for step in (2, 1, -1, -2):
            y = x.copy()
            y[idx] += step * h
            vals.append(fun(y))
        g[idx] = (-vals[0] + 8 * vals[1] - 8 * vals[2] + vals[3]) / (12 * h)
    return g


def _lame(young, nu):
    return young * nu / ((1 + nu) * (1 - 2 * nu)), young / (2 * (1 + nu))


def _vbd_energy(lam, mu, F):
    """Stable Neo-Hookean W = mu/2 (|F|^2 - 3) + lam/2 (J - alpha)^2
    (src/Dynamic/FEM/FEM_ContinuousMaterial.cpp:83-87)."""
    alpha = 1.0 + mu / lam
    J = np.linalg.det(F)
    return 0.5 * mu * (np.sum(F * F) - 3.0) + 0.5 * lam * (J - alpha) ** 2


def _mesh_energy(el, xyz, topo, P, lam, mu):
    ji, V = O.fem_constants(el, xyz, topo)
    dN = O.dN_table(el)
    E = 0.0
    for e in range(topo.shape[0]):
        X = P[topo[e]]
        for q in range(V.shape[1]):
            F = X.T @ dN[q] @ ji[e, q].reshape(3, 3).T
            E += _vbd_energy(lam, mu, F) * V[e, q]
    return E


def _vertex_hessian(energy_of_x, x, h=1e-5):
    """3x3 second derivative of a function of one vertex position by central differences."""
    I = np.eye(3)
    H = np.zeros((3, 3))
    for a in range(3):
        for b in range(3):
            H[a, b] = (energy_of_x(x + h * (I[a] + I[b])) - energy_of_x(x + h * (I[a] - I[b]))
                       - energy_of_x(x - h * (I[a] - I[b])) + energy_of_x(x - h * (I[a] + I[b]))) / (4 * h * h)
    return H


# ------------------------------------------- semi-implicit frame: free fall
@pytest.mark.parametrize("solver", [O.XPBD, O.EXPLICIT])
def test_semi_implicit_free_fall_closed_form(solver):
    """With no elastic force the frame v += g dt, p += v dt (then, for XPBD,
    v = (p - last_p)/dt) gives p_n = x_0 + g dt^2 n(n+1)/2 and v_n = n g dt."""
    xyz, topo = O.beam(O.TETRA10, (2, 1, 1))
    s = O.Scene(O.TETRA10, xyz, topo, material=O.STABLE_NEOHOOKE, solver=solver, young=1e-9, poisson=0.3)
    dt = 1e-3
    x0 = xyz.astype(np.float64)
    for n in range(1, 5):
        s.step(dt)
        assert np.abs(s.positions() - (x0 + G * dt * dt * n * (n + 1) / 2)).max() < 1e-13
        assert np.abs(s.get("v") - G * dt * n).max() < 1e-10


def test_xpbd_velocity_global_damping_closed_form():
    """kernel_velocity_update: |v| shrinks by min(1, k dt w): vertices with
    k dt w >= 1 stop, the others keep (1 - k dt w) of their speed.  Free fall,
    two substeps: v_1 = g dt c, p_2 = x_0 + g dt^2 (2 + c), v_2 = g dt c (1 + c),
    c = max(0, 1 - k dt w)."""
    xyz, topo = O.beam(O.TETRA10, (2, 1, 1), (0.2, 0.1, 0.1))
    m = O.masses(O.TETRA10, xyz, topo, 1000.0, True)
    dt = 1e-3
    w = 1.0 / m
    k = 1.0 / (dt * np.median(w))          # about half the vertices saturate
    c = np.maximum(0.0, 1.0 - k * dt * w)
    assert (c == 0).sum() > 3 and ((c > 0.05) & (c < 1)).sum() > 3
    s = O.Scene(O.TETRA10, xyz, topo, material=O.STABLE_NEOHOOKE, solver=O.XPBD, young=1e-9, poisson=0.3,
                global_damping=k)
    x0 = xyz.astype(np.float64)
    s.step(dt)
    assert np.abs(s.positions() - (x0 + G * dt * dt)).max() < 1e-13
    assert np.abs(s.get("v") - G[None, :] * dt * c[:, None]).max() < 1e-10
    s.step(dt)
    assert np.abs(s.positions() - (x0 + G[None, :] * dt * dt * (2 + c[:, None]))).max() < 1e-13
    assert np.abs(s.get("v") - G[None, :] * dt * (c * (1 + c))[:, None]).max() < 1e-10


# ------------------------------------------------ XPBD projections by hand
@pytest.mark.parametrize("material", [O.NEOHOOKE, O.STABLE_NEOHOOKE])
def test_xpbd_v0_single_tet_against_hand_derivation(material):
    """One stretched P1 tet, one substep without gravity: the lambda then the
    mu constraint, C = sqrt|E|, grad C = grad E / (2C) with grad E by finite
    differences of V e_m(Ds Dm^-1), dl = -C / (sum w |grad C|^2 + 1/(k dt^2)),
    x += dl w grad C."""
    X = np.array([[0, 0, 0], [0.1, 0, 0], [0, 0.1, 0], [0, 0, 0.1]], np.float64)
    topo = np.array([[0, 1, 2, 3]], np.int32)
    rng = np.random.default_rng(11)
    x = X * np.array([1.1, 1.0, 1.0]) + rng.uniform(-2e-3, 2e-3, X.shape)
    young, nu, rho, dt = 1e5, 0.3, 1000.0, 5e-3
    lam, mu = _lame(young, nu)
    w = np.full(4, 4.0 / (rho * 0.1 ** 3 / 6.0))     # FemShape: rho V / 4 per node
    s = O.Scene(O.TETRA, X.astype(np.float32), topo, material=material, solver=O.XPBD, xpbd_version=0,
                young=young, poisson=nu, gravity=(0, 0, 0))
    s.set("p", x)
    s.set("last_p", x)
    s.step(dt)
    X32 = X.astype(np.float32).astype(np.float64)
    p = x.copy()
    for m, k in ((0, lam), (1, mu)):
        E = _p1_energy(material, lam, mu, m, X32, p)
        gE = _fd_grad(lambda y: _p1_energy(material, lam, mu, m, X32, y), p)
        C = np.sqrt(abs(E))
        gC = gE / (2 * C)
        dl = -C / (np.sum(w[:, None] * gC * gC) + 1.0 / (k * dt * dt))
        p = p + dl * w[:, None] * gC
    disp = np.abs(p - x).max()
    assert disp > 1e-5
    assert np.abs(s.positions() - p).max() < 1e-7 * disp


def test_xpbd_v2_p2_element_against_hand_derivation():
    """V2 (the energy form) on one P2 element: beta = -2E / (sum w |grad E|^2 +
    4E / (k dt^2)), x += beta w grad E, grad E by finite differences of
    sum_q V_q e_m(F_q)."""
    xyz, topo = O.beam(O.TETRA10, (1, 1, 1), (0.2, 0.2, 0.2))
    topo = topo[:1]
    used = np.unique(topo)
    remap = -np.ones(xyz.shape[0], int)
    remap[used] = np.arange(used.size)
    X = xyz[used]
    t1 = remap[topo].astype(np.int32)
    rng = np.random.default_rng(12)
    x = X.astype(np.float64) * np.array([1.0, 1.08, 1.0]) + rng.uniform(-3e-3, 3e-3, X.shape)
    young, nu, dt = 1e5, 0.35, 2e-3
    lam, mu = _lame(young, nu)
    m_ = O.masses(O.TETRA10, X, t1, 1000.0, True)
    w = 1.0 / m_
    s = O.Scene(O.TETRA10, X, t1, material=O.STABLE_NEOHOOKE, solver=O.XPBD, xpbd_version=2, young=young,
                poisson=nu, gravity=(0, 0, 0))
    s.set("p", x)
    s.set("last_p", x)
    s.step(dt)
    ji, V = O.fem_constants(O.TETRA10, X, t1)
    dN = O.dN_table(O.TETRA10)

    def energy(m, y):
        e = 0.0
        for q in range(V.shape[1]):
            F = y[t1[0]].T @ dN[q] @ ji[0, q].reshape(3, 3).T
            e += _tet_energy(O.STABLE_NEOHOOKE, lam, mu, m, F) * V[0, q]
        return e

    p = x.copy()
    for m, k in ((0, lam), (1, mu)):
        E = energy(m, p)
        gE = _fd_grad(lambda y: energy(m, y), p)
        beta = -2 * E / (np.sum(w[:, None] * gE * gE) + 4 * E / (k * dt * dt))
        p = p + beta * w[:, None] * gE
    disp = np.abs(p - x).max()
    assert disp > 1e-5
    assert np.abs(s.positions() - p).max() < 1e-7 * disp


# ------------------------------------------------------------ VBD damping
def test_vbd_damped_fixed_point_solves_the_damped_local_equation():
    """With damping k the converged one-free-vertex VBD step satisfies
    -dW/dx - (k/dt) d2W/dx2 (x - x_last) - m/dt^2 (x - y) = 0, both derivatives
    by finite differences of the energy; the undamped equation does not hold
    there."""
    xyz, topo = O.beam(O.TETRA, (1, 1, 1), (0.1, 0.1, 0.1))
    free = 6
    fixed = [i for i in range(xyz.shape[0]) if i != free]
    young, nu, dt, kd = 1e5, 0.3, 0.02, 0.02
    lam, mu = _lame(young, nu)
    s = O.Scene(O.TETRA, xyz, topo, material=O.STABLE_NEOHOOKE, young=young, poisson=nu, iterations=80, damping=kd)
    s.fix(fixed)
    s.step(dt)
    x = s.positions()[free]
    P0 = xyz.astype(np.float64)
    x_last = P0[free]
    y = P0[free] + G * dt * dt
    m = O.masses(O.TETRA, xyz, topo, 1000.0, True)[free]

    def W(v):
        P = P0.copy()
        P[free] = v
        return _mesh_energy(O.TETRA, xyz, topo, P, lam, mu)

    gW = _fd_grad(W, x.copy())
    H = _vertex_hessian(W, x)
    inertia = m / dt ** 2 * (x - y)
    r_damped = -gW - (kd / dt) * H @ (x - x_last) - inertia
    r_plain = -gW - inertia
    scale = np.abs(inertia).max()
    assert np.abs(r_damped).max() < 1e-6 * scale
    assert np.abs(r_plain).max() > 1e-2 * scale


# ------------------------------------------------------- explicit damping
def test_explicit_damped_force_is_minus_gradient_minus_k_Hvv_v():
    """kernel_fem_eval_force: f_v = -dE/dx_v - k H_vv v_v with H_vv the vertex's
    own 3x3 block of d2E/dx2 (finite differences).  f is read back from the
    semi-implicit velocity change v_1 - v_0 = f w dt (no gravity)."""
    el = O.TETRA10
    xyz, topo = O.beam(el, (2, 1, 1), (0.2, 0.1, 0.1))
    rng = np.random.default_rng(13)
    p = xyz.astype(np.float64) + rng.uniform(-0.004, 0.004, xyz.shape)
    v0 = rng.uniform(-0.5, 0.5, xyz.shape)
    young, nu, kd, dt = 1e6, 0.45, 2e-3, 1e-4
    lam, mu = _lame(young, nu)
    s = O.Scene(el, xyz, topo, material=O.STABLE_NEOHOOKE, solver=O.EXPLICIT, young=young, poisson=nu,
                damping=kd, gravity=(0, 0, 0))
    s.set("p", p)
    s.set("last_p", p)
    s.set("v", v0)
    s.step(dt)
    m = O.masses(el, xyz, topo, 1000.0, True)
    f = (s.get("v") - v0) * m[:, None] / dt
    for vid in (0, 7, xyz.shape[0] - 1):
        def E(x):
            P = p.copy()
            P[vid] = x
            return _mesh_energy(el, xyz, topo, P, lam, mu)

        g = _fd_grad(E, p[vid].copy())
        H = _vertex_hessian(E, p[vid])
        damp = kd * H @ v0[vid]
        exp = -g - damp
        tol = 1e-5 * np.abs(exp).max()
        assert np.abs(damp).max() > 100 * tol
        assert np.abs(f[vid] - exp).max() < tol


# ------------------------------------------------- reference CPU XPBD path
def test_cpu_xpbd_free_fall_and_fixed_particles():
    """No elastic force: active particles fall as p_n = x_0 + g h^2 n(n+1)/2
    over n substeps h; fixed particles (RB_Fixation, inactive) stay put."""
    xyz, topo = O.beam(O.TETRA, (3, 1, 1))
    fixed = np.nonzero(xyz[:, 0] < 1e-6)[0]
    free = np.setdiff1d(np.arange(xyz.shape[0]), fixed)
    x0 = xyz.astype(np.float64)
    for prec in ("f", "d"):
        s = O.CpuXPBD(O.TETRA, xyz, topo, fixed, material=O.STABLE_NEOHOOKE, young=1e-9, poisson=0.3,
                      substeps=4, precision=prec)
        s.step(0.01)
        h = 0.01 / 4
        tol = 1e-13 if prec == "d" else 2e-6
        p = s.positions()
        assert np.abs(p[free] - (x0[free] + G * h * h * 10)).max() < tol
        assert np.array_equal(p[fixed], x0[fixed])


def test_cpu_xpbd_two_elements_insertion_order_by_hand():
    """Two P1 tets sharing a face, one of the shared vertices fixed.  Per
    substep: integrate the active particles, reset the multipliers, then for
    each iteration the constraints in insertion order (element 0 lambda,
    element 0 mu, element 1 lambda, element 1 mu) with accumulated
    multipliers dl = -(C + a l) / (sum w |grad C|^2 + a), a = 1/(k h^2),
    C = sign(E) sqrt|E|, grad C = grad E / (2C) (the fixed vertex moves too);
    RB_Fixation snaps it back; velocities from the positions with global
    damping.  Traced here with finite-difference gradients."""
    X = np.array([[0, 0, 0], [0.1, 0, 0], [0, 0.1, 0], [0, 0, 0.1], [0.1, 0.1, 0.1]], np.float64)
    topo = np.array([[0, 1, 2, 3], [1, 2, 3, 4]], np.int32)
    X32 = X.astype(np.float32)
    Xd = X32.astype(np.float64)
    fixed = np.array([1], np.int32)
    young, nu, rho, gd = 2e4, 0.3, 1000.0, 2.0
    dt, nsub, nit = 0.01, 2, 2
    lam, mu = _lame(young, nu)
    vols = [abs(np.linalg.det((Xd[t[1:]] - Xd[t[0]]).T)) / 6 for t in topo]
    m = np.zeros(5)
    for t, V in zip(topo, vols):
        m[t] += rho * V / 4
    w = 1.0 / m
    s = O.CpuXPBD(O.TETRA, X32, topo, fixed, material=O.NEOHOOKE, young=young, poisson=nu, density=rho,
                  iterations=nit, substeps=nsub, global_damping=gd, precision="d")
    s.step(dt)
    h = dt / nsub
    x = Xd.copy()
    v = np.zeros_like(x)
    active = np.ones(5, bool)
    active[fixed] = False
    for _ in range(nsub):
        last = x.copy()
        v[active] += G * h
        x[active] += v[active] * h
        lacc = np.zeros((2, 2))
        for _ in range(nit):
            for e, t in enumerate(topo):
                for mm, k in ((0, lam), (1, mu)):
                    fun = lambda y: _p1_energy(O.NEOHOOKE, lam, mu, mm, Xd[t], y)
                    E = fun(x[t])
                    gE = _fd_grad(fun, x[t].copy())
                    C = np.sign(E) * np.sqrt(abs(E))
                    gC = gE * (0.5 / C)
                    a = 1.0 / (k * h * h)
                    dl = -(C + a * lacc[e, mm]) / (np.sum(w[t][:, None] * gC * gC) + a)
                    lacc[e, mm] += dl
                    x[t] += dl * w[t][:, None] * gC
        x[fixed] = Xd[fixed]
        v[fixed] = 0
        v[active] = (x[active] - last[active]) / h
        for i in np.nonzero(active)[0]:
            n = np.linalg.norm(v[i])
            if n > 1e-12:
                v[i] += v[i] / n * (-n * min(1.0, gd * h * w[i]))
    disp = np.abs(x - Xd).max()
    assert disp > 1e-4
    assert np.abs(s.positions() - x).max() < 1e-7 * disp


def test_cpu_xpbd_float_tracks_double_on_the_16x4x4_beam():
    """The timed float path and its FP64 twin stay within float rounding of
    each other over 5 steps of the 16x4x4 P1 beam (Neo-Hookean, E = 1e6)."""
    xyz, topo = O.beam(O.TETRA, (16, 4, 4))
    fixed = np.nonzero(xyz[:, 0] < 1e-6)[0]
    a = O.CpuXPBD(O.TETRA, xyz, topo, fixed, precision="f")
    b = O.CpuXPBD(O.TETRA, xyz, topo, fixed, precision="d")
    for _ in range(5):
        a.step(0.01)
        b.step(0.01)
    disp = np.abs(b.positions() - xyz).max()
    assert disp > 1e-4
    assert np.abs(a.positions() - b.positions()).max() < 1e-3 * disp


# --------------------------------------------------- reference CPU VBD path
def test_cpu_vbd_free_fall_closed_form():
    xyz, topo = O.beam(O.TETRA10, (3, 1, 1))
    c = O.CpuVBD(O.TETRA10, xyz, topo, [], young=1e-9, poisson=0.3, iterations=3, precision="d")
    x0 = xyz.astype(np.float64)
    for n in range(1, 4):
        c.step(0.01)
        assert np.abs(c.positions() - (x0 + G * 1e-4 * n * (n + 1) / 2)).max() < 1e-13


@pytest.mark.parametrize("el,cells", [(O.TETRA10, (3, 2, 1)), (O.HEXA27, (2, 1, 1)), (O.TETRA, (4, 2, 2))])
def test_cpu_vbd_converges_to_the_gpu_path_step(el, cells):
    """Sequential Gauss-Seidel in id order (CPU) and colour-ordered sweeps (GPU
    path) minimise the same incremental potential with the same fixed set, so
    converged steps agree; after 3 iterations they do not yet."""
    xyz, topo = O.beam(el, cells)
    fixed = np.nonzero(xyz[:, 0] < 1e-6)[0]
    res = {}
    for it in (3, 300):
        c = O.CpuVBD(el, xyz, topo, fixed, young=1e5, poisson=0.4, iterations=it, precision="d")
        o = O.Scene(el, xyz, topo, young=1e5, poisson=0.4, iterations=it)
        o.fix(fixed)
        for _ in range(2):
            c.step(0.005)
            o.step(0.005)
        res[it] = np.abs(c.positions() - o.positions()).max()
        assert np.array_equal(c.positions()[fixed], xyz[fixed].astype(np.float64))
    assert res[300] < 1e-10
    assert res[3] > 1e3 * res[300]


def test_cpu_vbd_single_free_vertex_is_newton_minimiser():
    """One free vertex: the converged CPU VBD step minimises
    m/(2dt^2)|x-y|^2 + W(x); at that point the finite-difference gradient of
    the potential vanishes."""
    xyz, topo = O.beam(O.TETRA, (1, 1, 1), (0.1, 0.1, 0.1))
    free = 6
    fixed = [i for i in range(xyz.shape[0]) if i != free]
    young, nu, dt = 1e5, 0.3, 0.02
    lam, mu = _lame(young, nu)
    c = O.CpuVBD(O.TETRA, xyz, topo, fixed, young=young, poisson=nu, iterations=40, precision="d")
    c.step(dt)
    x = c.positions()[free]
    P0 = xyz.astype(np.float64)
    y = P0[free] + G * dt * dt
    m = O.masses(O.TETRA, xyz, topo, 1000.0, True)[free]

    def total(v):
        P = P0.copy()
        P[free] = v
        return m / (2 * dt * dt) * np.sum((v - y) ** 2) + _mesh_energy(O.TETRA, xyz, topo, P, lam, mu)

    g = _fd_grad(total, x.copy())
    assert np.abs(x - y).max() > 1e-5
    assert np.abs(g).max() < 1e-6 * (m / dt ** 2) * np.abs(x - y).max()


def test_cpu_vbd_chebyshev_keeps_the_converged_step():
    """Chebyshev extrapolation (rho = 0.5) changes the iterates, not the
    converged step."""
    xyz, topo = O.beam(O.TETRA10, (3, 2, 1))
    fixed = np.nonzero(xyz[:, 0] < 1e-6)[0]
    a = O.CpuVBD(O.TETRA10, xyz, topo, fixed, young=1e5, poisson=0.4, iterations=300, precision="d")
    b = O.CpuVBD(O.TETRA10, xyz, topo, fixed, young=1e5, poisson=0.4, iterations=300, rho=0.5, precision="d")
    c = O.CpuVBD(O.TETRA10, xyz, topo, fixed, young=1e5, poisson=0.4, iterations=4, rho=0.5, precision="d")
    d = O.CpuVBD(O.TETRA10, xyz, topo, fixed, young=1e5, poisson=0.4, iterations=4, precision="d")
    for s in (a, b, c, d):
        s.step(0.005)
    assert np.abs(a.positions() - b.positions()).max() < 1e-10
    assert np.abs(c.positions() - d.positions()).max() > 1e-9


def test_cpu_vbd_float_tracks_double():
    xyz, topo = O.beam(O.TETRA10, (8, 2, 2))
    fixed = np.nonzero(xyz[:, 0] < 1e-6)[0]
    a = O.CpuVBD(O.TETRA10, xyz, topo, fixed, precision="f")
    b = O.CpuVBD(O.TETRA10, xyz, topo, fixed, precision="d")
    for _ in range(3):
        a.step(0.005)
        b.step(0.005)
    disp = np.abs(b.positions() - xyz).max()
    assert disp > 1e-5
    assert np.abs(a.positions() - b.positions()).max() < 1e-3 * disp


def test_cpu_vbd_rejects_neohooke():
    """M_NeoHooke has no sub-Hessian on the CPU path."""
    xyz, topo = O.beam(O.TETRA, (1, 1, 1))
    with pytest.raises(ValueError):
        O.CpuVBD(O.TETRA, xyz, topo, [], material=O.NEOHOOKE)
