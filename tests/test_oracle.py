"""Pins of the FP64 oracle against what the mathematics and the reference's
own definitions fix (CPU only).  A plausible mistake - a wrong node order,
a dropped term, a transposed operand - fails one of these."""
import itertools
import math

import numpy as np
import pytest

import pyoracle as O
import ref_literal as R

ELEMS = [O.TETRA, O.HEXA, O.TETRA10, O.TETRA20, O.HEXA27]

# Reference-element node coordinates as listed by the reference's shape classes
# (include/Dynamic/FEM/FEM_Shape.h:53-55, :190-196, :240-246, :335-357, :519-557).
A, B = 1.0 / 3.0, 2.0 / 3.0
NODES = {
    O.TETRA: [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1)],
    O.HEXA: [(-1, -1, -1), (1, -1, -1), (1, 1, -1), (-1, 1, -1), (-1, -1, 1), (1, -1, 1), (1, 1, 1), (-1, 1, 1)],
    O.TETRA10: [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (.5, 0, 0), (.5, .5, 0), (0, .5, 0), (0, 0, .5),
                (.5, 0, .5), (0, .5, .5)],
    O.TETRA20: [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (A, 0, 0), (B, 0, 0), (B, A, 0), (A, B, 0), (0, B, 0),
                (0, A, 0), (0, 0, A), (0, 0, B), (B, 0, A), (A, 0, B), (0, B, A), (0, A, B), (A, 0, A), (A, A, A),
                (0, A, A), (A, A, 0)],
    O.HEXA27: [(-1, -1, -1), (1, -1, -1), (1, 1, -1), (-1, 1, -1), (-1, -1, 1), (1, -1, 1), (1, 1, 1), (-1, 1, 1),
               (0, -1, -1), (1, 0, -1), (0, 1, -1), (-1, 0, -1), (-1, -1, 0), (1, -1, 0), (1, 1, 0), (-1, 1, 0),
               (0, -1, 1), (1, 0, 1), (0, 1, 1), (-1, 0, 1), (0, 0, -1), (0, -1, 0), (1, 0, 0), (0, 1, 0),
               (-1, 0, 0), (0, 0, 1), (0, 0, 0)],
}


# ------------------------------------------------------------------ shapes
@pytest.mark.parametrize("el", ELEMS)
def test_shape_kronecker_and_partition(el):
    nodes = NODES[el]
    for j, x in enumerate(nodes):
        N, dN = O.shape_eval(el, x)
        e = np.zeros(len(nodes))
        e[j] = 1.0
        assert np.allclose(N, e, atol=1e-12), (el, j)
    rng = np.random.default_rng(1)
    for _ in range(5):
        x = rng.uniform(0.05, 0.3, 3)
        N, dN = O.shape_eval(el, x)
        assert abs(N.sum() - 1.0) < 1e-12
        assert np.abs(dN.sum(axis=0)).max() < 1e-11


@pytest.mark.parametrize("el", ELEMS)
def test_shape_gradients_fd(el):
    rng = np.random.default_rng(2)
    h = 1e-6
    for _ in range(3):
        x = rng.uniform(0.05, 0.25, 3)
        _, dN = O.shape_eval(el, x)
        for c in range(3):
            xp, xm = x.copy(), x.copy()
            xp[c] += h
            xm[c] -= h
            fd = (O.shape_eval(el, xp)[0] - O.shape_eval(el, xm)[0]) / (2 * h)
            assert np.abs(fd - dN[:, c]).max() < 1e-7


def _tet_monomial(a, b, c):
    return math.factorial(a) * math.factorial(b) * math.factorial(c) / math.factorial(a + b + c + 3)


@pytest.mark.parametrize("el,deg", [(O.TETRA, 1), (O.TETRA10, 2), (O.TETRA20, 4)])
def test_tet_quadrature_exactness(el, deg):
    qp, w = O.quadrature(el)
    assert abs(w.sum() - 1.0 / 6.0) < 1e-7
    for a, b, c in itertools.product(range(deg + 1), repeat=3):
        if a + b + c > deg:
            continue
        q = float(np.sum(w * qp[:, 0] ** a * qp[:, 1] ** b * qp[:, 2] ** c))
        assert abs(q - _tet_monomial(a, b, c)) < 2e-8, (a, b, c)


@pytest.mark.parametrize("el,deg", [(O.HEXA, 3), (O.HEXA27, 5)])
def test_hex_quadrature_exactness(el, deg):
    qp, w = O.quadrature(el)
    assert abs(w.sum() - 8.0) < 1e-5
    for a, b, c in itertools.product(range(deg + 1), repeat=3):
        exact = 1.0
        for k in (a, b, c):
            exact *= 0.0 if k % 2 else 2.0 / (k + 1)
        q = float(np.sum(w * qp[:, 0] ** a * qp[:, 1] ** b * qp[:, 2] ** c))
        assert abs(q - exact) < 1e-5, (a, b, c)


# ------------------------------------------------------------------ beams
def test_beam_counts_closed_form():
    assert O.beam(O.TETRA, (16, 4, 4))[0].shape[0] == 425
    assert O.beam(O.TETRA, (16, 4, 4))[1].shape[0] == 1536
    assert O.beam(O.TETRA10, (64, 16, 16))[0].shape[0] == 129 * 33 * 33 == 140481
    assert O.beam(O.TETRA10, (64, 16, 16))[1].shape[0] == 98304
    assert O.beam(O.HEXA27, (32, 8, 8))[0].shape[0] == 65 * 17 * 17 == 18785
    assert O.beam(O.TETRA20, (32, 8, 8))[0].shape[0] == 97 * 25 * 25 == 60625
    assert O.beam(O.HEXA, (4, 2, 3))[0].shape[0] == 5 * 3 * 4


def test_beam_grid_coordinates_and_volume():
    xyz, topo = O.beam(O.TETRA, (16, 4, 4))
    X, Y = 17, 5
    for i in (0, 1, 17, 17 * 5, 424):
        x, y, z = i % X, (i // X) % Y, i // (X * Y)
        assert np.allclose(xyz[i], [x / 16, y * 0.0625, z * 0.0625], atol=1e-7)
    p = xyz[topo].astype(np.float64)
    vol = np.einsum("ij,ij->i", np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0]), p[:, 3] - p[:, 0]) / 6
    assert abs(np.abs(vol).sum() - 0.0625) < 1e-9


@pytest.mark.parametrize("el,cells", [(O.TETRA10, (3, 2, 2)), (O.TETRA20, (2, 2, 1)), (O.HEXA27, (3, 2, 2))])
def test_high_order_nodes_at_reference_positions(el, cells):
    """Every element's nodes sit at the affine image of the reference nodes."""
    xyz, topo = O.beam(el, cells)
    ref = np.array(NODES[el], np.float64)
    if el == O.HEXA27:
        lin = (ref[:8] + 1) / 2
        bary = [np.prod([(1 - r[c]) / 2 if l[c] < 0 else (1 + r[c]) / 2 for c in range(3)]) for r in ref for l in ref[:8]]
        W = np.array(bary).reshape(27, 8)
    else:
        W = np.stack([1 - ref.sum(1), ref[:, 0], ref[:, 1], ref[:, 2]], 1)
    for e in range(topo.shape[0]):
        corners = xyz[topo[e, :W.shape[1]]].astype(np.float64)
        assert np.abs(W @ corners - xyz[topo[e]]).max() < 1e-6


def test_p2_first_encounter_numbering():
    """One cell: new P2 nodes are numbered in element order, edge order
    (0,1) (1,2) (0,2) (0,3) (1,3) (2,3) of each element, first encounter wins."""
    xyz, topo = O.beam(O.TETRA10, (1, 1, 1))
    lin = topo[:, :4]
    seen = {}
    nxt = 8
    edges = [(0, 1), (1, 2), (0, 2), (0, 3), (1, 3), (2, 3)]
    for e in range(6):
        for k, (a, b) in enumerate(edges):
            key = tuple(sorted((lin[e, a], lin[e, b])))
            if key not in seen:
                seen[key] = nxt
                nxt += 1
            assert topo[e, 4 + k] == seen[key]
    assert nxt == xyz.shape[0]


# ------------------------------------------------------------------ colourings
def test_dsat_known_answers():
    # one tetrahedron: all degrees 3, highest id first -> colours 3,2,1,0
    c, n = O.color_dsat(O.TETRA, np.array([[0, 1, 2, 3]]), 4)
    assert list(c) == [3, 2, 1, 0] and n == 4
    # two tetrahedra sharing the face (1,2,3): hand-traced
    c, n = O.color_dsat(O.TETRA, np.array([[0, 1, 2, 3], [1, 2, 3, 4]]), 5)
    assert list(c) == [3, 2, 1, 0, 3] and n == 4


@pytest.mark.parametrize("el,cells", [(O.TETRA, (4, 2, 2)), (O.TETRA10, (3, 2, 1)), (O.HEXA27, (2, 2, 1)),
                                      (O.HEXA, (3, 3, 2)), (O.TETRA20, (2, 1, 1))])
def test_dsat_matches_literal_restatement(el, cells):
    xyz, topo = O.beam(el, cells)
    adj = R.build_primal(topo, O.NODES[el])
    lit, nl = R.dsat(adj)
    c, n = O.color_dsat(el, topo, xyz.shape[0])
    assert list(c) == lit and n == nl
    assert O.primal_conflicts(el, topo, xyz.shape[0], c) == 0
    assert n >= O.NODES[el]          # an element is a clique


@pytest.mark.parametrize("el,cells,target", [(O.TETRA10, (3, 2, 1), 6), (O.HEXA27, (2, 1, 1), 20)])
def test_color_reduction_matches_literal(el, cells, target):
    xyz, topo = O.beam(el, cells)
    adj = R.build_primal(topo, O.NODES[el])
    lit, _ = R.dsat(adj)
    red = R.color_reduction(adj, lit, target)
    c, n = O.color_dsat(el, topo, xyz.shape[0], target)
    assert list(c) == red and n == target


@pytest.mark.parametrize("el,cells", [(O.TETRA, (4, 2, 2)), (O.TETRA10, (3, 2, 1)), (O.HEXA27, (3, 2, 2)),
                                      (O.TETRA20, (2, 1, 1))])
def test_slf_matches_literal_restatement(el, cells):
    xyz, topo = O.beam(el, cells)
    adj = R.build_dual_vertex(topo, O.NODES[el])
    lit, nl = R.greedy_slf(adj)
    c, n = O.color_slf(el, topo, xyz.shape[0])
    assert list(c) == lit and n == nl
    # proper element colouring: elements of one colour share no vertex
    for col in range(n):
        verts = topo[c == col].ravel()
        assert np.unique(verts).size == verts.size


# ------------------------------------------------------------------ FEM constants, masses
@pytest.mark.parametrize("el,cells", [(O.TETRA, (4, 2, 2)), (O.TETRA10, (2, 2, 2)), (O.TETRA20, (2, 1, 1)),
                                      (O.HEXA, (3, 2, 2)), (O.HEXA27, (2, 2, 2))])
def test_fem_constants(el, cells):
    xyz, topo = O.beam(el, cells)
    ji, V = O.fem_constants(el, xyz, topo)
    assert abs(V.sum() - 0.0625) < 1e-7
    dN = O.dN_table(el)
    for e in range(0, topo.shape[0], max(1, topo.shape[0] // 7)):
        for q in range(V.shape[1]):
            J = xyz[topo[e]].astype(np.float64).T @ dN[q]
            Ji = ji[e, q].reshape(3, 3).T    # column-major storage
            assert np.allclose(Ji @ J, np.eye(3), atol=1e-9)


@pytest.mark.parametrize("el", [O.TETRA, O.TETRA10, O.HEXA27])
def test_masses_sum(el):
    xyz, topo = O.beam(el, (4, 2, 2))
    for fs in (True, False):
        m = O.masses(el, xyz, topo, 1000.0, fs)
        assert abs(m.sum() - 62.5) < 62.5e-6   # float-rounded quadrature weights
        assert (m > 0).all()


# ------------------------------------------------------------------ materials
def _energy_vbd(mat, lam, mu, F):
    J = np.linalg.det(F)
    alpha = 1 + mu / lam if mat == O.STABLE_NEOHOOKE else 1.0
    return 0.5 * mu * (np.sum(F * F) - 3) + 0.5 * lam * (J - alpha) ** 2


@pytest.mark.parametrize("mat", [O.NEOHOOKE, O.STABLE_NEOHOOKE])
def test_vbd_material_derivatives(mat):
    """P = dW/dF and the vertex block H = d2W(F(x))/dx2 for F = F0 + x g^T,
    by central differences of the energy (W_nh with the reference's GPU
    convention: mu/2 (|F|^2-3) + lam/2 (J-1)^2; W_snh uses J - (1 + mu/lam))."""
    rng = np.random.default_rng(3)
    lam, mu = 8.2e6, 1.7e5
    for _ in range(4):
        F0 = np.eye(3) + 0.2 * rng.standard_normal((3, 3))
        g = rng.standard_normal(3)
        V = 0.7
        P, H = O.vbd_material(mat, lam, mu, F0, g, V)
        h = 1e-6
        Pfd = np.zeros((3, 3))
        for i in range(3):
            for j in range(3):
                E = np.zeros((3, 3))
                E[i, j] = h
                Pfd[i, j] = (_energy_vbd(mat, lam, mu, F0 + E) - _energy_vbd(mat, lam, mu, F0 - E)) / (2 * h)
        assert np.abs(P - Pfd).max() < 1e-4 * np.abs(P).max()
        Hfd = np.zeros((3, 3))
        hh = 1e-4
        for a in range(3):
            for b in range(3):
                def W(da, db):
                    x = np.zeros(3)
                    x[a] += da
                    x[b] += db
                    return _energy_vbd(mat, lam, mu, F0 + np.outer(x, g))
                Hfd[a, b] = (W(hh, hh) - W(hh, -hh) - W(-hh, hh) + W(-hh, -hh)) / (4 * hh * hh) * V
        assert np.abs(H - Hfd).max() < 1e-4 * np.abs(H).max()
    # rest state
    P, _ = O.vbd_material(O.STABLE_NEOHOOKE, lam, mu, np.eye(3), np.ones(3), 1.0)
    assert np.abs(P).max() < 1e-6 * mu
    P, _ = O.vbd_material(O.NEOHOOKE, lam, mu, np.eye(3), np.ones(3), 1.0)
    assert np.allclose(P, mu * np.eye(3))


@pytest.mark.parametrize("mat", [O.NEOHOOKE, O.STABLE_NEOHOOKE])
def test_xpbd_material_energies(mat):
    rng = np.random.default_rng(4)
    lam, mu = 3.1e6, 3.4e5
    alpha = 1 + mu / lam

    def e_of(m, F):
        J = np.linalg.det(F)
        if mat == O.NEOHOOKE:
            return (J - alpha) ** 2 if m == 0 else np.sum(F * F) - 3
        return (J - 1) ** 2 if m == 0 else np.sum(F * F) - 3 - 2 * (J - 1)

    for _ in range(3):
        F = np.eye(3) + 0.2 * rng.standard_normal((3, 3))
        for m in (0, 1):
            P, e = O.xpbd_material(mat, m, lam, mu, F)
            assert abs(e - e_of(m, F)) < 1e-12
            h = 1e-6
            for i in range(3):
                for j in range(3):
                    E = np.zeros((3, 3))
                    E[i, j] = h
                    fd = (e_of(m, F + E) - e_of(m, F - E)) / (2 * h)
                    assert abs(P[i, j] - fd) < 1e-6


# ------------------------------------------------------------------ steps
def test_vbd_free_fall():
    xyz, topo = O.beam(O.TETRA10, (3, 1, 1))
    s = O.Scene(O.TETRA10, xyz, topo, young=1e-9, poisson=0.3, iterations=3)
    dt = 0.01
    s.step(dt)
    exp = xyz.astype(np.float64).copy()
    exp[:, 1] -= 9.81 * dt * dt
    assert np.abs(s.positions() - exp).max() < 1e-12
    s.step(dt)
    exp[:, 1] -= 2 * 9.81 * dt * dt    # x2 = 2 x1 - x0 + g dt^2
    assert np.abs(s.positions() - exp).max() < 1e-12


def _beam_scene(iterations, **options):
    xyz, topo = O.beam(O.TETRA10, (3, 2, 1))
    s = O.Scene(O.TETRA10, xyz, topo, young=1e5, poisson=0.4, iterations=iterations)
    s.fix(np.nonzero(xyz[:, 0] < 1e-6)[0])
    s.box()
    if options:
        s.vbd_options(**options)
    return s, xyz, topo


def test_vbd_jacobi_equals_gauss_seidel_under_proper_colouring():
    """No two vertices of one colour share an element in a proper colouring,
    so solving a colour from frozen positions (v4) and in place (v3) give the
    same step, bit for bit."""
    a, _, _ = _beam_scene(4)
    b, _, _ = _beam_scene(4, jacobi=True)
    for _ in range(2):
        a.step(0.005)
        b.step(0.005)
    assert np.array_equal(a.positions(), b.positions())


@pytest.mark.parametrize("target", [6, 8])
def test_vbd_reduced_colouring_jacobi_converges_to_the_same_step(target):
    """ColorReduction to fewer colours leaves same-colour neighbours; the
    Jacobi sweep over that colouring still converges to the minimiser of the
    step's incremental potential that Gauss-Seidel reaches."""
    ref, xyz, topo = _beam_scene(300)
    ref.step(0.005)
    j, _, _ = _beam_scene(300, jacobi=True, target_color=target)
    assert j.ncolors == target
    assert O.primal_conflicts(O.TETRA10, topo, xyz.shape[0], j.colors()) > 0
    j.step(0.005)
    assert np.abs(j.positions() - ref.positions()).max() < 1e-12


def test_vbd_chebyshev_keeps_free_fall_and_accelerates():
    """Chebyshev extrapolation between iterations leaves the exact free-fall
    step unchanged, and with a spectral radius below the sweep's own it
    brings a 6-iteration step closer to the converged one."""
    xyz, topo = O.beam(O.TETRA10, (3, 2, 1))
    s = O.Scene(O.TETRA10, xyz, topo, young=1e-9, poisson=0.3, iterations=5)
    s.vbd_options(rho=0.9)
    s.step(0.01)
    exp = xyz.astype(np.float64).copy()
    exp[:, 1] -= 9.81 * 0.01 * 0.01
    assert np.abs(s.positions() - exp).max() < 1e-12
    conv, _, _ = _beam_scene(400)
    plain, _, _ = _beam_scene(6)
    cheb, _, _ = _beam_scene(6, rho=0.7)
    for sc in (conv, plain, cheb):
        sc.step(0.005)
        sc.step(0.005)
    e_plain = np.abs(plain.positions() - conv.positions()).max()
    e_cheb = np.abs(cheb.positions() - conv.positions()).max()
    assert e_cheb < 0.5 * e_plain


@pytest.mark.parametrize("el,cells", [(O.TETRA, (4, 2, 2)), (O.TETRA10, (3, 2, 1)), (O.HEXA27, (2, 1, 1))])
@pytest.mark.parametrize("mat", [O.NEOHOOKE, O.STABLE_NEOHOOKE])
def test_element_diagnostics_closed_forms(el, cells, mat):
    """Rest: no stress, volume = sum V, nothing inverted.  Uniaxial stretch
    F = diag(a, 1, 1): Cauchy stress sigma_xx - sigma_yy = mu (a - 1/a) for
    both Neo-Hookean forms (the lambda terms are equal on the diagonal), so
    stress = mu |a - 1/a| V_e, volume = a V_e.  A reflection inverts every
    element and keeps the volume."""
    xyz, topo = O.beam(el, cells)
    s = O.Scene(el, xyz, topo, material=mat, young=1e5, poisson=0.4)
    _, V = O.fem_constants(el, xyz, topo)
    Ve = V.sum(1)
    d = s.element_diagnostics()
    assert np.abs(d["stress"]).max() < 1e-9
    assert np.abs(d["volume"] - Ve).max() < 1e-15
    assert np.abs(d["volume_diff"]).max() < 1e-15 and d["inverted"].max() == 0
    a = 1.3
    mu = 1e5 / 2.8
    p = xyz.astype(np.float64).copy()
    p[:, 0] *= a
    s.set("p", p)
    d = s.element_diagnostics()
    assert np.allclose(d["stress"], mu * abs(a - 1 / a) * Ve, rtol=1e-12)
    assert np.allclose(d["volume"], a * Ve, rtol=1e-12)
    assert np.allclose(d["volume_diff"], (a - 1) * Ve, rtol=1e-12, atol=1e-18)
    p = xyz.astype(np.float64).copy()
    p[:, 0] *= -1
    s.set("p", p)
    d = s.element_diagnostics()
    assert d["inverted"].min() == 1
    assert np.allclose(d["volume"], Ve, rtol=1e-12)


def test_residual_vanishes_as_vbd_converges():
    """The residual force of the local problems is the gradient of the step's
    incremental potential: it shrinks with more iterations and vanishes at the
    converged step; fixed vertices report zero."""
    xyz, topo = O.beam(O.TETRA10, (3, 2, 1))
    fixed = np.nonzero(xyz[:, 0] < 1e-6)[0]
    res = []
    for it in (2, 20, 300):
        s = O.Scene(O.TETRA10, xyz, topo, young=1e5, poisson=0.4, iterations=it)
        s.fix(fixed)
        s.box()
        s.step(0.005)
        s.step(0.005)
        r = s.residual(0.005)
        assert np.abs(r[fixed]).max() == 0
        res.append(np.abs(r).max())
    assert res[0] > 1e-2 and res[1] < res[0] * 1e-2 and res[2] < 1e-9


def test_vbd_single_free_vertex_is_newton_minimiser():
    """Two fixed-by-constraint layers and one free vertex: VBD with many
    iterations converges to the minimiser of m/(2dt^2)|x-y|^2 + W(x), which is
    also found by brute-force Newton with finite-difference gradients."""
    xyz, topo = O.beam(O.TETRA, (1, 1, 1), (0.1, 0.1, 0.1))
    free = 6
    fixed = [i for i in range(xyz.shape[0]) if i != free]
    s = O.Scene(O.TETRA, xyz, topo, material=O.STABLE_NEOHOOKE, young=1e5, poisson=0.3, iterations=30)
    s.fix(fixed)
    dt = 0.02
    s.step(dt)
    x_vbd = s.positions()[free]
    # brute force: total energy of the free vertex position
    ji, V = O.fem_constants(O.TETRA, xyz, topo)
    dN = O.dN_table(O.TETRA)[0]
    m = O.masses(O.TETRA, xyz, topo, 1000.0, True)[free]
    lam = 1e5 * 0.3 / (1.3 * 0.4)
    mu = 1e5 / 2.6
    y = xyz[free].astype(np.float64) + np.array([0, -9.81 * dt * dt, 0])
    P0 = xyz.astype(np.float64)

    def total(x):
        P = P0.copy()
        P[free] = x
        E = 0.0
        for e in range(topo.shape[0]):
            F = P[topo[e]].T @ dN @ ji[e, 0].reshape(3, 3).T
            E += _energy_vbd(O.STABLE_NEOHOOKE, lam, mu, F) * V[e, 0]
        return m / (2 * dt * dt) * np.sum((x - y) ** 2) + E

    x = y.copy()
    h = 1e-7
    for _ in range(20):
        gr = np.array([(total(x + h * np.eye(3)[k]) - total(x - h * np.eye(3)[k])) / (2 * h) for k in range(3)])
        Hs = np.array([[(total(x + 1e-5 * (np.eye(3)[a] + np.eye(3)[b])) - total(x + 1e-5 * (np.eye(3)[a] - np.eye(3)[b]))
                         - total(x - 1e-5 * (np.eye(3)[a] - np.eye(3)[b])) + total(x - 1e-5 * (np.eye(3)[a] + np.eye(3)[b])))
                        / (4e-10) for b in range(3)] for a in range(3)])
        x = x - np.linalg.solve(Hs, gr)
    assert np.abs(x_vbd - x).max() < 1e-9


def test_xpbd_v0_equals_v2_when_energy_positive():
    """For one element with E > 0 on both constraints the sqrt form (V0) and
    the energy form (V2) give the same update (GPU_PBD_FEM.cu:49-58 vs :259-261)."""
    xyz, topo = O.beam(O.TETRA10, (1, 1, 1), (0.2, 0.2, 0.2))
    topo = topo[:1]
    used = np.unique(topo)
    remap = -np.ones(xyz.shape[0], int)
    remap[used] = np.arange(used.size)
    xyz1 = xyz[used]
    topo1 = remap[topo]
    rng = np.random.default_rng(5)
    p = xyz1 + rng.uniform(-0.02, 0.02, xyz1.shape)
    res = []
    for version in (0, 2):
        s = O.Scene(O.TETRA10, xyz1, topo1, material=O.STABLE_NEOHOOKE, solver=O.XPBD, xpbd_version=version,
                    young=1e5, poisson=0.3, gravity=(0, 0, 0))
        s.set("p", p)
        s.set("last_p", p)
        s.step(1e-3)
        res.append(s.positions())
    assert np.abs(res[0] - res[1]).max() < 1e-12
    assert np.abs(res[0] - p).max() > 1e-6


def _total_energy(el, xyz, topo, P, lam, mu, mat):
    ji, V = O.fem_constants(el, xyz, topo)
    dN = O.dN_table(el)
    E = 0.0
    for e in range(topo.shape[0]):
        X = P[topo[e]]
        for q in range(V.shape[1]):
            F = X.T @ dN[q] @ ji[e, q].reshape(3, 3).T
            E += _energy_vbd(mat, lam, mu, F) * V[e, q]
    return E


def test_explicit_force_is_minus_energy_gradient():
    el = O.TETRA10
    xyz, topo = O.beam(el, (2, 1, 1), (0.2, 0.1, 0.1))
    rng = np.random.default_rng(6)
    p = xyz.astype(np.float64) + rng.uniform(-0.004, 0.004, xyz.shape)
    lam, mu = 1e6 * 0.45 / (1.45 * 0.1), 1e6 / 2.9
    s = O.Scene(el, xyz, topo, material=O.STABLE_NEOHOOKE, solver=O.EXPLICIT, young=1e6, poisson=0.45,
                gravity=(0, 0, 0))
    s.set("p", p)
    s.set("last_p", p)
    s.step(1e-12)       # computes f at p, moves nothing measurable
    f = s.get("f")      # f is zeroed by the integrator: read it through a fresh eval below
    # re-evaluate: forces are consumed by the integration, so compare the
    # velocity change instead: v = f w dt
    v = s.get("v")
    m = O.masses(el, xyz, topo, 1000.0, True)
    f_from_v = v * m[:, None] / 1e-12
    h = 1e-7
    for vid in (0, 5, xyz.shape[0] - 1):
        for c in range(3):
            pp, pm = p.copy(), p.copy()
            pp[vid, c] += h
            pm[vid, c] -= h
            g = (_total_energy(el, xyz, topo, pp, lam, mu, O.STABLE_NEOHOOKE)
                 - _total_energy(el, xyz, topo, pm, lam, mu, O.STABLE_NEOHOOKE)) / (2 * h)
            assert abs(f_from_v[vid, c] + g) < 1e-5 * max(1.0, abs(g))
