"""Pins the CPU oracle before it is trusted as the checker.

* the reference's own regression decks (final U vs U.ref, ew/_cli/_run_tests_edelweissfe.py:96-109)
* element invariants: symmetry, rigid-body null space, patch test
* tangents against central finite differences of the residual (J2 plastic, Neo-Hooke TL)
* J2 closed form for linear hardening and yield-surface consistency
* Neo-Hooke homogeneous closed form and stress-free reference state
* CSR pattern: scipy canonical format vs brute-force dense assembly, and the cube closed form 9(3N-2)^3
"""
import numpy as np
import pytest
import scipy.sparse as sp

import decks
from oracle import oracle as O


@pytest.fixture(scope="module", autouse=True)
def _build():
    O.build()


@pytest.mark.parametrize("name", ["TensionBarHexa8", "CantileverBeamHexa8", "C3D20", "C3D8", "SimpleBeamHexa8",
                                  "WallShearHexa8", "WallShearHexa20", "LinearElasticIsotropic"])
def test_reference_decks(name):
    fx = decks.load(name)
    X, conn, sets = decks.mesh(fx, O.boxgen)
    r = O.solve_step(dict(el_kind=fx["el_type"], mat_kind=fx["material"], mp=fx["mp"], X=X, conn=conn),
                     decks.step(fx, sets))
    err = np.abs(r["U"] - np.array(fx["U_ref"])).max()
    assert err < 1e-6  # the reference's own bar
    assert err < 1e-7 * max(1.0, np.abs(fx["U_ref"]).max())


def _jittered_hex(nn, seed=0, amp=0.08):
    X, conn, _ = O.boxgen(1, 1, 1, nnodes=nn)
    rng = np.random.default_rng(seed)
    Xe = X[conn[0]] + rng.uniform(-amp, amp, size=(nn, 3))
    return Xe


@pytest.mark.parametrize("ek,nn", [("C3D8", 8), ("C3D20", 20), ("C3D8TL", 8), ("C3D20TL", 20), ("C3D20R", 20),
                                   ("C3D20RTL", 20)])
def test_symmetry_rigid_body_patch(ek, nn):
    Xe = _jittered_hex(nn)
    nd = 3 * nn
    mk, mp = ("neohookeanwa", [80.0, 200.0]) if "TL" in ek else ("linearelastic", [210000.0, 0.3])
    nq = O.NQP[O.EL_KINDS[ek]]
    st = np.zeros(nq * O.NSTATE[O.MAT_KINDS[mk]])
    K, P, _, _ = O.element(ek, mk, mp, Xe.ravel(), np.zeros(nd), np.zeros(nd), st)
    scale = np.abs(K).max()
    assert np.abs(K - K.T).max() < 1e-12 * scale
    w = np.linalg.eigvalsh(0.5 * (K + K.T))
    # six rigid-body modes; the 2x2x2 rule leaves a single 20-node element six more zero-energy modes
    zeros = 12 if ek.startswith("C3D20R") else 6
    assert np.all(np.abs(w[:zeros]) < 1e-10 * scale) and w[zeros] > 1e-3 * scale
    c = Xe.mean(axis=0)
    for R in (np.eye(3)[k] for k in range(3)):  # translations and rotations are in the null space
        assert np.abs(K @ np.tile(R, nn)).max() < 1e-10 * scale
        assert np.abs(K @ np.cross(R, Xe - c).ravel()).max() < 1e-10 * scale
    assert np.abs(P).max() < 1e-12 * scale
    if "TL" not in ek:  # patch test: linear field -> constant strain, K u = -P exactly
        G = np.array([[1e-3, 2e-4, 0], [0, -3e-4, 1e-4], [5e-4, 0, 2e-4]])
        u = (Xe @ G.T).ravel()
        K, P, st1, _ = O.element(ek, mk, mp, Xe.ravel(), u, u, st)
        assert np.abs(K @ u + P).max() < 1e-10 * np.abs(P).max()
        eps = np.array([G[0, 0], G[1, 1], G[2, 2], G[0, 1] + G[1, 0], G[0, 2] + G[2, 0], G[1, 2] + G[2, 1]])
        for q in range(nq):
            np.testing.assert_allclose(st1.reshape(nq, -1)[q, 6:12], eps, rtol=1e-10, atol=1e-16)


def _fd_tangent(ek, mk, mp, Xe, U0, dU0, st, h=1e-7):
    nd = len(U0)
    Kfd = np.zeros((nd, nd))
    for j in range(nd):
        e = np.zeros(nd)
        e[j] = h
        _, Pp, _, _ = O.element(ek, mk, mp, Xe, U0 + e, dU0 + e, st)
        _, Pm, _, _ = O.element(ek, mk, mp, Xe, U0 - e, dU0 - e, st)
        Kfd[:, j] = -(Pp - Pm) / (2 * h)
    return Kfd


def test_j2_tangent_vs_fd_plastic():
    Xe = _jittered_hex(8, seed=3)
    mp = [210000.0, 0.3, 250.0, 1000.0, 50.0, 20.0]
    rng = np.random.default_rng(1)
    G = np.array([[4e-3, 1e-3, 0], [0, -2e-3, 5e-4], [1e-3, 0, -1e-3]])
    u = (Xe @ G.T).ravel() + rng.uniform(-1e-4, 1e-4, 24)
    st = np.zeros(8 * 13)
    K, P, st1, pdt = O.element("C3D8", "vonmises", mp, Xe.ravel(), u, u, st)
    assert pdt > 1 and np.all(st1.reshape(8, 13)[:, 12] > 0)  # every QP yielded
    Kfd = _fd_tangent("C3D8", "vonmises", mp, Xe.ravel(), u, u, st, h=1e-9)
    assert np.abs(K - Kfd).max() < 1e-5 * np.abs(K).max()
    assert np.abs(K - K.T).max() < 1e-10 * np.abs(K).max()  # associative J2: symmetric tangent


@pytest.mark.parametrize("ek,nn", [("C3D8TL", 8), ("C3D20TL", 20), ("C3D20RTL", 20)])
def test_neohooke_tangent_vs_fd(ek, nn):
    Xe = _jittered_hex(nn, seed=5)
    rng = np.random.default_rng(2)
    u = rng.uniform(-0.08, 0.08, 3 * nn)
    nq = O.NQP[O.EL_KINDS[ek]]
    st = np.zeros(nq * 13)
    K, P, _, _ = O.element(ek, "neohookeanwa", [80.0, 200.0], Xe.ravel(), u, u, st)
    Kfd = _fd_tangent(ek, "neohookeanwa", [80.0, 200.0], Xe.ravel(), u, u, st, h=1e-6)
    assert np.abs(K - Kfd).max() < 1e-6 * np.abs(K).max()


@pytest.mark.parametrize("ek,nn", [("C3D8TL", 8), ("C3D20TL", 20), ("C3D20RTL", 20)])
def test_st_venant_kirchhoff_tangent_vs_fd(ek, nn):
    """St. Venant-Kirchhoff on the TL elements: the tangent (dtau/dF in the element's formula) against central
    differences of the forces, at a large deformation; symmetric (hyperelastic)."""
    Xe = _jittered_hex(nn, seed=6)
    rng = np.random.default_rng(4)
    u = rng.uniform(-0.1, 0.1, 3 * nn)
    nq = O.NQP[O.EL_KINDS[ek]]
    st = np.zeros(nq * 13)
    mp = [30000.0, 0.15]
    K, P, _, _ = O.element(ek, "stvenantkirchhoff", mp, Xe.ravel(), u, u, st)
    Kfd = _fd_tangent(ek, "stvenantkirchhoff", mp, Xe.ravel(), u, u, st, h=1e-6)
    assert np.abs(K - Kfd).max() < 1e-6 * np.abs(K).max()
    assert np.abs(K - K.T).max() < 1e-10 * np.abs(K).max()


def test_st_venant_kirchhoff_homogeneous_closed_form():
    """Homogeneous deformation gradient: every point holds tau = F S F^T with S = lambda tr(E) I + 2 mu E and
    the Green strain E = (F^T F - I) / 2 (engineering shear), and the state's energy lambda/2 tr(E)^2 + mu E:E;
    at rest the element is stress free and its tangent is the linear elastic one."""
    E_, nu = 30000.0, 0.15
    lam, mu = E_ * nu / ((1 + nu) * (1 - 2 * nu)), E_ / (2 * (1 + nu))
    Xe = _jittered_hex(8, seed=9, amp=0.05)
    Fh = np.array([[1.3, 0.1, 0.0], [0.05, 0.9, 0.02], [0.0, -0.1, 1.1]])
    u = (Xe @ (Fh - np.eye(3)).T).ravel()
    _, _, st1, _ = O.element("C3D8TL", "stvenantkirchhoff", [E_, nu], Xe.ravel(), u, u, np.zeros(8 * 13))
    Eg = 0.5 * (Fh.T @ Fh - np.eye(3))
    S = lam * np.trace(Eg) * np.eye(3) + 2 * mu * Eg
    T = Fh @ S @ Fh.T
    ref = np.array([T[0, 0], T[1, 1], T[2, 2], T[0, 1], T[1, 2], T[0, 2]])
    W = 0.5 * lam * np.trace(Eg) ** 2 + mu * (Eg * Eg).sum()
    for q in range(8):
        sq = st1.reshape(8, 13)[q]
        np.testing.assert_allclose(sq[:6], ref, rtol=1e-12, atol=1e-9)
        np.testing.assert_allclose(sq[6:12], [Eg[0, 0], Eg[1, 1], Eg[2, 2], 2 * Eg[0, 1], 2 * Eg[1, 2], 2 * Eg[0, 2]],
                                   rtol=1e-12, atol=1e-15)
        assert abs(sq[12] - W) < 1e-12 * W
    K0, P0, _, _ = O.element("C3D8TL", "stvenantkirchhoff", [E_, nu], Xe.ravel(), np.zeros(24), np.zeros(24),
                             np.zeros(8 * 13))
    Kle, _, _, _ = O.element("C3D8", "linearelastic", [E_, nu], Xe.ravel(), np.zeros(24), np.zeros(24),
                             np.zeros(8 * 12))
    assert np.abs(P0).max() == 0.0 and np.abs(K0 - Kle).max() < 1e-10 * np.abs(Kle).max()


def test_j2_linear_hardening_closed_form():
    """Homogeneous strain, one element: kappa follows the closed form and the returned stress sits on the
    yield surface (vonmises.py:121, 237-240)."""
    E, nu, fy0, H = 210000.0, 0.3, 250.0, 1000.0
    G = E / (2 * (1 + nu))
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    Xe = _jittered_hex(8, seed=7, amp=0.05)
    Gd = np.array([[3e-3, 0, 0], [0, -1e-3, 0], [0, 0, -1e-3]]) + np.array([[0, 1e-3, 0], [1e-3, 0, 0], [0, 0, 0]])
    u = (Xe @ Gd.T).ravel()
    K, P, st1, _ = O.element("C3D8", "vonmises", [E, nu, fy0, H, 0, 0], Xe.ravel(), u, u, np.zeros(8 * 13))
    eps = np.array([3e-3, -1e-3, -1e-3, 2e-3, 0, 0])
    tr = lam * eps[:3].sum() * np.array([1, 1, 1, 0, 0, 0]) + 2 * G * np.array([*eps[:3], *(eps[3:] / 2)])
    s = tr - np.array([1, 1, 1, 0, 0, 0]) * tr[:3].mean()
    ns = np.sqrt((s[:3] ** 2).sum() + 2 * (s[3:] ** 2).sum())
    dk = (ns - np.sqrt(2 / 3) * fy0) / (np.sqrt(6) * G + np.sqrt(2 / 3) * H)
    for q in range(8):
        sq = st1.reshape(8, 13)[q]
        assert abs(sq[12] - dk) < 1e-12 * dk
        sd = sq[:6] - np.array([1, 1, 1, 0, 0, 0]) * sq[:3].mean()
        nsd = np.sqrt((sd[:3] ** 2).sum() + 2 * (sd[3:] ** 2).sum())
        assert abs(nsd - np.sqrt(2 / 3) * (fy0 + H * sq[12])) < 1e-10 * fy0


def test_neohooke_homogeneous_closed_form():
    mu, Kb = 80.0, 200.0
    Xe = _jittered_hex(8, seed=9, amp=0.05)
    Fh = np.array([[1.3, 0.1, 0.0], [0.05, 0.9, 0.02], [0.0, -0.1, 1.1]])
    u = (Xe @ (Fh - np.eye(3)).T).ravel()
    _, _, st1, _ = O.element("C3D8TL", "neohookeanwa", [mu, Kb], Xe.ravel(), u, u, np.zeros(8 * 13))
    J = np.linalg.det(Fh)
    T = mu * Fh @ Fh.T + ((Kb - 2 * mu / 3) * (J * J - J) - mu) * np.eye(3)
    ref = np.array([T[0, 0], T[1, 1], T[2, 2], T[0, 1], T[1, 2], T[0, 2]])
    for q in range(8):
        np.testing.assert_allclose(st1.reshape(8, 13)[q, :6], ref, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("n,nn", [(2, 8), (3, 8), (2, 20)])
def test_csr_pattern_bruteforce(n, nn):
    X, conn, _ = O.boxgen(n, n, n, nnodes=nn)
    base, nDof, edofs = O.dof_numbering(len(X), conn)
    I, J = O.vij_pattern(edofs)
    indptr, indices, x = O.csr_pattern(I, J, nDof)
    dense = np.zeros((nDof, nDof), dtype=bool)
    for e in range(len(conn)):
        d = edofs[e]
        dense[np.ix_(d, d)] = True
    m = sp.csr_matrix(dense)
    assert np.array_equal(m.indptr, indptr) and np.array_equal(m.indices, indices)
    rows = np.repeat(np.arange(nDof), np.diff(indptr))
    assert np.array_equal(rows[x], I) and np.array_equal(indices[x], J)
    if nn == 8:
        # nnz of a structured hex mesh: 9 * prod over axes of (3*nodes - 2)
        assert len(indices) == 9 * (3 * (n + 1) - 2) ** 3


@pytest.mark.parametrize("nn", [8, 20])
def test_body_force_brick_closed_form(nn):
    """Consistent nodal loads of a uniform body force on rectangular bricks: V/8 at every C3D8 node;
    -V/8 at C3D20 corners and V/6 at mid-side nodes; the loads of a mesh sum to b times its volume."""
    L = (2.0, 3.0, 0.5)
    X, conn, _ = O.boxgen(1, 1, 1, *L, nnodes=nn)
    base, nDof, edofs = O.dof_numbering(len(X), conn)
    b = np.array([0.3, -1.7, 2.5])
    P = O.body_force("C3D8" if nn == 8 else "C3D20", conn, X, edofs, nDof, b).reshape(-1, 3)
    V = L[0] * L[1] * L[2]
    node_of_local = np.asarray(conn[0])
    rows = np.array([base[n] // 3 for n in node_of_local])
    w = np.full(nn, V / 8) if nn == 8 else np.r_[np.full(8, -V / 8), np.full(12, V / 6)]
    assert np.allclose(P[rows], w[:, None] * b[None, :], rtol=1e-13, atol=1e-14)
    # a jittered 3x2x2 mesh: total load = b x volume (partition of unity; the Gauss rule integrates detJ exactly)
    X2, conn2, _ = O.boxgen(3, 2, 2, *L, nnodes=nn)
    rng = np.random.default_rng(3)
    inner = np.all((X2 > 1e-9) & (X2 < np.array(L) - 1e-9), axis=1)
    X2 = X2 + inner[:, None] * rng.uniform(-0.05, 0.05, X2.shape)
    base2, nDof2, edofs2 = O.dof_numbering(len(X2), conn2)
    P2 = O.body_force("C3D8" if nn == 8 else "C3D20", conn2, X2, edofs2, nDof2, b).reshape(-1, 3)
    assert np.allclose(P2.sum(axis=0), b * V, rtol=1e-12)
    # element subset: the loads of the two halves add up to the whole
    half = np.arange(len(conn2)) % 2 == 0
    Pa = O.body_force("C3D8" if nn == 8 else "C3D20", conn2, X2, edofs2, nDof2, b, np.nonzero(half)[0])
    Pb = O.body_force("C3D8" if nn == 8 else "C3D20", conn2, X2, edofs2, nDof2, b, np.nonzero(~half)[0])
    assert np.allclose(Pa + Pb, P2.ravel(), rtol=1e-13, atol=1e-15)


def test_pressure_deck_pins_face_numbering():
    """LinearElasticIsotropic: one C3D8 (Marmot), point supports, pressure 0.15 * t**2 on face S4 in the second
    step. The linear solution is reproduced to rounding (U.ref holds 19 digits, displacements ~1e-5), which fixes
    the face's nodes, its normal's sign and the face quadrature of the restatement."""
    fx = decks.load("LinearElasticIsotropic")
    X, conn, sets = decks.mesh(fx, O.boxgen)
    r = O.solve_step(dict(el_kind=fx["el_type"], mat_kind=fx["material"], mp=fx["mp"], X=X, conn=conn),
                     decks.step(fx, sets))
    Uref = np.array(fx["U_ref"])
    assert np.abs(r["U"] - Uref).max() < 1e-12 * np.abs(Uref).max()


# outward unit normal and the local nodes (Abaqus order, 1-based in the comments) of faces 1..6 of a hexahedron
_FACES = {1: ((0, 0, -1), [0, 1, 2, 3], [8, 9, 10, 11]), 2: ((0, 0, 1), [4, 5, 6, 7], [12, 13, 14, 15]),
          3: ((0, -1, 0), [0, 1, 4, 5], [8, 12, 16, 17]), 4: ((1, 0, 0), [1, 2, 5, 6], [9, 13, 17, 18]),
          5: ((0, 1, 0), [2, 3, 6, 7], [10, 14, 18, 19]), 6: ((-1, 0, 0), [0, 3, 4, 7], [11, 15, 16, 19])}


@pytest.mark.parametrize("nn", [8, 20])
def test_surface_pressure_brick_closed_form(nn):
    """Consistent nodal loads of a pressure p on each face of a rectangular brick: -p A n / 4 at the four face
    nodes of a C3D8; +p A n / 12 at the face corners and -p A n / 3 at the face mid-side nodes of a C3D20
    (S1 = nodes 1-2-3-4 at zeta=-1, ..., S4 = 2-3-7-6 at xi=+1, ...); nothing on the other nodes."""
    L = (2.0, 3.0, 0.5)
    ek = "C3D8" if nn == 8 else "C3D20"
    xi = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1], [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1],
                   [0, -1, -1], [1, 0, -1], [0, 1, -1], [-1, 0, -1], [0, -1, 1], [1, 0, 1], [0, 1, 1], [-1, 0, 1],
                   [-1, -1, 0], [1, -1, 0], [1, 1, 0], [-1, 1, 0]], dtype=float)[:nn]  # Abaqus node order
    X = (xi + 1) / 2 * np.array(L) + np.array([0.4, -1.0, 2.0])
    conn = np.arange(nn, dtype=np.int32)[None, :]
    base, nDof, edofs = O.dof_numbering(len(X), conn)
    rows = np.array([base[n] // 3 for n in conn[0]])
    p = 1.7
    for face, (n, corners, mids) in _FACES.items():
        A = np.prod([L[i] for i in range(3) if n[i] == 0])
        P = O.surface_pressure(ek, conn, X, edofs, nDof, face, p).reshape(-1, 3)[rows]
        want = np.zeros((nn, 3))
        if nn == 8:
            want[corners] = -p * A * np.array(n) / 4
        else:
            want[corners] = p * A * np.array(n) / 12
            want[mids] = -p * A * np.array(n) / 3
        assert np.allclose(P, want, rtol=1e-13, atol=1e-13), face


@pytest.mark.parametrize("nn", [8, 20])
def test_surface_pressure_closed_surface_invariants(nn):
    """On a distorted element, a uniform pressure over all six faces is self-equilibrated: the face loads sum to
    zero force (the closed surface integral of n dA vanishes) and zero moment about the origin (sum_a X_a x P_a =
    integral of x x n dA = 0), which fails for any face with a flipped normal or mismatched node order; a subset
    of elements adds up with its complement."""
    ek = "C3D8" if nn == 8 else "C3D20"
    X, conn, _ = O.boxgen(2, 1, 2, 1.0, 1.5, 0.8, nnodes=nn)
    rng = np.random.default_rng(11)
    X = X + rng.uniform(-0.05, 0.05, X.shape)
    base, nDof, edofs = O.dof_numbering(len(X), conn)
    node_of_row = np.empty(nDof // 3, dtype=np.int64)
    node_of_row[base // 3] = np.arange(len(X))
    tot = np.zeros(nDof)
    for face in range(1, 7):
        tot += O.surface_pressure(ek, conn, X, edofs, nDof, face, 2.5)
    F = tot.reshape(-1, 3)
    assert np.abs(F.sum(axis=0)).max() < 1e-12
    M = np.cross(X[node_of_row], F).sum(axis=0)
    assert np.abs(M).max() < 1e-12
    one = O.surface_pressure(ek, conn, X, edofs, nDof, 4, 2.5).reshape(-1, 3)
    assert np.abs(one.sum(axis=0)).max() > 0.1  # a single face is not equilibrated
    sub = np.array([0, 3])
    rest = np.setdiff1d(np.arange(len(conn)), sub)
    a = O.surface_pressure(ek, conn, X, edofs, nDof, 2, 2.5, sub)
    b = O.surface_pressure(ek, conn, X, edofs, nDof, 2, 2.5, rest)
    assert np.allclose(a + b, O.surface_pressure(ek, conn, X, edofs, nDof, 2, 2.5), rtol=1e-14, atol=1e-15)


def test_exact_residual_is_correctly_rounded():
    """The oracle's refinement residual b - K x equals the exactly computed residual rounded once (rational
    arithmetic), also where the terms cancel to far below their size."""
    from fractions import Fraction

    import scipy.sparse as sp

    rng = np.random.default_rng(3)
    n = 40
    K = (sp.random(n, n, density=0.3, random_state=4, format="csr") * 1e4 + sp.eye(n)).tocsr()
    x = rng.standard_normal(n)
    b = K @ x + 1e-9 * rng.standard_normal(n)  # residuals ~1e-9 from terms ~1e4: heavy cancellation
    r = O.exact_residual(K, x, b)
    Kd = K.toarray()
    exact = [float(Fraction(b[i]) - sum(Fraction(Kd[i, j]) * Fraction(x[j]) for j in range(n))) for i in range(n)]
    assert np.array_equal(r, np.array(exact))
    assert not np.array_equal(r, b - K @ x)  # the plain residual is not that accurate here


def test_oracle_newton_with_several_blocks():
    """The oracle's Newton over several element blocks: a box split into two blocks, one linear elastic and one
    von Mises whose yield stress is never reached (the same elastic law), converges to the single-block
    linear-elastic solution; a J2 block that does yield changes it."""
    X, conn, sets = O.boxgen(4, 2, 3)
    step = dict(maxInc=0.5, minInc=1e-4, maxNumInc=100, maxIter=25,
                dirichlet=[(np.asarray(sets["gen_back"]), {0: 0.0, 1: 0.0, 2: 0.0}),
                           (np.asarray(sets["gen_front"]), {2: 4e-3})])
    le = [210000.0, 0.3]
    one = O.solve_step(dict(el_kind="C3D8", mat_kind="linearelastic", mp=le, X=X, conn=conn), step)
    half = len(conn) // 2

    def two(fy):
        blocks = [dict(el_kind="C3D8", mat_kind="linearelastic", mp=le, conn=conn[:half]),
                  dict(el_kind="C3D8", mat_kind="vonmises", mp=[210000.0, 0.3, fy, 1000.0, 0.0, 0.0],
                       conn=conn[half:])]
        return O.solve_step(dict(X=X, blocks=blocks), step)

    el = two(1e12)
    assert np.abs(el["U"] - one["U"]).max() < 1e-12 * np.abs(one["U"]).max()
    assert len(el["state"]) == 2 and np.all(el["state"][1].reshape(-1, 13)[:, 12] == 0.0)  # kappa stays 0
    pl = two(250.0)
    assert np.abs(pl["U"] - one["U"]).max() > 1e-3 * np.abs(one["U"]).max()
    assert np.all(pl["state"][1].reshape(-1, 13)[:, 12] > 0.0)
