"""Smoothed-aggregation multigrid PCG (ewfe_mg_* / ewfe_pcg_mg) on the GPU.

The linear solve has no bitwise counterpart in the reference (it factorises the row-eliminated matrix with a
direct solver, nonlinearimplicitstatic.py:727-751, or runs pyamg-preconditioned GMRES, linsolve/gmres/gmres.py);
what is pinned is the solution: scipy's sparse direct solve of the same row-eliminated matrix, 1e-9 relative
at a 1e-12 PCG tolerance, and the reference decks' U.ref through the full Newton loop with linsolver="amg".
"""
import numpy as np
import pytest

from gpu_util import GpuEval, jittered_box, rel
from test_gpu_parity import _deck_run

import decks
from oracle import oracle as O

pytestmark = pytest.mark.gpu

J2 = [210000.0, 0.3, 250.0, 1000.0, 0.0, 0.0]


def _plastic_system(n, tangent, elastic_modulus=None, dirichlet=True):
    """Fully plastic J2 tangent of an n^3 jittered cube under uniaxial strain with the bench's Dirichlet set (or, with
    `elastic_modulus`, the linear elastic tangent for that modulus); `tangent` is a value precision of newValues.
    With dirichlet=False the Dirichlet rows are returned but not applied."""
    import torch

    X, conn, sets = jittered_box(n, "C3D8", seed=3)
    if elastic_modulus is None:
        g = GpuEval("C3D8", "vonmises", J2, X, conn)
    else:
        g = GpuEval("C3D8", "linearelastic", [elastic_modulus, 0.3], X, conn)
    nd = g.dm.nDof
    base = g.dm.nodeDofBase
    u = np.zeros((len(X), 3))
    u[:, 2] = 5e-3 * X[:, 2]
    u[:, 0] = -0.15e-2 * X[:, 0]
    u[:, 1] = -0.15e-2 * X[:, 1]
    U = g.vec(u.ravel())
    vals = g.asm.newValues(tangent)
    P, F = g.dm.constructDofVector(), g.dm.constructDofVector()
    g.asm.evaluate(U, U, P, F, vals)
    org = np.where((X[:, 0] == 0) & (X[:, 1] == 0) & (X[:, 2] == 0))[0][0]
    xax = np.where((X[:, 0] == 1) & (X[:, 1] == 0) & (X[:, 2] == 0))[0][0]
    rows = np.unique(np.concatenate([base[sets["gen_back"]] + 2, base[sets["gen_front"]] + 2,
                                     [base[org], base[org] + 1, base[xax] + 1]])).astype(np.int32)
    rows_t = torch.as_tensor(rows, device="cuda")
    if dirichlet:
        g.asm.dirichletRows(rows_t, vals)
    rng = np.random.default_rng(5)
    b = rng.standard_normal(nd)
    b[rows] = 0.0
    b[base[sets["gen_front"]] + 2] = 1e-4
    return g, vals, rows, rows_t, b


@pytest.mark.parametrize("tangent", ["fp64", "fp32"])
def test_amg_pcg_matches_direct_solve(tangent):
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla

    print(f"amg pcg {tangent}", flush=True)
    g, vals, rows, rows_t, b = _plastic_system(20, tangent)
    nd = g.dm.nDof
    x = g.dm.constructDofVector()
    it, rr = g.asm.pcg(vals, g.vec(b), x, rows_t, rtol=1e-12, maxit=2000, check_every=4, precond="amg")
    print("levels", g.asm.mgLevels, "amg iterations", it, "relres", rr, flush=True)
    sizes = [lv[0] for lv in g.asm.mgLevels]
    # phase 1 of the aggregation tiles the 21^3 lattice with 3x3x3 blocks
    assert sizes[:2] == [21 ** 3, 7 ** 3] and sizes[-1] <= 200 and all(a > b for a, b in zip(sizes, sizes[1:]))
    ip, ix = g.asm.pattern()
    K = sp.csr_matrix((vals.cpu().numpy().astype(np.float64), ix, ip), shape=(nd, nd))
    xr = spla.spsolve(K.tocsc(), b)
    assert rr < 1e-12
    assert rel(x.cpu().numpy(), xr) < 1e-9
    # far fewer iterations than node-block Jacobi on the same system
    xj = g.dm.constructDofVector()
    itj, _ = g.asm.pcg(vals, g.vec(b), xj, rows_t, rtol=1e-12, maxit=20000, precond="jacobi")
    print("jacobi iterations", itj, flush=True)
    assert 0 < it < itj / 5
    # bitwise reproducible (fixed-order sums everywhere)
    x2 = g.dm.constructDofVector()
    it2, _ = g.asm.pcg(vals, g.vec(b), x2, rows_t, rtol=1e-12, maxit=2000, check_every=4, precond="amg")
    assert it2 == it and bool((x2 == x).all().item())


@pytest.mark.parametrize("tangent", ["fp64", "fp32"])
def test_amg_pcg_half_precision_fine_passes(tangent):
    """ewfe_mg_set_fine_precision(EWFE_F16): the V-cycle's fine-level smoothing and residual passes read a scaled FP16
    copy of the set-up matrix. The preconditioner changes, the solution does not: PCG to 1e-12 matches the sparse
    direct solve, in at most twice the iterations of the default (FP32 matrices: whose default V-cycle runs FP32
    vectors, so near 1e-12 either may need more); the same problem in other units (moduli and yield
    stress 1e6 times larger: entries far beyond FP16's range, 65504, unless the copy is scaled) is solved as well;
    switching back reproduces the default bitwise."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla

    print(f"amg pcg, half-precision fine passes, {tangent}", flush=True)
    g, vals, rows, rows_t, b = _plastic_system(20, tangent)
    nd = g.dm.nDof
    x0 = g.dm.constructDofVector()
    it0, _ = g.asm.pcg(vals, g.vec(b), x0, rows_t, rtol=1e-12, maxit=2000, check_every=4, precond="amg")
    g.asm.multigridSetup(vals, rows_t, finePrecision="fp16")
    x = g.dm.constructDofVector()
    it, rr = g.asm.pcg(vals, g.vec(b), x, rows_t, rtol=1e-12, maxit=2000, check_every=4, precond="amg", setup=False)
    print(f"iterations: default {it0}, fp16 fine passes {it}", flush=True)
    ip, ix = g.asm.pattern()
    K = sp.csr_matrix((vals.cpu().numpy().astype(np.float64), ix, ip), shape=(nd, nd))
    xr = spla.spsolve(K.tocsc(), b)
    assert rr < 1e-12 and rel(x.cpu().numpy(), xr) < 1e-9
    assert it <= 2 * it0
    # a linear elastic system in units with 1e6 times larger stresses (E = 2.1e11): modulus and the loads on the free
    # rows scale, the prescribed displacements (unit rows) do not, so the solution is that of E = 2.1e5
    ge, ke, _, rows_e, _ = _plastic_system(20, tangent, elastic_modulus=J2[0])
    Ke = sp.csr_matrix((ke.cpu().numpy().astype(np.float64), ix, ip), shape=(nd, nd))
    xe = spla.spsolve(Ke.tocsc(), b)
    gb, big, _, rows_b, _ = _plastic_system(20, tangent, elastic_modulus=J2[0] * 1e6)
    assert float(big.abs().max().item()) > 1e8
    bb = b.copy()
    free = np.ones(nd, dtype=bool)
    free[rows] = False
    bb[free] *= 1e6
    gb.asm.multigridSetup(big, rows_b, finePrecision="fp16")
    xb = gb.dm.constructDofVector()
    itb, rrb = gb.asm.pcg(big, gb.vec(bb), xb, rows_b, rtol=1e-12, maxit=2000, check_every=4, precond="amg",
                          setup=False)
    gb.asm.multigridSetup(big, rows_b)
    xd = gb.dm.constructDofVector()
    itd, _ = gb.asm.pcg(big, gb.vec(bb), xd, rows_b, rtol=1e-12, maxit=2000, check_every=4, precond="amg",
                        setup=False)
    print(f"E x 1e6: {itb} iterations ({itd} with the default passes), relres {rrb:.2e}, vs the default passes "
          f"{rel(xb.cpu().numpy(), xd.cpu().numpy()):.2e}, vs the direct solve at E {rel(xb.cpu().numpy(), xe):.2e}",
          flush=True)
    assert rrb < 1e-12 and rel(xb.cpu().numpy(), xd.cpu().numpy()) < 1e-9
    if tangent == "fp64":  # FP32 values of 1e6 K are not 1e6 times the FP32 values of K
        assert rel(xb.cpu().numpy(), xe) < 1e-9
    g.asm.multigridSetup(vals, rows_t)
    x2 = g.dm.constructDofVector()
    it2, _ = g.asm.pcg(vals, g.vec(b), x2, rows_t, rtol=1e-12, maxit=2000, check_every=4, precond="amg", setup=False)
    assert it2 == it0 and bool((x2 == x0).all().item())


@pytest.mark.parametrize("name", ["TensionBarHexa8", "C3D8"])
def test_reference_decks_with_amg(name):
    from edelweissfe_amd.solvers.nonlinearimplicitstaticgpu import NISTGPU

    print(f"deck {name} amg", flush=True)
    fx, U, solver = _deck_run(name, solver_options={"linsolver": "amg"})
    assert np.abs(U - np.array(fx["U_ref"])).max() < 1e-6
    X, conn, sets = decks.mesh(fx, O.boxgen)
    r = O.solve_step(dict(el_kind=fx["el_type"], mat_kind=fx["material"], mp=fx["mp"], X=X, conn=conn),
                     decks.step(fx, sets))
    assert rel(U, r["U"]) < 1e-9
    assert isinstance(solver, NISTGPU) and all(it > 0 for it, _ in solver.linearSolveIterations if it)


@pytest.mark.parametrize("tangent", ["fp64", "fp32"])
def test_multigrid_setup_per_increment(tangent):
    """The multigrid hierarchy set up on the first solve of each increment and reused for the increment's later
    corrections (multigridSetupPerIncrement): the plastic C3D8 cantilever deck (J2, several Newton iterations per
    increment) still reproduces U.ref and the oracle's Newton solution, with fewer setups than solves."""
    print(f"deck C3D8 amg, setup per increment, {tangent} tangent", flush=True)
    fx, U, solver = _deck_run("C3D8", tangent, solver_options={"linsolver": "amg",
                                                               "multigridSetupPerIncrement": True})
    assert np.abs(U - np.array(fx["U_ref"])).max() < 1e-6
    X, conn, sets = decks.mesh(fx, O.boxgen)
    r = O.solve_step(dict(el_kind=fx["el_type"], mat_kind=fx["material"], mp=fx["mp"], X=X, conn=conn),
                     decks.step(fx, sets))
    assert rel(U, r["U"]) < (1e-9 if tangent == "fp64" else 1e-7)
    solves = sum(1 for it, _ in solver.linearSolveIterations if it)
    increments_with_solves = sum(1 for h in solver.history if any(d > 0 for _, d in h))
    print(f"{solves} solves in {increments_with_solves} increments, {solver.multigridSetups} setups", flush=True)
    # one setup per increment that solved (and per discarded attempt), fewer than the solves
    assert increments_with_solves <= solver.multigridSetups < solves


def test_amg_pcg_on_20_node_mesh():
    """The multigrid on a C3D20 mesh (corner nodes couple with 81 nodes: the wide fine-level A P kernel): the
    solution of a Neo-Hooke tangent system matches scipy's direct solve, in far fewer iterations than block
    Jacobi, bitwise reproducibly."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    import torch

    print("amg pcg C3D20TL", flush=True)
    X, conn, sets = jittered_box(6, "C3D20TL", amp=0.03, seed=4)
    g = GpuEval("C3D20TL", "neohookeanwa", [80.0, 200.0], X, conn)
    nd = g.dm.nDof
    base = g.dm.nodeDofBase
    u = np.zeros((len(X), 3))
    u[:, 2] = 0.1 * X[:, 2]
    U = g.vec(u.ravel())
    vals = g.asm.newValues("fp64")
    P, F = g.dm.constructDofVector(), g.dm.constructDofVector()
    g.asm.evaluate(U, U, P, F, vals)
    rows = np.unique(np.concatenate([base[sets["gen_back"]][:, None] + np.arange(3)]).ravel()).astype(np.int32)
    rows_t = torch.as_tensor(rows, device="cuda")
    g.asm.dirichletRows(rows_t, vals)
    b = np.random.default_rng(6).standard_normal(nd)
    b[rows] = 0.0
    x = g.dm.constructDofVector()
    it, rr = g.asm.pcg(vals, g.vec(b), x, rows_t, rtol=1e-12, maxit=2000, check_every=4, precond="amg")
    xj = g.dm.constructDofVector()
    itj, _ = g.asm.pcg(vals, g.vec(b), xj, rows_t, rtol=1e-12, maxit=50000, precond="jacobi")
    print("levels", g.asm.mgLevels, "amg iterations", it, "jacobi", itj, flush=True)
    ip, ix = g.asm.pattern()
    K = sp.csr_matrix((vals.cpu().numpy(), ix, ip), shape=(nd, nd))
    xr = spla.spsolve(K.tocsc(), b)
    assert rr < 1e-12 and rel(x.cpu().numpy(), xr) < 1e-9
    assert 0 < it < itj / 3
    x2 = g.dm.constructDofVector()
    it2, _ = g.asm.pcg(vals, g.vec(b), x2, rows_t, rtol=1e-12, maxit=2000, check_every=4, precond="amg")
    assert it2 == it and bool((x2 == x).all().item())


def _host_csr(g, vals, nd):
    import scipy.sparse as sp

    ip, ix = g.asm.pattern()
    return sp.csr_matrix((g.asm.csrValues(vals).cpu().numpy().astype(np.float64), ix, ip), shape=(nd, nd))


def _with_diagonal_corrections(K, E):
    """K with E [n_nodes, 9] added to its 3x3 diagonal node blocks (node r: dofs 3r .. 3r + 2)."""
    C = K.tocoo()
    data = C.data.copy()
    dg = C.row // 3 == C.col // 3
    data[dg] += E[C.row[dg] // 3, 3 * (C.row[dg] % 3) + C.col[dg] % 3]
    return type(K)((data, (C.row, C.col)), shape=K.shape).tocsr()


def test_fp32_values_with_fp64_diagonal_blocks():
    """EWFE_F32_D64 ("fp32-d64" values): the FP32 values are bitwise the EWFE_F32 ones, and each node's FP64 diagonal
    correction makes its diagonal block minus the sum of its rows' off-diagonal FP32 values, so every row of the
    matrix sums to zero, as the rows of a tangent do for a rigid translation (checked against the same sums on the
    host, and the row sums of the FP64 tangent). SpMV, the compensated residual, Dirichlet rows and both CG variants
    use the corrections (against the host matrix with them); and the solution with them is much closer to the FP64
    tangent's than without."""
    import scipy.sparse.linalg as spla

    print("fp32 values with fp64 diagonal blocks, J2 8^3", flush=True)
    g64, v64, rows, rows_t, b = _plastic_system(8, "fp64", dirichlet=False)
    g32, v32, _, _, _ = _plastic_system(8, "fp32", dirichlet=False)
    gd, vd, _, _, _ = _plastic_system(8, "fp32-d64", dirichlet=False)
    nd = gd.dm.nDof
    nnz = gd.asm.nnz
    assert vd.numel() > nnz and vd.dtype == gd.torch.float32
    assert bool((gd.asm.csrValues(vd) == v32).all().item())
    K32 = _host_csr(g32, v32, nd)
    K64 = _host_csr(g64, v64, nd)
    offset = (4 * nnz + 255) // 256 * 256 // 4  # float index of the FP64 blocks (256-byte aligned)
    assert vd.numel() == offset + 18 * (nd // 3)
    E = vd[offset:].cpu().numpy().view(np.float64).reshape(-1, 9)
    # host: minus the off-diagonal row sums per column component, and the FP32 diagonal blocks
    C = K32.tocoo()
    offd = C.row // 3 != C.col // 3
    S = np.zeros((nd, 3))
    np.add.at(S, (C.row[offd], C.col[offd] % 3), C.data[offd])
    D32 = np.zeros((nd, 3))
    np.add.at(D32, (C.row[~offd], C.col[~offd] % 3), C.data[~offd])
    scale = np.abs(K64.data).max()
    assert np.abs(D32.reshape(-1, 9) + E + S.reshape(-1, 9)).max() <= 1e-15 * scale
    assert np.abs(E).max() <= 1e-5 * scale  # a rounding-sized correction
    Kd = _with_diagonal_corrections(K32, E)
    rs = np.zeros((nd, 3))
    Cd = Kd.tocoo()
    np.add.at(rs, (Cd.row, Cd.col % 3), Cd.data)
    rs64 = np.zeros((nd, 3))
    C64 = K64.tocoo()
    np.add.at(rs64, (C64.row, C64.col % 3), C64.data)
    rs32 = np.zeros((nd, 3))
    np.add.at(rs32, (C.row, C.col % 3), C.data)
    print(f"max row sum / max entry: fp64 {np.abs(rs64).max() / scale:.1e}, fp32 {np.abs(rs32).max() / scale:.1e}, "
          f"fp32 + fp64 blocks {np.abs(rs).max() / scale:.1e}", flush=True)
    assert np.abs(rs).max() <= 1e-14 * scale and np.abs(rs32).max() > 1e3 * np.abs(rs).max()
    # SpMV and residual with the blocks
    rng = np.random.default_rng(11)
    xh = rng.standard_normal(nd)
    y = gd.asm.spmv(vd, gd.vec(xh), gd.dm.constructDofVector()).cpu().numpy()
    assert rel(y, Kd @ xh) < 1e-14
    r = gd.asm.residual(vd, gd.vec(b), gd.vec(xh), gd.dm.constructDofVector()).cpu().numpy()
    assert np.abs(r - (b - Kd @ xh)).max() <= 1e-14 * np.abs(Kd @ xh).max()
    # Dirichlet rows: the FP32 values as for EWFE_F32 (unit rows), the corrections' rows of those dofs zero
    gd.asm.dirichletRows(rows_t, vd)
    g32.asm.dirichletRows(rows_t, v32)
    g64.asm.dirichletRows(rows_t, v64)
    assert bool((gd.asm.csrValues(vd) == v32).all().item())
    Ed = vd[offset:].cpu().numpy().view(np.float64).reshape(-1, 9)
    Eexp = E.copy()
    for row in rows:
        Eexp[row // 3, 3 * (row % 3): 3 * (row % 3) + 3] = 0.0
    assert np.array_equal(Ed, Eexp)
    Kd = _with_diagonal_corrections(_host_csr(g32, v32, nd), Ed)
    xd = spla.spsolve(Kd.tocsc(), b)
    for precond in ("jacobi", "amg"):
        x = gd.dm.constructDofVector()
        it, rr = gd.asm.pcg(vd, gd.vec(b), x, rows_t, rtol=1e-12, maxit=20000, check_every=4, precond=precond)
        print(f"pcg {precond}: {it} iterations, relres {rr:.1e}", flush=True)
        assert rr < 1e-12 and rel(x.cpu().numpy(), xd) < 1e-9
    # closer to the FP64 tangent's solution
    x64 = spla.spsolve(_host_csr(g64, v64, nd).tocsc(), b)
    x32 = spla.spsolve(_host_csr(g32, v32, nd).tocsc(), b)
    e32, ed = rel(x32, x64), rel(xd, x64)
    print(f"solution vs FP64 tangent: fp32 {e32:.2e}, fp32 + fp64 blocks {ed:.2e}", flush=True)
    assert ed < e32 / 4


@pytest.mark.parametrize("case", ["C3D8 reversed element order", "C3D20", "C3D20 colour path", "C3D8TL",
                                  "C3D8 4 pipeline chunks"])
def test_fp32_d64_corrections_on_other_assembly_paths(case):
    """The diagonal corrections of EWFE_F32_D64 from the separate rebuild pass (meshes the structured 8-node gather does
    not take, 20-node elements through their gather or the colour path) and from the structured gather of a chunked
    pipeline (each chunk's rows on the second stream): the FP32 values bitwise the EWFE_F32 ones, diagonal block +
    correction = minus the sum of the off-diagonal FP32 blocks of its rows."""
    from edelweissfe_amd._lib import check, lib
    print(f"fp32-d64 corrections: {case}", flush=True)
    el = case.split()[0]
    X, conn, _ = jittered_box(4, el, seed=7)
    if "reversed" in case:
        conn = conn[::-1].copy()
    mat, mp = ("neohookeanwa", [80.0, 200.0]) if el.endswith("TL") else ("vonmises", J2)
    out = {}
    for dt in ("fp32", "fp32-d64"):
        g = GpuEval(el, mat, mp, X, conn, pipelineChunks=4 if "chunks" in case else 0)
        if "colour" in case:
            check(lib().ewfe_plan_set_row_gather(g.asm._plan, 0))
        if "chunks" in case:
            assert g.asm.pipelineChunks == 4 and g.asm.structuredMesh
        nd = g.dm.nDof
        u = np.zeros((len(X), 3))
        u[:, 2] = 5e-3 * X[:, 2]
        u[:, 0] = -0.15e-2 * X[:, 0]
        U = g.vec(u.ravel())
        vals = g.asm.newValues(dt)
        P, F = g.dm.constructDofVector(), g.dm.constructDofVector()
        g.asm.evaluate(U, U, P, F, vals)
        g.torch.cuda.synchronize()
        out[dt] = (g, vals)
    g, vd = out["fp32-d64"]
    assert bool((g.asm.csrValues(vd) == out["fp32"][1]).all().item())
    nnz = g.asm.nnz
    E = vd[(4 * nnz + 255) // 256 * 64:].cpu().numpy().view(np.float64).reshape(-1, 9)
    C = _host_csr(g, vd, nd).tocoo()
    offd = C.row // 3 != C.col // 3
    S = np.zeros((nd, 3))
    np.add.at(S, (C.row[offd], C.col[offd] % 3), C.data[offd])
    D32 = np.zeros((nd, 3))
    np.add.at(D32, (C.row[~offd], C.col[~offd] % 3), C.data[~offd])
    scale = np.abs(C.data).max()
    err = np.abs(D32.reshape(-1, 9) + E + S.reshape(-1, 9)).max() / scale
    print(f"|D32 + E + S| / max|K| = {err:.1e}, max|E| / max|K| = {np.abs(E).max() / scale:.1e}", flush=True)
    assert err <= 1e-15 and 0 < np.abs(E).max() <= 1e-5 * scale
