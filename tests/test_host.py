"""CPU tests of the product's host side: the C ABI library loads and exports every declared symbol, the
host-built CSR pattern / slot map / colouring equal the oracle's scipy restatement bit for bit, boxGen,
DofManager and the time stepper follow the reference's ordering and semantics."""
import ctypes
import os
import re

import numpy as np
import pytest

import decks
from edelweissfe_amd import _lib
from edelweissfe_amd.model import FEModel, boxGenArrays, jitterNodes
from edelweissfe_amd.stepping import IncrementController, ReachedMaxIncrements, ReachedMinIncrementSize
from oracle import oracle as O

EWFE_EINVAL = -1  # include/ewfe.h

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "ewfe.h")).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[\w\s\*]+?\b(ewfe_\w+)\s*\(", txt, flags=re.M)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    syms = declared_symbols()
    assert len(syms) >= 14
    for s in syms:
        assert hasattr(L, s), s
        assert s in _lib.SIGNATURES, s
    assert L.ewfe_version() >= 1
    assert L.ewfe_state_size(_lib.EWFE_VON_MISES) == 7


def host_plan(kind, conn, node_dof, n_dof):
    L = _lib.lib()
    plan = ctypes.c_void_p()
    nnz = ctypes.c_int64()
    conn = np.ascontiguousarray(conn, dtype=np.int32)
    node_dof = np.ascontiguousarray(node_dof, dtype=np.int64)
    _lib.check(L.ewfe_plan_create_host(ctypes.byref(plan), kind, len(conn), _lib.ptr(conn), len(node_dof),
                                       _lib.ptr(node_dof), n_dof, ctypes.byref(nnz), 0))
    indptr = np.zeros(n_dof + 1, dtype=np.int32)
    indices = np.zeros(nnz.value, dtype=np.int32)
    _lib.check(L.ewfe_plan_pattern(plan, _lib.ptr(indptr), _lib.ptr(indices)))
    order = np.zeros(len(conn), dtype=np.int64)
    nc = ctypes.c_int32()
    _lib.check(L.ewfe_plan_element_order(plan, _lib.ptr(order), ctypes.byref(nc)))
    L.ewfe_plan_destroy(plan)
    return indptr, indices, order, nc.value


def _check_pattern(kind, X, conn):
    base, nDof, edofs = O.dof_numbering(len(X), conn)
    I, J = O.vij_pattern(edofs)
    ip_ref, ix_ref, _ = O.csr_pattern(I, J, nDof)
    ip, ix, order, ncol = host_plan(kind, conn, base, nDof)
    assert np.array_equal(ip, ip_ref) and np.array_equal(ix, ix_ref)
    assert np.array_equal(np.sort(order), np.arange(len(conn)))
    return order, ncol


@pytest.mark.parametrize("el,n", [("C3D8", 3), ("C3D8", (5, 2, 4)), ("C3D20", 2), ("C3D20", (1, 3, 2))])
def test_pattern_bitexact_boxgen(el, n):
    n = (n, n, n) if isinstance(n, int) else n
    X, conn, _ = O.boxgen(*n, nnodes=8 if el == "C3D8" else 20)
    order, ncol = _check_pattern(_lib.ELEMENT_KINDS[el], X, conn)
    if el == "C3D8":  # 8-node elements are gathered per row node: no colouring (plan order = evaluation order)
        assert ncol == 1


@pytest.mark.parametrize("deck", ["C3D8", "C3D20", "CantileverBeamHexa8"])
def test_pattern_bitexact_explicit_meshes(deck):
    fx = decks.load(deck)
    X, conn, _ = decks.mesh(fx, O.boxgen)
    _check_pattern(_lib.ELEMENT_KINDS[fx["el_type"]], X, conn)


def test_pattern_shuffled_numbering_and_unused_nodes():
    X, conn, _ = O.boxgen(3, 2, 2)
    rng = np.random.default_rng(4)
    perm = rng.permutation(len(X))  # new node order
    inv = np.argsort(perm)
    X2 = np.concatenate([X[perm], np.zeros((3, 3))])  # three trailing nodes without elements
    conn2 = inv[conn].astype(np.int32)
    conn2 = conn2[rng.permutation(len(conn2))]
    _check_pattern(_lib.EWFE_C3D8, X2, conn2)


def test_colouring_is_conflict_free():
    X, conn, _ = O.boxgen(4, 3, 5, nnodes=20)
    base, nDof, _ = O.dof_numbering(len(X), conn)
    _, _, order, ncol = host_plan(_lib.EWFE_C3D20, conn, base, nDof)
    # recover colour boundaries: the plan keeps each colour contiguous and in element order
    sizes = []
    start = 0
    for p in range(1, len(order) + 1):
        if p == len(order) or order[p] < order[p - 1]:
            sizes.append(p - start)
            start = p
    assert len(sizes) == ncol
    s = 0
    for sz in sizes:
        nodes = conn[order[s: s + sz]].ravel()
        assert len(np.unique(nodes)) == len(nodes)
        s += sz


def test_plan_rejects_bad_input():
    L = _lib.lib()
    plan = ctypes.c_void_p()
    nnz = ctypes.c_int64()
    conn = np.array([[0, 1, 2, 3, 4, 5, 6, 99]], dtype=np.int32)
    node_dof = np.arange(0, 24, 3, dtype=np.int64)
    rc = L.ewfe_plan_create_host(ctypes.byref(plan), 0, 1, _lib.ptr(conn), 8, _lib.ptr(node_dof), 24,
                                 ctypes.byref(nnz), 0)
    assert rc == -1 and b"outside" in L.ewfe_last_error()
    with pytest.raises(_lib.EwfeError):
        _lib.check(rc)


def test_plan_rejects_empty_and_ragged_input():
    """No elements, no nodes or no dofs fail with a message; a node numbered past n_nodes is refused."""
    L = _lib.lib()
    plan = ctypes.c_void_p()
    nnz = ctypes.c_int64()
    conn = np.zeros((0, 8), dtype=np.int32)
    node_dof = np.arange(0, 24, 3, dtype=np.int64)
    for n_el, n_nodes, n_dof in ((0, 8, 24), (1, 0, 24), (1, 8, 0)):
        c = conn if n_el == 0 else np.arange(8, dtype=np.int32).reshape(1, 8)
        rc = L.ewfe_plan_create_host(ctypes.byref(plan), 0, n_el, _lib.ptr(c), n_nodes, _lib.ptr(node_dof), n_dof,
                                     ctypes.byref(nnz), 0)
        assert rc != 0 and L.ewfe_last_error()
    c = np.arange(8, dtype=np.int32).reshape(1, 8)
    rc = L.ewfe_plan_create_host(ctypes.byref(plan), 0, 1, _lib.ptr(c), 7, _lib.ptr(node_dof), 24, ctypes.byref(nnz), 0)
    assert rc != 0


def test_material_table_rejects_bad_indices():
    """ewfe_plan_set_materials checks its arguments before any device work: an index outside [0, n_mat) or a
    missing table fails with EWFE_EINVAL and a message; n_mat = 0 (no table) succeeds."""
    L = _lib.lib()
    X, conn, _, _ = boxGenArrays(2, 2, 2, elType="C3D8")
    base, nDof, _ = O.dof_numbering(len(X), conn)
    plan = ctypes.c_void_p()
    nnz = ctypes.c_int64()
    conn = np.ascontiguousarray(conn, dtype=np.int32)
    node_dof = np.ascontiguousarray(base, dtype=np.int64)
    _lib.check(L.ewfe_plan_create_host(ctypes.byref(plan), _lib.EWFE_C3D8, len(conn), _lib.ptr(conn), len(node_dof),
                                       _lib.ptr(node_dof), nDof, ctypes.byref(nnz), 0))
    try:
        table = np.zeros((2, 8))
        idx = np.zeros(len(conn), dtype=np.int32)
        idx[3] = 2  # outside [0, 2)
        rc = L.ewfe_plan_set_materials(plan, 2, _lib.ptr(table), _lib.ptr(idx))
        assert rc == EWFE_EINVAL and b"outside" in L.ewfe_last_error()
        idx[3] = -1
        assert L.ewfe_plan_set_materials(plan, 2, _lib.ptr(table), _lib.ptr(idx)) == EWFE_EINVAL
        assert L.ewfe_plan_set_materials(plan, 2, None, _lib.ptr(idx)) == EWFE_EINVAL
        assert L.ewfe_plan_set_materials(plan, -1, _lib.ptr(table), _lib.ptr(idx)) == EWFE_EINVAL
        assert L.ewfe_plan_set_materials(plan, 0, None, None) == 0
    finally:
        L.ewfe_plan_destroy(plan)


def test_block_materials_of_one_kind():
    """ElementBlock.setMaterials: several materials of one kind with one index per element; the per-element
    parameter sets follow the index, the per-element first parameter overrides them; mixed kinds and bad
    indices are rejected."""
    from edelweissfe_amd.model import ElementBlock, Material

    X, conn, _, _ = boxGenArrays(2, 1, 1, elType="C3D8")
    b = ElementBlock("C3D8", conn)
    b.setMaterials([Material("linearelastic", [30000.0, 0.2]), Material("linearelastic", [5000.0, 0.49])], [1, 0])
    mp = b.elementMaterialProperties()
    assert mp.shape == (2, 8) and np.array_equal(mp[:, :2], [[5000.0, 0.49], [30000.0, 0.2]])
    b.elParams = np.array([7.0, 8.0])
    assert np.array_equal(b.elementMaterialProperties()[:, 0], [7.0, 8.0])
    with pytest.raises(NotImplementedError):
        b.setMaterials([Material("linearelastic", [1.0, 0.2]), Material("vonmises", [1.0, 0.3, 1.0, 0.0])], [0, 1])
    with pytest.raises(ValueError):
        b.setMaterials([Material("linearelastic", [1.0, 0.2]), Material("linearelastic", [2.0, 0.2])], [0, 2])


@pytest.mark.parametrize("el,n", [("C3D8", (3, 2, 4)), ("C3D20", (2, 3, 1))])
def test_boxgen_matches_reference_order(el, n):
    npe = 8 if el == "C3D8" else 20
    Xo, conno, setso = O.boxgen(*n, lX=2.0, lY=3.0, lZ=0.5, nnodes=npe)
    X, conn, sets, grid = boxGenArrays(*n, lX=2.0, lY=3.0, lZ=0.5, elType=el)
    assert np.array_equal(X, Xo) and np.array_equal(conn, conno)
    for k, v in setso.items():
        if k != "all":
            assert np.array_equal(sets[k], v), k


def test_jitter_keeps_faces_planar():
    X, conn, sets, grid = boxGenArrays(4, 4, 4, elType="C3D8")
    Xj = jitterNodes(X, grid, (5, 5, 5), np.array([0.25] * 3), 0.05, seed=0)
    assert np.abs(Xj - X).max() <= 0.05 * 0.25 + 1e-15
    assert np.all(Xj[sets["gen_back"], 2] == 0.0) and np.all(Xj[sets["gen_front"], 2] == 1.0)
    assert np.all(Xj[sets["gen_left"], 0] == 0.0) and np.all(Xj[sets["gen_right"], 0] == 1.0)


def test_dofmanager_matches_oracle_numbering():
    from edelweissfe_amd.dofmanager import DofManager
    from edelweissfe_amd.model import ElementBlock

    fx = decks.load("C3D8")
    X, conn, sets = decks.mesh(fx, O.boxgen)
    m = FEModel()
    m.addNodes(X)
    m.addElementBlock(ElementBlock("C3D8", conn))
    m.nodeSets.update({k: np.asarray(v) for k, v in sets.items()})
    dm = DofManager(m, device="cpu")
    base, nDof, _ = O.dof_numbering(len(X), conn)
    assert dm.nDof == nDof and np.array_equal(dm.nodeDofBase, base)
    assert dm.nAccumulatedNodalFluxesFieldwise["displacement"] == 3 * conn.size
    idx = dm.idcsOfFieldsOnNodeSetsInDofVector["displacement"]["nSetRight"]
    assert np.array_equal(idx, (base[sets["nSetRight"]][:, None] + np.arange(3)).ravel())


def test_time_stepper_matches_oracle_with_cutbacks():
    seq_a, seq_b = [], []
    ts = IncrementController(0.0, 2.0, 0.3, 0.25, 1e-3, 50)
    ora = O.time_steps(2.0, 0.3, 0.25, 1e-3, 50)
    ga, gb = iter(ts), ora.gen()
    for k in range(40):
        try:
            a = next(ga)
        except StopIteration:
            a = None
        try:
            b = next(gb)
        except StopIteration:
            b = None
        if a is None or b is None:
            assert a is None and b is None
            break
        seq_a.append((a.number, a.stepProgressIncrement, a.stepProgress, a.timeIncrement))
        seq_b.append((b["number"], b["stepProgressIncrement"], b["stepProgress"], b["timeIncrement"]))
        if k in (3, 4, 9):
            ts.discard(0.25)
            ora.discard(0.25)
        if k == 6:
            ts.holdSize()
            ora.allowInc = False
    assert seq_a == seq_b and len(seq_a) > 8
    t2 = IncrementController(0.0, 1.0, 1.0, 1.0, 1.0, 10)
    g = iter(t2)
    next(g)
    with pytest.raises(ReachedMinIncrementSize):
        t2.discard(0.5)
    # a size already at the minimum cannot shrink; more than maxCount increments stop the step
    t3 = IncrementController(0.0, 1.0, 0.1, 0.1, 0.1, 100)
    next(t3), next(t3)
    with pytest.raises(ReachedMinIncrementSize):
        t3.discard(0.5)
    t4 = IncrementController(5.0, 1.0, 0.01, 0.01, 0.01, 3)
    steps = []
    with pytest.raises(ReachedMaxIncrements):
        for s in t4:
            steps.append((s.number, s.totalTime))
    assert [n for n, _ in steps] == [0, 1, 2, 3, 4] and abs(steps[-1][1] - 5.04) < 1e-12


def _pipeline_orders(conn, node_dof, n_dof, chunks):
    L = _lib.lib()
    plan = ctypes.c_void_p()
    nnz = ctypes.c_int64()
    conn = np.ascontiguousarray(conn, dtype=np.int32)
    node_dof = np.ascontiguousarray(node_dof, dtype=np.int64)
    _lib.check(L.ewfe_plan_create_host(ctypes.byref(plan), _lib.EWFE_C3D8, len(conn), _lib.ptr(conn), len(node_dof),
                                       _lib.ptr(node_dof), n_dof, ctypes.byref(nnz), chunks))
    nch, st = ctypes.c_int32(), ctypes.c_int32()
    _lib.check(L.ewfe_plan_pipeline_info(plan, ctypes.byref(nch), ctypes.byref(st)))
    rows = np.unique(node_dof[np.unique(conn)])  # row nodes: nodes with elements, ascending first dof
    pe, pr = np.zeros(nch.value + 1, np.int64), np.zeros(nch.value + 1, np.int64)
    eo, ro = np.zeros(len(conn), np.int32), np.zeros(len(rows), np.int32)
    _lib.check(L.ewfe_plan_pipeline_orders(plan, _lib.ptr(pe), _lib.ptr(pr), _lib.ptr(eo), _lib.ptr(ro)))
    L.ewfe_plan_destroy(plan)
    return nch.value, bool(st.value), pe, pr, eo, ro, rows


def _check_pipeline(conn, node_dof, n_dof, chunks):
    K, structured, pe, pr, eo, ro, rows = _pipeline_orders(conn, node_dof, n_dof, chunks)
    assert 1 <= K and pe[0] == 0 and pe[-1] == len(conn) and pr[0] == 0 and pr[-1] == len(rows)
    assert np.all(np.diff(pe) >= 0) and np.all(np.diff(pr) >= 0)
    assert np.array_equal(np.sort(eo), np.arange(len(conn)))  # every element evaluated once
    assert np.array_equal(np.sort(ro), np.arange(len(rows)))  # every row node gathered once
    el_chunk = np.empty(len(conn), np.int64)
    for c in range(K):
        el_chunk[eo[pe[c]:pe[c + 1]]] = c
    # a row node is gathered with the chunk that evaluates the last of its elements
    node_of_row = {int(d): i for i, d in enumerate(rows)}
    last = np.zeros(len(rows), np.int64)
    for e, nodes in enumerate(conn):
        for nd in nodes:
            r = node_of_row[int(node_dof[nd])]
            last[r] = max(last[r], el_chunk[e])
    for c in range(K):
        assert np.all(last[ro[pr[c]:pr[c + 1]]] == c)
    return K, structured, pe, eo


@pytest.mark.parametrize("n,chunks", [((4, 3, 5), 1), ((4, 3, 5), 3), ((7, 20, 18), 4), ((6, 2, 2), 6),
                                      ((2, 2, 2), 9)])
def test_pipeline_orders_boxgen(n, chunks):
    X, conn, _ = O.boxgen(*n)
    base, nDof, _ = O.dof_numbering(len(X), conn)
    K, structured, pe, eo = _check_pipeline(conn, base, nDof, chunks)
    assert structured and K == min(chunks, n[0])
    # structured chunks are slabs of element layers along the slowest axis
    layer = eo // (n[1] * n[2])
    for c in range(K):
        lay = np.unique(layer[pe[c]:pe[c + 1]])
        assert np.array_equal(lay, np.arange(lay[0], lay[-1] + 1))


@pytest.mark.parametrize("chunks", [1, 2, 5])
def test_pipeline_orders_unstructured(chunks):
    X, conn, _ = O.boxgen(4, 3, 3)
    rng = np.random.default_rng(7)
    perm = rng.permutation(len(X))
    inv = np.argsort(perm)
    conn2 = inv[conn].astype(np.int32)[rng.permutation(len(conn))]
    base, nDof, _ = O.dof_numbering(len(X), conn2)
    K, structured, _, _ = _check_pipeline(conn2, base, nDof, chunks)
    assert not structured and K == chunks


_CORNER_LUT = [0, 1, 4, 5, 3, 2, 7, 6]  # corner offset index 4x + 2y + z -> C3D8 local node (boxgen.py:171-185)


def _axis_prefix(i, n):
    return 0 if i == 0 else (3 * n - 2 if i >= n else 3 * i - 1)


def _structured_gather_plan(E, r):
    """The structured gather kernel's index arithmetic for row node r of an E[0] x E[1] x E[2] boxGen grid:
    CSR offset, row length, and per neighbour rank m the neighbour node and the contributing
    (element, local a, local b) in summation order."""
    N = [e + 1 for e in E]
    i = [r // (N[1] * N[2]), (r // N[2]) % N[1], r % N[2]]
    lo = [int(i[k] > 0) for k in range(3)]
    c = [1 + lo[k] + int(i[k] < N[k] - 1) for k in range(3)]
    deg = c[0] * c[1] * c[2]
    base = 9 * (_axis_prefix(i[0], N[0]) * (3 * N[1] - 2) * (3 * N[2] - 2)
                + c[0] * _axis_prefix(i[1], N[1]) * (3 * N[2] - 2) + c[0] * c[1] * _axis_prefix(i[2], N[2]))
    blocks = []
    for m in range(deg):
        d = [m // (c[1] * c[2]) - lo[0], (m // c[2]) % c[1] - lo[1], m % c[2] - lo[2]]
        nb = ((i[0] + d[0]) * N[1] + i[1] + d[1]) * N[2] + i[2] + d[2]
        contrib = []
        for o in range(8):
            oo = [o >> 2, (o >> 1) & 1, o & 1]
            e = [i[k] - 1 + oo[k] for k in range(3)]
            if all((d[k] == 0 or d[k] == 2 * oo[k] - 1) and 0 <= e[k] < E[k] for k in range(3)):
                a = _CORNER_LUT[4 * (1 - oo[0]) + 2 * (1 - oo[1]) + (1 - oo[2])]
                b = _CORNER_LUT[4 * (1 - oo[0] + d[0]) + 2 * (1 - oo[1] + d[1]) + (1 - oo[2] + d[2])]
                contrib.append(((e[0] * E[1] + e[1]) * E[2] + e[2], a, b))
        blocks.append((nb, contrib))
    return base, 3 * deg, blocks


@pytest.mark.parametrize("E", [(1, 1, 1), (3, 2, 4), (2, 5, 1), (4, 4, 4)])
def test_structured_gather_index_math_matches_pattern(E):
    X, conn, _ = O.boxgen(*E)
    base_dof, nDof, _ = O.dof_numbering(len(X), conn)
    ip, ix, _, _ = host_plan(_lib.EWFE_C3D8, conn, base_dof, nDof)
    conn = np.asarray(conn)
    inc = [[] for _ in range(len(X))]
    for e, nodes in enumerate(conn):
        for a, nd in enumerate(nodes):
            inc[nd].append((e, a))
    for r in range(len(X)):
        base, L, blocks = _structured_gather_plan(E, r)
        assert base == ip[3 * r] and L == ip[3 * r + 1] - ip[3 * r] == ip[3 * r + 3] - ip[3 * r + 2]
        assert len(blocks) * 3 == L
        for m, (nb, contrib) in enumerate(blocks):
            assert ix[base + 3 * m] == 3 * nb  # column of neighbour block m
            # contributions: every element holding both nodes, ascending element, with their local indices
            ref = [(e, a, list(conn[e]).index(nb)) for e, a in inc[r] if nb in conn[e]]
            assert contrib == sorted(ref)


def test_reference_shaped_model_adapter():
    """A model in the reference's object shape (nodes, elements, sections, sets, NodeField) becomes one element
    block per element type and material kind, with the elements' order kept, materials of one kind in a block's
    table, sets mapped to indices, and results written back into the NodeField's rows."""
    import torch

    import refshape as R
    from edelweissfe_amd.adapter import ReferenceModelAdapter
    from edelweissfe_amd.dofmanager import DofManager

    X, conn, sets, _ = boxGenArrays(4, 2, 3, elType="C3D8")
    le1 = {"name": "LinearElastic", "properties": np.array([210000.0, 0.3])}
    j2 = {"name": "VonMises", "properties": np.array([210000.0, 0.3, 250.0, 1000.0, 0.0, 0.0])}
    le2 = {"name": "LinearElastic", "properties": np.array([30000.0, 0.2])}
    ref = R.model_from_arrays(X, [("C3D8", conn[:8], le1), ("C3D8", conn[8:16], j2), ("C3D8", conn[16:], le2)],
                              {"bottom": sets["gen_back"], "top": sets["gen_front"]})
    nf = ref.nodeFields["displacement"]
    nf.nodes.reverse()  # rows of the NodeField in another order than the dofs
    nf._indicesOfNodesInArray = {n: i for i, n in enumerate(nf.nodes)}
    ad = ReferenceModelAdapter(ref)
    gm = ad.model
    assert np.allclose(gm.nodeCoordinates, X)
    assert [(b.elType, b.material.name) for b in gm.elementBlocks] == [("C3D8", "linearelastic"), ("C3D8", "vonmises")]
    b0, b1 = gm.elementBlocks
    assert np.array_equal(b0.conn, conn[np.r_[0:8, 16:len(conn)]]) and np.array_equal(b1.conn, conn[8:16])
    assert np.array_equal(b0.labels, np.r_[1:9, 17:len(conn) + 1])
    assert np.array_equal(b0.materialIndex, np.repeat([0, 1], [8, len(conn) - 16]))
    assert np.array_equal(b0.materialParameterTable()[1, :2], [30000.0, 0.2])
    assert np.array_equal(gm.nodeSets["top"], sets["gen_front"])
    assert gm.elementSets["block1"][0] == 1 and np.array_equal(gm.elementSets["block1"][1], np.arange(8))
    assert gm.elementSets["block2"][0] == 0 and np.array_equal(gm.elementSets["block2"][1], np.arange(8, 16))
    dm = DofManager(gm, device="cpu")
    U = torch.arange(dm.nDof, dtype=torch.float64)
    ad.writeBack(dm, U, 2 * U, 3 * U, 1.5)
    for r, n in enumerate(dm.nodesWithDofs):
        row = nf._indicesOfNodesInArray[list(ref.nodes.values())[n]]
        assert np.array_equal(nf["U"][row], [3 * r, 3 * r + 1, 3 * r + 2])
        assert np.array_equal(nf["dU"][row], 3 * nf["U"][row])
    assert ref.time == 1.5
    assert np.array_equal(ad.readDisplacement(dm), U.numpy())


def test_value_array_sizes_and_fp32_solver_options():
    """ewfe_vals_bytes: 8 nnz (FP64), 4 nnz (FP32, matrix-core FP32), and for FP32 values with FP64 diagonal
    corrections the FP32 values up to a 256-byte boundary plus 9 doubles per node; other dtypes are rejected. The GPU
    solver's FP32 options: corrected diagonal blocks by default, the FP32 PCG tolerance derived (None) unless given,
    given as a number or text, unknown options rejected."""
    L = _lib.lib()
    X, conn, _ = O.boxgen(3, 2, 4)
    base, nDof, _ = O.dof_numbering(len(X), conn)
    plan, nnz = ctypes.c_void_p(), ctypes.c_int64()
    _lib.check(L.ewfe_plan_create_host(ctypes.byref(plan), 0, len(conn), _lib.ptr(np.ascontiguousarray(conn)),
                                       len(X), _lib.ptr(np.ascontiguousarray(base)), nDof, ctypes.byref(nnz), 0))
    try:
        n = nnz.value
        assert L.ewfe_vals_bytes(plan, _lib.EWFE_F64) == 8 * n
        assert L.ewfe_vals_bytes(plan, _lib.EWFE_F32) == 4 * n == L.ewfe_vals_bytes(plan, _lib.EWFE_F32_MFMA)
        assert L.ewfe_vals_bytes(plan, _lib.EWFE_F32_D64) == (4 * n + 255) // 256 * 256 + 72 * len(X)
        assert L.ewfe_vals_bytes(plan, _lib.EWFE_F16) < 0 and L.ewfe_vals_bytes(plan, 9) < 0
    finally:
        L.ewfe_plan_destroy(plan)

    from edelweissfe_amd.solvers.nonlinearimplicitstaticgpu import NISTGPU

    s = NISTGPU()
    assert s.options["fp32TangentDiagonal"] == "rebuilt" and s.options["pcgToleranceFP32"] is None
    s = NISTGPU(pcgToleranceFP32="1e-5", fp32TangentDiagonal="assembled")
    assert s.options["pcgToleranceFP32"] == 1e-5 and s.options["fp32TangentDiagonal"] == "assembled"
    assert NISTGPU(pcgToleranceFP32=2e-6).options["pcgToleranceFP32"] == 2e-6
    assert NISTGPU(pcgToleranceFP32="none").options["pcgToleranceFP32"] is None
    with pytest.raises(AttributeError):
        NISTGPU(fp32Diagonal="rebuilt")
