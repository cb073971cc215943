"""The input-deck reader (edelweissfe_amd.inputfile) on the reference's own regression decks.

CPU: the models it builds equal those of the golden fixtures (parsed independently by
tests/golden/make_fixtures.py), and the oracle's Newton loop run on them reproduces the decks' U.ref (1e-6
absolute, the reference's bar, ew/_cli/_run_tests_edelweissfe.py:96-109). GPU: runReferenceTest (parse, build,
NISTGPU, compare with U.ref) on every deck.
"""
import os

import numpy as np
import pytest

import decks
from edelweissfe_amd import inputfile as IF
from oracle import oracle as O

DECK_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "decks")
FIXTURE_DECKS = ["TensionBarHexa8", "SimpleBeamHexa8", "CantileverBeamHexa8", "C3D8", "C3D20", "WallShearHexa8",
                 "WallShearHexa20", "LinearElasticIsotropic"]
# decks read only through the reader: BodyForce3D (*include of the mesh), BoxGen (C3D20R, findClosestNode),
# AnalyticalFieldsDirichlet (C3D20R torsion bar, Dirichlet values scaled by scalarExpression fields)
# WallShearHexa8GMRES (C3D8N, the reference's pyamg-GMRES solver: the multigrid PCG here)
# AnalyticalFieldsSetField (J2 with kappa set from a scalarExpression field at the points)
# and SetField (kappa set uniformly, then the stress set at every point after a support is released);
# C3D8TL (a C3D8TL beam of Marmot's StVenantKirchhoffIsotropic bent to a 20 mm tip deflection in 50 increments)
ALL_DECKS = FIXTURE_DECKS + ["BodyForce3D", "BoxGen", "AnalyticalFieldsDirichlet", "WallShearHexa8GMRES",
                             "AnalyticalFieldsSetField", "SetField", "SetDef", "StatusFile", "ComputeTimeMonitor",
                             "PipeGen", "CubitGen", "ADLinearElastic", "C3D8TL"]


def _deck(name):
    return IF.parseInputFile(os.path.join(DECK_DIR, name, "test.inp"))


def _oracle_step(model, step):
    """The oracle's step description of a step built by the reader (loads of earlier steps are not carried:
    decks whose earlier steps only set boundary conditions)."""
    a = step.actions
    st = {k: getattr(step, f) for k, f in (("maxInc", "maxIncrementSize"), ("minInc", "minIncrementSize"),
                                          ("maxNumInc", "maxNumberIncrements"), ("maxIter", "maxIter"),
                                          ("stepLength", "length"), ("startInc", "startIncrementSize"))}
    opts = a["options"]["NISTSolver"].options if "NISTSolver" in a["options"] else {}
    st["extrapolation"] = opts.get("extrapolation", "linear")
    st["dirichlet"] = [(d.nSet, {c: d.delta[:, j] * d.active for j, c in enumerate(d.components)})
                       for d in a["dirichlet"].values()]
    st["nodeforces"] = [(f.nodeSet, list(f.nodeForcesDelta[0])) for f in a["nodeforces"].values()]
    st["bodyforces"] = []
    for b in a["bodyforce"].values():
        assert b.amplitude(0.5) == 0.5  # the oracle scales body forces linearly with the step progress
        assert len(b.elementSet) == len(model.elementBlocks[0])
        st["bodyforces"].append(list(b.delta))
    st["pressures"] = []
    for d in a.get("distributedload", {}).values():
        assert d.loadType == "pressure" and np.all(d.magnitudeAtStepStart == 0)
        for face, (_, els) in d.surface.items():
            st["pressures"].append((els, face, float(d.delta[0]), d.amplitude))
    return st


def _steps(model, inp):
    """Every step of the deck, with the step-end actions applied between steps as the solver does."""
    for step, solverName in IF.createStepsFromInputFile(model, inp):
        yield step, solverName
        for typ in step.actions.values():
            for action in typ.values():
                action.applyAtStepEnd(model)


def test_parser_structure():
    d = _deck("BodyForce3D")  # *include of mesh.inp in the middle of the deck
    assert len(d["*node"]) == 1 and len(d["*element"]) == 1 and d["*element"][0]["type"] == "C3D8"
    assert d["*step"][0]["maxinc"] == "1e-1" and len(d["*step"][0]["data"]) == 3
    assert IF.splitLineAtCommas("bodyforce, name=b, forceVector='0.0, 10.0, 20.0', f(t)=t") == [
        "bodyforce", "name=b", "forceVector=0.0, 10.0, 20.0", "f(t)=t"]
    amp = IF._Amplitude("t**2 + 0.5*sin(pi*t)")
    assert abs(amp(0.3) - (0.09 + 0.5 * np.sin(np.pi * 0.3))) < 1e-15
    with pytest.raises(ValueError):
        IF._Amplitude("__import__('os')")


@pytest.mark.parametrize("name", FIXTURE_DECKS)
def test_reader_builds_the_fixture_models(name):
    model = IF.createModelFromInputFile(_deck(name))
    fx = decks.load(name)
    X, conn, sets = decks.mesh(fx, O.boxgen)
    block = model.elementBlocks[0]
    assert block.elType == fx["el_type"]
    assert np.array_equal(model.nodeCoordinates, X) and np.array_equal(block.conn, conn)
    assert block.material.name == fx["material"] and np.allclose(block.material.materialProperties[: len(fx["mp"])],
                                                                 fx["mp"])
    for act in fx["step"]["dirichlet"] + fx["step"]["nodeforces"]:
        assert np.array_equal(np.sort(model.nodeSets[act["nSet"]]), np.sort(sets[act["nSet"]]))
    for step, _ in _steps(model, _deck(name)):
        st = _oracle_step(model, step)  # the last step's description, taken before its step-end actions
    ref = decks.step(fx, sets)
    assert st["extrapolation"] == ref["extrapolation"] and st["maxInc"] == ref["maxInc"]
    for (n1, c1), (n2, c2) in zip(st["dirichlet"], ref["dirichlet"]):
        assert np.array_equal(np.sort(n1), np.sort(n2)) and sorted(c1) == sorted(c2)
        assert all(np.all(np.asarray(c1[k]) == c2[k]) for k in c1)
    assert len(st["dirichlet"]) == len(ref["dirichlet"]) and len(st["nodeforces"]) == len(ref["nodeforces"])
    assert st["bodyforces"] == ref["bodyforces"]
    assert len(st["pressures"]) == len(ref.get("pressures", []))
    for (e1, f1, m1, a1), (e2, f2, m2, a2) in zip(st["pressures"], ref.get("pressures", [])):
        assert np.array_equal(np.sort(e1), np.sort(e2)) and f1 == f2 and m1 == m2 and a1(0.3) == a2(0.3)


@pytest.mark.parametrize("name", ALL_DECKS)
def test_reader_and_oracle_reproduce_uref(name):
    """Deck -> reader -> oracle Newton loop -> U.ref (CPU)."""
    inp = _deck(name)
    model = IF.createModelFromInputFile(inp)
    U = _oracle_run(model, inp)
    Uref = np.loadtxt(os.path.join(DECK_DIR, name, "U.ref"))
    assert np.abs(U - Uref).max() < 1e-6
    assert np.abs(U - Uref).max() < 1e-7 * max(1.0, np.abs(Uref).max())


@pytest.mark.gpu
@pytest.mark.parametrize("name", ALL_DECKS)
def test_run_reference_decks_from_input_files(name):
    """Deck -> reader -> NISTGPU -> U.ref, like the reference's test runner."""
    print(f"deck {name}", flush=True)
    passed, err = IF.runReferenceTest(os.path.join(DECK_DIR, name), solverOptions={"pcgTolerance": 1e-13})
    print(f"deck {name}: max |U - U.ref| = {err:.3e}", flush=True)
    assert passed


@pytest.mark.gpu
def test_element_result_collector_on_tension_bar():
    """Quadrature-point stresses gathered like the reference's ElementResultCollector: away from the ends of
    the TensionBarHexa8 bar (5 x 5 cross-section; 100 N in -z on the z = 0 face, the z = 500 face fixed: tension),
    sigma_zz = 4 at every point of a layer."""
    from edelweissfe_amd.results import ElementResultCollector

    model = IF.finiteElementSimulation(_deck("TensionBarHexa8"), solverOptions={"pcgTolerance": 1e-13})
    block = model.elementBlocks[0]
    layer = np.nonzero(np.arange(len(block)) % 30 == 15)[0]  # boxGen: z fastest, 30 layers -> mid-length layer
    col = ElementResultCollector(model.solver, model, layer, range(8), "stress")
    sig = col.getCurrentResults()
    assert sig.shape == (16, 8, 6)
    assert np.abs(sig[:, :, 2] - 4.0).max() < 1e-3
    one = ElementResultCollector(model.solver, model, "all", slice(0, 1), "strain").getCurrentResults()
    assert one.shape == (len(block), 6)


@pytest.mark.parametrize("elType", ["C3D8", "C3D20"])
def test_boxgen_surfaces_carry_pressure_on_their_faces(tmp_path, elType):
    """boxGen's element sets and surfaces (boxgen.py:376-446) read from a deck: a pressure p on each of the six
    surfaces of a 3 x 4 x 2 box (5 x 6 x 2.5) sums to -p A n (outward normal n of that face of the box) and loads
    only the nodes on that face; the surface's element set is the layer of elements along it."""
    deck = tmp_path / "box.inp"
    deck.write_text(f"""*material, name=LinearElastic, id=mat
30000.0, 0.15
*job, name=job, domain=3d
*solver, solver=NIST, name=theSolver
*modelGenerator, generator=boxGen, name=gen
nX=3
nY=4
nZ=2
lX=5
lY=6
lZ=2.5
elType={elType}
*section, name=section1, material=mat, type=solid
all
*step, solver=theSolver, maxInc=1, minInc=1, maxNumInc=1, maxIter=10
dirichlet, name=fix, nSet=gen_back, field=displacement, 1=0, 2=0, 3=0
distributedload, name=p, surface=gen_top, type=pressure, magnitude=2.0
""")
    inp = IF.parseInputFile(str(deck))
    model = IF.createModelFromInputFile(inp)
    b = model.elementBlocks[0]
    X = model.nodeCoordinates
    base, nDof, edofs = O.dof_numbering(len(X), b.conn)
    L = np.array([5.0, 6.0, 2.5])
    normals = {"left": (0, -1), "right": (0, 1), "bottom": (1, -1), "top": (1, 1), "back": (2, -1), "front": (2, 1)}
    centroids = X[b.conn].mean(axis=1)
    for side, (axis, sign) in normals.items():
        face = list(model.surfaces[f"gen_{side}"].keys())[0]
        els = model.elementSets[model.surfaces[f"gen_{side}"][face]][1]
        h = L[axis] / (3, 4, 2)[axis]
        layer = np.nonzero((centroids[:, axis] > L[axis] - h) if sign > 0 else (centroids[:, axis] < h))[0]
        assert np.array_equal(np.sort(els), layer)
        P = O.surface_pressure(elType, b.conn, X, edofs, nDof, face, 2.0, els).reshape(-1, 3)
        area = np.prod(np.delete(L, axis))
        want = np.zeros(3)
        want[axis] = -2.0 * area * sign
        assert np.allclose(P.sum(axis=0), want, rtol=1e-12, atol=1e-12), side
        on_face = np.isclose(X[:, axis], L[axis] if sign > 0 else 0.0)
        rows_off_face = np.asarray(base[~on_face] // 3, dtype=np.int64)  # P rows of the nodes off the face
        assert np.abs(P[rows_off_face]).max() == 0.0, side
    (step, _), = list(IF.createStepsFromInputFile(model, inp))
    act = step.actions["distributedload"]["p"]
    (face, (blockIndex, els)), = act.surface.items()
    assert face == 2 and blockIndex == 0 and np.array_equal(els, model.elementSets["gen_top"][1])
    assert act.loadType == "pressure"


_FIELD_DECK = """*material, name=LinearElastic, id=mat
30000.0, 0.15
*job, name=job, domain=3d
*solver, solver=NISTParallel, name=theSolver
*modelGenerator, generator=boxGen, name=gen
nX=10
nY=10
nZ=1
lX=50
lY=50
lZ=10
elType=C3D8
*section, name=section1, material=mat, type=solid
materialParameterFromField, index=0, field=F1, type=settovalue
materialParameterFromField, index=0, field=F2, type=scale, "f(p,f)=1 + f + 0*p"
all
*analyticalField, name=F1, type=scalarExpression
"f(x,y,z)" = "30000*(1 + 0.5*x/50)"
*analyticalField, name=F2, type=scalarExpression
"f(x,y,z)" = "0.1*np.sin(np.pi*y/100)"
*step, solver=theSolver, maxInc=1, minInc=1, maxNumInc=10, maxIter=25
dirichlet, name=x0u, nSet=gen_left, field=displacement, 1=0.
dirichlet, name=y0u, nSet=gen_bottom, field=displacement, 2=0.
dirichlet, name=z0u, nSet=gen_back, field=displacement, 3=0.
*step, solver=theSolver, maxInc=1, minInc=1e-2, maxNumInc=100, maxIter=25
distributedload, name=p, surface=gen_top, type=pressure, magnitude=100
"""


def test_boolean_element_sets(tmp_path):
    """*elSet with boolean = difference / union / intersection (abqmodelconstructor.py:153-171) on a boxGen
    mesh, into new sets and in place."""
    deck = tmp_path / "b.inp"
    deck.write_text("""*modelGenerator, generator=boxGen, name=gen
nX=4
nY=1
nZ=1
elType=C3D8
*elSet, elSet=gen_all, boolean=difference, newElSet=inner
1, 4
*elSet, elSet=inner, boolean=union, newElSet=grown
1
*elSet, elSet=gen_all, boolean=intersection, newElSet=ends
1, 4, 9
*elSet, elSet=grown, boolean=difference
2
*material, name=linearelastic, id=m
100., 0.3
*section, name=s, material=m, type=solid
all
""")
    model = IF.createModelFromInputFile(IF.parseInputFile(str(deck)))
    labels = model.elementBlocks[0].labels
    got = {k: sorted(int(labels[i]) for i in model.elementSets[k][1]) for k in ("inner", "grown", "ends")}
    assert got == {"inner": [2, 3], "grown": [1, 3], "ends": [1, 4]}


def test_pipegen_closed_ring_and_radii():
    """pipeGen beyond the PipeGen deck (a 90 degree C3D20 segment): a closed 360 degree C3D8 ring wraps its last
    element onto the first node column, every node lies on its layer's radius, every element has positive
    volume, and the ring's inner and outer node sets are the radii R - lT and R."""
    from edelweissfe_amd.model import pipeGenArrays, quadraturePoints, shapeFunctionDerivatives

    R = lambda h: 3.0 + 0.5 * h  # noqa: E731
    nT, nY, nC, lT = 2, 3, 12, 0.6
    X, conn, sets, esets, surf = pipeGenArrays(nT, nY, nC, R, lT=lT, lY=2.0, phi=360.0, x0=1.0, y0=-1.0, z0=2.0,
                                               elType="C3D8")
    assert len(X) == (nY + 1) * (nT + 1) * nC and len(conn) == nT * nY * nC
    last, first = conn[nC - 1], conn[0]  # element (it=0, iy=0, ic=nC-1) closes onto column 0
    assert last[1] == first[0] and last[2] == first[3] and last[5] == first[4] and last[6] == first[7]
    r = np.hypot(X[:, 0] - 1.0, X[:, 2] - 2.0)
    h = X[:, 1] + 1.0
    assert np.allclose(r[sets["gen_outer"]], R(h[sets["gen_outer"]]), atol=1e-13)
    assert np.allclose(r[sets["gen_inner"]], R(h[sets["gen_inner"]]) - lT, atol=1e-13)
    dN = shapeFunctionDerivatives(8, quadraturePoints(8))  # [q, a, k]
    J = np.einsum("qak,eai->eqik", dN, X[conn])
    assert np.linalg.det(J).min() > 0
    assert set(esets) >= {"gen_all", "gen_inner", "gen_outer", "gen_top", "gen_bottom"}
    assert surf["gen_inner"] == {3: "gen_inner"} and surf["gen_bottom"] == {1: "gen_bottom"}


def _field_deck(tmp_path):
    deck = tmp_path / "fields.inp"
    deck.write_text(_FIELD_DECK)
    return IF.parseInputFile(str(deck))


def test_material_parameter_from_scalar_fields(tmp_path):
    """sectionbase.py:86-128: per-element Young's modulus from two scalarExpression fields at the element
    centres (setToValue, then scale by f(p, f)); on the regular 10 x 10 x 1 grid the centre of element
    (ix, iy, iz) is ((ix + 1/2) 5, (iy + 1/2) 5, 5). The deck then runs through the oracle with those
    parameters (two steps, pressure on boxGen's gen_top)."""
    inp = _field_deck(tmp_path)
    model = IF.createModelFromInputFile(inp)
    b = model.elementBlocks[0]
    ix, iy = np.meshgrid(np.arange(10), np.arange(10), indexing="ij")
    xc, yc = ((ix + 0.5) * 5).ravel(), ((iy + 0.5) * 5).ravel()
    E = 30000 * (1 + 0.5 * xc / 50) * (1 + 0.1 * np.sin(np.pi * yc / 100))
    assert np.allclose(b.elParams, E, rtol=1e-14)
    bad = IF.parseInputFile(str(tmp_path / "fields.inp"))
    bad["*section"][0]["data"][0] = bad["*section"][0]["data"][0].replace("index=0", "index=1")
    with pytest.raises(NotImplementedError):
        IF.createModelFromInputFile(bad)
    U = _oracle_run(model, inp)
    assert np.isfinite(U).all() and np.abs(U).max() > 0


_ORACLE_STATE_INDEX = {"kappa": [12], "stress": list(range(6))}  # the oracle's J2 state: stress 6, strain 6, kappa


def _oracle_run(model, inp):
    b = model.elementBlocks[0]
    om = dict(el_kind=b.elType, mat_kind=b.material.name, mp=b.material.materialProperties,
              X=model.nodeCoordinates, conn=b.conn, el_mp=b.elParams,
              el_mp8=b.elementMaterialProperties() if b.materialIndex is not None else None)
    nq, ns = O.NQP[O.EL_KINDS[b.elType]], O.NSTATE[O.MAT_KINDS[b.material.name]]
    U, state = None, np.zeros(len(b) * nq * ns)
    for step, _ in _steps(model, inp):
        for act in step.actions.get("setfield", {}).values():  # setfield.py: at step start
            if act.active:
                els, qps, vals = act.qpValues(model)
                st = state.reshape(len(b), nq, ns)
                for c, k in enumerate(_ORACLE_STATE_INDEX[act.result]):
                    st[np.asarray(els)[:, None], qps[None, :], k] = vals[:, :, c]
        r = O.solve_step(om, _oracle_step(model, step), U0=U, state_n=state)
        U, state = r["U"], r["state"]
    return U


@pytest.mark.gpu
def test_material_parameter_fields_deck_on_gpu(tmp_path):
    """The field deck above through NISTGPU against the oracle's Newton with the same per-element parameters."""
    inp = _field_deck(tmp_path)
    model = IF.finiteElementSimulation(inp, solverOptions={"pcgTolerance": 1e-13})
    Ug = model.U.cpu().numpy()
    Uo = _oracle_run(IF.createModelFromInputFile(inp), inp)
    err = np.abs(Ug - Uo).max() / np.abs(Uo).max()
    print(f"field deck: GPU vs oracle {err:.2e}", flush=True)
    assert err < 1e-12  # measured 6e-16


_ACTIONS_DECK = """*material, name=LinearElastic, id=mat
1000.0, 0.3
*job, name=job, domain=3d
*solver, solver=NIST, name=theSolver
*modelGenerator, generator=boxGen, name=gen
nX=2
nY=2
nZ=4
lX=1
lY=1
lZ=4
elType=C3D8
*section, name=section1, material=mat, type=solid
all
*analyticalField, name=g, type=scalarExpression
"f(x,y,z)" = "1 + 0.5*x"
*step, solver=theSolver, maxInc=0.5, minInc=0.5, maxNumInc=10, maxIter=25
dirichlet, name=fix, nSet=gen_back, field=displacement, 1=0, 2=0, 3=0
dirichlet, name=pull, nSet=gen_front, field=displacement, 3=0.01, analyticalField=g
*step, solver=theSolver, maxInc=0.5, minInc=0.5, maxNumInc=10, maxIter=25
dirichlet, name=pull, nSet=gen_front, field=displacement, components='[x, x, 0.01]'
nodeforces, name=side, nSet=gen_top, field=displacement, components='[0, 0.5, x]'
"""


def _actions_deck(tmp_path):
    deck = tmp_path / "actions.inp"
    deck.write_text(_ACTIONS_DECK)
    return IF.parseInputFile(str(deck))


def test_dirichlet_fields_components_and_redefinition(tmp_path):
    """dirichlet.py:80-112 and nodeforces.py:85-108: a prescribed value scaled node by node by a scalarExpression
    field (analyticalField), a redefinition in the next step (components=[x, x, v]: the next step's increment),
    and nodal forces from components=[..]; the oracle's Newton over both steps ends with uz on the loaded face
    equal to 0.01 (1 + x / 2) + 0.01."""
    inp = _actions_deck(tmp_path)
    model = IF.createModelFromInputFile(inp)
    front = np.asarray(model.nodeSets["gen_front"])
    U = _oracle_run(model, inp)
    uz = U.reshape(-1, 3)[front, 2]
    # to the direct solve's rounding on the prescribed rows (~3e-15)
    assert np.allclose(uz, 0.01 * (1 + 0.5 * model.nodeCoordinates[front, 0]) + 0.01, rtol=0, atol=1e-13)
    model2 = IF.createModelFromInputFile(inp)
    steps = list(IF.createStepsFromInputFile(model2, inp))
    side = steps[1][0].actions["nodeforces"]["side"]
    assert np.array_equal(side.nodeForcesDelta[0], [0.0, 0.5, 0.0])


@pytest.mark.gpu
def test_dirichlet_fields_deck_on_gpu(tmp_path):
    """The same deck through NISTGPU against the oracle's Newton."""
    inp = _actions_deck(tmp_path)
    model = IF.finiteElementSimulation(inp, solverOptions={"pcgTolerance": 1e-13})
    Ug = model.U.cpu().numpy()
    Uo = _oracle_run(IF.createModelFromInputFile(inp), inp)
    err = np.abs(Ug - Uo).max() / np.abs(Uo).max()
    print(f"actions deck: GPU vs oracle {err:.2e}", flush=True)
    assert err < 1e-12  # measured 3e-16
