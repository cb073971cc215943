"""Turns the golden deck fixtures (tests/golden/*.json) into model/step descriptions.

The same description drives the oracle's Newton loop and the product's NISTGPU solver, so both are
compared against the reference's U.ref vector of the deck (1e-6 absolute,
ew/_cli/_run_tests_edelweissfe.py:96-109).
"""
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with open(os.path.join(GOLDEN, f"{name}.json")) as f:
        return json.load(f)


def mesh(fx, boxgen):
    """Returns X [nNodes,3] (model insertion order), conn [nEl,npe] 0-based, node sets {name: node idx}.
    ``boxgen`` is the generator to use for boxGen decks (the oracle's or the product's)."""
    if "generator" in fx:
        g = fx["generator"]
        npe = 8 if fx["el_type"].startswith("C3D8") else 20
        X, conn, sets = boxgen(int(g["nX"]), int(g["nY"]), int(g["nZ"]), float(g.get("lX", 1)),
                               float(g.get("lY", 1)), float(g.get("lZ", 1)), nnodes=npe, name="gen")
        return np.asarray(X), np.asarray(conn), sets
    labels = [int(n[0]) for n in fx["nodes"]]
    pos = {lab: i for i, lab in enumerate(labels)}
    X = np.array([n[1:4] for n in fx["nodes"]], dtype=np.float64)
    conn = np.array([[pos[l] for l in e[1:]] for e in fx["elements"]], dtype=np.int32)
    sets = {k: np.array([pos[l] for l in v], dtype=np.int64) for k, v in fx["nsets"].items()}
    # nodes that no element uses carry no dofs and are not part of U (nodefield.py:87)
    return X, conn, sets


def step(fx, sets):
    st = dict(fx["step"])
    st["dirichlet"] = [(sets[d["nSet"]], {int(k): v for k, v in d["components"].items()}) for d in st["dirichlet"]]
    st["nodeforces"] = [(sets[d["nSet"]], d["vector"]) for d in st["nodeforces"]]
    for d in st.get("bodyforces", []):
        if d["elSet"] != "all":
            raise NotImplementedError("deck body forces on element sets other than 'all'")
    st["bodyforces"] = [d["vector"] for d in st.get("bodyforces", [])]
    if st.get("pressures"):
        if "elements" not in fx:
            raise NotImplementedError("deck pressures on generated meshes")
        pos = {e[0]: i for i, e in enumerate(fx["elements"])}
        st["pressures"] = [(np.array([pos[lab] for lab in d["elements"]], dtype=np.int64), d["face"], d["magnitude"],
                            amplitude(d["f(t)"])) for d in st["pressures"]]
    return st


def amplitude(expr):
    """f(t) of a fixture's step action (an arithmetic expression in t written by make_fixtures.py from the
    deck; linear when absent)."""
    if not expr:
        return lambda t: t
    code = compile(expr, "<f(t)>", "eval")
    return lambda t: float(eval(code, {"__builtins__": {}}, {"t": float(t)}))


def used_node_order(X, conn):
    used = np.zeros(len(X), dtype=bool)
    used[np.asarray(conn).ravel()] = True
    return np.nonzero(used)[0]
