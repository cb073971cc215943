"""Writes the golden fixtures under tests/golden/ from the reference's own regression decks.

Run in the build container (the only place the reference tree exists):

    python tests/golden/make_fixtures.py /root/reference/testfiles

For each deck it reads ``test.inp`` (input data) and ``U.ref`` (expected final nodal displacements,
compared at 1e-6 absolute by ew/_cli/_run_tests_edelweissfe.py:96-109) and stores both as plain data in
``<deck>.json``: mesh or boxGen options, material, node sets used by the step actions, step
controls and the reference displacement vector. Nothing of the reference is imported or executed:
the decks are parsed as text with the small keyword reader below.
"""

import json
import os
import sys

DECKS = ["TensionBarHexa8", "SimpleBeamHexa8", "CantileverBeamHexa8", "C3D8", "C3D20", "WallShearHexa8",
         "WallShearHexa20", "LinearElasticIsotropic"]


def read_keywords(path):
    blocks = []
    cur = None
    with open(path) as f:
        for raw in f:
            line = raw.strip()
            if not line or line.startswith("**"):
                continue
            if line.startswith("*"):
                parts = [p.strip() for p in line[1:].split(",")]
                opts = {}
                for p in parts[1:]:
                    if "=" in p:
                        k, v = p.split("=", 1)
                        opts[k.strip().lower()] = v.strip()
                cur = {"kw": parts[0].lower(), "opts": opts, "data": []}
                blocks.append(cur)
            elif cur is not None:
                cur["data"].append(line)
    return blocks


def parse_line_opts(line):
    """'dirichlet, name=x, nSet=y, 1=0.0' -> ('dirichlet', {...}); quoted values may hold commas."""
    out, buf, q = [], "", False
    for ch in line:
        if ch == "'":
            q = not q
            continue
        if ch == "," and not q:
            out.append(buf.strip())
            buf = ""
        else:
            buf += ch
    out.append(buf.strip())
    kind = out[0].lower()
    opts = {}
    for p in out[1:]:
        if "=" in p:
            k, v = p.split("=", 1)
            opts[k.strip()] = v.strip()
    return kind, opts


def numbers(lines):
    vals = []
    for ln in lines:
        vals += [float(t) for t in ln.replace(",", " ").split() if t]
    return vals


def convert(deck_dir, name):
    blocks = read_keywords(os.path.join(deck_dir, "test.inp"))
    fx = {"name": name}
    nodes, elements, nsets, elsets, surfaces = [], [], {}, {}, {}
    actions = {}  # step actions persist across steps, keyed by (kind, name) (stepmanager.py:133-200)
    n_steps = 0
    for b in blocks:
        kw, o = b["kw"], b["opts"]
        if kw == "material":
            fx["material"] = o["name"].lower()
            fx["mp"] = numbers(b["data"])
        elif kw == "node":
            for ln in b["data"]:
                v = numbers([ln])
                nodes.append([int(v[0])] + v[1:4])
        elif kw == "element":
            fx["el_type"] = o["type"].upper()
            vals = numbers(b["data"])
            npe = 8 if fx["el_type"].startswith("C3D8") else 20
            for i in range(0, len(vals), npe + 1):
                elements.append([int(t) for t in vals[i: i + npe + 1]])
        elif kw == "nset":
            vals = [int(v) for v in numbers(b["data"])]
            if o.get("generate", "").lower() == "true":
                vals = list(range(vals[0], vals[1] + 1, vals[2]))
            nsets[o["nset"]] = vals
        elif kw == "elset":
            vals = [int(v) for v in numbers(b["data"])]
            if o.get("generate", "").lower() == "true":
                vals = list(range(vals[0], vals[1] + 1, vals[2]))
            elsets[o["elset"]] = vals
        elif kw == "surface":
            surfaces[o["name"]] = [(ln.split(",")[0].strip(), int(ln.split(",")[1].strip().upper().replace("S", "")))
                                   for ln in b["data"]]
        elif kw == "modelgenerator":
            g = {}
            for ln in b["data"]:
                k, v = ln.split("=")
                g[k.strip()] = v.strip()
            fx["generator"] = g
            fx["el_type"] = g["elType"].upper()
        elif kw == "step":
            st = {k: o[k.lower()] for k in ["maxInc", "minInc", "maxNumInc", "maxIter", "stepLength", "startInc"]
                  if k.lower() in o}
            st = {k: (int(float(v)) if k in ("maxNumInc", "maxIter") else float(v)) for k, v in st.items()}
            st["extrapolation"] = "linear"
            loaded = any(k[0] != "dirichlet" or any(v["components"].values()) for k, v in actions.items())
            if n_steps and loaded:
                raise NotImplementedError(f"{name}: only the last step of a deck may carry loads")
            n_steps += 1
            for ln in b["data"]:
                kind, lo = parse_line_opts(ln)
                key = (kind, lo.get("name", kind))
                if kind == "options" and "extrapolation" in lo:
                    st["extrapolation"] = lo["extrapolation"]
                elif kind == "dirichlet":
                    comps = {str(int(k) - 1): float(v) for k, v in lo.items() if k in ("1", "2", "3")}
                    actions[key] = {"nSet": lo["nSet"], "components": comps}
                elif kind == "nodeforces":
                    vec = [float(t) for t in lo["components"].split(",")]
                    actions[key] = {"nSet": lo["nSet"], "vector": vec}
                elif kind == "bodyforce":
                    vec = [float(t) for t in lo["forceVector"].split(",")]
                    actions[key] = {"elSet": lo["elSet"], "vector": vec}
                elif kind == "distributedload":
                    if lo["type"].lower() != "pressure":
                        raise NotImplementedError(f"{name}: distributed load type {lo['type']}")
                    for elSet, face in surfaces[lo["surface"]]:
                        labs = [e[0] for e in elements] if elSet == "all" else elsets[elSet]
                        actions[key + (face,)] = {"surface": lo["surface"], "elements": labs, "face": face,
                                                  "magnitude": float(lo["magnitude"]), "f(t)": lo.get("f(t)")}
            st["dirichlet"] = [v for k, v in actions.items() if k[0] == "dirichlet"]
            st["nodeforces"] = [v for k, v in actions.items() if k[0] == "nodeforces"]
            st["bodyforces"] = [v for k, v in actions.items() if k[0] == "bodyforce"]
            if any(k[0] == "distributedload" for k in actions):
                st["pressures"] = [v for k, v in actions.items() if k[0] == "distributedload"]
            fx["step"] = st
    if nodes:
        fx["nodes"] = nodes
        fx["elements"] = elements
        fx["nsets"] = nsets
    with open(os.path.join(deck_dir, "U.ref")) as f:
        fx["U_ref"] = [float(t) for t in f.read().split()]
    return fx


def main():
    root = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/testfiles"
    here = os.path.dirname(os.path.abspath(__file__))
    for d in DECKS:
        fx = convert(os.path.join(root, d), d)
        with open(os.path.join(here, f"{d}.json"), "w") as f:
            json.dump(fx, f)
        print(d, len(fx["U_ref"]), fx.get("el_type"), fx.get("material"), fx.get("mp"))


if __name__ == "__main__":
    main()
