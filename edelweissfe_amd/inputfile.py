"""Input decks in the reference's Abaqus-like syntax, run on the GPU path.

The subset of the reference's input language that the GPU path can run, with the reference's structure:

  * ``parseInputFile`` (edelweissfe/utils/inputfileparser.py:255-363): keyword lines ``*keyword, opt=val`` split
    at commas outside quotes, data lines appended to the last keyword, ``**`` comments, ``*include`` parsed in
    place; keywords and option names are case-insensitive (stored lower case here);
  * ``createModelFromInputFile`` (edelweissfe/helpers/inputfilehelpers.py:169-211 and
    edelweissfe/generators/abqmodelconstructor.py:67-265): boxGen generators first, then ``*node`` (label order
    of definition), ``*element``, ``*nSet`` / ``*elSet`` (labels, ``generate=True``, or names of other sets), the
    implicit sets ``all``, materials, sections (with ``materialParameterFromField`` over scalarExpression
    ``*analyticalField`` definitions for the material's first property) and element-based ``*surface`` definitions;
  * ``createStepsFromInputFile`` (inputfilehelpers.py:214-244, edelweissfe/steps/stepmanager.py:133-200): step
    actions live across steps, keyed by module and name; a redefinition in a later step updates the action;
  * ``finiteElementSimulation`` (edelweissfe/drivers/inputfiledrivensimulation.py:56-217) without output
    managers: every ``*step`` is solved by NISTGPU, whatever CPU solver (NIST, NISTParallel, ...) the deck names;
    the solver of the last step stays reachable as ``model.solver`` (quadrature-point results);
  * ``runReferenceTest`` (edelweissfe/_cli/_run_tests_edelweissfe.py:88-109): the final nodal displacements
    against ``U.ref``, 1e-6 absolute.

Decks outside the subset (other element types or providers' features, several element blocks or materials,
constraints, distributed loads other than pressure, analytical fields other than scalarExpression, ...) raise
NotImplementedError.
"""

from __future__ import annotations

import ast
import math
import operator
import os
import shlex

import numpy as np

from .model import (BOXGEN_SURFACES, ELEMENT_LIBRARY, ElementBlock, FEModel, Material, boxGenArrays,
                    boxGenElementSets, elementType, pipeGenArrays)

# keyword -> option names it accepts; the output keywords are read and not acted on
_KEYWORDS = {
    "*node": {"nset"},
    "*element": {"elset", "type", "provider"},
    "*elset": {"elset", "generate", "boolean", "newelset"},
    "*nset": {"nset", "generate"},
    "*surface": {"type", "name"},
    "*section": {"name", "thickness", "material", "type"},
    "*material": {"name", "id", "statevars", "provider"},
    "*fieldoutput": set(),
    "*output": {"name", "type", "jobname"},
    "*job": {"name", "domain", "solver", "starttime"},
    "*solver": {"name", "solver"},
    "*step": {"steplength", "startinc", "maxinc", "mininc", "maxnuminc", "maxiter", "type", "solver", "criticaliter",
              "cutbackfactor", "jobname", "maxgrowiter"},
    "*updateconfiguration": {"configuration"},
    "*modelgenerator": {"generator", "name", "executeaftermanualgeneration"},
    "*include": {"input"},
    "*analyticalfield": {"name", "type"},
    "*constraint": {"name", "type"},
    "*advancedmaterial": {"name", "id", "provider"},
}

# material names of the reference's providers -> materials of the GPU path
# material names a deck may use -> the GPU path's laws; ADLinearElastic is Marmot's isotropic linear elasticity
# with an automatically differentiated tangent (the same law: [E, nu])
_MATERIALS = {"linearelastic": "linearelastic", "adlinearelastic": "linearelastic", "vonmises": "vonmises",
              "neohookeanwa": "neohookeanwa", "neohookeanpencegouformulationa": "neohookeanwa",
              "stvenantkirchhoffisotropic": "stvenantkirchhoff", "stvenantkirchhoff": "stvenantkirchhoff"}
# solvers a deck may name; all run as NISTGPU here (edelweissfe/config/solvers.py:48-76)
_SOLVERS = {"nist", "nistparallel", "nistparallelformarmotelements", "nistgpu"}


def splitLineAtCommas(line: str) -> list:
    """edelweissfe/utils/misc.py:45-65: split at commas outside quotes, strip the parts."""
    lexer = shlex.shlex(line, posix=True)
    lexer.whitespace_split = True
    lexer.whitespace = ","
    return [x.strip() for x in lexer]


def _assignments(parts) -> dict:
    """misc.py:68-91: 'a=b' -> {a: b}; a bare word maps to 'True'."""
    out = {}
    for entry in parts:
        k, *v = [x.strip() for x in entry.split("=")]
        out[k] = "=".join(v) if v else "True"
    return out


def parseInputFile(fileName: str, currentKeyword: str | None = None, existingFileDict: dict | None = None) -> dict:
    """Parsed deck: lower-case keyword -> list of entries (dict: lower-case options, 'data' lines, 'inputFile')."""
    fileDict = existingFileDict if existingFileDict is not None else {k: [] for k in _KEYWORDS}
    keyword = currentKeyword
    with open(fileName) as f:
        for raw in f:
            line = raw.strip()
            if not line or line.startswith("**"):
                continue
            if line.startswith("*"):
                parts = splitLineAtCommas(line)
                lastKeyword, keyword = keyword, parts[0].lower()
                if keyword not in _KEYWORDS:
                    raise KeyError(f"keyword {parts[0]} is not part of the input language of the GPU path")
                entry = {"data": [], "inputFile": fileName}
                for ass in parts[1:]:
                    if not ass:
                        continue
                    k, _, v = ass.partition("=")
                    k = k.strip().lower()
                    if k not in _KEYWORDS[keyword]:
                        raise KeyError(f'option "{k}" not valid for {keyword}')
                    entry[k] = v.strip()
                if keyword == "*include":
                    parseInputFile(os.path.join(os.path.dirname(fileName), entry["input"]), lastKeyword, fileDict)
                    keyword = lastKeyword
                else:
                    fileDict[keyword].append(entry)
            else:
                if keyword is None:
                    raise KeyError(f"data line before any keyword in {fileName}")
                fileDict[keyword][-1]["data"].append(line)
    return fileDict


def _numbers(lines, dtype=float):
    return np.array([dtype(t) for ln in lines for t in splitLineAtCommas(ln) if t], dtype=dtype)


def _is_int(s):
    try:
        int(s)
        return True
    except ValueError:
        return False


def _truthy(v):
    return str(v).strip().lower() in ("true", "yes", "1", "on")


class _Deck:
    """Geometry collected from a deck before it becomes an FEModel (node labels in definition order)."""

    def __init__(self):
        self.nodeLabels, self.coords = [], []
        self.nodeIndex = {}
        self.elType, self.elLabels, self.elNodes = None, [], []
        self.nodeSets, self.elementSets, self.surfaces = {}, {}, {}

    def addNodes(self, labels, coords):
        for lab, x in zip(labels, coords):
            if lab in self.nodeIndex:  # a redefinition keeps the node's place (dict.update semantics)
                self.coords[self.nodeIndex[lab]] = x
            else:
                self.nodeIndex[lab] = len(self.nodeLabels)
                self.nodeLabels.append(lab)
                self.coords.append(x)

    def addElements(self, elType, labels, nodeLists):
        et = elType.upper()
        if self.elType not in (None, et):
            raise NotImplementedError("decks with several element types need several element blocks")
        self.elType = et
        self.elLabels += list(labels)
        self.elNodes += list(nodeLists)


def _boxgen(deck: _Deck, gen: dict):
    """edelweissfe/generators/boxgen.py:91-448 through model.boxGenArrays; node labels continue after the
    highest existing label, element labels likewise."""
    opts = _assignments(gen["data"])  # 'nX = 4' lines
    name = gen.get("name", "boxGen")
    g = lambda k, d: opts.get(k, d)  # noqa: E731
    elType = g("elType", None)
    if elType is None:
        raise KeyError("boxGen needs elType")
    X, conn, sets, _ = boxGenArrays(int(g("nX", 1)), int(g("nY", 1)), int(g("nZ", 1)), float(g("lX", 1.0)),
                                    float(g("lY", 1.0)), float(g("lZ", 1.0)), float(g("x0", 0.0)), float(g("y0", 0.0)),
                                    float(g("z0", 0.0)), elType=elType, name=name)
    n0 = max(deck.nodeLabels, default=0)
    labels = list(range(n0 + 1, n0 + 1 + len(X)))
    deck.addNodes(labels, list(X))
    e0 = max(deck.elLabels, default=0)
    first = len(deck.elLabels)
    deck.addElements(elType, range(e0 + 1, e0 + 1 + len(conn)), [[labels[i] for i in row] for row in conn])
    for k, idx in sets.items():
        deck.nodeSets[k] = [labels[i] for i in idx]
    for k, idx in boxGenElementSets(int(g("nX", 1)), int(g("nY", 1)), int(g("nZ", 1)), name).items():
        deck.elementSets[k] = [deck.elLabels[first + i] for i in idx]
    for suffix, face in BOXGEN_SURFACES.items():
        deck.surfaces[f"{name}_{suffix}"] = {face: f"{name}_{suffix}"}


def _pipegen(deck: _Deck, gen: dict):
    """edelweissfe/generators/pipegen.py:91-357 through model.pipeGenArrays; node and element labels continue
    after the highest existing ones."""
    opts = _assignments(gen["data"])
    name = gen.get("name", "cylBoxGen")
    g = lambda k, d: opts.get(k, d)  # noqa: E731
    elType = g("elType", None)
    if elType is None:
        raise KeyError("pipeGen needs elType")
    Ro, Ri = str(g("Ro(y)", "1.0")), str(g("Ri(y)", "1.0"))
    if Ri == "1.0":
        expr, outer = Ro, True
    elif Ro == "1.0":
        expr, outer = Ri, False
    else:
        raise Exception("Only the outer or the inner radius can be defined as a function. Please use 'lT' option "
                        "instead!")
    phi, lT, lY, y0 = float(g("phi", 360)), float(g("lT", 0.5)), float(g("lY", 1.0)), float(g("y0", 0.0))
    if abs(phi) > 360:
        raise Exception("The angle can't be higher than 360° or lower than -360°.")
    nT, nY, nC = int(g("nT", 1)), int(g("nY", 4)), int(g("nC", 4))
    f = _Expression(expr, ("y",))
    npe = ELEMENT_LIBRARY[elementType(elType)]["nNodes"]
    heights = np.linspace(y0, y0 + lY, (nY if npe == 8 else 2 * nY) + 1) - y0
    R = np.array([f(h) for h in heights])
    if (np.any(R - lT < 1e-5) and outer) or (np.any(R < 1e-5) and not outer):
        raise Exception("The radius must be > 1e-5 at every point.")
    # the option exG is used as given: a deck's 'exG = False' is a non-empty text and, as in the generator,
    # counts as set (nodes on the exact geometry)
    exG = g("exG", True)
    X, conn, sets, esets, surfaces = pipeGenArrays(
        nT, nY, nC, lambda h: f(h), lT=lT, lY=lY, phi=phi, x0=float(g("x0", 0.0)), y0=y0, z0=float(g("z0", 0.0)),
        outer=outer, reduceMidRadius=not exG, elType=elType, name=name)
    n0 = max(deck.nodeLabels, default=0)
    labels = list(range(n0 + 1, n0 + 1 + len(X)))
    deck.addNodes(labels, list(X))
    e0 = max(deck.elLabels, default=0)
    first = len(deck.elLabels)
    deck.addElements(elType, range(e0 + 1, e0 + 1 + len(conn)), [[labels[i] for i in row] for row in conn])
    for k, idx in sets.items():
        deck.nodeSets[k] = [labels[i] for i in idx]
    for k, idx in esets.items():
        deck.elementSets[k] = [deck.elLabels[first + i] for i in idx]
    deck.surfaces.update(surfaces)


def _geometry(deck: _Deck, inputFile: dict):
    """Nodes, elements and sets of a parsed deck or mesh file (abqmodelconstructor.py:67-212)."""
    for nd in inputFile["*node"]:
        rows = [splitLineAtCommas(ln) for ln in nd["data"]]
        labels = [int(r[0]) for r in rows]
        deck.addNodes(labels, [np.array([float(t) for t in r[1:4]] + [0.0] * (4 - len(r))) for r in rows])
        if "nset" in nd:
            deck.nodeSets[nd["nset"]] = labels
    for el in inputFile["*element"]:
        rows = [[int(t) for t in splitLineAtCommas(ln) if t] for ln in el["data"]]
        deck.addElements(el["type"], [r[0] for r in rows], [r[1:] for r in rows])
        if "elset" in el:
            deck.elementSets[el["elset"]] = [r[0] for r in rows]
    for es in inputFile["*elset"]:
        data = [t for ln in es["data"] for t in splitLineAtCommas(ln) if t]
        if "boolean" in es:  # abqmodelconstructor.py:153-171: the set combined with the listed labels
            op = es["boolean"].lower()
            nums = [int(t) for t in data]
            base = deck.elementSets[es["elset"]]
            if op == "difference":
                labs = [n for n in base if n not in set(nums)]
            elif op == "union":
                labs = list(base) + nums
            elif op == "intersection":
                labs = [n for n in nums if n in set(base)]
            else:
                raise Exception("Undefined boolean operation!")
            # stored as newElSet (the set itself is replaced when newElSet names it or is not given)
            deck.elementSets[es.get("newelset", es["elset"])] = sorted(set(labs))
            continue
        if data and _is_int(data[0]):
            nums = [int(t) for t in data]
            deck.elementSets[es["elset"]] = (list(range(nums[0], nums[1] + 1, nums[2])) if _truthy(es.get("generate", "")) else nums)
        else:
            deck.elementSets[es["elset"]] = [lab for s in data for lab in deck.elementSets[s]]
    for ns in inputFile["*nset"]:
        data = [t for ln in ns["data"] for t in splitLineAtCommas(ln) if t]
        if data and _is_int(data[0]):
            nums = [int(t) for t in data]
            deck.nodeSets[ns["nset"]] = (list(range(nums[0], nums[1] + 1, nums[2])) if _truthy(ns.get("generate", "")) else nums)
        else:
            deck.nodeSets[ns["nset"]] = [lab for s in data for lab in deck.nodeSets[s]]


def _cubit(deck: _Deck, gen: dict):
    """edelweissfe/generators/cubit.py:26-100: the mesh Cubit exported (outFile, default mesh.inc, beside the deck),
    element types replaced by elType or per element set by elTypePerBlock, added like a deck's own geometry. Cubit
    itself is not part of this build: the exported file must exist (the reference re-exports it when overwrite is
    set; its content is the same mesh)."""
    opts = _assignments(gen["data"])
    outFile = os.path.join(os.path.dirname(gen.get("inputFile", "")), opts.get("outFile", "mesh.inc").strip("'\""))
    if not os.path.exists(outFile):
        raise NotImplementedError(f"running Cubit is not part of the GPU path and {outFile} does not exist")
    mesh = parseInputFile(outFile)
    elType = opts.get("elType")
    perBlock = opts.get("elTypePerBlock")
    for el in mesh["*element"]:
        if elType:
            el["type"] = elType.strip("'\"")
        if perBlock:
            pairs = dict(t.split("=", 1) for t in perBlock.strip("'\"").replace(" ", "").split(",") if t)
            el["type"] = pairs[el["elset"]]
    _geometry(deck, mesh)


def _findClosestNode(deck: _Deck, gen: dict):
    """edelweissfe/generators/findclosestnode.py:51-73: the node nearest to 'location' (the first of equally near
    ones, in node order) becomes the node set 'storeIn'."""
    opts = _assignments([t for ln in gen["data"] for t in splitLineAtCommas(ln) if t])
    loc = np.array([float(t) for t in splitLineAtCommas(opts["location"]) if t])
    if len(loc) != 3:
        raise ValueError("findClosestNode: the location needs 3 coordinates")
    if opts["storeIn"] in deck.nodeSets:
        raise Exception(f"Nodeset {opts['storeIn']} already exists")
    coords = np.asarray(deck.coords, dtype=np.float64).reshape(-1, 3)
    deck.nodeSets[opts["storeIn"]] = [deck.nodeLabels[int(np.linalg.norm(coords - loc, axis=1).argmin())]]


def createModelFromInputFile(inputFile: dict) -> FEModel:
    """FEModel of a parsed deck: nodes, one element block with its material, node and element sets."""
    for job in inputFile["*job"]:
        if job.get("domain", "3d").lower() != "3d":
            raise NotImplementedError("the GPU path covers 3D solid elements only")
    deck = _Deck()
    gens = inputFile["*modelgenerator"]
    generators = {"boxgen": _boxgen, "pipegen": _pipegen, "cubit": _cubit, "findclosestnode": _findClosestNode}
    for gen in gens:
        if gen["generator"].lower() not in generators:
            raise NotImplementedError(f"model generator {gen['generator']} is not part of the GPU path")
    for gen in (g for g in gens if not _truthy(g.get("executeaftermanualgeneration", "false"))):
        generators[gen["generator"].lower()](deck, gen)
    _geometry(deck, inputFile)
    for gen in (g for g in gens if _truthy(g.get("executeaftermanualgeneration", "false"))):
        generators[gen["generator"].lower()](deck, gen)
    if deck.elType is None:
        raise ValueError("the deck defines no elements")

    model = FEModel()
    model.addNodes(np.array(deck.coords, dtype=np.float64).reshape(-1, 3), np.array(deck.nodeLabels))
    elIndex = {lab: i for i, lab in enumerate(deck.elLabels)}
    conn = np.array([[deck.nodeIndex[n] for n in nodes] for nodes in deck.elNodes], dtype=np.int32)
    block = ElementBlock(deck.elType, conn, labels=np.array(deck.elLabels))
    model.addElementBlock(block, "all")
    for name, labs in deck.elementSets.items():
        model.elementSets[name] = (0, np.array(sorted(elIndex[lab] for lab in labs), dtype=np.int64))
    for name, labs in deck.nodeSets.items():
        model.nodeSets[name] = np.array([deck.nodeIndex[lab] for lab in labs], dtype=np.int64)
    model.nodeSets["all"] = np.arange(model.nNodes)

    # materials (abqmodelconstructor.py:241-265) and sections (edelweissfe/sections/solid.py:45-64)
    materials = {}
    for m in inputFile["*material"]:
        key = m["name"].lower()
        if key not in _MATERIALS:
            raise NotImplementedError(f"material {m['name']} is not available on the GPU path")
        materials[m.get("id", m["name"])] = Material(_MATERIALS[key], _numbers(m["data"]))
    fields = _analyticalFields(inputFile)
    model.analyticalFields = fields
    model.fieldOutputs = _fieldOutputs(inputFile)
    # every section assigns its material to the elements of its sets (sectionbase.py:84-112); the block keeps
    # the distinct materials (of one kind) and every element's index
    assigned = np.zeros(len(block), dtype=bool)
    used, matIndex = [], np.zeros(len(block), dtype=np.int32)
    elParams = None
    for sec in inputFile["*section"]:
        if sec.get("type", "solid").lower() != "solid":
            raise NotImplementedError(f"section type {sec.get('type')} is not part of the GPU path")
        mat = materials[sec["material"]]
        if not any(mat is m for m in used):
            used.append(mat)
        mi = next(i for i, m in enumerate(used) if m is mat)
        defs, setNames = _sectionLines(sec, fields)
        for setName in setNames:
            _, idx = model.elementSets[setName]
            assigned[idx] = True
            matIndex[idx] = mi
            if defs:  # sectionbase.py:86-128: the parameter of each element from the fields at its centre
                if elParams is None:
                    elParams = np.full(len(block), np.nan)
                centres = model.nodeCoordinates[block.conn[idx]].mean(axis=1)  # getCoordinatesAtCenter
                for e, c in zip(idx, centres):
                    value = float(mat.materialProperties[0])
                    for typ, field, expr in defs:
                        f = field(*c)
                        value = expr(value, f) if typ == "settovalue" else value * expr(value, f)
                    elParams[e] = value
    if not used or not assigned.all():
        raise ValueError("every element needs a section with a material")
    block.setMaterials(used, matIndex if len(used) > 1 else None)
    if elParams is not None:
        unset = np.isnan(elParams)  # elements of sections without fields keep their material's value
        elParams[unset] = block.materialParameterTable()[matIndex[unset], 0]
        block.elParams = np.ascontiguousarray(elParams)

    # element-based surfaces: boxGen's, then data lines 'elSet, face' with face 1..6 or S1..S6
    # (abqmodelconstructor.py:213-225)
    model.surfaces.update(deck.surfaces)
    for sf in inputFile["*surface"]:
        if sf.get("type", "element").lower() != "element":
            raise NotImplementedError(f"surface type {sf.get('type')} is not part of the GPU path")
        surface = {}
        for ln in sf["data"]:
            elSet, face = splitLineAtCommas(ln)
            if elSet not in model.elementSets:
                raise KeyError(f"surface {sf['name']}: unknown element set {elSet}")
            surface[int(face.upper().replace("S", ""))] = elSet
        model.surfaces[sf["name"]] = surface
    return model


def _analyticalFields(inputFile):
    """*analyticalField definitions (edelweissfe/analyticalfields/scalarexpression.py): name -> f(x, y, z).
    The GPU path reads the scalarExpression type; randomScalar needs gstools, which it does not carry."""
    fields = {}
    for af in inputFile["*analyticalfield"]:
        typ = af.get("type", "").lower()
        if typ != "scalarexpression":
            raise NotImplementedError(f"analytical field type {af.get('type')} is not part of the GPU path")
        opts = _assignments([t for ln in af["data"] for t in splitLineAtCommas(ln) if t])
        fields[af["name"]] = _Expression(opts["f(x,y,z)"], ("x", "y", "z"))
    return fields


def _sectionLines(sec, fields):
    """A *section's data lines (sectionbase.py:31-75): materialParameterFromField definitions and the element
    set names. Returns ([(type, field, f(p, f))], set names)."""
    defs, setNames = [], []
    for ln in sec["data"]:
        parts = [t for t in splitLineAtCommas(ln) if t]
        if not parts:
            continue
        key = parts[0].lower()
        if key == "materialparameterfromfield":
            d = _assignments(parts[1:])
            typ = d.get("type", "").lower()
            if typ not in ("settovalue", "scale"):
                raise KeyError(f"{sec.get('name')}: materialParameterFromField type must be 'setToValue' or 'scale'")
            if int(d["index"]) != 0:
                raise NotImplementedError("per-element parameters cover the material's first property on the GPU path")
            defs.append((typ, fields[d["field"]], _Expression(d.get("f(p,f)", "f"), ("p", "f"))))
        elif key == "writematerialpropertiestofile":
            raise NotImplementedError("writeMaterialPropertiesToFile (an output) is not part of the GPU path")
        else:
            setNames += parts
    return defs, setNames


class _Expression:
    """An arithmetic expression in the given variables (numbers, + - * / **, pi, sqrt, exp, sin, cos, tan, log,
    abs, min, max, also as np.<name> or math.<name>), evaluated without executing anything from the deck (the
    reference evaluates such strings with Python's eval or lambdifies them with sympy)."""

    _ops = {ast.Add: operator.add, ast.Sub: operator.sub, ast.Mult: operator.mul, ast.Div: operator.truediv,
            ast.Pow: operator.pow, ast.USub: operator.neg, ast.UAdd: operator.pos}
    _funcs = {"sqrt": math.sqrt, "exp": math.exp, "sin": math.sin, "cos": math.cos, "tan": math.tan,
              "log": math.log, "abs": abs, "min": min, "max": max}
    _modules = ("np", "numpy", "math")

    def __init__(self, expr: str, names=("t",)):
        self.expr = expr
        self.names = tuple(names)
        self.tree = ast.parse(expr, mode="eval").body
        self(*([0.5] * len(self.names)))  # reject unsupported expressions early

    def _func(self, f):
        if isinstance(f, ast.Name) and f.id in self._funcs:
            return self._funcs[f.id]
        if (isinstance(f, ast.Attribute) and isinstance(f.value, ast.Name) and f.value.id in self._modules
                and f.attr in self._funcs):
            return self._funcs[f.attr]
        return None

    def _eval(self, node, env):
        if isinstance(node, ast.Constant) and isinstance(node.value, (int, float)):
            return node.value
        if isinstance(node, ast.Name) and node.id in env:
            return env[node.id]
        if isinstance(node, ast.Name) and node.id == "pi":
            return math.pi
        if (isinstance(node, ast.Attribute) and isinstance(node.value, ast.Name) and node.value.id in self._modules
                and node.attr == "pi"):
            return math.pi
        if isinstance(node, ast.BinOp) and type(node.op) in self._ops:
            return self._ops[type(node.op)](self._eval(node.left, env), self._eval(node.right, env))
        if isinstance(node, ast.UnaryOp) and type(node.op) in self._ops:
            return self._ops[type(node.op)](self._eval(node.operand, env))
        if isinstance(node, ast.Call) and not node.keywords and self._func(node.func) is not None:
            return self._func(node.func)(*(self._eval(a, env) for a in node.args))
        raise ValueError(f"unsupported expression {self.expr!r}")

    def __call__(self, *values):
        return float(self._eval(self.tree, {k: float(v) for k, v in zip(self.names, values)}))


class _Amplitude(_Expression):
    """f(t) of a step action (an expression in the step progress t)."""

    def __init__(self, expr: str):
        super().__init__(expr, ("t",))


def _vector(s):
    return [float(t) for t in splitLineAtCommas(s) if t]


def _componentList(s, unset):
    """'[x, 0.5, 0]' -> [unset, 0.5, 0.0] ('x' marks a component that is not set)."""
    out = []
    for t in s.strip().strip("[]()").split(","):
        t = t.strip()
        if t:
            out.append(unset if t.lower() == "x" else float(t))
    return out


def _dirichletValues(kw, model):
    """dirichlet.py:80-103: prescribed values from '1=..', '2=..', '3=..' or 'components=[..]' (x: free), scaled
    node by node by the analytical field at the node when 'analyticalField' is given."""
    comps = {}
    if "components" in kw:
        comps.update({i: v for i, v in enumerate(_componentList(kw["components"], None)) if v is not None})
    comps.update({int(k) - 1: float(v) for k, v in kw.items() if _is_int(k)})
    scale = None
    if "analyticalField" in kw:
        f = model.analyticalFields[kw["analyticalField"]]
        scale = [f(*model.nodeCoordinates[n]) for n in np.asarray(model.nodeSets[kw["nSet"]])]
    return comps, scale


def _nodeLoad(kw):
    """nodeforces.py:85-108: 'components=[..]' (x: 0) or '1=..', '2=..', '3=..'."""
    if "components" in kw:
        return (_componentList(kw["components"], 0.0) + [0.0, 0.0, 0.0])[:3]
    return [float(kw.get(str(i + 1), 0.0)) for i in range(3)]


def _fieldOutputs(inputFile):
    """*fieldOutput 'create=perElement' definitions (edelweissfe/outputmanagers/ and utils/fieldoutput.py): name ->
    (result, element set, quadrature points); the GPU path uses them for the setfield action only."""
    out = {}
    for fo in inputFile["*fieldoutput"]:
        for ln in fo["data"]:
            d = _assignments([t for t in splitLineAtCommas(ln) if t])
            if d.get("create", "").lower() != "perelement":
                continue
            qp = d.get("quadraturePoint", "0")
            if ":" in qp:
                a, b = (int(t) if t else None for t in qp.split(":"))
                qps = slice(a, b)
            else:
                qps = np.array([int(t) for t in qp.split(",")])
            out[d["name"]] = (d["result"], d.get("elSet", "all"), qps)
    return out


def _makeAction(module, name, kw, model):
    from .stepping import BodyForce, Dirichlet, DistributedLoad, NodeForces, Options, SetField

    amp = _Amplitude(kw["f(t)"]) if "f(t)" in kw else None
    field = kw.get("field", "displacement")
    if module == "dirichlet":
        comps, scale = _dirichletValues(kw, model)
        return Dirichlet(name, model, kw["nSet"], comps, field=field, amplitude=amp, scale=scale)
    if module == "nodeforces":
        return NodeForces(name, model, kw["nSet"], _nodeLoad(kw), field=field, amplitude=amp)
    if module == "bodyforce":
        return BodyForce(name, model, kw["elSet"], _vector(kw["forceVector"]), amplitude=amp)
    if module == "distributedload":
        return DistributedLoad(name, model, kw["surface"], kw["type"], _vector(kw["magnitude"]), amplitude=amp)
    if module == "setfield":
        result, elSet, qps = model.fieldOutputs[kw["fieldOutput"]]
        kind = kw["type"]
        value = model.analyticalFields[kw["value"]] if kind == "analyticalField" else _vector(kw["value"])
        return SetField(name, model, result, elSet, qps, kind, value)
    if module == "options":
        return Options(name, kw.pop("category", name), **kw)
    raise NotImplementedError(f"step action {module} is not part of the GPU path")


def createStepsFromInputFile(model: FEModel, inputFile: dict):
    """Yields (AdaptiveStep, solver name) per *step; actions persist across steps and are updated when a later
    step defines them again (stepmanager.py:133-200). Must be consumed step by step: a step's actions are
    created against the model as it is when the step starts."""
    from .stepping import AdaptiveStep

    actions = {}
    for number, st in enumerate(inputFile["*step"]):
        opts = {k: v for k, v in st.items() if k not in ("data", "inputFile", "type", "solver", "jobname")}
        if st.get("type", "AdaptiveStep").lower() != "adaptivestep":
            raise NotImplementedError(f"step type {st['type']} is not part of the GPU path")
        seen = set()
        for line in st["data"]:
            module, *definition = splitLineAtCommas(line)
            module = module.lower()
            kw = _assignments(definition)
            name = kw.pop("name", None) or kw.get("category") or module
            if name in seen:
                raise Exception(f"StepAction {name} has multiple definitions in step {number}")
            seen.add(name)
            typ = actions.setdefault(module, {})
            if name in typ:
                act = typ[name]
                if module == "bodyforce":
                    act.updateStepAction(forceVector=_vector(kw["forceVector"]) if "forceVector" in kw else None,
                                         delta=_vector(kw["delta"]) if "delta" in kw else None,
                                         amplitude=_Amplitude(kw["f(t)"]) if "f(t)" in kw else None)
                elif module == "distributedload":
                    act.updateStepAction(magnitude=_vector(kw["magnitude"]) if "magnitude" in kw else None,
                                         delta=_vector(kw["delta"]) if "delta" in kw else None,
                                         amplitude=_Amplitude(kw["f(t)"]) if "f(t)" in kw else None)
                elif module == "options":
                    act.options.update({k: v for k, v in kw.items() if k != "category"})
                elif module == "dirichlet":
                    comps, scale = _dirichletValues(kw, model)
                    act.updateStepAction(comps, amplitude=_Amplitude(kw["f(t)"]) if "f(t)" in kw else None,
                                         scale=scale)
                elif module == "nodeforces":
                    act.updateStepAction(_nodeLoad(kw), amplitude=_Amplitude(kw["f(t)"]) if "f(t)" in kw else None)
                elif module == "setfield":
                    act.updateStepAction()
                else:
                    raise NotImplementedError(f"updating a {module} action in a later step is not part of the GPU path")
            else:
                typ[name] = _makeAction(module, name, dict(kw), model)
        yield AdaptiveStep(number, model, actions, **opts), st.get("solver", "default")


def finiteElementSimulation(inputFile: dict, verbose: bool = False, solverOptions: dict | None = None):
    """Builds the model of a parsed deck and solves its steps with NISTGPU (inputfiledrivensimulation.py:56-217,
    without output managers). Returns the model; model.U holds the final displacements (dof order)."""
    from .solvers.nonlinearimplicitstaticgpu import NISTGPU, _Journal

    model = createModelFromInputFile(inputFile)
    solverDefs = {s.get("name", "default"): s for s in inputFile["*solver"]}
    for s in solverDefs.values():
        if s["solver"].lower() not in _SOLVERS:
            raise NotImplementedError(f"solver {s['solver']} is not part of the GPU path")
    solvers = {}
    for step, solverName in createStepsFromInputFile(model, inputFile):
        if solverName not in solvers:
            s = solverDefs.get(solverName)
            opts = dict(_assignments(t for ln in (s["data"] if s else []) for t in splitLineAtCommas(ln)))
            cfg = opts.get("linsolverConfigFile")
            if cfg and not os.path.isabs(cfg):  # relative to the deck that names it
                opts["linsolverConfigFile"] = os.path.join(os.path.dirname(s["inputFile"]), cfg)
            opts.update(solverOptions or {})
            solvers[solverName] = NISTGPU(journal=_Journal(verbose=verbose), **opts)
        solvers[solverName].solveStep(step, model)
        model.solver = solvers[solverName]
    return model


def runReferenceTest(directory: str, verbose: bool = False, solverOptions: dict | None = None):
    """_run_tests_edelweissfe.py:88-109: runs <directory>/test.inp and compares the final displacements with
    <directory>/U.ref. Returns (passed at 1e-6 absolute, max absolute difference)."""
    inputFile = parseInputFile(os.path.join(directory, "test.inp"))
    model = finiteElementSimulation(inputFile, verbose=verbose, solverOptions=solverOptions)
    U = model.U.detach().cpu().numpy().ravel()
    Uref = np.loadtxt(os.path.join(directory, "U.ref"))
    err = float(np.abs(U - Uref).max())
    return err < 1e-6, err
