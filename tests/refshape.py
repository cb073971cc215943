"""Stand-ins shaped like the reference's model objects (edelweissfe/models/femodel.py:45-297, points/node.py,
sets/nodeset.py, sets/elementset.py, sections/solid.py, fields/nodefield.py, stepactions/dirichlet.py and
nodeforces.py): only the attributes the solver reads. They let the tests hand NISTGPU a model and step in the
reference's object shape without importing the reference."""
import numpy as np


class Node:
    def __init__(self, label, coordinates):
        self.label = label
        self.coordinates = np.asarray(coordinates, dtype=np.float64)
        self.fields = {}


class Element:
    def __init__(self, elType, elNumber, nodes):
        self.elType = elType
        self.elNumber = elNumber
        self.nodes = list(nodes)


class NamedSet(list):
    """NodeSet / ElementSet: an ordered collection of objects with a name."""

    def __init__(self, name, items):
        super().__init__(items)
        self.name = name

    def __hash__(self):
        return id(self)


class Section:
    def __init__(self, material, elSets):
        self.material = material
        self.elSets = list(elSets)


class NodeField:
    """fields/nodefield.py: value arrays per entry ("U", "P", "dU") with a row per node."""

    def __init__(self, nodes):
        self.nodes = list(nodes)
        self._indicesOfNodesInArray = {n: i for i, n in enumerate(self.nodes)}
        self._values = {k: np.zeros((len(self.nodes), 3)) for k in ("U", "P", "dU")}

    def __getitem__(self, key):
        return self._values[key]


class Model:
    def __init__(self):
        self.time = 0.0
        self.domainSize = 3
        self.nodes = {}
        self.elements = {}
        self.nodeSets = {}
        self.elementSets = {}
        self.sections = {}
        self.surfaces = {}
        self.nodeFields = {}
        self.constraints = {}
        self.scalarVariables = {}


class Dirichlet:
    """stepactions/dirichlet.py: nSet (a NodeSet), 0-based components, delta [nNodes, nComponents] over the step."""

    def __init__(self, nSet, values: dict):
        self.field = "displacement"
        self.nSet = nSet
        self._components = sorted(values)
        self.delta = np.tile([float(values[c]) for c in self._components], (len(nSet), 1))
        self.active = True

    @property
    def components(self):
        return self._components

    def applyAtStepStart(self, model):
        pass

    def applyAtIncrementStart(self, model, timeStep):
        pass

    def applyAtStepEnd(self, model):
        self.active = False

    def getDelta(self, timeStep):
        if self.active:
            return self.delta * timeStep.stepProgressIncrement
        return self.delta * 0.0


class NodeForces:
    """stepactions/nodeforces.py: nodeSet (a NodeSet), a force vector per node scaled by the step progress."""

    def __init__(self, nodeSet, vector):
        self.field = "displacement"
        self.nodeSet = nodeSet
        self.vector = np.asarray(vector, dtype=np.float64)

    def applyAtStepStart(self, model):
        pass

    def applyAtIncrementStart(self, model, timeStep):
        pass

    def applyAtStepEnd(self, model):
        pass

    def getCurrentLoad(self, timeStep):
        return np.tile(self.vector, (len(self.nodeSet), 1)) * timeStep.stepProgress


def model_from_arrays(X, blocks, nodeSets):
    """A reference-shaped model: nodes labelled 1.., elements labelled 1.. in block order; blocks =
    [(elType, conn, material)] with material a Marmot-provider dict {"name", "properties"}, one section per block
    on an element set named "block<i>"; nodeSets {name: node indices}."""
    m = Model()
    nodes = [Node(i + 1, x) for i, x in enumerate(X)]
    for n in nodes:
        m.nodes[n.label] = n
    label = 1
    for b, (elType, conn, material) in enumerate(blocks):
        els = []
        for row in conn:
            el = Element(elType, label, [nodes[i] for i in row])
            m.elements[label] = el
            els.append(el)
            label += 1
        es = NamedSet(f"block{b}", els)
        m.elementSets[es.name] = es
        m.sections[f"section{b}"] = Section(material, [es])
    for name, idx in nodeSets.items():
        m.nodeSets[name] = NamedSet(name, [nodes[i] for i in idx])
    used = sorted({i for _, conn, _ in blocks for i in np.asarray(conn).ravel()})
    m.nodeFields["displacement"] = NodeField([nodes[i] for i in used])
    return m
