"""A model in the reference's object shape, run on the GPU path.

The reference's solver reads its model through objects (edelweissfe/models/femodel.py:45-297):
``model.nodes`` (label -> Node with ``coordinates``), ``model.elements`` (label -> element with ``elType``,
``elNumber`` and ``nodes``), ``model.sections`` (each with ``material`` and ``elSets``: the section assigns its
material to the elements of its sets, ew/sections/solid.py:45-64, sectionbase.py:84-112), ``model.nodeSets`` /
``model.elementSets`` / ``model.surfaces`` (named sets of those objects), ``model.nodeFields`` (the field values
the solver writes back per node, nonlinearimplicitstatic.py:280-287) and ``model.time``.

``ReferenceModelAdapter`` turns such a model into the GPU path's ``FEModel``: node coordinates in
``model.nodes`` order, one ``ElementBlock`` per element type and material kind (elements keep their order in
``model.elements`` inside a block; several materials of one kind share a block through its material table), node
sets as node index arrays, element sets and surfaces by block. After a converged increment it writes U, P and dU
back into the reference's ``NodeField`` of the displacement field and advances ``model.time``.

Materials are read as the reference's providers hold them: a dict {"name", "properties"} (Marmot provider,
abqmodelconstructor.py:255-262) or a material object with ``materialProperties`` (LinearElasticMaterial,
VonMisesMaterial, NeoHookeanWaMaterial).
"""

from __future__ import annotations

from collections import OrderedDict

import numpy as np

from .model import ElementBlock, FEModel, Material, elementType

# material names of the reference's providers -> the GPU path's laws
_MATERIAL_NAMES = {"linearelastic": "linearelastic", "adlinearelastic": "linearelastic", "vonmises": "vonmises",
                   "neohookeanwa": "neohookeanwa", "neohookeanpencegouformulationa": "neohookeanwa",
                   "linearelasticmaterial": "linearelastic", "vonmisesmaterial": "vonmises",
                   "neohookeanwamaterial": "neohookeanwa", "stvenantkirchhoffisotropic": "stvenantkirchhoff",
                   "stvenantkirchhoff": "stvenantkirchhoff"}


def _material(mat):
    """(law name, properties) of a section's material in either provider's shape."""
    if isinstance(mat, dict):
        name, props = str(mat["name"]), mat["properties"]
    else:
        name, props = type(mat).__name__, getattr(mat, "materialProperties")
    key = name.lower()
    if key not in _MATERIAL_NAMES:
        raise NotImplementedError(f"material {name} is not available on the GPU path")
    return _MATERIAL_NAMES[key], np.asarray(props, dtype=np.float64).ravel()


class ReferenceModelAdapter:
    def __init__(self, refModel):
        self.refModel = refModel
        if getattr(refModel, "domainSize", 3) != 3:
            raise NotImplementedError("the GPU path covers 3D solid elements only")
        nodes = list(refModel.nodes.values())
        self.nodeIndex = {id(n): i for i, n in enumerate(nodes)}
        coords = np.array([np.asarray(n.coordinates, dtype=np.float64).ravel()[:3] for n in nodes]).reshape(-1, 3)
        gm = FEModel()
        gm.addNodes(coords, np.array([n.label for n in nodes], dtype=np.int64))
        gm.time = float(getattr(refModel, "time", 0.0))

        # every section assigns its material to the elements of its sets
        materialOf = {}
        for section in getattr(refModel, "sections", {}).values():
            for elSet in section.elSets:
                for el in elSet:
                    materialOf[id(el)] = section.material
        # blocks: (element type, material law) in order of first appearance; elements keep their order
        groups = OrderedDict()
        for el in refModel.elements.values():
            if id(el) not in materialOf:
                raise ValueError(f"element {el.elNumber} has no section with a material")
            law, props = _material(materialOf[id(el)])
            groups.setdefault((elementType(el.elType), law), []).append((el, props))
        self.elementLocation = {}  # id(element) -> (block index, index in the block)
        for (et, law), items in groups.items():
            conn = np.array([[self.nodeIndex[id(n)] for n in el.nodes] for el, _ in items], dtype=np.int32)
            block = ElementBlock(et, conn, labels=np.array([el.elNumber for el, _ in items], dtype=np.int64))
            distinct, index = [], np.zeros(len(items), dtype=np.int32)
            for i, (_, props) in enumerate(items):
                for j, q in enumerate(distinct):
                    if len(q) == len(props) and np.array_equal(q, props):
                        index[i] = j
                        break
                else:
                    index[i] = len(distinct)
                    distinct.append(props)
            block.setMaterials([Material(law, q) for q in distinct], index if len(distinct) > 1 else None)
            b = len(gm.elementBlocks)
            gm.addElementBlock(block)
            for i, (el, _) in enumerate(items):
                self.elementLocation[id(el)] = (b, i)

        for name, nSet in getattr(refModel, "nodeSets", {}).items():
            gm.nodeSets[name] = np.array([self.nodeIndex[id(n)] for n in nSet], dtype=np.int64)
        self.multiBlockElementSets = {}
        for name, elSet in getattr(refModel, "elementSets", {}).items():
            parts = self.elementSetParts(elSet)
            if len(parts) == 1:
                gm.elementSets[name] = parts[0]
            else:
                self.multiBlockElementSets[name] = parts
        for name, surface in getattr(refModel, "surfaces", {}).items():
            gm.surfaces[name] = {int(face): getattr(elSet, "name", None) or elSet for face, elSet in surface.items()}
        self.model = gm

        # the displacement NodeField: row of every node with dofs in its value arrays
        self.nodeField = getattr(refModel, "nodeFields", {}).get("displacement")

    def elementSetParts(self, elSet):
        """[(block index, element indices)] of a set of reference elements, one entry per block it touches."""
        byBlock = OrderedDict()
        for el in elSet:
            b, i = self.elementLocation[id(el)]
            byBlock.setdefault(b, []).append(i)
        return [(b, np.array(sorted(idx), dtype=np.int64)) for b, idx in byBlock.items()]

    def _fieldRows(self, dofManager):
        """(dof-vector node rows, NodeField rows) of the nodes that carry dofs and appear in the NodeField."""
        key = id(dofManager)
        if getattr(self, "_rowsKey", None) != key:
            nf = self.nodeField
            where = getattr(nf, "_indicesOfNodesInArray", None)
            nodes = list(self.refModel.nodes.values())
            src, dst = [], []
            for r, n in enumerate(dofManager.nodesWithDofs):
                node = nodes[n]
                row = where[node] if where is not None else r
                src.append(r)
                dst.append(row)
            self._rows = (np.array(src, dtype=np.int64), np.array(dst, dtype=np.int64))
            self._rowsKey = key
        return self._rows

    def readDisplacement(self, dofManager):
        """The NodeField's "U" as a dof vector (host), or None without a NodeField."""
        nf = self.nodeField
        if nf is None:
            return None
        src, dst = self._fieldRows(dofManager)
        U = np.zeros((dofManager.nDof // 3, 3))
        U[src] = np.asarray(nf["U"], dtype=np.float64).reshape(-1, 3)[dst]
        return U.ravel()

    def writeBack(self, dofManager, U, P, dU, time):
        """nonlinearimplicitstatic.py:280-296: U, P, dU into the NodeField, the model's time advanced."""
        nf = self.nodeField
        if nf is not None:
            src, dst = self._fieldRows(dofManager)
            for name, vec in (("U", U), ("P", P), ("dU", dU)):
                values = vec.detach().cpu().numpy().reshape(-1, 3)
                nf[name][dst] = values[src]
        self.refModel.time = time
