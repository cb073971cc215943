"""Dof numbering for the GPU path (edelweissfe/numerics/dofmanager.py:129-590, displacement field).

* Only nodes carrying the displacement field -- nodes used by an element (femodel.py:75-85,
  nodefield.py:87) -- get dofs, in node insertion order; dof = 3 * rank + component
  (dofmanager.py:259-300).
* ``idcsOfFieldsOnNodeSetsInDofVector[field][nSetName]``: the dofs of a node set, node-major
  (dofmanager.py:558-590).
* ``nAccumulatedNodalFluxesFieldwise``: sum over elements of nodes * 3 (dofmanager.py:354-411), the
  denominator of the averaged flux in the convergence test.
* Dof vectors are float64 device tensors.
"""

from __future__ import annotations

import numpy as np


class DofManager:
    def __init__(self, model, device="cuda"):
        import torch

        self.device = device
        self._torch = torch
        used = np.zeros(model.nNodes, dtype=bool)
        nAcc = 0
        for b in model.elementBlocks:
            used[b.conn.ravel()] = True
            nAcc += 3 * b.conn.size
        self.nodeDofBase = np.full(model.nNodes, -1, dtype=np.int64)
        self.nodeDofBase[used] = 3 * np.arange(int(used.sum()))
        self.nodesWithDofs = np.nonzero(used)[0]
        self.nDof = int(3 * used.sum())
        self.idcsOfFieldsInDofVector = {"displacement": slice(0, self.nDof)}
        self.idcsOfScalarVariablesInDofVector = {}
        self.nAccumulatedNodalFluxesFieldwise = {"displacement": nAcc}
        self.idcsOfFieldsOnNodeSetsInDofVector = {"displacement": {}}
        for name, nodes in model.nodeSets.items():
            base = self.nodeDofBase[np.asarray(nodes, dtype=np.int64)]
            base = base[base >= 0]
            self.idcsOfFieldsOnNodeSetsInDofVector["displacement"][name] = (
                base[:, None] + np.arange(3)[None, :]).ravel()

    def constructDofVector(self):
        return self._torch.zeros(self.nDof, dtype=self._torch.float64, device=self.device)

    def coordinatesAsDofVector(self, model):
        """Reference coordinates laid out like a dof vector: X[dof(node) + c] = x_c(node)."""
        X = np.zeros(self.nDof)
        X.reshape(-1, 3)[:] = model.nodeCoordinates[self.nodesWithDofs]
        return self._torch.as_tensor(X, dtype=self._torch.float64, device=self.device)

    def getNodeForIndexInDofVector(self, index):
        return int(self.nodesWithDofs[int(index) // 3])

    def writeDofVectorToNodeField(self, vec):
        """Returns the [nNodesWithField, 3] host array of a dof vector (NodeField values)."""
        return vec.detach().cpu().numpy().reshape(-1, 3)
