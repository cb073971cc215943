"""Quadrature-point results of the GPU path, gathered like the reference's ElementResultCollector
(edelweissfe/utils/elementresultcollector.pyx:37-127): one array [elements, quadrature points, 6] (or
[elements, 6] for a single point) of the per-point "stress" or "strain" that the reference's elements keep in
their state (displacementelement/element.py:222-235, displacementtlelement/element.py:377-412). The values are
computed on the device from the model's current displacements and converged state whenever they are requested.
"""

from __future__ import annotations

import numpy as np


class ElementResultCollector:
    """``ElementResultCollector(elements, quadraturePoints, result)`` of the reference for a model solved by
    NISTGPU: ``elements`` are element indices of the model's element block (or the name of an element set),
    ``quadraturePoints`` a range or slice of point numbers, ``result`` "stress" or "strain".
    ``getCurrentResults()`` returns a host array; the caller copies it to keep it (the reference's contract)."""

    def __init__(self, solver, model, elements, quadraturePoints, result: str):
        self.solver, self.model, self.result = solver, model, result
        if isinstance(elements, str):
            _, elements = model.elementSets[elements]
        self.elements = np.asarray(elements, dtype=np.int64)
        nInt = model.elementBlocks[0].nInt
        qps = range(nInt)[quadraturePoints] if isinstance(quadraturePoints, slice) else quadraturePoints
        self.quadraturePoints = np.asarray(list(qps), dtype=np.int64)
        if len(self.elements) == 0 or len(self.quadraturePoints) == 0:
            raise ValueError("ElementResultCollector needs elements and quadrature points")
        if self.quadraturePoints.min() < 0 or self.quadraturePoints.max() >= nInt:
            raise IndexError(f"quadrature points outside 0..{nInt - 1}")
        self.resultsTable = None

    def update(self):
        import torch

        asm = self.solver.assembly
        U = self.model.U
        full = asm.qpResults(self.result, U, which="converged")
        e = torch.as_tensor(self.elements, device=full.device)
        q = torch.as_tensor(self.quadraturePoints, device=full.device)
        self.resultsTable = full.index_select(0, e).index_select(1, q).cpu().numpy()
        if len(self.quadraturePoints) == 1:
            self.resultsTable = self.resultsTable.reshape(len(self.elements), -1)

    def getCurrentResults(self) -> np.ndarray:
        self.update()
        return self.resultsTable
