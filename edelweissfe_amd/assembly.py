"""GPU element evaluation + CSR assembly of one element block through libewfe.so.

``ElementAssembly`` is the device-side counterpart of the reference's per-step CSRGenerator plus the
element loop of NISTParallel:
  * construction = DofManager VIJ pattern + CSRGenerator.__init__ (dofmanager.py:522-557,
    csrgenerator.pyx:47-98): ewfe_plan_create builds the canonical CSR pattern, the element -> slot map and
    the colouring;
  * ``evaluate`` = computeElements + assembleStiffnessCSR (nonlinearimplicitstatic.py:794-849, 753-769,
    mk2.pyx:111-198, csrgenerator.pyx:100-115) in one stream-ordered call;
  * ``acceptLastState`` = MarmotElementWrapper.acceptLastState (element.pyx:278-281).
Quadrature-point state is kept on the device in plan order, [component][qp][element].
"""

from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import EWFE_F16, EWFE_F32, EWFE_F32_D64, EWFE_F32_MFMA, EWFE_F32_MFMA_D64, EWFE_F64, check, lib, ptr


class PlanLinearAlgebra:
    """The linear-algebra calls on a plan's CSR pattern (Dirichlet rows, SpMV, residual, PCG, multigrid), shared
    by the element plans of ElementAssembly and the system plan of several blocks (SystemAssembly)."""

    def _init_solver_state(self):
        self._pcg_ws = None
        self._mg = None
        self.mgLevels = None

    def __del__(self):
        try:
            if getattr(self, "_mg", None):
                lib().ewfe_mg_destroy(self._mg)
                self._mg = None
            if getattr(self, "_plan", None):
                lib().ewfe_plan_destroy(self._plan)
                self._plan = None
        except Exception:
            pass

    def pattern(self):
        """Host copy of the CSR pattern (indptr, indices), int32."""
        indptr = np.zeros(self.nDof + 1, dtype=np.int32)
        indices = np.zeros(self.nnz, dtype=np.int32)
        check(lib().ewfe_plan_pattern(self._plan, ptr(indptr), ptr(indices)))
        return indptr, indices

    def newValues(self, dtype="fp64"):
        """A zeroed CSR value array: "fp64" (float64 [nnz]), "fp32" (float32 [nnz]) or "fp32-d64" (float32 array of
        ewfe_vals_bytes / 4 entries: the nnz FP32 values, then FP64 diagonal blocks, 9 per node, rebuilt from the
        rounded off-diagonal values by every evaluation so that each row sums to zero as the FP64 matrix's does;
        ``csrValues`` gives the nnz values)."""
        if dtype == "fp64":
            return self.torch.zeros(self.nnz, dtype=self.torch.float64, device=self.device)
        if dtype == "fp32":
            return self.torch.zeros(self.nnz, dtype=self.torch.float32, device=self.device)
        if dtype == "fp32-d64":
            return self.torch.zeros(self._d64Length(), dtype=self.torch.float32, device=self.device)
        raise ValueError(f"unknown value precision {dtype!r} (fp64, fp32, fp32-d64)")

    def _d64Length(self):
        n = getattr(self, "_d64len", None)
        if n is None:
            nbytes = lib().ewfe_vals_bytes(self._plan, EWFE_F32_D64)
            check(int(min(nbytes, 0)))
            n = self._d64len = nbytes // 4
        return n

    def csrValues(self, vals):
        """The nnz CSR values of a value array (a view)."""
        return vals[: self.nnz]

    def _dtype_code(self, vals):
        if vals.dtype == self.torch.float64 and vals.numel() == self.nnz:
            return EWFE_F64
        if vals.dtype == self.torch.float32:
            if vals.numel() == self.nnz:
                return EWFE_F32
            if vals.numel() == self._d64Length():
                return EWFE_F32_D64
        raise TypeError("CSR values must come from newValues of this plan (float64 or float32 [nnz], or fp32-d64)")

    def _stream(self):
        return ctypes.c_void_p(self.torch.cuda.current_stream(self.device).cuda_stream)

    def dirichletRows(self, rows_dev, vals):
        check(lib().ewfe_dirichlet_rows(self._plan, ptr(rows_dev), rows_dev.numel(), ptr(vals),
                                        self._dtype_code(vals), self._stream()))

    def spmv(self, vals, x, y):
        check(lib().ewfe_spmv(self._plan, ptr(vals), self._dtype_code(vals), ptr(x), ptr(y), self._stream()))
        return y

    def residual(self, vals, b, x, r):
        """r = b - K x with row sums accurate to twice the working precision (refinement residuals)."""
        check(lib().ewfe_residual(self._plan, ptr(vals), self._dtype_code(vals), ptr(b), ptr(x), ptr(r),
                                  self._stream()))
        return r

    def multigrid(self, degree=2, ratio=20.0):
        """The smoothed-aggregation hierarchy of this plan (built on first use; patterns only)."""
        if self._mg is None:
            mg = ctypes.c_void_p()
            nl = ctypes.c_int()
            check(lib().ewfe_mg_create(self._plan, int(degree), float(ratio), ctypes.byref(mg), ctypes.byref(nl)))
            self._mg = mg
            sizes = []
            for lvl in range(nl.value):
                nn, nb = ctypes.c_int64(), ctypes.c_int64()
                check(lib().ewfe_mg_level_info(mg, lvl, ctypes.byref(nn), ctypes.byref(nb)))
                sizes.append((nn.value, nb.value))
            self.mgLevels = sizes
        return self._mg

    def multigridSetup(self, vals, rows_dev, finePrecision=None):
        """Numeric multigrid setup for the row-eliminated matrix `vals` (Dirichlet rows `rows_dev`). With
        ``finePrecision="fp16"`` the V-cycle's fine-level passes read a half-precision copy of `vals` made here
        (ewfe_mg_set_fine_precision); otherwise the values of each apply."""
        check(lib().ewfe_mg_set_fine_precision(self.multigrid(), EWFE_F16 if finePrecision == "fp16" else EWFE_F64))
        n_rows = 0 if rows_dev is None else rows_dev.numel()
        check(lib().ewfe_mg_setup(self.multigrid(), ptr(vals), self._dtype_code(vals), ptr(rows_dev), n_rows,
                                  self._stream()))

    def pcg(self, vals, b, x, rows_dev, rtol=1e-10, maxit=20000, check_every=32, precond="jacobi", setup=True):
        """PCG for the row-eliminated system, preconditioned by node-block Jacobi (``precond="jacobi"``) or by
        the smoothed-aggregation V-cycle (``precond="amg"``, set up here for `vals` unless ``setup=False``);
        returns (iterations, relative residual)."""
        if self._pcg_ws is None:
            nbytes = lib().ewfe_pcg_workspace_size(self._plan)
            self._pcg_ws = self.torch.empty((nbytes + 7) // 8, dtype=self.torch.float64, device=self.device)
        it = ctypes.c_int()
        rr = ctypes.c_double()
        n_rows = 0 if rows_dev is None else rows_dev.numel()
        if precond == "amg":
            if setup:
                self.multigridSetup(vals, rows_dev)
            mg = self.multigrid()
            check(lib().ewfe_pcg_mg(mg, ptr(vals), self._dtype_code(vals), ptr(b), ptr(x), ptr(rows_dev), n_rows,
                                    float(rtol), int(maxit), int(check_every), ptr(self._pcg_ws), ctypes.byref(it),
                                    ctypes.byref(rr), self._stream()))
        elif precond == "jacobi":
            check(lib().ewfe_pcg(self._plan, ptr(vals), self._dtype_code(vals), ptr(b), ptr(x), ptr(rows_dev),
                                 n_rows, float(rtol), int(maxit), int(check_every), ptr(self._pcg_ws),
                                 ctypes.byref(it), ctypes.byref(rr), self._stream()))
        else:
            raise ValueError(f"unknown preconditioner {precond!r}")
        return it.value, rr.value


class ElementAssembly(PlanLinearAlgebra):
    def __init__(self, block, dofManager, model, device="cuda", pipelineChunks=0):
        import torch

        self.torch = torch
        self.device = torch.device(device)
        if self.device.type != "cuda":
            raise RuntimeError("ElementAssembly runs on a HIP device only")
        if self.device.index is None:
            self.device = torch.device("cuda", torch.cuda.current_device())
        torch.cuda.set_device(self.device)
        self.block = block
        self.nDof = dofManager.nDof
        L = lib()
        plan = ctypes.c_void_p()
        nnz = ctypes.c_int64()
        conn = np.ascontiguousarray(block.conn, dtype=np.int32)
        node_dof = np.ascontiguousarray(dofManager.nodeDofBase, dtype=np.int64)
        check(L.ewfe_plan_create_chunked(ctypes.byref(plan), block.kind, len(conn), ptr(conn), len(node_dof),
                                         ptr(node_dof), self.nDof, ctypes.byref(nnz), int(pipelineChunks)))
        self._plan = plan
        self.nnz = nnz.value
        nch, structured = ctypes.c_int32(), ctypes.c_int32()
        check(L.ewfe_plan_pipeline_info(plan, ctypes.byref(nch), ctypes.byref(structured)))
        self.pipelineChunks = nch.value
        self.structuredMesh = bool(structured.value)
        order = np.zeros(len(conn), dtype=np.int64)
        ncol = ctypes.c_int32()
        check(L.ewfe_plan_element_order(plan, ptr(order), ctypes.byref(ncol)))
        self.elementOrder = order
        self.nColors = ncol.value
        ip, ix = ctypes.c_void_p(), ctypes.c_void_p()
        check(L.ewfe_plan_pattern_dev(plan, ctypes.byref(ip), ctypes.byref(ix)))
        self.X = dofManager.coordinatesAsDofVector(model)
        self.matKind = block.material.kind
        mp = np.zeros(8)
        mp[: len(block.material.materialProperties)] = block.material.materialProperties[:8]
        self.matParams = mp
        self.elParams = (None if block.elParams is None else
                         torch.as_tensor(block.elParams, dtype=torch.float64, device=self.device))
        if block.materialIndex is not None:  # several materials: the plan's per-element parameter sets
            table = np.ascontiguousarray(block.materialParameterTable())
            check(L.ewfe_plan_set_materials(plan, len(table), ptr(table),
                                            ptr(np.ascontiguousarray(block.materialIndex, dtype=np.int32))))
        self.nState = L.ewfe_state_size(self.matKind)
        nInt = block.nInt
        if self.nState > 0:
            block.stateVars = torch.zeros(self.nState * nInt * len(conn), dtype=torch.float64, device=self.device)
            block.stateVarsTemp = torch.zeros_like(block.stateVars)
        else:
            block.stateVars = block.stateVarsTemp = None
        block.planOrder = order
        self.pnewdt = torch.empty(1, dtype=torch.float64, device=self.device)
        self._init_solver_state()

    # ------------------------------------------------------------------------------------------------
    def evaluate(self, U_np, dU, P, F, vals, mfma=False):
        """Fills P (= sum Pe), F (= sum |Pe|), the CSR values and the trial state. Returns pNewDT (device
        tensor, 1e36 or < 1 when a quadrature point asks for a cutback). FP32 values: ``mfma=True`` forms the
        tangents in FP32 on the matrix cores (8-node elements, and 20-node elements assembled through the row-node
        gather), otherwise in FP32 arithmetic on the vector units (8-node elements) or with FP64 FMAs rounded
        when stored (20-node elements); fp32-d64 values (``newValues``) then also
        receive their FP64 diagonal blocks."""
        for t in (U_np, dU, P, F):
            if t.numel() != self.nDof or t.dtype != self.torch.float64 or not t.is_contiguous():
                raise ValueError("dof vectors must be contiguous float64 tensors of length nDof")
        code = self._dtype_code(vals)
        b = self.block
        check(lib().ewfe_eval_assemble(self._plan, self.matKind, ptr(self.matParams), ptr(self.elParams),
                                       ptr(self.X), ptr(U_np), ptr(dU), ptr(b.stateVars), ptr(b.stateVarsTemp),
                                       ptr(P), ptr(F), ptr(vals),
                                       {EWFE_F32: EWFE_F32_MFMA, EWFE_F32_D64: EWFE_F32_MFMA_D64}.get(code, code)
                                       if mfma else code, ptr(self.pnewdt),
                                       self._stream()))
        return self.pnewdt

    def bodyForce(self, PExt, force, elements=None):
        """PExt += the body force `force` (3 values per unit volume) integrated over `elements` (element indices
        of this block; all when None), reference configuration (displacementelement/element.py:348-371)."""
        if PExt.numel() != self.nDof or PExt.dtype != self.torch.float64 or not PExt.is_contiguous():
            raise ValueError("PExt must be a contiguous float64 tensor of length nDof")
        mask = None
        if elements is not None and len(elements) != len(self.block):
            if getattr(self, "_maskKey", None) is not elements:
                m = np.zeros(len(self.block), dtype=np.uint8)
                m[np.asarray(elements, dtype=np.int64)] = 1
                self._mask = self.torch.as_tensor(m, device=self.device)
                self._maskKey = elements
            mask = self._mask
        f = np.ascontiguousarray(np.asarray(force, dtype=np.float64).ravel()[:3])
        check(lib().ewfe_body_force(self._plan, ptr(self.X), ptr(mask), ptr(f), ptr(PExt), self._stream()))

    def surfacePressure(self, PExt, elements, face, pressure):
        """PExt += the pressure `pressure` (positive pushes into the element) on face `face` (Abaqus numbering 1..6)
        of `elements` (element indices of this block), reference configuration; the distributed load of type
        pressure of the Marmot solid elements (nonlinearimplicitstatic.py:460-501). The device copy of `elements`
        is kept per array object: pass a new array rather than changing one in place."""
        if PExt.numel() != self.nDof or PExt.dtype != self.torch.float64 or not PExt.is_contiguous():
            raise ValueError("PExt must be a contiguous float64 tensor of length nDof")
        if int(face) not in range(1, 7):
            raise ValueError(f"face {face}: hexahedra have faces 1..6")
        els = np.asarray(elements, dtype=np.int64).ravel()
        if len(els) and (els.min() < 0 or els.max() >= len(self.block) or len(np.unique(els)) != len(els)):
            raise ValueError("elements must be distinct element indices of this block")
        # device copies of the element lists of the surfaces in use: id -> (the caller's array, its device copy),
        # at most 16 (a surface load's lists are the same objects in every iteration)
        cache = self.__dict__.setdefault("_faceSets", {})
        hit = cache.get(id(elements))
        if hit is None or hit[0] is not elements:
            if len(cache) >= 16:
                cache.pop(next(iter(cache)))
            hit = cache[id(elements)] = (elements, self.torch.as_tensor(els.astype(np.int32), device=self.device))
        check(lib().ewfe_surface_pressure(self._plan, ptr(self.X), ptr(hit[1]), len(els), int(face),
                                          float(pressure), ptr(PExt), self._stream()))

    def qpResults(self, result, U, which="converged"):
        """[nEl, nInt, 6] device tensor (element order) of the quadrature-point "stress" or "strain" at the
        displacement U, with the converged ("converged") or trial ("trial") state for J2 stresses."""
        if result not in ("stress", "strain"):
            raise KeyError(f"result {result} is not available on the GPU path (stress, strain)")
        if U.numel() != self.nDof or U.dtype != self.torch.float64 or not U.is_contiguous():
            raise ValueError("U must be a contiguous float64 tensor of length nDof")
        b = self.block
        state = b.stateVars if which == "converged" else b.stateVarsTemp
        out = self.torch.empty((len(b), b.nInt, 6), dtype=self.torch.float64, device=self.device)
        check(lib().ewfe_qp_results(self._plan, self.matKind, ptr(self.matParams), ptr(self.elParams), ptr(self.X),
                                    ptr(U), ptr(state), 0 if result == "stress" else 1, ptr(out), self._stream()))
        return out

    def stateInElementOrder(self, which="trial"):
        """[nEl, nInt, nState] host array of the state in original element order."""
        t = self.block.stateVarsTemp if which == "trial" else self.block.stateVars
        if t is None:
            return np.zeros((len(self.block), self.block.nInt, 0))
        a = t.detach().cpu().numpy().reshape(self.nState, self.block.nInt, -1)  # [s][q][p]
        out = np.empty((len(self.block), self.block.nInt, self.nState))
        out[self.elementOrder] = np.transpose(a, (2, 1, 0))
        return out

    def setStateFromElementOrder(self, state):
        """Inverse of stateInElementOrder for the converged state."""
        if self.nState == 0:
            return
        s = np.asarray(state, dtype=np.float64)[self.elementOrder]  # [p][q][s]
        t = self.torch.as_tensor(np.ascontiguousarray(np.transpose(s, (2, 1, 0))).ravel(), device=self.device)
        self.block.stateVars.copy_(t)


class SystemAssembly(PlanLinearAlgebra):
    """The linear system of a model with several element blocks: the system plan of ewfe_plan_create_union over the
    blocks' plans (the row-wise union of their patterns, the reference's single VIJ pattern over all elements,
    dofmanager.py:522-557, csrgenerator.pyx:47-98), on which the Dirichlet rows, SpMV, residual, PCG and multigrid
    run; ``accumulate`` adds a block's CSR values into the system's."""

    def __init__(self, assemblies, dofManager, device="cuda"):
        import torch

        self.torch = torch
        self.device = assemblies[0].device
        self.nDof = dofManager.nDof
        plans = (ctypes.c_void_p * len(assemblies))(*[a._plan.value for a in assemblies])
        node_dof = np.ascontiguousarray(dofManager.nodeDofBase, dtype=np.int64)
        plan, nnz = ctypes.c_void_p(), ctypes.c_int64()
        check(lib().ewfe_plan_create_union(ctypes.byref(plan), plans, len(assemblies), len(node_dof), ptr(node_dof),
                                           ctypes.byref(nnz)))
        self._plan = plan
        self.nnz = nnz.value
        self._init_solver_state()

    def accumulate(self, block, blockValues, vals):
        """vals += the CSR values of block `block` (its plan's pattern), mapped into the system pattern. With fp32-d64
        system values the system's FP64 diagonal blocks are rebuilt from its summed values after each block (the
        block's own FP64 blocks are not read)."""
        if blockValues.dtype != vals.dtype:
            raise TypeError("block and system values must have one precision")
        check(lib().ewfe_csr_accumulate(self._plan, int(block), ptr(blockValues), ptr(vals), self._dtype_code(vals),
                                        self._stream()))


_ = _lib  # re-exported for callers that need the enum values
