"""NISTGPU: the nonlinear implicit static Newton solver with element evaluation, CSR assembly, Dirichlet
rows and the linear solve on an MI355X.

Method for method the same contract as the reference's NIST solver
(edelweissfe/solvers/nonlinearimplicitstatic.py:60-1084):

  solveStep(step, model, fieldOutputController, outputmanagers)                    :110-332
  solveIncrement(U_n, dU, P, K, stepActions, model, timeStep, prevTimeStep,
                 extrapolation, maxIter, maxGrowingIter)                          :334-458
  computeElements(elements, U_np, dU, P, K, F, timeStep) -> (P, K, F)              :794-849
  assembleStiffnessCSR(K)                                                          :753-769
  applyDirichletK(K, dirichlets) / applyDirichlet(timeStep, R, dirichlets)         :559-623
  checkConvergence(R, ddU, F, iterationCounter, residualHistory)                   :625-725
  linearSolve(A, b)                                                                :727-751
  computeSpatialAveragedFluxes(F)                                                  :771-792
  extrapolateLastIncrement(...) / checkDivergingSolution(...)                       :947-1006
  findDirichletIndices(dirichlet)                                                  :1077-1084

Differences that are part of the design (not of the semantics):
  * dof vectors are float64 device tensors; "K" is the CSR value array of the plan's pattern (the VIJ
    array is never materialised), filled by computeElements in the same launch as P and F, so the
    "CSR generation" phase is fused into "elements" and its timer stays 0;
  * the linear solve is device PCG on the row-eliminated matrix, preconditioned by node-block Jacobi
    (option ``linsolver="pcg"``) or by a smoothed-aggregation multigrid V-cycle (``linsolver="amg"``, the
    device counterpart of the reference's pyamg-preconditioned "gmres" solver; the default ``"auto"`` takes
    the multigrid from 20,000 dofs on when the mesh allows it, block Jacobi otherwise), converged to
    ``pcgTolerance`` relative residual; the Newton loop's own flux/correction tolerances decide convergence
    exactly as in the reference;
  * ``tangentPrecision="fp32"`` stores the tangent in FP32 (element tangents of 8-node elements in FP32 arithmetic;
    of 20-node elements from FP64 FMAs, rounded when stored), ``"fp32-mfma"`` forms the element tangents in FP32 on
    the matrix cores; residual, return mapping and update stay FP64 in both. With ``fp32TangentDiagonal="rebuilt"``
    (the default) the FP32 tangent carries FP64 corrections of its diagonal node blocks (``newValues("fp32-d64")``)
    that make every row of the matrix sum to zero like the FP64 tangent's (rigid translations): rounding to FP32
    otherwise limits Newton's contraction to ~1e-4..1e-3 per iteration at these condition numbers (the J2 cube at
    128^3: 10-12 instead of 7 solves), with the corrections it converges in the FP64 tangent's iterations.
    ``"assembled"`` keeps the plain FP32 matrix. Newton converges to the same FP64 residual tolerance either way,
    extra iterations counted; the FP32 tangent's linear systems are solved to ``pcgToleranceFP32`` (default: the
    FP64 ``pcgTolerance`` with rebuilt diagonals, 1e-4 with assembled ones, where a tighter solve buys nothing),
    the convergence test (FP64 residual and correction) is unchanged;
  * the 1e-7 absolute floor of the flux test is the option ``fluxResidualFloor`` (default 1e-7);
  * ``multigridSetupPerIncrement=True`` sets the multigrid hierarchy up on the first solve of an increment only: the
    later corrections of the increment are preconditioned by that hierarchy (its prolongator, coarse operators,
    block-diagonal inverse and spectral bound), with the current matrix in the fine-level smoother and in CG;
  * ``multigridFinePrecision="fp16"`` runs the V-cycle's fine-level smoothing and residual passes on a half-precision
    copy of the matrix the hierarchy was set up for (scaled by a power of two; "matrix", the default, uses the
    current matrix in its own precision); CG, its residuals and the Newton loop keep the assembled values;
  * ``linsolverRefinement=k`` adds k steps of iterative refinement to every FP64 linear solve: the residual
    r = b - K x is formed with row sums in twice the working precision (ewfe_residual) and K d = r is solved to
    the same PCG tolerance for the correction x += d. A PCG stopped at a relative residual of 1e-14 leaves an
    error of up to cond(K) * 1e-14 in x, a direct factorisation with refinement in working precision (the
    reference's solvers) up to cond(K) * 1e-16; with the extended residuals the refined x is accurate to
    working precision whenever cond(K) * 1e-16 < 1.
"""

from __future__ import annotations

import json
import time as _time

import numpy as np

from .._lib import EwfeError
from ..adapter import ReferenceModelAdapter
from ..assembly import ElementAssembly, SystemAssembly
from ..dofmanager import DofManager
from ..model import FEModel
from ..stepping import (ConditionalStop, CutbackRequest, DivergingSolution, ReachedMaxIncrements,
                        ReachedMaxIterations, ReachedMinIncrementSize, StepFailed)

TIMING_KEYS = ["elements", "constraints", "distributed loads", "body forces", "dirichlet R", "CSR generation",
               "dirichlet K", "linear solve", "convergence check", "output"]  # edelweissfe/config/timing.py:31-46

DEFAULT_JOB_INFO = {  # edelweissfe/config/phenomena.py:59-93, displacement field
    "fieldCorrectionTolerance": {"displacement": 1e-8},
    "fluxResidualTolerance": {"displacement": 1e-8},
    "fluxResidualToleranceAlternative": {"displacement": 5e-3},
}


def _setName(action, attribute):
    """The node set of a step action: its name (the GPU path's actions) or the NodeSet object of a reference
    action (dirichlet.py: ``nSet``, nodeforces.py: ``nodeSet``), whose name keys the DofManager's index lists."""
    name = getattr(action, "nSetName", None)
    return name if name is not None else getattr(action, attribute).name


class _Journal:
    """Minimal stand-in for the reference's Journal (ew/journal/journal.py): keeps the lines, optionally
    prints them (to ``stream``, default stdout)."""

    def __init__(self, verbose=False, stream=None):
        self.verbose = verbose
        self.stream = stream
        self.lines = []

    def message(self, msg, ident="", level=0):
        self.lines.append(msg)
        if self.verbose:
            print(f"{ident:>20}: {msg}", file=self.stream, flush=True)

    def errorMessage(self, msg, ident=""):
        self.message("ERROR " + msg, ident)


class NISTGPU:
    identification = "NISTGPUSolver"

    NISTOptions = {
        "defaultMaxIter": 10,
        "defaultCriticalIter": 5,
        "defaultMaxGrowingIter": 10,
        "extrapolation": "linear",
        "linsolver": "auto",
        "linsolverConfigFile": "",
        "tangentPrecision": "fp64",
        "pcgTolerance": 1e-10,
        "pcgToleranceFP32": None,
        "pcgMaxIter": 50000,
        "linsolverRefinement": 0,
        "multigridSetupPerIncrement": False,
        "multigridFinePrecision": "matrix",
        "fp32TangentDiagonal": "rebuilt",
        "fluxResidualFloor": 1e-7,
        "device": "cuda",
    }

    AUTO_AMG_MIN_DOFS = 20000  # linsolver="auto": the multigrid from this many dofs on
    # the reference's linear solver names (edelweissfe/config/linsolve.py:46-86): its direct solvers map to the
    # default choice, its algebraic-multigrid Krylov solvers to the multigrid PCG
    REFERENCE_LINSOLVERS = {"superlu": "auto", "umfpack": "auto", "pardiso": "auto", "panuapardiso": "auto",
                            "klu": "auto", "petsclu": "auto", "mumps": "auto", "gmres": "amg", "amgcl": "amg"}

    def __init__(self, jobInfo=None, journal=None, **kwargs):
        jobInfo = jobInfo or DEFAULT_JOB_INFO
        self.journal = journal or _Journal()
        self.fieldCorrectionTolerances = jobInfo["fieldCorrectionTolerance"]
        self.fluxResidualTolerances = jobInfo["fluxResidualTolerance"]
        self.fluxResidualTolerancesAlt = jobInfo["fluxResidualToleranceAlternative"]
        self.options = self.NISTOptions.copy()
        self._updateOptions(kwargs, self.journal)
        self.assemblies = {}
        self._adapters = {}
        self._systemKey = None
        self.system = None
        self.history = []  # per converged increment: list of (|R|inf, |ddU|inf) per iteration
        self.linearSolveIterations = []
        self.multigridSetups = 0  # numeric multigrid setups so far
        self._mgSetUpThisIncrement = False

    def _updateOptions(self, updatedOptions, journal):
        for k, v in updatedOptions.items():
            if k in self.NISTOptions:
                default = self.NISTOptions[k]
                # options without a default value (None) are tolerances: floats, or None for the default rule
                self.options[k] = (None if v is None or (isinstance(v, str) and v.lower() == "none") else float(v)) \
                    if default is None else type(default)(v)
            else:
                raise AttributeError("Invalid option {:} for {:}".format(k, self.identification))

    # ------------------------------------------------------------------------------------------------
    def _sync(self):
        self.torch.cuda.synchronize()

    def solveStep(self, step, model, fieldOutputController=None, outputmanagers=()):
        import torch

        self.torch = torch
        try:
            self._updateOptions(step.actions["options"]["NISTSolver"].options, self.journal)
        except KeyError:
            pass
        if self.options["linsolverConfigFile"]:
            with open(self.options["linsolverConfigFile"]) as f:
                cfg = json.load(f)
            self._updateOptions({k: v for k, v in cfg.items() if k in self.NISTOptions}, self.journal)
        if self.options["linsolver"].lower() not in ("pcg", "amg", "auto") + tuple(self.REFERENCE_LINSOLVERS):
            raise AttributeError("invalid linear solver {:} requested".format(self.options["linsolver"]))
        # a model in the reference's object shape (model.nodes / model.elements / model.sections ...) runs through
        # its GPU-path image; step actions and results stay with the given model
        refModel = model
        if isinstance(model, FEModel):
            self._adapter = None
        else:
            self._adapter = self._adapters.get(id(model))
            if self._adapter is None or self._adapter.refModel is not model:
                self._adapter = self._adapters[id(model)] = ReferenceModelAdapter(model)
            model = self._adapter.model
        self.gpuModel = model
        device = self.options["device"]
        tic_setup = _time.perf_counter()
        self.theDofManager = DofManager(model, device=device)
        self.journal.message("total size of eq. system: {:}".format(self.theDofManager.nDof), self.identification)
        for block in model.elementBlocks:
            if not block.hasMaterial:
                raise Exception("No material was assigned to the element block")
        self.blockAssemblies = []
        for block in model.elementBlocks:
            asm = self.assemblies.get(id(block))
            if asm is None or asm.block is not block or asm.nDof != self.theDofManager.nDof:
                asm = self.assemblies[id(block)] = ElementAssembly(block, self.theDofManager, model, device=device)
            self.blockAssemblies.append(asm)
        self.assembly = self.blockAssemblies[0]
        if len(self.blockAssemblies) == 1:
            self.system = self.assembly
        else:  # one linear system over every block's elements (the system plan of the blocks' plans)
            key = tuple(id(a) for a in self.blockAssemblies)
            if self._systemKey != key:
                self.system = SystemAssembly(self.blockAssemblies, self.theDofManager, device=device)
                self._systemKey = key
        self._blockValues = None
        self._linsolver = self.REFERENCE_LINSOLVERS.get(self.options["linsolver"].lower(),
                                                        self.options["linsolver"].lower())
        if self._linsolver == "auto":
            # the multigrid from AUTO_AMG_MIN_DOFS dofs on, where its setup pays; block Jacobi below and on meshes
            # the multigrid does not take (dofs that are not 3 per node)
            self._linsolver = "pcg"
            if self.theDofManager.nDof >= self.AUTO_AMG_MIN_DOFS:
                try:
                    self.system.multigrid()
                    self._linsolver = "amg"
                except EwfeError as e:
                    self.journal.message(f"multigrid not available ({e}); node-block Jacobi PCG", self.identification)
        if self._linsolver == "amg":
            self.system.multigrid()  # aggregation and sparsity patterns: once per mesh
        self.computationTimes = dict.fromkeys(TIMING_KEYS + ["multigrid setup", "setup"], 0.0)
        # dof numbering, CSR pattern, element plan, multigrid patterns: once per step (mesh), not per increment
        self.computationTimes["setup"] = _time.perf_counter() - tic_setup
        tp = self.options["tangentPrecision"].lower()
        if tp not in ("fp64", "fp32", "fp32-mfma"):
            raise AttributeError("invalid tangentPrecision {:} (fp64, fp32, fp32-mfma)".format(tp))
        self._mfmaTangent = tp == "fp32-mfma"
        dg = self.options["fp32TangentDiagonal"].lower()
        if dg not in ("assembled", "rebuilt"):
            raise AttributeError("invalid fp32TangentDiagonal {:} (assembled, rebuilt)".format(dg))
        K = self.system.newValues("fp64" if tp == "fp64" else ("fp32-d64" if dg == "rebuilt" else "fp32"))

        extrapolation = self.options["extrapolation"]
        maxIter, criticalIter = step.maxIter, step.criticalIter
        maxGrowingIter, cutbackFactor = step.maxGrowIter, step.cutbackFactor

        U = getattr(model, "U", None)
        U = self.theDofManager.constructDofVector() if U is None else U.clone()
        if self._adapter is not None:  # nonlinearimplicitstatic.py:194-195: U from the NodeField
            U0 = self._adapter.readDisplacement(self.theDofManager)
            if U0 is not None:
                U.copy_(torch.as_tensor(U0, dtype=torch.float64, device=U.device))
        P = self.theDofManager.constructDofVector()
        dU = self.theDofManager.constructDofVector()
        self._dirichletRowsCache = {}
        prevTimeStep = None
        self.applyStepActionsAtStepStart(refModel, step.actions)
        try:
            for timeStep in step.getTimeStep():
                try:
                    U, dU, P, iterationCounter, incrementResidualHistory = self.solveIncrement(
                        U, dU, P, K, step.actions, refModel, timeStep, prevTimeStep, extrapolation, maxIter,
                        maxGrowingIter)
                except CutbackRequest as e:
                    self.journal.message(str(e), self.identification, 1)
                    step.discardAndChangeIncrement(max(e.cutbackSize, cutbackFactor))
                    prevTimeStep = None
                except (ReachedMaxIterations, DivergingSolution) as e:
                    self.journal.message(str(e), self.identification, 1)
                    step.discardAndChangeIncrement(cutbackFactor)
                    prevTimeStep = None
                else:
                    prevTimeStep = timeStep
                    if iterationCounter >= criticalIter:
                        step.preventIncrementIncrease()
                    model.U, model.P, model.dU = U, P, dU
                    model.advanceToTime(timeStep.totalTime)
                    if self._adapter is not None:
                        self._adapter.writeBack(self.theDofManager, U, P, dU, timeStep.totalTime)
                    self.history.append(list(self._currentHistory))
                    self.journal.message("Converged in {:} iteration(s)".format(iterationCounter),
                                         self.identification, 1)
        except (ReachedMaxIncrements, ReachedMinIncrementSize):
            self.journal.errorMessage("Incrementation failed", self.identification)
            raise StepFailed()
        except ConditionalStop:
            self.applyStepActionsAtStepEnd(refModel, step.actions)
        else:
            self.applyStepActionsAtStepEnd(refModel, step.actions)
        return True, refModel

    def solveIncrement(self, U_n, dU, P, K, stepActions, model, timeStep, prevTimeStep, extrapolation, maxIter,
                       maxGrowingIter):
        iterationCounter = 0
        self._mgSetUpThisIncrement = False
        incrementResidualHistory = dict.fromkeys(self.theDofManager.idcsOfFieldsInDofVector, (0.0, 0))
        self._currentHistory = []
        dm = self.theDofManager
        R, F, PExt, U_np = (dm.constructDofVector() for _ in range(4))
        ddU = None
        dirichlets = list(stepActions["dirichlet"].values())
        nodeforces = list(stepActions["nodeforces"].values())
        bodyforces = list(stepActions.get("bodyforce", {}).values())
        distributedloads = list(stepActions.get("distributedload", {}).values())
        self.applyStepActionsAtIncrementStart(model, timeStep, stepActions)
        dU, isExtrapolatedIncrement = self.extrapolateLastIncrement(extrapolation, timeStep, dU, dirichlets,
                                                                    prevTimeStep, model)
        elements = self.gpuModel.elementBlocks
        while True:
            U_np.copy_(U_n)
            U_np += dU
            PExt.zero_()
            P, K, F = self.computeElements(elements, U_np, dU, P, K, F, timeStep)
            PExt, K = self.assembleLoads(nodeforces, distributedloads, bodyforces, U_np, PExt, K, timeStep)
            R.copy_(P)
            R += PExt
            if iterationCounter == 0 and not isExtrapolatedIncrement and dirichlets:
                R = self.applyDirichlet(timeStep, R, dirichlets)
            else:
                for dirichlet in dirichlets:
                    R[self._rows(dirichlet)] = 0.0
                converged, _ = self.checkConvergence(R, ddU, F, iterationCounter, incrementResidualHistory)
                if converged:
                    break
                if self.checkDivergingSolution(incrementResidualHistory, maxGrowingIter):
                    raise DivergingSolution("Residual grew {:} times, cutting back".format(maxGrowingIter))
                if iterationCounter == maxIter:
                    raise ReachedMaxIterations("Reached max. iterations in current increment, cutting back")
            K_ = self.assembleStiffnessCSR(K)
            K_ = self.applyDirichletK(K_, dirichlets)
            ddU = self.linearSolve(K_, R)
            dU += ddU
            iterationCounter += 1
        return U_np, dU, P, iterationCounter, incrementResidualHistory

    # ------------------------------------------------------------------------------------------------
    def computeElements(self, elements, U_np, dU, P, K, F, timeStep):
        tic = _time.perf_counter()
        mfma = getattr(self, "_mfmaTangent", False)
        if len(self.blockAssemblies) == 1:
            pnewdt = self.assembly.evaluate(U_np, dU, P, F, K, mfma=mfma)
        else:
            # block by block (blocks hold consecutive runs of the model's elements): P, F and the system's CSR
            # values are the sums of the blocks' in block order
            if self._blockValues is None or self._blockValues[0][0].dtype != K.dtype:
                dm = self.theDofManager
                self._blockValues = [(a.newValues("fp64" if K.dtype == self.torch.float64 else "fp32"),
                                      dm.constructDofVector(), dm.constructDofVector()) for a in self.blockAssemblies]
            P.zero_()
            F.zero_()
            K.zero_()
            pnewdt = None
            for b, (asm, (Kb, Pb, Fb)) in enumerate(zip(self.blockAssemblies, self._blockValues)):
                pb = asm.evaluate(U_np, dU, Pb, Fb, Kb, mfma=mfma)
                P += Pb
                F += Fb
                self.system.accumulate(b, Kb, K)
                pnewdt = pb.clone() if pnewdt is None else self.torch.minimum(pnewdt, pb)
        pdt = float(pnewdt.item())  # synchronises
        self.computationTimes["elements"] += _time.perf_counter() - tic
        if pdt < 1.0:
            raise CutbackRequest("An element requests for a cutback", pdt)
        return P, K, F

    def assembleStiffnessCSR(self, K):
        # the CSR values were accumulated by computeElements (same launch); nothing left to do
        return K

    def _rows(self, dirichlet):
        key = id(dirichlet)
        if key not in self._dirichletRowsCache:
            idx = self.findDirichletIndices(dirichlet)
            self._dirichletRowsCache[key] = self.torch.as_tensor(idx, dtype=self.torch.int64,
                                                                 device=self.system.device)
        return self._dirichletRowsCache[key]

    def _allRows32(self, dirichlets):
        key = ("all",) + tuple(id(d) for d in dirichlets)
        if key not in self._dirichletRowsCache:
            rows = np.concatenate([self.findDirichletIndices(d) for d in dirichlets]) if dirichlets else \
                np.zeros(0, dtype=np.int64)
            self._dirichletRowsCache[key] = self.torch.as_tensor(rows.astype(np.int32), device=self.system.device)
        return self._dirichletRowsCache[key]

    def applyDirichletK(self, K, dirichlets):
        if dirichlets:
            tic = _time.perf_counter()
            self.system.dirichletRows(self._allRows32(dirichlets), K)
            self.computationTimes["dirichlet K"] += _time.perf_counter() - tic
        return K

    def applyDirichlet(self, timeStep, R, dirichlets):
        tic = _time.perf_counter()
        for dirichlet in dirichlets:
            delta = dirichlet.getDelta(timeStep)
            R[self._rows(dirichlet)] = self.torch.as_tensor(delta.flatten(), dtype=self.torch.float64,
                                                            device=R.device)
        self.computationTimes["dirichlet R"] += _time.perf_counter() - tic
        return R

    def computeSpatialAveragedFluxes(self, F):
        out = {}
        for field, nAcc in self.theDofManager.nAccumulatedNodalFluxesFieldwise.items():
            sl = self.theDofManager.idcsOfFieldsInDofVector[field]
            out[field] = max(1e-10, float(F[sl].abs().sum().item()) / nAcc)
        return out

    def checkConvergence(self, R, ddU, F, iterationCounter, residualHistory):
        tic = _time.perf_counter()
        convergedAtAll = True
        nodesWithLargestResidual = {}
        avg = self.computeSpatialAveragedFluxes(F)
        tolFlux = self.fluxResidualTolerances if iterationCounter < 15 else self.fluxResidualTolerancesAlt
        floor = self.options["fluxResidualFloor"]
        for field, sl in self.theDofManager.idcsOfFieldsInDofVector.items():
            absR = R[sl].abs()
            idx = int(absR.argmax().item())
            fluxResidual = float(absR[idx].item())
            nodesWithLargestResidual[field] = self.theDofManager.getNodeForIndexInDofVector(idx)
            fieldCorrection = float(ddU[sl].abs().max().item()) if ddU is not None else 0.0
            convergedCorrection = fieldCorrection < self.fieldCorrectionTolerances[field]
            convergedFlux = fluxResidual <= max(tolFlux[field] * avg[field], floor)
            prev, nGrew = residualHistory[field]
            if fluxResidual > prev:
                nGrew += 1
            residualHistory[field] = (fluxResidual, nGrew)
            self._currentHistory.append((fluxResidual, fieldCorrection))
            self.journal.message("{:11.2e}{:1}{:11.2e}{:1}".format(fluxResidual, "v" if convergedFlux else " ",
                                                                   fieldCorrection,
                                                                   "v" if convergedCorrection else " "),
                                 self.identification, 2)
            convergedAtAll = convergedAtAll and convergedCorrection and convergedFlux
        self.computationTimes["convergence check"] += _time.perf_counter() - tic
        return convergedAtAll, nodesWithLargestResidual

    def linearSolve(self, A, b):
        tic = _time.perf_counter()
        x = self.theDofManager.constructDofVector()
        rows = self._allRows32(self._activeDirichlets) if self._activeDirichlets else None
        if not bool(b.any().item()):  # zero right-hand side (e.g. the zero-size first increment): x = 0
            self.linearSolveIterations.append((0, 0.0))
            self.computationTimes["linear solve"] += _time.perf_counter() - tic
            return x
        # a plain FP32 tangent (entries rounded to ~6e-8) makes Newton contract linearly, by ~1e-4..1e-3 per
        # iteration at these condition numbers: its corrections are solved only to 1e-4 unless pcgToleranceFP32 says
        # otherwise; with the diagonal corrections the FP32 tangent contracts like the FP64 one
        rtol = self.options["pcgTolerance"]
        if A.dtype == self.torch.float32:
            rtol = self.options["pcgToleranceFP32"]
            if rtol is None:
                rtol = self.options["pcgTolerance"] if A.numel() != self.system.nnz else 1e-4
        amg = self._linsolver == "amg"
        if amg and not (self.options["multigridSetupPerIncrement"] and self._mgSetUpThisIncrement):
            self.system.multigridSetup(A, rows, finePrecision=self.options["multigridFinePrecision"])
            self.multigridSetups += 1
            self._mgSetUpThisIncrement = True
            self.torch.cuda.synchronize(self.system.device)
            toc = _time.perf_counter()
            self.computationTimes["multigrid setup"] += toc - tic
            tic = toc
        precond = "amg" if amg else "jacobi"
        it, rel = self.system.pcg(A, b, x, rows, rtol=rtol, maxit=self.options["pcgMaxIter"],
                                    check_every=4 if amg else 32, precond=precond, setup=False)
        if A.dtype == self.torch.float64:
            for _ in range(self.options["linsolverRefinement"]):
                # Dirichlet rows are unit rows with x = b, so r = 0 there
                r = self.system.residual(A, b, x, self.theDofManager.constructDofVector())
                if not bool(r.any().item()):
                    break
                d = self.theDofManager.constructDofVector()
                it2, _ = self.system.pcg(A, r, d, rows, rtol=rtol, maxit=self.options["pcgMaxIter"],
                                           check_every=4 if amg else 32, precond=precond, setup=False)
                it += it2
                x += d
        self.linearSolveIterations.append((it, rel))
        self.journal.message("linear solve: {:} PCG iterations, relative residual {:.2e}".format(it, rel),
                             self.identification, 2)
        self.computationTimes["linear solve"] += _time.perf_counter() - tic
        if not np.isfinite(rel) or bool(self.torch.isnan(x).any().item()):
            raise DivergingSolution("Obtained NaN in linear solve")
        return x

    def assembleLoads(self, nodeForces, distributedLoads, bodyForces, U_np, PExt, K, timeStep):
        """nonlinearimplicitstatic.py:903-945: nodal forces, then distributed loads, then body forces."""
        for cLoad in nodeForces:
            idx = self.theDofManager.idcsOfFieldsOnNodeSetsInDofVector[cLoad.field][_setName(cLoad, "nodeSet")]
            load = cLoad.getCurrentLoad(timeStep).flatten()
            PExt.index_add_(0, self.torch.as_tensor(idx, device=PExt.device),
                            self.torch.as_tensor(load, dtype=self.torch.float64, device=PExt.device))
        PExt, K = self.computeDistributedLoads(distributedLoads, U_np, PExt, K, timeStep)
        PExt, K = self.computeBodyForces(bodyForces, U_np, PExt, K, timeStep)
        return PExt, K

    def computeDistributedLoads(self, distributedLoads, U_np, PExt, K, timeStep):
        """nonlinearimplicitstatic.py:460-501 on the device: every distributed load adds, face by face of its
        surface, the load on the faces of the surface's elements to PExt. The GPU path carries the load type
        pressure, which acts on the reference configuration and adds nothing to K."""
        if not distributedLoads:
            return PExt, K
        tic = _time.perf_counter()
        for dLoad in distributedLoads:
            if str(dLoad.loadType).lower() != "pressure":
                raise NotImplementedError(f"distributed load type {dLoad.loadType} is not on the GPU path (pressure)")
            load = np.asarray(dLoad.getCurrentLoad(timeStep), dtype=np.float64).ravel()
            for faceID, elSet in dLoad.surface.items():
                for blockIndex, elements in self._elementSetParts(elSet):
                    self.blockAssemblies[blockIndex].surfacePressure(PExt, elements, int(faceID), float(load[0]))
        self.torch.cuda.synchronize(self.system.device)
        self.computationTimes["distributed loads"] += _time.perf_counter() - tic
        return PExt, K

    def computeBodyForces(self, bodyForces, U_np, PExt, K, timeStep):
        """nonlinearimplicitstatic.py:503-557 on the device: every body force adds its integral over its
        element set (reference configuration) to PExt."""
        if not bodyForces:
            return PExt, K
        tic = _time.perf_counter()
        for bForce in bodyForces:
            parts = ([(bForce.blockIndex, bForce.elementSet)] if hasattr(bForce, "blockIndex") else
                     self._elementSetParts(bForce.elementSet))
            for blockIndex, elements in parts:
                self.blockAssemblies[blockIndex].bodyForce(PExt, bForce.getCurrentLoad(timeStep), elements)
        self.torch.cuda.synchronize(self.system.device)
        self.computationTimes["body forces"] += _time.perf_counter() - tic
        return PExt, K

    def extrapolateLastIncrement(self, extrapolation, timeStep, dU, dirichlets, prevTimeStep, model):
        if extrapolation == "linear" and prevTimeStep and prevTimeStep.timeIncrement:
            dU *= timeStep.stepProgressIncrement / prevTimeStep.stepProgressIncrement
            dU = self.applyDirichlet(timeStep, dU, dirichlets)
            isExtrapolatedIncrement = True
        else:
            isExtrapolatedIncrement = False
            dU.zero_()
        return dU, isExtrapolatedIncrement

    def checkDivergingSolution(self, incrementResidualHistory, maxGrowingIter):
        for _, nGrew in incrementResidualHistory.values():
            if nGrew > maxGrowingIter:
                return True
        return False

    def _elementSetParts(self, elSet):
        """[(block index, element indices)] of an element set given as (block, indices), as a set of reference
        element objects (models in the reference's shape) or by name."""
        if isinstance(elSet, tuple) and len(elSet) == 2 and not hasattr(elSet, "name"):
            return [elSet]
        if isinstance(elSet, str):
            if elSet in self.gpuModel.elementSets:
                return [self.gpuModel.elementSets[elSet]]
            return self._adapter.multiBlockElementSets[elSet]
        return self._adapter.elementSetParts(elSet)

    def findDirichletIndices(self, dirichlet):
        fieldIndices = self.theDofManager.idcsOfFieldsOnNodeSetsInDofVector[dirichlet.field][_setName(dirichlet,
                                                                                                       "nSet")]
        return fieldIndices.reshape((len(dirichlet.nSet), -1))[:, dirichlet.components].flatten()

    def applyStepActionsAtStepStart(self, model, stepActions):
        self._activeDirichlets = list(stepActions["dirichlet"].values())
        for typ in stepActions.values():
            for action in typ.values():
                action.applyAtStepStart(model)

    def applyStepActionsAtStepEnd(self, model, stepActions):
        for typ in stepActions.values():
            for action in typ.values():
                action.applyAtStepEnd(model)

    def applyStepActionsAtIncrementStart(self, model, timeStep, stepActions):
        for typ in stepActions.values():
            for action in typ.values():
                action.applyAtIncrementStart(model, timeStep)
