"""Shared helpers of the GPU tests: build a one-block model, evaluate it through libewfe.so, and the same
inputs through the oracle."""
import numpy as np

from edelweissfe_amd.model import ElementBlock, FEModel, Material
from oracle import oracle as O


def one_block_model(el, mat, mp, X, conn, elParams=None, sets=None, materials=None):
    """materials = (list of parameter lists, material index per element): several materials of kind mat."""
    m = FEModel()
    m.addNodes(X)
    b = ElementBlock(el, conn, elParams=elParams)
    if materials is None:
        b.setMaterial(Material(mat, mp))
    else:
        b.setMaterials([Material(mat, q) for q in materials[0]], materials[1])
    m.addElementBlock(b, "all")
    if sets:
        m.nodeSets.update({k: np.asarray(v) for k, v in sets.items()})
    return m, b


class GpuEval:
    def __init__(self, el, mat, mp, X, conn, elParams=None, pipelineChunks=0, materials=None):
        import torch

        from edelweissfe_amd.assembly import ElementAssembly
        from edelweissfe_amd.dofmanager import DofManager

        self.torch = torch
        self.model, self.block = one_block_model(el, mat, mp, X, conn, elParams, materials=materials)
        self.dm = DofManager(self.model)
        self.asm = ElementAssembly(self.block, self.dm, self.model, pipelineChunks=pipelineChunks)

    def vec(self, a):
        return self.torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device="cuda")

    def __call__(self, U_np, dU, dtype="fp64"):
        """dtype: "fp64", "fp32" (FP32 values; FP32 tangent arithmetic for 8-node elements, FP64 FMAs rounded
        when stored for 20-node elements) or "fp32-mfma" (tangent on the matrix cores)."""
        P, F = self.dm.constructDofVector(), self.dm.constructDofVector()
        vals = self.asm.newValues("fp64" if dtype == "fp64" else "fp32")
        pdt = self.asm.evaluate(self.vec(U_np), self.vec(dU), P, F, vals, mfma=dtype == "fp32-mfma")
        self.torch.cuda.synchronize()
        return dict(P=P.cpu().numpy(), F=F.cpu().numpy(), data=vals.double().cpu().numpy(),
                    state=self.asm.stateInElementOrder("trial"), pNewDT=float(pdt.item()))


def oracle_eval(el, mat, mp, X, conn, U_np, dU, state_n=None, el_mp=None, el_mp8=None):
    base, nDof, edofs = O.dof_numbering(len(X), conn)
    I, J = O.vij_pattern(edofs)
    indptr, indices, x = O.csr_pattern(I, J, nDof)
    nq = O.NQP[O.EL_KINDS[el]]
    if state_n is None:
        state_n = np.zeros(len(conn) * nq * O.NSTATE[O.MAT_KINDS[mat]])
    out = O.eval_assemble(el, mat, mp, conn, X, edofs, U_np, dU, state_n, x, len(indices), el_mp=el_mp,
                          el_mp8=el_mp8)
    out["indptr"], out["indices"] = indptr, indices
    return out


def jittered_box(n, el="C3D8", amp=0.05, seed=0, L=(1.0, 1.0, 1.0)):
    from edelweissfe_amd.model import boxGenArrays, jitterNodes

    n = (n, n, n) if isinstance(n, int) else n
    X, conn, sets, grid = boxGenArrays(*n, *L, elType=el)
    mult = 1 if el in ("C3D8", "C3D8TL") else 2
    h = np.array(L) / np.array(n)
    shape = tuple(mult * k + 1 for k in n)
    Xj = jitterNodes(X, grid, shape, h, amp, seed)
    return Xj, conn, sets


def j2_plastic_strain_from_oracle(state, E, nu):
    """Oracle state (stress 6, strain 6, kappa) per QP -> (plastic strain in engineering Voigt, kappa)."""
    st = state.reshape(-1, 13)
    sig, eps = st[:, :6], st[:, 6:12]
    f = E / ((1 + nu) * (1 - 2 * nu))
    C = np.zeros((6, 6))
    C[:3, :3] = f * nu
    C[np.arange(3), np.arange(3)] = f * (1 - nu)
    C[np.arange(3, 6), np.arange(3, 6)] = f * (1 - 2 * nu) / 2
    ep = eps - np.linalg.solve(C, sig.T).T
    return ep, st[:, 12]


def rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300)
