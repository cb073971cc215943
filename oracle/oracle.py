"""CPU oracle for the element-evaluation + assembly path -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of bench.py may
import this module; the product (``edelweissfe_amd``) never does.

It restates, independently of the product code:

* boxGen node/element/node-set ordering           -- ew/generators/boxgen.py:125-371
* DofManager dof numbering and the VIJ I/J pattern -- ew/numerics/dofmanager.py:259-300, 445-471, 522-557
* CSRGenerator pattern + VIJ->CSR slot map         -- ew/numerics/csrgenerator.pyx:47-98 (scipy's canonical
  COO->CSR conversion is the library routine the reference itself calls)
* element kernels, P/F scatter and CSR update      -- ewfe_oracle.c (C, see its header)
* the NIST Newton loop (Dirichlet rows, convergence, zero first increment,
  extrapolation, cutbacks)                         -- ew/solvers/nonlinearimplicitstatic.py:110-458, 559-792,
  947-1006; ew/timesteppers/adaptivetimestepper.py:90-183; ew/steps/adaptivestep.py:56-156;
  ew/stepactions/dirichlet.py:87-112; ew/stepactions/nodeforces.py:113-131; ew/config/phenomena.py:59-93

Parity status of each function is recorded next to it ("pinned by ...").
"""

from __future__ import annotations

import ctypes
import math
import os
import subprocess

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libewfe_oracle.so")

EL_KINDS = {"C3D8": 0, "C3D20": 1, "C3D8TL": 2, "C3D20TL": 3, "C3D20R": 4, "C3D20RTL": 5}
MAT_KINDS = {"linearelastic": 0, "vonmises": 1, "neohookeanwa": 2, "stvenantkirchhoff": 3}
NNODES = {0: 8, 1: 20, 2: 8, 3: 20, 4: 20, 5: 20}
NQP = {0: 8, 1: 27, 2: 8, 3: 27, 4: 8, 5: 8}
# per-QP state sizes of the reference elements: stress 6 + strain 6 + material state
NSTATE = {0: 12, 1: 13, 2: 13, 3: 13}

_lib = None


def build(force: bool = False) -> str:
    """Compile ewfe_oracle.c with gcc (OpenMP) into oracle/libewfe_oracle.so."""
    src = os.path.join(HERE, "ewfe_oracle.c")
    if force or not os.path.exists(LIB_PATH) or os.path.getmtime(LIB_PATH) < os.path.getmtime(src):
        tmp = LIB_PATH + ".tmp"
        subprocess.check_call(
            ["gcc", "-O2", "-fPIC", "-shared", "-fopenmp", "-std=c99", "-D_POSIX_C_SOURCE=199309L", src, "-o", tmp,
             "-lm"]
        )
        os.replace(tmp, LIB_PATH)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        d = ctypes.POINTER(ctypes.c_double)
        i32 = ctypes.POINTER(ctypes.c_int32)
        i64 = ctypes.POINTER(ctypes.c_int64)
        L.oracle_element.restype = ctypes.c_int
        L.oracle_element.argtypes = [ctypes.c_int, ctypes.c_int, d, d, d, d, d, d, ctypes.c_int, d, d, d]
        L.oracle_eval_assemble.restype = ctypes.c_int64
        L.oracle_eval_assemble.argtypes = [
            ctypes.c_int, ctypes.c_int, d, d, d, ctypes.c_int64, i32, d, i64, d, d, d, d, ctypes.c_int, d, d, d, i32,
            ctypes.c_int64, d, ctypes.c_int, d, d,
        ]
        L.oracle_body_force.restype = None
        L.oracle_body_force.argtypes = [ctypes.c_int, ctypes.c_int64, i32, d, i64, ctypes.POINTER(ctypes.c_uint8), d,
                                        d]
        L.oracle_surface_pressure.restype = None
        L.oracle_surface_pressure.argtypes = [ctypes.c_int, ctypes.c_int64, i32, d, i64, ctypes.POINTER(ctypes.c_uint8),
                                              ctypes.c_int, ctypes.c_double, d]
        _lib = L
    return _lib


def _p(a, ct=ctypes.c_double):
    if a is None:
        return None
    return a.ctypes.data_as(ctypes.POINTER(ct))


# ----------------------------------------------------------------------------------------------
# boxGen  (ew/generators/boxgen.py:125-371).  Pinned by: node/element counts, the closed-form
# nnz 9*(3N-2)^3 of the CSR pattern on cubes, and the U.ref decks generated with boxGen.
# ----------------------------------------------------------------------------------------------
def boxgen(nX, nY, nZ, lX=1.0, lY=1.0, lZ=1.0, x0=0.0, y0=0.0, z0=0.0, nnodes=8, name="gen"):
    """Returns (X [nNodes,3] in model insertion order, conn [nEl,nnodes] 0-based, nodeSets, elGridIdx)."""
    if nnodes == 8:
        nx, ny, nz = nX + 1, nY + 1, nZ + 1
    else:
        nx, ny, nz = 2 * nX + 1, 2 * nY + 1, 2 * nZ + 1
    xs, ys, zs = np.linspace(x0, x0 + lX, nx), np.linspace(y0, y0 + lY, ny), np.linspace(z0, z0 + lZ, nz)
    grid_to_model = -np.ones((nx, ny, nz), dtype=np.int64)
    coords = []
    for ix in range(nx):
        for iy in range(ny):
            for iz in range(nz):
                if nnodes == 8 or (ix % 2 + iy % 2 + iz % 2) < 2:
                    grid_to_model[ix, iy, iz] = len(coords)
                    coords.append((xs[ix], ys[iy], zs[iz]))
    X = np.array(coords, dtype=np.float64)
    conn = []
    for ix in range(nX):
        for iy in range(nY):
            for iz in range(nZ):
                if nnodes == 8:
                    g = [(ix, iy, iz), (ix, iy, iz + 1), (ix + 1, iy, iz + 1), (ix + 1, iy, iz),
                         (ix, iy + 1, iz), (ix, iy + 1, iz + 1), (ix + 1, iy + 1, iz + 1), (ix + 1, iy + 1, iz)]
                else:
                    a, b, c = 2 * ix, 2 * iy, 2 * iz
                    g = [(a, b, c), (a, b, c + 2), (a + 2, b, c + 2), (a + 2, b, c),
                         (a, b + 2, c), (a, b + 2, c + 2), (a + 2, b + 2, c + 2), (a + 2, b + 2, c),
                         (a, b, c + 1), (a + 1, b, c + 2), (a + 2, b, c + 1), (a + 1, b, c),
                         (a, b + 2, c + 1), (a + 1, b + 2, c + 2), (a + 2, b + 2, c + 1), (a + 1, b + 2, c),
                         (a, b + 1, c), (a, b + 1, c + 2), (a + 2, b + 1, c + 2), (a + 2, b + 1, c)]
                conn.append([grid_to_model[t] for t in g])
    conn = np.array(conn, dtype=np.int32)

    def filt(sl):
        sub = grid_to_model[sl]
        return sub[sub >= 0].ravel()  # boolean-mask order = C order of the slice

    sets = {}
    for nm, sl in [("top", np.s_[:, -1, :]), ("bottom", np.s_[:, 0, :]), ("right", np.s_[-1, :, :]),
                   ("left", np.s_[0, :, :]), ("front", np.s_[:, :, -1]), ("back", np.s_[:, :, 0]),
                   ("bottomRight", np.s_[-1, 0, :]), ("bottomLeft", np.s_[0, 0, :]),
                   ("bottomFront", np.s_[:, 0, -1]), ("bottomBack", np.s_[:, 0, 0]),
                   ("topRight", np.s_[-1, -1, :]), ("topLeft", np.s_[0, -1, :]), ("topFront", np.s_[:, -1, -1]),
                   ("topBack", np.s_[:, -1, 0]), ("rightBack", np.s_[-1, :, 0]), ("rightFront", np.s_[-1, :, -1]),
                   ("leftBack", np.s_[0, :, 0]), ("leftFront", np.s_[0, :, -1])]:
        sets[f"{name}_{nm}"] = filt(sl)
    for nm, (i, j, k) in [("bottomRightFront", (-1, 0, -1)), ("bottomRightBack", (-1, 0, 0)),
                          ("bottomLeftFront", (0, 0, -1)), ("bottomLeftBack", (0, 0, 0)),
                          ("topRightFront", (-1, -1, -1)), ("topRightBack", (-1, -1, 0)),
                          ("topLeftFront", (0, -1, -1)), ("topLeftBack", (0, -1, 0))]:
        sets[f"{name}_{nm}"] = np.array([grid_to_model[i, j, k]])
    sets["all"] = np.arange(len(X))
    return X, conn, sets


# ----------------------------------------------------------------------------------------------
# DofManager / VIJ / CSRGenerator.  Pinned by scipy (the routine the reference calls) and by
# brute-force dense assembly on tiny meshes in tests/test_oracle.py.
# ----------------------------------------------------------------------------------------------
def dof_numbering(n_nodes, conn):
    """dofmanager.py:259-300: only nodes carrying the displacement field get dofs, in node order;
    dof = base + 3*rank + component. Returns (node_dof_base [-1 if no dofs], nDof, edofs [nEl, 3*npe])."""
    used = np.zeros(n_nodes, dtype=bool)
    used[conn.ravel()] = True
    base = -np.ones(n_nodes, dtype=np.int64)
    base[used] = 3 * np.arange(used.sum())
    nDof = int(3 * used.sum())
    edofs = (base[conn][:, :, None] + np.arange(3)[None, None, :]).reshape(conn.shape[0], -1)
    return base, nDof, edofs.astype(np.int64)


def vij_pattern(edofs):
    """dofmanager.py:543-555: I = tile(idcs).flatten(), J = flatten('F')."""
    nEl, n = edofs.shape
    I = np.tile(edofs[:, None, :], (1, n, 1)).reshape(-1)  # I[off + a*n + b] = idcs[b]
    J = np.repeat(edofs, n, axis=1).reshape(-1)  # J[off + a*n + b] = idcs[a]
    return I, J


def csr_pattern(I, J, nDof):
    """csrgenerator.pyx:68-97: scipy's canonical CSR of the COO pattern + the slot of every VIJ entry."""
    m = sp.csr_matrix((np.zeros(len(I)), (I, J)), shape=(nDof, nDof))
    if not m.has_canonical_format:
        m.sum_duplicates()
    indptr, indices = m.indptr.astype(np.int32), m.indices.astype(np.int32)
    # the reference's binary search finds the unique slot of (row, col) in the sorted row
    row_start = indptr[I]
    key = I.astype(np.int64) * nDof + J
    flat = np.repeat(np.arange(nDof, dtype=np.int64), np.diff(indptr)) * nDof + indices
    x = np.searchsorted(flat, key).astype(np.int32)
    assert np.all(flat[x] == key) and np.all(x >= row_start)
    return indptr, indices, x


# ----------------------------------------------------------------------------------------------
# Element evaluation + assembly (C).
# ----------------------------------------------------------------------------------------------
def element(el_kind, mat_kind, mp, Xe, Ue, dUe, state_n):
    ek, mk = EL_KINDS[el_kind], MAT_KINDS[mat_kind]
    nd = 3 * NNODES[ek]
    nstate = NSTATE[mk]
    mp8 = np.zeros(8)
    mp8[: len(mp)] = mp
    Xe, Ue, dUe = (np.ascontiguousarray(v, dtype=np.float64) for v in (Xe, Ue, dUe))
    state_n = np.ascontiguousarray(state_n, dtype=np.float64)
    state_t = np.zeros_like(state_n)
    Ke, Pe, pdt = np.zeros(nd * nd), np.zeros(nd), np.array([1e36])
    lib().oracle_element(ek, mk, _p(mp8), _p(Xe), _p(Ue), _p(dUe), _p(state_n), _p(state_t), nstate, _p(Ke), _p(Pe),
                         _p(pdt))
    return Ke.reshape(nd, nd), Pe, state_t, pdt[0]


def eval_assemble(el_kind, mat_kind, mp, conn, X, edofs, U_np, dU, state_n, x_slot=None, nnz=0, el_mp=None,
                  n_threads=1, return_vij=False, el_mp8=None):
    """NISTParallel.computeElements + CSRGenerator.updateCSR. Returns dict(P, F, V?, data, state, pNewDT,
    failed, t_elements, t_csr). el_mp8 [nEl, 8]: per-element parameter sets replacing mp (several materials);
    el_mp [nEl]: per-element first parameter."""
    ek, mk = EL_KINDS[el_kind], MAT_KINDS[mat_kind]
    nEl, nn = conn.shape
    nd = 3 * nn
    nstate = NSTATE[mk]
    mp8 = np.zeros(8)
    mp8[: len(mp)] = mp
    nDof = len(U_np)
    P, F = np.zeros(nDof), np.zeros(nDof)
    V = np.zeros(nEl * nd * nd)
    state_t = np.zeros_like(state_n)
    data = np.zeros(nnz) if x_slot is not None else None
    pdt, times = np.array([1e36]), np.zeros(2)
    conn = np.ascontiguousarray(conn, dtype=np.int32)
    edofs = np.ascontiguousarray(edofs, dtype=np.int64)
    el_mp = None if el_mp is None else np.ascontiguousarray(el_mp, dtype=np.float64)
    if el_mp8 is not None:
        full = np.zeros((nEl, 8))
        el_mp8 = np.asarray(el_mp8, dtype=np.float64)
        full[:, : el_mp8.shape[1]] = el_mp8[:, :8]
        el_mp8 = np.ascontiguousarray(full)
    failed = lib().oracle_eval_assemble(
        ek, mk, _p(mp8), _p(el_mp8), _p(el_mp), nEl, _p(conn, ctypes.c_int32), _p(np.ascontiguousarray(X, dtype=np.float64)),
        _p(edofs, ctypes.c_int64), _p(np.ascontiguousarray(U_np)), _p(np.ascontiguousarray(dU)),
        _p(np.ascontiguousarray(state_n)), _p(state_t), nstate, _p(P), _p(F), _p(V),
        _p(x_slot, ctypes.c_int32) if x_slot is not None else None, nnz, _p(data), n_threads, _p(pdt), _p(times))
    out = dict(P=P, F=F, data=data, state=state_t, pNewDT=pdt[0], failed=int(failed), t_elements=times[0],
               t_csr=times[1])
    if return_vij:
        out["V"] = V
    return out


def body_force(el_kind, conn, X, edofs, nDof, b, elements=None):
    """DisplacementElement.computeBodyForce + NIST.computeBodyForces (element.py:348-371,
    nonlinearimplicitstatic.py:503-557): the external load vector of a body force b on `elements` (all when
    None). Pinned by brick closed forms and the WallShear decks (tests/test_oracle.py)."""
    conn = np.ascontiguousarray(conn, dtype=np.int32)
    mask = None
    if elements is not None:
        mask = np.zeros(len(conn), dtype=np.uint8)
        mask[np.asarray(elements, dtype=np.int64)] = 1
    PExt = np.zeros(nDof)
    lib().oracle_body_force(EL_KINDS[el_kind], len(conn), _p(conn, ctypes.c_int32),
                            _p(np.ascontiguousarray(X, dtype=np.float64)),
                            _p(np.ascontiguousarray(edofs, dtype=np.int64), ctypes.c_int64),
                            _p(mask, ctypes.c_uint8) if mask is not None else None,
                            _p(np.ascontiguousarray(b, dtype=np.float64)), _p(PExt))
    return PExt


def surface_pressure(el_kind, conn, X, edofs, nDof, face, pressure, elements=None):
    """Distributed load of type pressure (nonlinearimplicitstatic.py:460-501 calling the Marmot element's
    computeDistributedLoad): the external load vector of the pressure `pressure` on face `face` (Abaqus numbering
    1..6) of `elements` (all when None), reference configuration. Pinned by flat-face closed forms and the
    LinearElasticIsotropic deck (tests/test_oracle.py)."""
    conn = np.ascontiguousarray(conn, dtype=np.int32)
    mask = None
    if elements is not None:
        mask = np.zeros(len(conn), dtype=np.uint8)
        mask[np.asarray(elements, dtype=np.int64)] = 1
    PExt = np.zeros(nDof)
    lib().oracle_surface_pressure(EL_KINDS[el_kind], len(conn), _p(conn, ctypes.c_int32),
                                  _p(np.ascontiguousarray(X, dtype=np.float64)),
                                  _p(np.ascontiguousarray(edofs, dtype=np.int64), ctypes.c_int64),
                                  _p(mask, ctypes.c_uint8) if mask is not None else None, int(face), float(pressure),
                                  _p(PExt))
    return PExt


# ----------------------------------------------------------------------------------------------
# Newton loop (NIST).  Pinned by the reference's U.ref decks (tests/golden, 1e-6 abs).
# ----------------------------------------------------------------------------------------------
class CutbackRequest(Exception):
    def __init__(self, msg, size):
        super().__init__(msg)
        self.cutbackSize = size


class ReachedMaxIterations(Exception):
    pass


class DivergingSolution(Exception):
    pass


class StepFailed(Exception):
    pass


def _two_product(a, b):
    """a * b = p + e exactly (Dekker's product with Veltkamp's split; numpy rounds every operation to nearest)."""
    p = a * b
    c = 134217729.0  # 2^27 + 1
    t = c * a
    ah = t - (t - a)
    al = a - ah
    t = c * b
    bh = t - (t - b)
    bl = b - bh
    return p, ((ah * bh - p) + ah * bl + al * bh) + al * bl


def exact_residual(K, x, b):
    """b - K x, every row correctly rounded: the products split exactly into two doubles, the row sums taken by
    math.fsum (exactly rounded summation)."""
    Kc = K.tocsr()
    p, e = _two_product(Kc.data, x[Kc.indices])
    r = np.empty(Kc.shape[0])
    ip = Kc.indptr
    pl, el = (-p).tolist(), (-e).tolist()
    bl = b.tolist()
    for i in range(Kc.shape[0]):
        r[i] = math.fsum([bl[i]] + pl[ip[i]:ip[i + 1]] + el[ip[i]:ip[i + 1]])
    return r


def refined_solve(K, b, steps=3):
    """K x = b by sparse LU with iterative refinement on correctly rounded residuals: the refined x is accurate to
    working precision while cond(K) * 1e-16 < 1 (the row-eliminated systems of slender decks reach cond ~ 1e10,
    where a plain direct solve, even refined in working precision, is off by up to ~cond * 1e-16 * |x|)."""
    lu = spla.splu(K.tocsc())
    x = lu.solve(b)
    for _ in range(steps):
        x = x + lu.solve(exact_residual(K, x, b))
    return x


def time_steps(stepLength=1.0, startInc=1.0, maxInc=1.0, minInc=1e-4, maxNumInc=1000, increaseFactor=1.1):
    """AdaptiveTimeStepper.generateTimeStep (adaptivetimestepper.py:90-134) as an object with the same
    discard / prevent hooks (:136-183)."""

    class TS:
        def __init__(s):
            s.nGood, s.counter, s.finished = 0, 0, 0.0
            s.increment = min(startInc, maxInc)
            s.allowInc = True

        def gen(s):
            while s.finished < 1.0 - 1e-15:
                rem = 1.0 - s.finished
                if rem < s.increment:
                    s.increment = rem
                inc = s.increment if s.counter > 0 else 0.0
                dT = stepLength * inc
                s.finished += inc
                yield dict(number=s.counter, stepProgressIncrement=inc, stepProgress=s.finished, timeIncrement=dT,
                           stepTime=stepLength * s.finished, totalTime=stepLength * s.finished)
                if s.counter > maxNumInc:
                    raise StepFailed("max increments")
                if s.nGood >= 3 and s.allowInc:
                    s.increment = min(s.increment * increaseFactor, maxInc)
                s.allowInc = True
                s.counter += 1
                s.nGood += 1

        def discard(s, factor):
            if s.counter == 0:
                raise StepFailed("failed zero increment")
            s.finished -= s.increment
            s.counter -= 1
            s.nGood = 0
            if s.increment == minInc:
                raise StepFailed("min increment")
            new = s.increment * factor
            s.increment = maxInc if new > maxInc else (minInc if new < minInc else new)

    return TS()


def _amp_delta(tstep):
    """dirichlet.py:105-112 with the linear amplitude: amp(p) - amp(p - dp)."""
    p = tstep["stepProgress"]
    return p - (p - tstep["stepProgressIncrement"])


def solve_step(model, step, tol_flux=1e-8, tol_corr=1e-8, tol_flux_alt=5e-3, flux_floor=1e-7, log=None,
               state_n=None, U0=None):
    """NIST.solveStep with Dirichlet + nodal-force actions (and, for one block, element loads).

    model: dict(el_kind, mat_kind, mp, X, conn, el_mp?, el_mp8?), or dict(X, blocks=[dict(el_kind, mat_kind, mp,
           conn, el_mp?, el_mp8?), ...]) for several element blocks (element types / material kinds) over X
    step:  dict(maxInc, minInc, maxNumInc, maxIter, criticalIter=5, maxGrowIter=10, cutbackFactor=0.25,
                startInc=1.0, stepLength=1.0, extrapolation='linear',
                dirichlet=[(node_idx, {comp: value})], nodeforces=[(node_idx, vec3)], bodyforces=[vec3],
                pressures=[(element idx, face 1..6, magnitude, f(stepProgress) or None)])
    Returns dict(U, P, state, history=[per increment: list of (|R|inf, |ddU|inf)], iterations=[...]).
    """
    X = model["X"]
    blocks = model.get("blocks")
    single = blocks is None
    if single:
        blocks = [dict(el_kind=model["el_kind"], mat_kind=model["mat_kind"], mp=model["mp"], conn=model["conn"],
                       el_mp=model.get("el_mp"), el_mp8=model.get("el_mp8"))]
    elif step.get("pressures") or step.get("bodyforces"):
        raise NotImplementedError("element loads of models with several blocks")
    # one DofManager and one VIJ pattern over the elements of every block, in block order (femodel.py keeps
    # them in one dict; dofmanager.py:259-300, 522-557)
    used = np.zeros(len(X), dtype=bool)
    for b in blocks:
        used[np.asarray(b["conn"]).ravel()] = True
    base = -np.ones(len(X), dtype=np.int64)
    base[used] = 3 * np.arange(used.sum())
    nDof = int(3 * used.sum())
    bedofs = [(base[np.asarray(b["conn"])][:, :, None] + np.arange(3)[None, None, :]).reshape(len(b["conn"]), -1)
              for b in blocks]
    vij = [vij_pattern(e) for e in bedofs]
    I, J = np.concatenate([v[0] for v in vij]), np.concatenate([v[1] for v in vij])
    indptr, indices, xall = csr_pattern(I, J, nDof)
    bx = np.split(xall, np.cumsum([len(v[0]) for v in vij])[:-1])
    nnz = len(indices)
    conn, edofs = blocks[0]["conn"], bedofs[0]  # element loads: single-block models only
    if state_n is None:
        bstate = [np.zeros(len(b["conn"]) * NQP[EL_KINDS[b["el_kind"]]] * NSTATE[MAT_KINDS[b["mat_kind"]]])
                  for b in blocks]
    else:
        bstate = [state_n] if single else list(state_n)
    U = np.zeros(nDof) if U0 is None else U0.copy()
    dU = np.zeros(nDof)
    P = np.zeros(nDof)
    nAcc = sum(3 * np.asarray(b["conn"]).size for b in blocks)  # dofmanager.py:354-379: elements x nodes x dim

    def evaluate(U_np, dU):
        """computeElements over every block in block order: P, F and the CSR values summed block by block."""
        outs = [eval_assemble(b["el_kind"], b["mat_kind"], b["mp"], b["conn"], X, e, U_np, dU, st, xb, nnz,
                              el_mp=b.get("el_mp"), el_mp8=b.get("el_mp8"))
                for b, e, st, xb in zip(blocks, bedofs, bstate, bx)]
        out = dict(outs[0])
        for o in outs[1:]:
            out["P"] = out["P"] + o["P"]
            out["F"] = out["F"] + o["F"]
            out["data"] = out["data"] + o["data"]
            out["failed"] = out["failed"] or o["failed"]
            out["pNewDT"] = min(out["pNewDT"], o["pNewDT"])
        out["states"] = [o["state"] for o in outs]
        return out

    dirichlets = []
    for nodes, comps in step.get("dirichlet", []):
        if not comps:  # a Dirichlet action redefined without components constrains nothing (dirichlet.py:80-93)
            continue
        cs = sorted(comps)
        idx = (base[nodes][:, None] + np.array(cs)[None, :]).ravel()
        # a value is one number for all nodes or one per node (dirichlet.py:96-99, analyticalField)
        delta = np.column_stack([np.broadcast_to(np.asarray(comps[c], dtype=float), (len(nodes),)) for c in cs]).ravel()
        dirichlets.append((idx, delta))
    nodeforces = []
    for nodes, vec in step.get("nodeforces", []):
        idx = (base[nodes][:, None] + np.arange(3)[None, :]).ravel()
        nodeforces.append((idx, np.tile(np.asarray(vec, float), len(nodes))))

    ts = time_steps(step.get("stepLength", 1.0), step.get("startInc", 1.0), step["maxInc"], step["minInc"],
                    step["maxNumInc"])
    maxIter, criticalIter = step["maxIter"], step.get("criticalIter", 5)
    maxGrow, cutbackFactor = step.get("maxGrowIter", 10), step.get("cutbackFactor", 0.25)
    extrap = step.get("extrapolation", "linear")
    history, iterations = [], []
    prev = None
    for tstep in ts.gen():
        try:
            # extrapolateLastIncrement (nonlinearimplicitstatic.py:947-986)
            if extrap == "linear" and prev and prev["timeIncrement"]:
                dU *= tstep["stepProgressIncrement"] / prev["stepProgressIncrement"]
                for idx, delta in dirichlets:
                    dU[idx] = delta * _amp_delta(tstep)
                isExtrap = True
            else:
                isExtrap = False
                dU[:] = 0.0
            it, ddU, hist = 0, None, []
            resHist = (0.0, 0)
            while True:
                U_np = U + dU
                out = evaluate(U_np, dU)
                if out["failed"]:
                    raise CutbackRequest("element requested cutback", out["pNewDT"])
                Pv, Fv = out["P"], out["F"]
                PExt = np.zeros(nDof)
                for idx, load in nodeforces:
                    PExt[idx] += load * tstep["stepProgress"]
                for els, face, mag, amp in step.get("pressures", []):
                    PExt += surface_pressure(blocks[0]["el_kind"], conn, X, edofs, nDof, face,
                                             mag * (amp(tstep["stepProgress"]) if amp else tstep["stepProgress"]),
                                             els)
                for bvec in step.get("bodyforces", []):
                    PExt += body_force(blocks[0]["el_kind"], conn, X, edofs, nDof,
                                       np.asarray(bvec, float) * tstep["stepProgress"])
                R = Pv + PExt
                if it == 0 and not isExtrap and dirichlets:
                    for idx, delta in dirichlets:
                        R[idx] = delta * _amp_delta(tstep)
                else:
                    for idx, _ in dirichlets:
                        R[idx] = 0.0
                    avgFlux = max(1e-10, np.linalg.norm(Fv, 1) / nAcc)
                    tolF = tol_flux if it < 15 else tol_flux_alt
                    fluxRes = np.max(np.abs(R))
                    corr = np.linalg.norm(ddU, np.inf) if ddU is not None else 0.0
                    hist.append((fluxRes, corr))
                    conv = corr < tol_corr and fluxRes <= max(tolF * avgFlux, flux_floor)
                    nGrew = resHist[1] + (1 if fluxRes > resHist[0] else 0)
                    resHist = (fluxRes, nGrew)
                    if log:
                        log(f"inc {tstep['number']} it {it}: |R|={fluxRes:.3e} |ddU|={corr:.3e}")
                    if conv:
                        break
                    if nGrew > maxGrow:
                        raise DivergingSolution()
                    if it == maxIter:
                        raise ReachedMaxIterations()
                K = sp.csr_matrix((out["data"].copy(), indices.copy(), indptr.copy()), shape=(nDof, nDof))
                if dirichlets:
                    rows = np.concatenate([d[0] for d in dirichlets])
                    for r in rows:
                        K.data[K.indptr[r]: K.indptr[r + 1]] = 0.0
                    diag = K.diagonal()
                    diag[rows] = 1.0
                    K.setdiag(diag)
                    K.eliminate_zeros()
                ddU = refined_solve(K, R)
                if np.isnan(ddU).any():
                    raise DivergingSolution()
                dU += ddU
                it += 1
        except CutbackRequest as e:
            ts.discard(max(e.cutbackSize, cutbackFactor))
            prev = None
            continue
        except (ReachedMaxIterations, DivergingSolution):
            ts.discard(cutbackFactor)
            prev = None
            continue
        prev = tstep
        if it >= criticalIter:
            ts.allowInc = False
        U = U_np
        P = Pv
        bstate = out["states"]
        history.append(hist)
        iterations.append(it)
    return dict(U=U, P=P, state=bstate[0] if single else bstate, history=history, iterations=iterations, nDof=nDof,
                edofs=edofs if single else bedofs, indptr=indptr, indices=indices)
