"""ewfe_pcg_noalloc (the PCG with its workspace in the library's fixed device scratch) against ewfe_pcg with a
caller's workspace: the same system gives bitwise the same solution and iteration count, on two calls in a row,
the second on a smaller system (the scratch keeps the first call's contents)."""
import ctypes

import numpy as np
import pytest

from gpu_util import GpuEval, jittered_box

pytestmark = pytest.mark.gpu

LE = [210000.0, 0.3]


def _system(n, seed):
    import torch

    X, conn, sets = jittered_box(n, "C3D8", seed=seed)
    g = GpuEval("C3D8", "linearelastic", LE, X, conn)
    vals = g.asm.newValues("fp64")
    P, F = g.dm.constructDofVector(), g.dm.constructDofVector()
    z = g.vec(np.zeros(g.dm.nDof))
    g.asm.evaluate(z, z, P, F, vals)
    rows = (g.dm.nodeDofBase[sets["gen_back"]][:, None] + np.arange(3)).ravel().astype(np.int32)
    rows_t = torch.as_tensor(rows, device="cuda")
    g.asm.dirichletRows(rows_t, vals)
    b = np.random.default_rng(seed).standard_normal(g.dm.nDof)
    b[rows] = 1e-3
    return g, vals, rows_t, g.vec(b)


def _solve(g, vals, rows_t, b, noalloc):
    from edelweissfe_amd._lib import EWFE_F64, check, lib, ptr

    L = lib()
    x = g.dm.constructDofVector()
    it, rr = ctypes.c_int(), ctypes.c_double()
    stream = g.asm._stream()
    if noalloc:
        check(L.ewfe_pcg_noalloc(g.asm._plan, ptr(vals), EWFE_F64, ptr(b), ptr(x), ptr(rows_t), rows_t.numel(),
                                 1e-12, 5000, 16, ctypes.byref(it), ctypes.byref(rr), stream))
    else:
        nbytes = L.ewfe_pcg_workspace_size(g.asm._plan)
        ws = g.torch.empty((nbytes + 7) // 8, dtype=g.torch.float64, device="cuda")
        check(L.ewfe_pcg(g.asm._plan, ptr(vals), EWFE_F64, ptr(b), ptr(x), ptr(rows_t), rows_t.numel(), 1e-12, 5000,
                         16, ptr(ws), ctypes.byref(it), ctypes.byref(rr), stream))
    g.torch.cuda.synchronize()
    return x.cpu().numpy(), it.value, rr.value


def test_pcg_noalloc_matches_pcg():
    for n, seed in ((7, 1), (4, 2)):  # the second call is on a smaller system
        print(f"case pcg_noalloc C3D8 linear elastic n={n}", flush=True)
        g, vals, rows_t, b = _system(n, seed)
        xa, ita, rra = _solve(g, vals, rows_t, b, noalloc=True)
        xb, itb, rrb = _solve(g, vals, rows_t, b, noalloc=False)
        print(f"noalloc: {ita} iterations, relres {rra:.2e}; workspace: {itb} iterations, relres {rrb:.2e}",
              flush=True)
        assert rra < 1e-12 and ita == itb and rra == rrb
        assert np.array_equal(xa, xb)
