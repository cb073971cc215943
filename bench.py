"""Benchmark of the element evaluation + CSR assembly path on one MI355X.

Workload (default): C3D8 small-strain J2 plasticity on a 128x128x128 structured hex mesh of the unit cube
(2,097,152 elements, 8 quadrature points each, state = plastic strain 6 + kappa per point, structure of
arrays), nodes jittered by U(-5%, +5%) of h (seed 0; faces kept planar), E=210000, nu=0.3, fy0=250,
H=1000. The state is taken half way through the uniaxial loading to 1 % strain: every quadrature point is
plastic and runs the radial return. One "step" = one Newton iteration's element evaluation + assembly
(the reference's timing keys "elements" + "CSR generation"): P, F, trial state and all CSR values of the
6,440,067-dof system, inputs resident in HBM. The metric is elements per second.

Also reported: the FP32-tangent variant of the same step, the roofline fraction of the step (algorithmic
bytes per element: connectivity + nodal X/U/dU + state read/write + P/F + CSR values) and the CPU
baseline: the oracle (C, OpenMP) on a bounded slab of the same workload, on this host's cores.
"""

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6  # MI355X dense FP64 (vector and matrix)
J2 = [210000.0, 0.3, 250.0, 1000.0, 0.0, 0.0]


def uniaxial(X, strain):
    u = np.zeros_like(X)
    u[:, 2] = strain * X[:, 2]
    u[:, 0] = -0.5 * strain * X[:, 0]
    u[:, 1] = -0.5 * strain * X[:, 1]
    return u.ravel()


def build_workload(n, seed=0):
    from edelweissfe_amd.model import ElementBlock, FEModel, Material, boxGenArrays, jitterNodes

    X, conn, sets, grid = boxGenArrays(n, n, n, elType="C3D8")
    X = jitterNodes(X, grid, (n + 1,) * 3, np.full(3, 1.0 / n), 0.05, seed)
    m = FEModel()
    m.addNodes(X)
    b = ElementBlock("C3D8", conn)
    b.setMaterial(Material("vonmises", J2))
    m.addElementBlock(b, "all")
    m.nodeSets.update(sets)
    return m, b, X, conn


def algorithmic_bytes_per_element(n_el, n_nodes, nnz, n_state, n_qp, val_bytes, npe=8):
    return (4 * npe + 72.0 * n_nodes / n_el + 16.0 * n_state * n_qp + 48.0 * n_nodes / n_el
            + val_bytes * nnz / n_el)


def newton_increment(n, tangent, strain=2e-3, pcg_tol=1e-10, seed=0, linsolver="amg", pcg_tol_fp32=1e-6):
    """One implicit static increment of the same J2 cube from the virgin state to `strain` (into the plastic
    range), solved by NISTGPU: uz = 0 on z = 0, uz = strain on z = 1, ux = uy = 0 at (0,0,0), uy = 0 at
    (1,0,0). Newton to the reference's criteria |R|inf <= 1e-8 * averaged flux (without the 1e-7 absolute
    floor) and |ddU|inf < 1e-8; returns timings per phase, iterations and PCG iterations."""
    import torch

    from edelweissfe_amd.solvers.nonlinearimplicitstaticgpu import NISTGPU
    from edelweissfe_amd.stepping import AdaptiveStep, Dirichlet

    m, b, X, conn = build_workload(n, seed)
    acts = {"dirichlet": {
        "zbottom": Dirichlet("zbottom", m, "gen_back", {2: 0.0}),
        "ztop": Dirichlet("ztop", m, "gen_front", {2: strain}),
        "origin": Dirichlet("origin", m, "gen_bottomLeftBack", {0: 0.0, 1: 0.0}),
        "xaxis": Dirichlet("xaxis", m, "gen_bottomRightBack", {1: 0.0}),
    }}
    step = AdaptiveStep(1, m, acts, maxInc=1.0, startInc=1.0, minInc=1e-3, maxIter=20)
    jobInfo = {"fieldCorrectionTolerance": {"displacement": 1e-8},
               "fluxResidualTolerance": {"displacement": 1e-8},
               "fluxResidualToleranceAlternative": {"displacement": 1e-8}}
    from edelweissfe_amd.solvers.nonlinearimplicitstaticgpu import _Journal

    solver = NISTGPU(jobInfo, _Journal(verbose=True, stream=sys.stderr), tangentPrecision=tangent,
                     pcgTolerance=pcg_tol, pcgToleranceFP32=pcg_tol_fp32, fluxResidualFloor=0.0,
                     extrapolation="off", linsolver=linsolver)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    solver.solveStep(step, m)
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    its = [len(h) for h in solver.history]
    newton_iters = sum(its)  # linear solves = convergence checks of the increments that converged
    pcg_its = [it for it, _ in solver.linearSolveIterations]
    U = m.U.detach().cpu().numpy()
    levels = solver.assembly.mgLevels
    t_setup = solver.computationTimes["setup"]
    return dict(seconds=t - t_setup, setup_seconds=t_setup, linsolver=linsolver, multigrid_levels=levels,
                pcg_tolerance=pcg_tol_fp32 if tangent == "fp32" else pcg_tol,
                newton_iterations=newton_iters,
                iterations_per_increment=its, pcg_iterations=pcg_its,
                ms_per_newton_iteration=1e3 * (t - t_setup) / max(1, newton_iters),
                phases_s={k: v for k, v in solver.computationTimes.items() if v > 0},
                residual_history=[h for h in solver.history], U=U)


def cpu_baseline(X, conn, U_n, dU, layers, n_per_layer, target_s=12.0):
    """The oracle (C + OpenMP restatement of NISTParallel's element loop, serial P/F scatter and serial
    VIJ->CSR scatter) on the first `layers` x-layers of the same mesh and inputs, repeated until about
    `target_s` seconds of evaluation time have been measured (median reported); plus one single-thread run."""
    from oracle import oracle as O

    O.build()
    ns = layers * n_per_layer
    c = conn[:ns]
    used = np.unique(c)
    remap = -np.ones(len(X), dtype=np.int64)
    remap[used] = np.arange(len(used))
    cs = remap[c].astype(np.int32)
    Xs = X[used]
    base, nDof, edofs = O.dof_numbering(len(Xs), cs)
    I, J = O.vij_pattern(edofs)
    indptr, indices, x = O.csr_pattern(I, J, nDof)
    del I, J
    Un = U_n.reshape(-1, 3)[used].ravel()
    dUs = dU.reshape(-1, 3)[used].ravel()
    threads = min(os.cpu_count() or 1, 16)
    st0 = np.zeros(ns * 8 * 13)
    r0 = O.eval_assemble("C3D8", "vonmises", J2, cs, Xs, edofs, Un, Un, st0, x, len(indices), n_threads=threads)
    def one(nt):
        r = O.eval_assemble("C3D8", "vonmises", J2, cs, Xs, edofs, Un + dUs, dUs, r0["state"], x, len(indices),
                            n_threads=nt)
        return r["t_elements"] + r["t_csr"], r["t_elements"], r["t_csr"]

    ts = [one(threads)]
    while sum(t[0] for t in ts) < target_s and len(ts) < 1000:
        ts.append(one(threads))
    t, te, tc = (float(np.median([a[k] for a in ts])) for k in range(3))
    t1 = one(1)
    return dict(value=ns / t, unit="elements/s", cores=threads, kind="port",
                sample=f"first {layers} x-layers ({ns} elements) of the same jittered 128^3 J2 mesh and state, "
                       f"{len(ts)} timed evaluations ({sum(a[0] for a in ts):.1f} s), median "
                       f"'elements' {te:.4f} s + 'CSR generation' {tc:.4f} s",
                single_thread_value=ns / t1[0],
                single_thread_split_s={"elements": t1[1], "CSR generation": t1[2]})


def copy_calibration(dev, nbytes=1 << 30):
    """A fill and a plain device copy of `nbytes` each (torch's elementwise kernels). Under rocprofv3 --pmc their
    FETCH_SIZE / WRITE_SIZE per dispatch, read against the known byte counts, give the scale that turns the same
    counters of the assembly kernels into HBM bytes."""
    import torch

    src = torch.empty(nbytes // 8, dtype=torch.float64, device=dev)
    src.fill_(1.0)
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst
    torch.cuda.empty_cache()


def _timed_evals(asm, U, dU, P, F, K, steps, warmup):
    import torch

    for _ in range(warmup):
        asm.evaluate(U, dU, P, F, K)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    for _ in range(steps):
        asm.evaluate(U, dU, P, F, K)
    ev1.record(stream)
    torch.cuda.synchronize()
    return ev0.elapsed_time(ev1) / steps


def workload_c3d20_tl(args):
    """Finite-strain total-Lagrangian compressible Neo-Hooke (mu=80, K=200) on 40^3 C3D20 hexahedra (64,000
    elements, 27 quadrature points, 60 dofs each), nodes jittered by U(-3%, 3%) of h (seed 0): one evaluation +
    CSR assembly at the state of a 1.25 homogeneous stretch along z (lambda = 1.5 is reached in 20
    increments; this is increment 10), FP64 and FP32 values."""
    import torch

    from edelweissfe_amd.assembly import ElementAssembly
    from edelweissfe_amd.dofmanager import DofManager
    from edelweissfe_amd.model import ElementBlock, FEModel, Material, boxGenArrays, jitterNodes

    n = args.n if args.n != 128 else 40
    X, conn, sets, grid = boxGenArrays(n, n, n, elType="C3D20TL")
    X = jitterNodes(X, grid, (2 * n + 1,) * 3, np.full(3, 1.0 / n), 0.03, 0)
    m = FEModel()
    m.addNodes(X)
    b = ElementBlock("C3D20TL", conn)
    b.setMaterial(Material("neohookeanwa", [80.0, 200.0]))
    m.addElementBlock(b, "all")
    dm = DofManager(m)
    asm = ElementAssembly(b, dm, m)
    u = np.zeros_like(X)
    u[:, 2] = 0.25 * X[:, 2]
    u[:, 0] = -0.1 * X[:, 0]
    u[:, 1] = -0.1 * X[:, 1]
    U = torch.as_tensor(u.ravel(), device="cuda")
    P, F = dm.constructDofVector(), dm.constructDofVector()
    K64, K32 = asm.newValues("fp64"), asm.newValues("fp32")
    ms64 = _timed_evals(asm, U, U, P, F, K64, args.steps, args.warmup)
    ms32 = _timed_evals(asm, U, U, P, F, K32, args.steps, args.warmup)
    n_el = len(conn)
    B64 = 4 * 20 + 120.0 * m.nNodes / n_el + 8.0 * asm.nnz / n_el  # no quadrature-point state
    B32 = 4 * 20 + 120.0 * m.nNodes / n_el + 4.0 * asm.nnz / n_el
    ach = B64 * n_el / (ms64 * 1e-3) / 1e9
    # FLOP bound: the dense B^T C B count per element, 2 (6 6 n + n n 6) per point with n = 60 dofs, 27 points
    # (the kernel itself executes the rank-structured form of the same tangent, about a quarter of that)
    flop_el = 27 * 2 * (6 * 6 * 60 + 60 * 60 * 6)
    tflops = flop_el * n_el / (ms64 * 1e-3) / 1e12
    return {
        "metric": "C3D20-TL Neo-Hooke elem evals+assembly/s", "value": n_el / (ms64 * 1e-3), "unit": "elements/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms64, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"C3D20TL Neo-Hooke Wa {n}^3 jittered cube, eval + CSR assembly", "elements": n_el,
                   "dofs": dm.nDof, "nnz": asm.nnz},
        "roofline": {"bound": "mfma", "achieved": tflops, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": tflops / FP64_PEAK_TFLOPS, "traffic": None, "algorithmic_flop_per_element": flop_el,
                     "scope": "fill + colour launches of k_hex20; FP64 vector and matrix peaks are the same",
                     "hbm_view": {"achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
                                  "algorithmic_bytes_per_element": B64}},
        "cpu_baseline": None,
        "fp32_tangent": {"value": n_el / (ms32 * 1e-3), "ms_per_step": ms32,
                         "algorithmic_bytes_per_element": B32},
    }


def workload_adjoint(args):
    """48^3 C3D8TL Neo-Hooke block (mu_e = 80 exp(0.2 z_e), z_e ~ N(0,1) seed 1, K=200), bottom (z=0) fixed,
    top pulled by nodal forces; converged forward solve, then the adjoint gradient dJ/dmu_e of
    J = 1/2 |u_top - u_target|^2 (u_target from seed 2): tangent at the converged state, PCG for lambda, and
    the adjoint kernel. Reports the forward solve, the adjoint linear solve and the adjoint kernel times."""
    import torch

    from edelweissfe_amd.adjoint import parameter_gradient
    from edelweissfe_amd.model import ElementBlock, FEModel, Material, boxGenArrays, jitterNodes
    from edelweissfe_amd.solvers.nonlinearimplicitstaticgpu import NISTGPU, _Journal
    from edelweissfe_amd.stepping import AdaptiveStep, Dirichlet, NodeForces

    n = args.n if args.n != 128 else 48
    X, conn, sets, grid = boxGenArrays(n, n, n, elType="C3D8TL")
    X = jitterNodes(X, grid, (n + 1,) * 3, np.full(3, 1.0 / n), 0.05, 0)
    mu = 80.0 * np.exp(0.2 * np.random.default_rng(1).standard_normal(len(conn)))
    m = FEModel()
    m.addNodes(X)
    b = ElementBlock("C3D8TL", conn, elParams=mu)
    b.setMaterial(Material("neohookeanwa", [80.0, 200.0]))
    m.addElementBlock(b, "all")
    m.nodeSets.update(sets)
    top = sets["gen_front"]
    fz = 8.0 / len(top)  # ~10 % stretch of the unit cube
    acts = {"dirichlet": {"fix": Dirichlet("fix", m, "gen_back", {0: 0.0, 1: 0.0, 2: 0.0})},
            "nodeforces": {"pull": NodeForces("pull", m, "gen_front", [0.0, 0.0, fz])}}
    step = AdaptiveStep(1, m, acts, maxInc=0.5, startInc=0.5, maxIter=20)
    solver = NISTGPU(None, _Journal(verbose=True, stream=sys.stderr), pcgTolerance=1e-10, fluxResidualFloor=0.0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    solver.solveStep(step, m)
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    base = solver.theDofManager.nodeDofBase
    top_dofs = torch.as_tensor((base[top][:, None] + np.arange(3)).ravel(), device="cuda")
    target = m.U[top_dofs] * (1 + 0.1 * torch.as_tensor(np.random.default_rng(2).standard_normal(len(top_dofs)),
                                                        device="cuda"))
    dJdU = torch.zeros_like(m.U)
    dJdU[top_dofs] = m.U[top_dofs] - target
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    grad, lam, (it, rel) = parameter_gradient(solver, m, dJdU, rtol=1e-10)
    torch.cuda.synchronize()
    t_adj = time.perf_counter() - t0
    # the adjoint kernel alone
    from edelweissfe_amd._lib import check, lib, ptr

    asm = solver.assembly
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    stream = torch.cuda.current_stream()
    ev0.record(stream)
    for _ in range(args.steps):
        check(lib().ewfe_adjoint_param_grad(asm._plan, asm.matKind, ptr(asm.matParams), ptr(asm.elParams),
                                            ptr(asm.X), ptr(m.U), ptr(lam), ptr(grad), asm._stream()))
    ev1.record(stream)
    torch.cuda.synchronize()
    ms_kernel = ev0.elapsed_time(ev1) / args.steps
    n_el = len(conn)
    B = 4 * 8 + 72.0 * m.nNodes / n_el + 8.0  # X, U, lambda gathered + one gradient value
    ach = B * n_el / (ms_kernel * 1e-3) / 1e9
    return {
        "metric": "adjoint dJ/dmu_e gradient evaluations/s (C3D8TL Neo-Hooke)", "value": 1e3 / ms_kernel,
        "unit": "gradients/s", "n_gpus": 1, "steps": args.steps, "warmup": 0, "ms_per_step": ms_kernel,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"C3D8TL Neo-Hooke {n}^3, per-element mu field ({n_el} parameters), adjoint kernel",
                   "forward_solve_s": t_fwd, "adjoint_solve_and_kernel_s": t_adj, "adjoint_pcg_iterations": it,
                   "adjoint_pcg_relres": rel, "newton_iterations": [len(h) for h in solver.history],
                   "grad_norm": float(grad.norm().item())},
        "roofline": {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
                     "traffic": None, "algorithmic_bytes_per_element": B, "scope": "k_adjoint8"},
        "cpu_baseline": None,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=128, help="elements per cube edge")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-layers", type=int, default=4)
    ap.add_argument("--no-newton", action="store_true")
    ap.add_argument("--newton-n", type=int, default=None, help="mesh of the Newton increment (default --n)")
    ap.add_argument("--linsolver", default="amg", choices=["amg", "pcg"],
                    help="Newton increment's linear solver: multigrid- or node-block-Jacobi-preconditioned CG")
    ap.add_argument("--pcg-tol-fp32", type=float, default=1e-4,
                    help="relative PCG tolerance of the FP32-tangent increment (NISTGPU's default)")
    ap.add_argument("--no-newton-controls", action="store_true",
                    help="skip the control increments (FP32 tangent at 1e-6, FP64 at the FP32 tolerance)")
    ap.add_argument("--workload", default="c3d8-j2", choices=["c3d8-j2", "c3d20-nh-tl", "c3d8-nh-adjoint"])
    ap.add_argument("--chunks", type=int, default=0,
                    help="pipeline chunks of the 8-node evaluation (0 = the library's automatic choice)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only; multi-GPU runs are not part of it "
              f"(--gpus {args.gpus} requested)", file=sys.stderr)
        sys.exit(2)

    import torch

    from edelweissfe_amd.assembly import ElementAssembly
    from edelweissfe_amd.dofmanager import DofManager

    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    if args.workload != "c3d8-j2":
        fn = workload_c3d20_tl if args.workload == "c3d20-nh-tl" else workload_adjoint
        print(json.dumps(fn(args)), flush=True)
        return
    dev = torch.device("cuda:0")
    n = args.n
    copy_calibration(dev)
    t0 = time.time()
    m, b, X, conn = build_workload(n)
    dm = DofManager(m, device="cuda")
    asm = ElementAssembly(b, dm, m, device="cuda", pipelineChunks=args.chunks)
    t_setup = time.time() - t0
    print(f"setup {n}^3: {len(conn)} elements, {dm.nDof} dofs, nnz {asm.nnz}, colours {asm.nColors}, "
          f"pipeline chunks {asm.pipelineChunks} (structured mesh: {asm.structuredMesh}), {t_setup:.1f}s",
          file=sys.stderr, flush=True)

    U_n = uniaxial(X, 5e-3)
    dU = uniaxial(X, 1e-3)
    U_n_t = torch.as_tensor(U_n, device=dev)
    dU_t = torch.as_tensor(dU, device=dev)
    U_np_t = U_n_t + dU_t
    P, F = dm.constructDofVector(), dm.constructDofVector()
    K64 = asm.newValues("fp64")
    K32 = asm.newValues("fp32")
    # converged state at U_n (first loading from the virgin state), then accepted
    asm.evaluate(U_n_t, U_n_t, P, F, K64)
    b.acceptLastState()
    torch.cuda.synchronize()
    kappa = b.stateVars.view(7, 8, -1)[6]
    plastic_frac = float((kappa > 0).double().mean().item())

    stream = torch.cuda.current_stream(dev)

    def timed(K, steps, warmup, mfma=False):
        for _ in range(warmup):
            asm.evaluate(U_np_t, dU_t, P, F, K, mfma=mfma)
        torch.cuda.synchronize()
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t_wall = time.perf_counter()
        ev0.record(stream)
        for _ in range(steps):
            asm.evaluate(U_np_t, dU_t, P, F, K, mfma=mfma)
        ev1.record(stream)
        torch.cuda.synchronize()
        t_wall = time.perf_counter() - t_wall
        return ev0.elapsed_time(ev1) / steps, t_wall * 1e3 / steps

    ms64, wall64 = timed(K64, args.steps, args.warmup)
    pdt = float(asm.pnewdt.item())
    ms32, wall32 = timed(K32, args.steps, args.warmup)
    ms32m, wall32m = timed(K32, args.steps, args.warmup, mfma=True)
    print(f"fp64 step {ms64:.3f} ms (wall {wall64:.3f}), fp32-tangent step {ms32:.3f} ms (wall {wall32:.3f}), "
          f"fp32-mfma tangent step {ms32m:.3f} ms (wall {wall32m:.3f}), plastic QPs {plastic_frac:.3f}, "
          f"pNewDT {pdt:g}", file=sys.stderr, flush=True)

    n_el = len(conn)
    B64 = algorithmic_bytes_per_element(n_el, m.nNodes, asm.nnz, 7, 8, 8)
    B32 = algorithmic_bytes_per_element(n_el, m.nNodes, asm.nnz, 7, 8, 4)
    nnz = asm.nnz
    ach64 = B64 * n_el / (ms64 * 1e-3) / 1e9
    ach32 = B32 * n_el / (ms32 * 1e-3) / 1e9

    # HBM bytes per evaluation from the committed rocprofv3 --pmc summary of this workload (FETCH_SIZE and
    # WRITE_SIZE, copy-calibrated); counters cannot be read live here
    traffic64, traffic_src = None, None
    tfile = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "traffic_c3d8_j2_128cube.json")
    if n == 128 and os.path.exists(tfile):
        with open(tfile) as f:
            tj = json.load(f)
        traffic64 = tj["bytes_per_evaluation"]["fp64"]["total"]
        traffic_src = tj["source"] + " (k_hex8_elem + k_gather8s, bytes per evaluation)"

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(X, conn, U_n, dU, args.cpu_layers, n * n)
        print(f"cpu baseline {cpu['value']:.4g} el/s on {cpu['cores']} threads", file=sys.stderr, flush=True)

    nnz, ncolors, n_chunks = asm.nnz, asm.nColors, asm.pipelineChunks
    newton = None
    if not args.no_newton:
        del asm, K64, K32
        torch.cuda.empty_cache()
        nn_ = args.newton_n or n
        res = {}
        # untimed warm-up of both variants on a small mesh: first-call costs (library loading, torch kernels)
        # stay out of the timed increments
        for tan in ("fp64", "fp32"):
            newton_increment(8, tan, linsolver=args.linsolver)
        torch.cuda.empty_cache()
        # headline: plain FP64 (PCG 1e-10, the reference's exact-solve semantics) vs FP32 tangent at its
        # precision-matched PCG tolerance; controls separate the precision effect from the tolerance effect
        runs = [("fp64", 1e-10, "fp64"), ("fp32", args.pcg_tol_fp32, "fp32")]
        if not args.no_newton_controls:
            runs += [("fp32", 1e-6, "fp32_tol1e-6"), ("fp64", args.pcg_tol_fp32, "fp64_inexact")]
        controls = {}
        for tan, tol, key in runs:
            kw = {"pcg_tol_fp32": tol} if tan == "fp32" else {"pcg_tol": tol}
            r = newton_increment(nn_, tan, linsolver=args.linsolver, **kw)
            print(f"newton {key} (pcg tol {r['pcg_tolerance']:g}): {r['seconds']:.3f}s, {r['newton_iterations']} "
                  f"iterations, pcg {r['pcg_iterations']}, phases {r['phases_s']}", file=sys.stderr, flush=True)
            if key in ("fp64", "fp32"):
                res[tan] = r
            else:
                Uc = r.pop("U")
                controls[key] = {k: r[k] for k in ("pcg_tolerance", "seconds", "newton_iterations", "pcg_iterations")}
                controls[key]["Uc"] = Uc
            torch.cuda.empty_cache()
        U64, U32 = res["fp64"].pop("U"), res["fp32"].pop("U")
        for c in controls.values():
            c["max_rel_displacement_difference_vs_fp64"] = float(np.abs(c.pop("Uc") - U64).max() / np.abs(U64).max())
            c["speedup_vs_fp64"] = res["fp64"]["seconds"] / c["seconds"]
        newton = {"mesh": f"{nn_}^3", "load": "one increment from the virgin state to 0.2% uniaxial strain",
                  "fp64": res["fp64"], "fp32_tangent": res["fp32"],
                  "increment_speedup_fp32_vs_fp64": res["fp64"]["seconds"] / res["fp32"]["seconds"],
                  "controls": controls,
                  "max_rel_displacement_difference_fp32_vs_fp64":
                      float(np.abs(U32 - U64).max() / np.abs(U64).max())}

    line = {
        "metric": "C3D8 J2 elem evals+assembly/s",
        "value": n_el / (ms64 * 1e-3),
        "unit": "elements/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms64,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"C3D8 J2 small strain {n}^3 jittered unit cube, one Newton iteration's element "
                               f"evaluation + CSR assembly, FP64", "elements": n_el, "dofs": dm.nDof,
                   "nnz": nnz, "qp_per_element": 8, "state_per_qp": 7, "plastic_qp_fraction": plastic_frac,
                   "pipeline_chunks": n_chunks},
        "roofline": {"bound": "hbm", "achieved": ach64, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": ach64 / HBM_PEAK_GBS, "traffic": traffic64,
                     "traffic_source": traffic_src,
                     "algorithmic_bytes_per_element": B64,
                     "scope": "one evaluation call: k_hex8_elem (element records) + k_gather8s (row-node CSR "
                              "assembly on the boxGen grid; k_gather8 on other meshes)"},
        "cpu_baseline": cpu,
        "fp32_tangent": {"value": n_el / (ms32 * 1e-3), "ms_per_step": ms32, "speedup_vs_fp64": ms64 / ms32,
                         "roofline_achieved_GBs": ach32, "roofline_frac": ach32 / HBM_PEAK_GBS,
                         "algorithmic_bytes_per_element": B32,
                         "tangent": "FP64 FMAs, rounded to FP32 when the element records are stored"},
        "fp32_tangent_mfma": {"value": n_el / (ms32m * 1e-3), "ms_per_step": ms32m, "speedup_vs_fp64": ms64 / ms32m,
                              "roofline_frac": B32 * n_el / (ms32m * 1e-3) / 1e9 / HBM_PEAK_GBS,
                              "tangent": "FP32 on the matrix cores (v_mfma_f32_16x16x4_f32)"},
        "setup_s": t_setup,
        "newton_increment": newton,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
