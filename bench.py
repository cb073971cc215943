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


def newton_schedule(n, tangent, pcg_tol, increments=10, strain=0.01, seed=0, linsolver="amg", setup_per_increment=True,
                    fine_precision="matrix", diagonal="rebuilt"):
    """The J2 cube's load case as configured: from the virgin state to `strain` uniaxial strain in `increments`
    equal increments (after the reference's zero-size first increment, adaptivetimestepper.py:90-134; NIST's default
    linear extrapolation), solved by NISTGPU: uz = 0 on z = 0, uz = strain on z = 1, ux = uy = 0 at (0,0,0), uy = 0
    at (1,0,0). Newton to the reference's criteria |R|inf <= 1e-8 * averaged flux (without the 1e-7 absolute
    floor) and |ddU|inf < 1e-8, linear solves to the relative PCG tolerance `pcg_tol`; with `setup_per_increment`
    the multigrid hierarchy is set up on the first solve of each increment and reused for its later corrections
    (otherwise on every solve); `fine_precision="fp16"` runs the V-cycle's fine-level passes on a half-precision copy of
    the matrix; `diagonal="rebuilt"` (NISTGPU's default) gives the FP32 tangent FP64 corrections of its diagonal node
    blocks that make every row sum to zero, "assembled" keeps the plain FP32 matrix (fp32TangentDiagonal). Returns
    timings per phase, iterations per increment and PCG iterations; plan and multigrid-pattern
    construction (once per mesh) are reported apart."""
    import torch

    from edelweissfe_amd.solvers.nonlinearimplicitstaticgpu import NISTGPU, _Journal
    from edelweissfe_amd.stepping import AdaptiveStep, Dirichlet

    m, b, X, conn = build_workload(n, seed)
    acts = {"dirichlet": {
        "zbottom": Dirichlet("zbottom", m, "gen_back", {2: 0.0}),
        "ztop": Dirichlet("ztop", m, "gen_front", {2: strain}),
        "origin": Dirichlet("origin", m, "gen_bottomLeftBack", {0: 0.0, 1: 0.0}),
        "xaxis": Dirichlet("xaxis", m, "gen_bottomRightBack", {1: 0.0}),
    }}
    inc = 1.0 / increments
    step = AdaptiveStep(1, m, acts, maxInc=inc, startInc=inc, minInc=1e-4, maxIter=25)
    jobInfo = {"fieldCorrectionTolerance": {"displacement": 1e-8},
               "fluxResidualTolerance": {"displacement": 1e-8},
               "fluxResidualToleranceAlternative": {"displacement": 1e-8}}
    solver = NISTGPU(jobInfo, _Journal(verbose=False, stream=sys.stderr), tangentPrecision=tangent,
                     pcgTolerance=pcg_tol, pcgToleranceFP32=pcg_tol, fluxResidualFloor=0.0, linsolver=linsolver,
                     multigridSetupPerIncrement=setup_per_increment, multigridFinePrecision=fine_precision,
                     fp32TangentDiagonal=diagonal)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    solver.solveStep(step, m)
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    its = [len(h) for h in solver.history]
    newton_iters = sum(its)
    solves = [it for it, _ in solver.linearSolveIterations]
    U = m.U.detach().cpu().numpy()
    t_setup = solver.computationTimes["setup"]
    return dict(seconds=t - t_setup, setup_seconds=t_setup, pcg_tolerance=pcg_tol,
                multigrid_setup="per increment" if setup_per_increment else "per solve",
                multigrid_fine_passes=fine_precision, fp32_diagonal_blocks=diagonal, increments=len(its) - 1,
                newton_iterations=newton_iters, iterations_per_increment=its, linear_solves=len(solves),
                pcg_iterations_total=int(sum(solves)),
                ms_per_newton_iteration=1e3 * (t - t_setup) / max(1, newton_iters),
                phases_s={k: round(v, 4) for k, v in solver.computationTimes.items() if v > 0},
                multigrid_levels=solver.system.mgLevels, U=U, history=solver.history)


def host_cores():
    """The host cores this process may use and how that was determined: the CPU affinity mask, capped by a cgroup
    CPU quota when one is set; plus what nproc, OMP_NUM_THREADS and the CPU model say."""
    info = {"os_cpu_count": os.cpu_count(), "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}
    cores = os.cpu_count() or 1
    try:
        cores = len(os.sched_getaffinity(0))
        info["affinity"] = cores
    except (AttributeError, OSError):
        pass
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            quota, period = f.read().split()[:2]
        if quota != "max":
            q = max(1, int(float(quota) / float(period)))
            info["cgroup_quota_cores"] = q
            cores = min(cores, q)
    except (OSError, ValueError):
        pass
    try:
        with open("/proc/cpuinfo") as f:
            info["cpu_model"] = next(ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name"))
    except (OSError, StopIteration):
        info["cpu_model"] = None
    info["cores_used"] = cores
    return cores, info


def _slab(X, conn, ns):
    """The first ns elements of a mesh with its nodes renumbered compactly (the oracle's inputs)."""
    c = conn[:ns]
    used = np.unique(c)
    remap = -np.ones(len(X), dtype=np.int64)
    remap[used] = np.arange(len(used))
    return remap[c].astype(np.int32), X[used], used


def cpu_baseline(el, mat, mp, X, conn, U_np, dU, ns, desc, target_s=12.0, state_from=None):
    """The oracle (C + OpenMP restatement of NISTParallel's element loop, serial P/F scatter and serial VIJ->CSR
    scatter, nonlinearimplicitstaticparallelmk2.pyx:111-198, csrgenerator.pyx:100-115) on the first `ns` elements
    of the mesh (`desc` names them), on every host core this process may use, repeated until about `target_s`
    seconds of evaluation time were measured (median reported); plus one single-thread run. state_from: the
    displacement from which a converged state is first computed (J2)."""
    from oracle import oracle as O

    O.build()
    cs, Xs, used = _slab(X, conn, ns)
    base, nDof, edofs = O.dof_numbering(len(Xs), cs)
    I, J = O.vij_pattern(edofs)
    indptr, indices, x = O.csr_pattern(I, J, nDof)
    del I, J
    Un = U_np.reshape(-1, 3)[used].ravel()
    dUs = dU.reshape(-1, 3)[used].ravel()
    threads, info = host_cores()
    nq = O.NQP[O.EL_KINDS[el]]
    st = np.zeros(ns * nq * O.NSTATE[O.MAT_KINDS[mat]])
    if state_from is not None:
        S = state_from.reshape(-1, 3)[used].ravel()
        st = O.eval_assemble(el, mat, mp, cs, Xs, edofs, S, S, st, x, len(indices), n_threads=threads)["state"]

    def one(nt):
        r = O.eval_assemble(el, mat, mp, cs, Xs, edofs, Un, dUs, st, x, len(indices), n_threads=nt)
        return r["t_elements"] + r["t_csr"], r["t_elements"], r["t_csr"]

    ts = [one(threads)]
    while sum(t[0] for t in ts) < target_s and len(ts) < 1000:
        ts.append(one(threads))
    t, te, tc = (float(np.median([a[k] for a in ts])) for k in range(3))
    t1 = one(1)
    return dict(value=ns / t, unit="elements/s", cores=threads, kind="port",
                sample=f"{desc}: {ns} elements, {len(ts)} timed evaluations ({sum(a[0] for a in ts):.1f} s), median "
                       f"'elements' {te:.4f} s + 'CSR generation' {tc:.4f} s",
                host=info, single_thread_value=ns / t1[0],
                single_thread_split_s={"elements": t1[1], "CSR generation": t1[2]})


def plain_10cube_newton():
    """The plain CPU-runnable configuration in full: a 10^3 C3D8 unit cube, linear elastic E = 210000, nu = 0.3,
    bottom face fixed, top face uz = +1e-3 in one increment, Newton to a 1e-10 flux tolerance: the oracle's Newton
    (element loop + serial scatters in C, scipy LU with refinement; one host thread) against NISTGPU."""
    import torch

    from edelweissfe_amd.model import ElementBlock, FEModel, Material, boxGenArrays
    from edelweissfe_amd.solvers.nonlinearimplicitstaticgpu import DEFAULT_JOB_INFO, NISTGPU
    from edelweissfe_amd.stepping import AdaptiveStep, Dirichlet
    from oracle import oracle as O

    LE = [210000.0, 0.3]
    X, conn, sets, _ = boxGenArrays(10, 10, 10, elType="C3D8")
    jobInfo = dict(DEFAULT_JOB_INFO, fluxResidualTolerance={"displacement": 1e-10})

    def gpu_run():
        m = FEModel()
        m.addNodes(X)
        b = ElementBlock("C3D8", conn)
        b.setMaterial(Material("linearelastic", LE))
        m.addElementBlock(b, "all")
        m.nodeSets.update(sets)
        acts = {"dirichlet": {"bottom": Dirichlet("bottom", m, "gen_back", {0: 0.0, 1: 0.0, 2: 0.0}),
                              "top": Dirichlet("top", m, "gen_front", {2: 1e-3})}}
        solver = NISTGPU(jobInfo, pcgTolerance=1e-13, linsolverRefinement=2)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        solver.solveStep(AdaptiveStep(1, m, acts, maxInc=1.0, startInc=1.0, maxIter=25), m)
        torch.cuda.synchronize()
        return time.perf_counter() - t0, m.U.cpu().numpy(), [len(h) for h in solver.history]

    gpu_run()  # first-call costs (plan code paths, torch kernels) stay out of the timing
    t_gpu, Ug, its = gpu_run()
    step = dict(maxInc=1.0, minInc=1e-4, maxNumInc=1000, maxIter=25,
                dirichlet=[(np.asarray(sets["gen_back"]), {0: 0.0, 1: 0.0, 2: 0.0}),
                           (np.asarray(sets["gen_front"]), {2: 1e-3})])
    t0 = time.perf_counter()
    r = O.solve_step(dict(el_kind="C3D8", mat_kind="linearelastic", mp=LE, X=X, conn=conn), step, tol_flux=1e-10)
    t_cpu = time.perf_counter() - t0
    return {"workload": "C3D8 linear elastic 10^3 unit cube (1,331 nodes), bottom fixed, top uz = 1e-3, one "
                        "increment, Newton to a 1e-10 flux tolerance",
            "gpu_seconds": t_gpu, "cpu_seconds": t_cpu, "cpu_kind": "port (oracle Newton, one host thread)",
            "iterations_per_increment": its,
            "max_rel_displacement_difference_gpu_vs_oracle": float(np.abs(Ug - r["U"]).max() / np.abs(r["U"]).max())}


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


def _timed_evals(asm, U, dU, P, F, K, steps, warmup, mfma=False):
    import torch

    for _ in range(warmup):
        asm.evaluate(U, dU, P, F, K, mfma=mfma)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    for _ in range(steps):
        asm.evaluate(U, dU, P, F, K, mfma=mfma)
    ev1.record(stream)
    torch.cuda.synchronize()
    return ev0.elapsed_time(ev1) / steps


def c3d8tl_measure(n, steps, warmup):
    """Finite-strain total-Lagrangian Neo-Hooke (mu=80, K=200) on the n^3 C3D8 mesh of the adjoint workload (jitter
    +-5 % h, seed 0) at a homogeneous 5 % stretch along z: one evaluation + CSR assembly with FP64 values and with the
    FP32 tangent (FP32 arithmetic for this material)."""
    import torch

    from edelweissfe_amd.assembly import ElementAssembly
    from edelweissfe_amd.dofmanager import DofManager
    from edelweissfe_amd.model import ElementBlock, FEModel, Material, boxGenArrays, jitterNodes

    X, conn, _, grid = boxGenArrays(n, n, n, elType="C3D8TL")
    X = jitterNodes(X, grid, (n + 1,) * 3, np.full(3, 1.0 / n), 0.05, 0)
    m = FEModel()
    m.addNodes(X)
    b = ElementBlock("C3D8TL", conn)
    b.setMaterial(Material("neohookeanwa", [80.0, 200.0]))
    m.addElementBlock(b, "all")
    dm = DofManager(m, device="cuda")
    asm = ElementAssembly(b, dm, m, device="cuda")
    U = torch.as_tensor(uniaxial(X, 0.05), device="cuda")
    P, F = dm.constructDofVector(), dm.constructDofVector()
    ms = {}
    for prec in ("fp64", "fp32"):
        K = asm.newValues(prec)
        ms[prec] = _timed_evals(asm, U, U, P, F, K, steps, warmup)
        del K
    n_el = len(conn)
    B64 = algorithmic_bytes_per_element(n_el, m.nNodes, asm.nnz, 0, 8, 8)
    B32 = algorithmic_bytes_per_element(n_el, m.nNodes, asm.nnz, 0, 8, 4)
    del asm
    return {"elements": n_el, "ms_per_evaluation_fp64": ms["fp64"], "ms_per_evaluation_fp32": ms["fp32"],
            "fp32_speedup": ms["fp64"] / ms["fp32"], "value": n_el / (ms["fp64"] * 1e-3), "unit": "elements/s",
            "roofline_frac": B64 * n_el / (ms["fp64"] * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "roofline_frac_fp32": B32 * n_el / (ms["fp32"] * 1e-3) / 1e9 / HBM_PEAK_GBS}


def c3d20_tl_measure(n, steps, warmup, cpu=True):
    """Finite-strain total-Lagrangian compressible Neo-Hooke (mu=80, K=200) on n^3 C3D20 hexahedra (27 quadrature
    points, 60 dofs each), nodes jittered by U(-3%, 3%) of h (seed 0): one evaluation + CSR assembly at the state
    of a 1.25 homogeneous stretch along z (lambda = 1.5 is reached in 20 increments; this is increment 10), FP64 and
    FP32 values; the oracle on the first 2 x-layers as the CPU baseline."""
    import torch

    from edelweissfe_amd.assembly import ElementAssembly
    from edelweissfe_amd.dofmanager import DofManager
    from edelweissfe_amd.model import ElementBlock, FEModel, Material, boxGenArrays, jitterNodes

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
    from edelweissfe_amd._lib import check, lib

    ms64 = _timed_evals(asm, U, U, P, F, K64, steps, warmup)
    ms32 = _timed_evals(asm, U, U, P, F, K32, steps, warmup)
    ms32m = _timed_evals(asm, U, U, P, F, K32, steps, warmup, mfma=True)
    # the colour-ordered read-modify-write path of the same plan, for comparison
    check(lib().ewfe_plan_set_row_gather(asm._plan, 0))
    ms64c = _timed_evals(asm, U, U, P, F, K64, steps, warmup)
    ms32c = _timed_evals(asm, U, U, P, F, K32, steps, warmup)
    check(lib().ewfe_plan_set_row_gather(asm._plan, 1))
    n_el = len(conn)
    B64 = 4 * 20 + 120.0 * m.nNodes / n_el + 8.0 * asm.nnz / n_el  # no quadrature-point state
    B32 = 4 * 20 + 120.0 * m.nNodes / n_el + 4.0 * asm.nnz / n_el
    # executed FLOPs of the rank-structured tangent and the forces: per point and node pair (a <= b) ~27 FMAs per
    # 3x3 block, 210 pairs, plus ~3 NN^2 operations for the gradients and forces; the dense B^T C B count
    # 2 (6 6 n + n n 6) per point, n = 60, is what a dense formulation would execute
    flop_dense = 27 * 2 * (6 * 6 * 60 + 60 * 60 * 6)
    flop_exec = 27 * (210 * 27 * 2 + 20 * 20 * 3 * 2)
    out = dict(n_el=n_el, dofs=dm.nDof, nnz=asm.nnz, ms64=ms64, ms32=ms32, ms32_mfma=ms32m, ms64_colour=ms64c,
               ms32_colour=ms32c,
               B64=B64, B32=B32, flop_dense=flop_dense,
               flop_exec=flop_exec, cpu=None)
    if cpu:
        out["cpu"] = cpu_baseline("C3D20TL", "neohookeanwa", [80.0, 200.0], X, conn, u.ravel(), u.ravel(),
                                  2 * n * n, f"first 2 x-layers of the jittered {n}^3 C3D20TL mesh at the same state",
                                  target_s=8.0)
    return out


def workload_c3d20_tl(args):
    """The C3D20TL Neo-Hooke 40^3 evaluation + assembly as its own bench line (c3d20_tl_measure)."""
    n = args.n if args.n != 128 else 40
    r = c3d20_tl_measure(n, args.steps, args.warmup, cpu=not args.no_cpu_baseline)
    ms64, n_el = r["ms64"], r["n_el"]
    ach = r["B64"] * n_el / (ms64 * 1e-3) / 1e9
    tflops = r["flop_exec"] * n_el / (ms64 * 1e-3) / 1e12
    # HBM bytes per evaluation from the committed rocprofv3 --pmc summary of this workload (copy-calibrated)
    traffic, traffic_src = None, None
    tfile = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "traffic_c3d20tl_neohooke_40cube.json")
    if n == 40 and os.path.exists(tfile):
        with open(tfile) as f:
            tj = json.load(f)
        traffic = tj["bytes_per_evaluation"]["fp64"]["total"]
        traffic_src = tj["source"] + " (k_hex20 + k_gather20, bytes per evaluation)"
    return {
        "metric": "C3D20-TL Neo-Hooke elem evals+assembly/s", "value": n_el / (ms64 * 1e-3), "unit": "elements/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms64, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"C3D20TL Neo-Hooke Wa {n}^3 jittered cube, eval + CSR assembly", "elements": n_el,
                   "dofs": r["dofs"], "nnz": r["nnz"]},
        "roofline": {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
                     "traffic": traffic, "traffic_source": traffic_src, "algorithmic_bytes_per_element": r["B64"],
                     "scope": "element pass (records) + row-node gather, k_hex20 + k_gather20",
                     "flop_view": {"executed_flop_per_element": r["flop_exec"], "achieved": tflops,
                                   "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": tflops / FP64_PEAK_TFLOPS,
                                   "dense_btcb_flop_per_element": r["flop_dense"]}},
        "cpu_baseline": r["cpu"],
        "fp32_tangent": {"value": n_el / (r["ms32"] * 1e-3), "ms_per_step": r["ms32"],
                         "algorithmic_bytes_per_element": r["B32"]},
        "fp32_tangent_mfma": {"value": n_el / (r["ms32_mfma"] * 1e-3), "ms_per_step": r["ms32_mfma"],
                              "tangent": "FP32 on the matrix cores (v_mfma_f32_16x16x4_f32)"},
        "colour_path": {"ms_per_step_fp64": r["ms64_colour"], "ms_per_step_fp32": r["ms32_colour"]},
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
        "cpu_baseline": None if args.no_cpu_baseline else cpu_baseline(
            "C3D8TL", "neohookeanwa", [80.0, 200.0], X, conn, m.U.cpu().numpy(), m.U.cpu().numpy(), 4 * n * n,
            f"forward evaluation + assembly of the first 4 x-layers of the {n}^3 C3D8TL mesh at the converged state "
            "(mu = 80 for all elements)", target_s=8.0),
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
    ap.add_argument("--newton-n", type=int, default=None, help="mesh of the Newton schedule (default --n)")
    ap.add_argument("--linsolver", default="amg", choices=["amg", "pcg"],
                    help="Newton schedule's linear solver: multigrid- or node-block-Jacobi-preconditioned CG")
    ap.add_argument("--pcg-tol-fp32", type=float, default=2e-6,
                    help="relative PCG tolerance of the FP32-tangent schedule and of its controls")
    ap.add_argument("--no-newton-controls", action="store_true",
                    help="skip the matrix-core FP32 tangent schedule")
    ap.add_argument("--newton-increments", type=int, default=10, help="increments of the 1%% uniaxial schedule")
    ap.add_argument("--no-legs", action="store_true", help="skip the 64^3, C3D20TL 40^3 and 10^3 Newton legs")
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
        copy_calibration(torch.device("cuda:0"))
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
    K32 = asm.newValues("fp32-d64")  # the FP32 values and their FP64 diagonal corrections (NISTGPU's default)
    K32p = asm.newValues("fp32")  # plain FP32 values
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
    ms32p, _ = timed(K32p, args.steps, args.warmup)
    ms32m, wall32m = timed(K32, args.steps, args.warmup, mfma=True)
    print(f"fp64 step {ms64:.3f} ms (wall {wall64:.3f}), fp32-tangent step {ms32:.3f} ms (wall {wall32:.3f}; plain "
          f"FP32 values {ms32p:.3f} ms), "
          f"fp32-mfma tangent step {ms32m:.3f} ms (wall {wall32m:.3f}), plastic QPs {plastic_frac:.3f}, "
          f"pNewDT {pdt:g}", file=sys.stderr, flush=True)

    n_el = len(conn)
    B64 = algorithmic_bytes_per_element(n_el, m.nNodes, asm.nnz, 7, 8, 8)
    B32 = algorithmic_bytes_per_element(n_el, m.nNodes, asm.nnz, 7, 8, 4) + 72.0 * m.nNodes / n_el  # + corrections
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
        cpu = cpu_baseline("C3D8", "vonmises", J2, X, conn, U_n + dU, dU, args.cpu_layers * n * n,
                           f"first {args.cpu_layers} x-layers of the same jittered {n}^3 J2 mesh and state",
                           state_from=U_n)
        print(f"cpu baseline {cpu['value']:.4g} el/s on {cpu['cores']} threads ({cpu['host']})", file=sys.stderr,
              flush=True)

    nnz, ncolors, n_chunks = asm.nnz, asm.nColors, asm.pipelineChunks
    del asm, K64, K32, K32p
    torch.cuda.empty_cache()

    # the other workloads' legs: the 64^3 J2 evaluation, the 40^3 C3D20TL evaluation and the plain 10^3
    # linear elastic cube's full Newton, each with its CPU baseline
    legs = None
    if not args.no_legs:
        legs = {}
        m2, b2, X2, conn2 = build_workload(64)
        dm2 = DofManager(m2, device="cuda")
        asm2 = ElementAssembly(b2, dm2, m2, device="cuda")
        U2n, dU2 = uniaxial(X2, 5e-3), uniaxial(X2, 1e-3)
        P2, F2 = dm2.constructDofVector(), dm2.constructDofVector()
        K2 = asm2.newValues("fp64")
        t2n = torch.as_tensor(U2n, device=dev)
        asm2.evaluate(t2n, t2n, P2, F2, K2)
        b2.acceptLastState()
        ms2 = _timed_evals(asm2, torch.as_tensor(U2n + dU2, device=dev), torch.as_tensor(dU2, device=dev), P2, F2, K2,
                           args.steps, args.warmup)
        B2 = algorithmic_bytes_per_element(len(conn2), m2.nNodes, asm2.nnz, 7, 8, 8)
        legs["c3d8_j2_64cube"] = {
            "elements": len(conn2), "ms_per_evaluation_fp64": ms2, "value": len(conn2) / (ms2 * 1e-3),
            "unit": "elements/s", "roofline_frac": B2 * len(conn2) / (ms2 * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "cpu_baseline": None if args.no_cpu_baseline else cpu_baseline(
                "C3D8", "vonmises", J2, X2, conn2, U2n + dU2, dU2, 4 * 64 * 64,
                "first 4 x-layers of the jittered 64^3 J2 mesh and state", target_s=8.0, state_from=U2n)}
        del asm2, K2
        torch.cuda.empty_cache()
        r4 = c3d20_tl_measure(40, args.steps, args.warmup, cpu=not args.no_cpu_baseline)
        legs["c3d20tl_neohooke_40cube"] = {
            "elements": r4["n_el"], "ms_per_evaluation_fp64": r4["ms64"], "ms_per_evaluation_fp32": r4["ms32"],
            "ms_per_evaluation_fp32_mfma": r4["ms32_mfma"], "ms_per_evaluation_fp64_colour_path": r4["ms64_colour"],
            "value": r4["n_el"] / (r4["ms64"] * 1e-3), "unit": "elements/s",
            "roofline_frac": r4["B64"] * r4["n_el"] / (r4["ms64"] * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "executed_flop_frac": r4["flop_exec"] * r4["n_el"] / (r4["ms64"] * 1e-3) / 1e12 / FP64_PEAK_TFLOPS,
            "cpu_baseline": r4["cpu"]}
        torch.cuda.empty_cache()
        legs["c3d8tl_neohooke_48cube"] = c3d8tl_measure(48, args.steps, args.warmup)
        torch.cuda.empty_cache()
        legs["plain_c3d8_le_10cube_newton"] = plain_10cube_newton()
        print(f"legs: {json.dumps({k: {kk: vv for kk, vv in v.items() if kk != 'cpu_baseline'} for k, v in legs.items()})}",
              file=sys.stderr, flush=True)
        torch.cuda.empty_cache()

    newton = None
    if not args.no_newton:
        nn_ = args.newton_n or n
        # untimed warm-up of every variant on a small mesh: first-call costs stay out of the timed schedules
        for tan in ("fp64", "fp32", "fp32-mfma"):
            newton_schedule(8, tan, 1e-4, increments=2, linsolver=args.linsolver)
        torch.cuda.empty_cache()
        # plain FP64 and the FP32 tangent (with its diagonal corrections, NISTGPU's default) at several PCG tolerances
        # (the fastest run of each precision that meets the Newton tolerance is compared), the plain FP32 matrix
        # (assembled diagonal blocks) at its own best tolerance, all with the multigrid set up once per increment.
        # Controls: a setup on every solve; the matrix-core FP32 tangent; both precisions with a half-precision copy
        # of the matrix in the V-cycle's fine-level passes (a mixed-precision preconditioner, so not the plain FP64
        # path)
        t32 = args.pcg_tol_fp32
        runs = [("fp64", 1e-10, True, "matrix", "rebuilt"), ("fp64", 1e-6, True, "matrix", "rebuilt"),
                ("fp64", 2e-6, True, "matrix", "rebuilt"), ("fp64", 1e-4, True, "matrix", "rebuilt"),
                ("fp32", 1e-6, True, "matrix", "rebuilt"), ("fp32", 2e-6, True, "matrix", "rebuilt"),
                ("fp32", 1e-4, True, "matrix", "assembled")]
        if t32 not in (1e-6, 2e-6):
            runs.append(("fp32", t32, True, "matrix", "rebuilt"))
        if not args.no_newton_controls:
            runs += [("fp64", t32, False, "matrix", "rebuilt"), ("fp32", t32, False, "matrix", "rebuilt"),
                     ("fp32-mfma", t32, True, "matrix", "rebuilt"), ("fp64", t32, True, "fp16", "rebuilt"),
                     ("fp32", t32, True, "fp16", "rebuilt")]
        res = {}
        for tan, tol, per_inc, fine, dg in runs:
            r = newton_schedule(nn_, tan, tol, increments=args.newton_increments, linsolver=args.linsolver,
                                setup_per_increment=per_inc, fine_precision=fine, diagonal=dg)
            print(f"newton {tan} (pcg tol {tol:g}, multigrid setup {r['multigrid_setup']}, fine-level passes {fine}, "
                  f"fp32 diagonal {dg}): {r['seconds']:.3f}s, iterations per increment "
                  f"{r['iterations_per_increment']}, {r['linear_solves']} solves, {r['pcg_iterations_total']} PCG "
                  f"iterations, phases {r['phases_s']}", file=sys.stderr, flush=True)
            key = (f"{tan}_pcg{tol:g}" + ("" if per_inc else "_setup_per_solve") +
                   ("_fp16_fine" if fine == "fp16" else "") +
                   ("_assembled_diagonal" if tan != "fp64" and dg == "assembled" else ""))
            if tan == "fp64":
                r.pop("fp32_diagonal_blocks")
            res[key] = r
            torch.cuda.empty_cache()
        ref64 = res["fp64_pcg1e-10"]
        U64 = ref64["U"]
        for r in res.values():
            r["max_rel_displacement_difference_vs_fp64_pcg1e-10"] = float(np.abs(r["U"] - U64).max() /
                                                                          np.abs(U64).max())
        for r in res.values():
            r.pop("U")
            r.pop("history")
        plain = lambda k: not k.endswith(("_fp16_fine", "_setup_per_solve", "_assembled_diagonal"))
        best64 = min((k for k in res if k.startswith("fp64_") and plain(k)), key=lambda k: res[k]["seconds"])
        fp32_key = min((k for k in res if k.startswith("fp32_") and plain(k)), key=lambda k: res[k]["seconds"])
        newton = {"mesh": f"{nn_}^3", "load": f"virgin state to 1% uniaxial strain in {args.newton_increments} "
                                              "increments (after the zero-size first increment), linear extrapolation",
                  "runs": res, "fastest_fp64": best64, "fastest_fp32": fp32_key,
                  "speedup_fp32_vs_fastest_fp64": res[best64]["seconds"] / res[fp32_key]["seconds"],
                  "speedup_fp32_assembled_diagonal_vs_fastest_fp64": (res[best64]["seconds"] /
                                                                      res["fp32_pcg0.0001_assembled_diagonal"]["seconds"]),
                  "speedup_fp32_vs_fp64_pcg1e-10": ref64["seconds"] / res[fp32_key]["seconds"],
                  "newton_step_speedup_fp32_vs_fastest_fp64": (res[best64]["ms_per_newton_iteration"] /
                                                               res[fp32_key]["ms_per_newton_iteration"]),
                  "newton_step_ms": {k: v["ms_per_newton_iteration"] for k, v in res.items()}}
        if not args.no_newton_controls:
            newton["speedup_fp32mfma_vs_fastest_fp64"] = (res[best64]["seconds"] /
                                                          res[f"fp32-mfma_pcg{args.pcg_tol_fp32:g}"]["seconds"])
            # the same comparison when both precisions use the half-precision fine-level V-cycle passes
            newton["speedup_fp32_vs_fp64_both_fp16_fine"] = (res[f"fp64_pcg{args.pcg_tol_fp32:g}_fp16_fine"]["seconds"] /
                                                             res[f"fp32_pcg{args.pcg_tol_fp32:g}_fp16_fine"]["seconds"])

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
                         "algorithmic_bytes_per_element": B32, "ms_per_step_plain_fp32_values": ms32p,
                         "tangent": "FP32 arithmetic on the vector units, FP32 CSR values with FP64 corrections of "
                                    "the diagonal node blocks (every row sums to zero, as the FP64 tangent's)"},
        "fp32_tangent_mfma": {"value": n_el / (ms32m * 1e-3), "ms_per_step": ms32m, "speedup_vs_fp64": ms64 / ms32m,
                              "roofline_frac": B32 * n_el / (ms32m * 1e-3) / 1e9 / HBM_PEAK_GBS,
                              "tangent": "FP32 on the matrix cores (v_mfma_f32_16x16x4_f32), values and diagonal "
                                         "corrections as fp32_tangent"},
        "setup_s": t_setup,
        "legs": legs,
        "newton": newton,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
