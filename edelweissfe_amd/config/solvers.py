"""Solver registry (edelweissfe/config/solvers.py:48-76): ``*solver, solver=NISTGPU`` selects the GPU path
the same way ``solver=NISTParallel`` selects the OpenMP path in the reference."""

solverLibrary = {
    "NISTGPU": ("edelweissfe_amd.solvers.nonlinearimplicitstaticgpu", "NISTGPU"),
}


def getSolverByName(name: str):
    import importlib

    try:
        mod, cls = solverLibrary[name]
    except KeyError:
        raise AttributeError(f"solver {name} is not available (available: {', '.join(solverLibrary)})")
    return getattr(importlib.import_module(mod), cls)
