"""MI355X-native element evaluation + assembly path of EdelweissFE's nonlinear implicit static solver.

Public entry points: ``model`` (FEModel, ElementBlock, Material, boxGen), ``stepping`` (AdaptiveStep,
Dirichlet, NodeForces), ``solvers.nonlinearimplicitstaticgpu.NISTGPU`` and the C ABI in include/ewfe.h
(libewfe.so, loaded by ``_lib``).
"""
