"""shadowfax_amd -- MI355X (gfx950) face-flux engine for the Shadowfax
moving-mesh hydrodynamics code: exact and HLLC Riemann solvers and the
VorFace/VorCell flux pass as HIP kernels behind a C ABI
(include/shadowfax_amd.h)."""
from ._capi import SfxError, lib  # noqa: F401
from .riemann import (ExactRiemannSolver, HLLCRiemannSolver,  # noqa: F401
                      RiemannSolver, RiemannSolverException, RiemannSolverFactory)

__all__ = ["ExactRiemannSolver", "HLLCRiemannSolver", "RiemannSolver",
           "RiemannSolverException", "RiemannSolverFactory", "SfxError", "lib"]
