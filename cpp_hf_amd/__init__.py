"""cpp_hf_amd — the Hartree-Fock hot path of cpp_hf on MI355X GPUs.

Drop-in for the reference package's solver surface: the same
``HartreeFockKernel``, ``solve`` / ``solve_direct_minimization``,
``solve_scf``, ``build_fock``, ``hf_energy``, ``free_energy``,
``occupation_entropy``, ``resample_kgrid`` and ``linalg.eigh``, computed by
hand-written HIP kernels for gfx950 (libhfgpu) behind a C ABI and a pybind11
module ``cpp_hf_amd._native`` with the reference module's signatures.
"""

from __future__ import annotations

__version__ = "0.1.0"

from .problem import HartreeFockKernel
from .solver import SolverConfig, SolveResult, solve, solve_direct_minimization
from .reference_scf import SCFConfig, SCFResult, solve_scf
from .fock import build_fock, hf_energy, free_energy, occupation_entropy
from .utils import resample_kgrid
from .sweep import solve_many
from .continuation import ContinuationResult, solve_continuation
from .symmetry import make_project_fn
from .deflation import DeflatedResult, solve_deflated
from .adjoint import build_fock_vjp, energy_gradients, free_energy_derivative
from .superlattice import (ExtendedGridLayout, SuperlatticeHartreeFockKernel,
                           build_extended_layout, superlattice_fock)

__all__ = [
    "HartreeFockKernel",
    "SolverConfig",
    "SolveResult",
    "solve",
    "solve_direct_minimization",
    "SCFConfig",
    "SCFResult",
    "solve_scf",
    "build_fock",
    "hf_energy",
    "free_energy",
    "occupation_entropy",
    "resample_kgrid",
    "solve_many",
    "ContinuationResult",
    "solve_continuation",
    "make_project_fn",
    "DeflatedResult",
    "solve_deflated",
    "build_fock_vjp",
    "energy_gradients",
    "free_energy_derivative",
    "ExtendedGridLayout",
    "build_extended_layout",
    "superlattice_fock",
    "SuperlatticeHartreeFockKernel",
]
