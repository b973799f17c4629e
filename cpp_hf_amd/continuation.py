"""Coarse-to-fine k-mesh continuation (reference: src/cpp_hf/continuation.py).

Solve on a coarse mesh, carry the converged density to the fine mesh with
the periodic linear resampling of :func:`cpp_hf_amd.resample_kgrid` (a
bit-exact GPU index map, utils.hpp:120-176), hermitise it, and solve again on
the fine mesh.  Either stage runs :func:`solve` (``SolverConfig`` or None)
or :func:`solve_scf` (``SCFConfig``).
"""

from __future__ import annotations

from typing import Any, NamedTuple, Union

import numpy as np

from .problem import HartreeFockKernel
from .reference_scf import SCFConfig, SCFResult, solve_scf
from .solver import SolverConfig, SolveResult, solve
from .utils import resample_kgrid

StageResult = Union[SolveResult, SCFResult]


class ContinuationResult(NamedTuple):
    coarse: StageResult
    fine: StageResult
    P0_fine: np.ndarray | None


def _check_pair(coarse: HartreeFockKernel, fine: HartreeFockKernel) -> None:
    """The two kernels must describe the same model (continuation.py:106-130)."""
    if coarse.h.shape[-2:] != fine.h.shape[-2:]:
        raise ValueError(
            "coarse_kernel and fine_kernel must share orbital dimensions "
            f"(..., nb, nb); got {tuple(coarse.h.shape[-2:])} vs {tuple(fine.h.shape[-2:])}.")
    for flag in ("include_hartree", "include_exchange"):
        if getattr(coarse, flag) != getattr(fine, flag):
            raise ValueError(f"coarse_kernel and fine_kernel must agree on {flag}.")
    for name, k in (("coarse_kernel", coarse), ("fine_kernel", fine)):
        if k.h.shape[0] != k.h.shape[1]:
            raise ValueError(f"{name} must have a square k-mesh, got shape {tuple(k.h.shape[:2])}.")


def _stage(kernel, P0, n_electrons, config: Any, which: str) -> StageResult:
    if isinstance(config, SCFConfig):
        return solve_scf(kernel, P0, float(n_electrons), config=config)
    if config is None or isinstance(config, SolverConfig):
        return solve(kernel, P0, float(n_electrons), config=config)
    raise TypeError(f"{which}_config must be SolverConfig, SCFConfig, or None; "
                    f"got {type(config).__name__}.")


def _final_density(result: StageResult) -> np.ndarray:
    P = result.density_matrix if isinstance(result, SCFResult) else result.density
    if P is None:
        kind = "SCF" if isinstance(result, SCFResult) else "direct-minimization"
        field = "density_matrix" if isinstance(result, SCFResult) else "density"
        raise ValueError(f"coarse {kind} result did not retain {field}.")
    return P


def solve_continuation(coarse_kernel: HartreeFockKernel, fine_kernel: HartreeFockKernel,
                       P0_coarse: np.ndarray, n_electrons_coarse: float,
                       n_electrons_fine: float, *,
                       coarse_config: SolverConfig | SCFConfig | None = None,
                       fine_config: SolverConfig | SCFConfig | None = None,
                       retain_P0_fine: bool = True) -> ContinuationResult:
    """Coarse solve, resample the density to the fine mesh, fine solve."""
    _check_pair(coarse_kernel, fine_kernel)
    P0c = np.asarray(P0_coarse, dtype=coarse_kernel.h.dtype)
    if P0c.shape != coarse_kernel.h.shape:
        raise ValueError(f"P0_coarse shape {tuple(P0c.shape)} does not match coarse_kernel "
                         f"hamiltonian shape {tuple(coarse_kernel.h.shape)}.")
    coarse = _stage(coarse_kernel, P0c, n_electrons_coarse, coarse_config, "coarse")
    P = resample_kgrid(_final_density(coarse), int(fine_kernel.h.shape[0]), method="linear")
    P0f = (0.5 * (P + np.conj(np.swapaxes(P, -1, -2)))).astype(fine_kernel.h.dtype, copy=False)
    fine = _stage(fine_kernel, P0f, n_electrons_fine, fine_config, "fine")
    return ContinuationResult(coarse=coarse, fine=fine, P0_fine=P0f if retain_P0_fine else None)


__all__ = ["ContinuationResult", "solve_continuation"]
