"""Deflated trust-region Newton: several distinct HF solutions of one model
(reference: src/cpp_hf/deflation.py:59-187).

Every solve runs on the GPU through :func:`cpp_hf_amd.solve` with
``optimizer="newton"``.  After the first (unbiased) solution, each further
solve adds the Gaussian bias ``sigma * sum_i exp(-d_i^2 / (2 L^2))`` around
the densities found so far (``SolverConfig.deflation_*``; the bias itself is
evaluated on the device), starts from ``P0`` plus a Hermitian perturbation,
and is re-solved without the bias so that every returned density is a
stationary point of the plain free energy.  The steps, their order, the
defaults and the random stream are the reference's, so a seed selects the
same perturbations.
"""

from __future__ import annotations

from dataclasses import replace
from typing import NamedTuple

import numpy as np

from .solver import SolveResult, SolverConfig, solve


class DeflatedResult(NamedTuple):
    solutions: list          # distinct SolveResults, ascending energy
    energies: np.ndarray     # their energies, ascending
    best: SolveResult        # the lowest-energy solution
    n_found: int             # len(solutions)


def _weighted_sq(M: np.ndarray, w2d: np.ndarray) -> float:
    """sum_k w_k ||M_k||_F^2 (the metric of the device-side bias)."""
    return float(np.sum(np.asarray(w2d) * np.sum(np.abs(M) ** 2, axis=(-2, -1))))


def _weighted_distance(A: np.ndarray, B: np.ndarray, w2d: np.ndarray) -> float:
    return float(np.sqrt(max(_weighted_sq(np.asarray(A) - np.asarray(B), w2d), 0.0)))


def _random_hermitian(shape, scale: float, rng) -> np.ndarray:
    X = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex128)
    return scale * 0.5 * (X + np.conj(np.swapaxes(X, -1, -2)))


def solve_deflated(kernel, P0, n_electrons: float, *, base_config: SolverConfig | None = None,
                   n_solutions: int = 4, sigma: float | None = None,
                   length: float | None = None, distinct_tol: float | None = None,
                   max_retries: int = 4, sigma_growth: float = 3.0, perturb: float = 1e-2,
                   repolish: bool = True, seed: int = 0) -> DeflatedResult:
    """Up to ``n_solutions`` distinct HF solutions by deflated Newton solves.

    ``length`` defaults to half the weighted norm of the first density,
    ``sigma`` to ``max(|E_0|, 1)`` and ``distinct_tol`` to ``length / 4``;
    each new solution gets ``max_retries`` biased attempts with ``sigma``
    multiplied by ``sigma_growth`` after each miss, the seed perturbation
    growing as ``perturb * (attempt + 1)``.  The search stops early when no
    attempt finds a new basin.
    """
    w2d = np.asarray(kernel.w2d)
    rng = np.random.default_rng(seed)
    cfg0 = base_config if base_config is not None else SolverConfig()
    plain = replace(cfg0, optimizer="newton", return_density=True, deflation_targets=None,
                    deflation_sigma=0.0)
    ne = float(n_electrons)
    P0 = np.ascontiguousarray(P0, dtype=np.complex128)

    found = [solve(kernel, P0, ne, config=plain)]
    if length is None:
        norm0 = float(np.sqrt(max(_weighted_sq(np.asarray(found[0].density), w2d), 0.0)))
        length = 0.5 * norm0 if norm0 > 0.0 else 1.0
    if sigma is None:
        sigma = max(abs(float(found[0].energy)), 1.0)
    if distinct_tol is None:
        distinct_tol = 0.25 * length

    while len(found) < n_solutions:
        height = float(sigma)
        new = None
        for attempt in range(max_retries):
            biased = replace(plain, deflation_targets=np.stack([np.asarray(s.density)
                                                                for s in found]),
                             deflation_sigma=height, deflation_length=float(length))
            start = np.ascontiguousarray(
                P0 + _random_hermitian(P0.shape, perturb * (attempt + 1), rng),
                dtype=np.complex128)
            r = solve(kernel, start, ne, config=biased)
            if repolish:
                r = solve(kernel, np.ascontiguousarray(r.density, dtype=np.complex128), ne,
                          config=plain)
            if bool(r.converged) and all(
                    _weighted_distance(r.density, s.density, w2d) > distinct_tol for s in found):
                new = r
                break
            height *= sigma_growth
        if new is None:
            break
        found.append(new)

    found.sort(key=lambda s: float(s.energy))
    energies = np.array([float(s.energy) for s in found])
    return DeflatedResult(solutions=found, energies=energies, best=found[0], n_found=len(found))


__all__ = ["DeflatedResult", "solve_deflated"]
