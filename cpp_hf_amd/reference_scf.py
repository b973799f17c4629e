"""Roothaan SCF front-end (reference: src/cpp_hf/reference_scf.py).

Same config, result type and native call as the reference; the loop
(linear, DIIS or ODA mixing, optional trust radius) runs on the GPU.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Any, Callable

import numpy as np

from ._compat import _native, native_required
from .solver import _kernel_args_for_native


@dataclass(frozen=True)
class SCFConfig:
    max_iter: int = 200
    density_tol: float = 1e-7
    comm_tol: float = 1e-6
    mixing: float = 0.5
    level_shift: float = 0.0
    project_fn: Callable[[Any], Any] | None = None
    block_sizes: tuple[int, ...] | None = None
    acceleration: str = "linear"
    diis_size: int = 6
    diis_start: int = 2
    diis_damping: float = 1.0
    trust_radius: float = 0.0
    return_density: bool = True
    return_fock: bool = True
    # Mixed precision (no counterpart in the reference; "fp64" keeps its
    # behaviour): "mixed" builds the Fock matrix with the FP32 exchange
    # workspace while the density residual is above mixed_switch, then runs
    # FP64 builds to convergence.
    exchange_precision: str = "fp64"
    mixed_switch: float = 1e-6

    def __post_init__(self) -> None:
        if self.max_iter <= 0:
            raise ValueError("max_iter must be positive")


@dataclass(frozen=True)
class SCFResult:
    density_matrix: Any | None
    fock_matrix: Any | None
    energy: Any
    chemical_potential: Any
    iterations: int
    converged: bool
    message: str
    history: dict[str, Any] = field(default_factory=dict)


def _mixed_flag(precision: str) -> bool:
    p = str(precision).lower()
    if p not in ("fp64", "mixed"):
        raise ValueError(f"exchange_precision must be 'fp64' or 'mixed', got {precision!r}")
    return p == "mixed"


def solve_scf(kernel, P0, n_electrons: float, *, config: SCFConfig | None = None) -> SCFResult:
    """Reference SCF solve on the GPU (reference_scf.py:114-165)."""
    native_required()
    if config is None:
        config = SCFConfig()
    P0 = np.ascontiguousarray(P0, dtype=np.complex128)
    density, fock, energy, mu, iterations, converged, hE, hD, hC = _native.solve_scf(
        _kernel_args_for_native(kernel), P0, float(n_electrons), int(config.max_iter),
        float(config.density_tol), float(config.comm_tol), float(config.mixing),
        float(config.level_shift), config.project_fn,
        [int(s) for s in (config.block_sizes or ())], str(config.acceleration),
        int(config.diis_size), int(config.diis_start), float(config.diis_damping),
        float(config.trust_radius), bool(config.return_density), bool(config.return_fock),
        exchange_fp32=_mixed_flag(config.exchange_precision), fp32_switch=float(config.mixed_switch),
    )
    n_iter = int(iterations)
    is_conv = bool(converged)
    if is_conv:
        message = f"converged in {n_iter} iterations"
    else:
        dr = float(hD[n_iter - 1]) if n_iter > 0 else float("nan")
        cr = float(hC[n_iter - 1]) if n_iter > 0 else float("nan")
        message = (f"stopped after {n_iter} iterations "
                   f"(density_res={dr:.3e}, comm_res={cr:.3e})")
    return SCFResult(
        density_matrix=density, fock_matrix=fock, energy=energy, chemical_potential=mu,
        iterations=n_iter, converged=is_conv, message=message,
        history={"E": hE, "density_residual": hD, "commutator_residual": hC},
    )
