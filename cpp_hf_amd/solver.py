"""Direct-minimisation solver front-end (reference: src/cpp_hf/solver.py).

Packs the kernel into the reference's ``kernel_args`` dict, validates the
inputs and unpacks the tuple returned by ``_native.solve_dm``, which runs the
preconditioned Riemannian CG loop (or, with ``optimizer="newton"``, the
trust-region Newton loop) on the GPU.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Any, NamedTuple

import numpy as np

from ._compat import _native, native_required
from .utils import validate_electron_count


@dataclass(frozen=True)
class SolverConfig:
    """Direct-minimisation settings; fields and defaults of solver.py:19-64.

    ``optimizer="newton"`` selects the trust-region Newton solver
    (solver_rtr.hpp), which alone honours ``deflation_*``; ``block_sizes``
    applies to the CG loop (per-block eigh when the problem is
    block-diagonal), as in the reference.
    """

    max_iter: int = 200
    tol_E: float = 1e-7
    tol_grad: float = 0.0
    plateau_window: int = 5
    denom_scale: float = 1e-3
    max_step: float = 0.6
    cg_restart: int = 10
    bt_shrink: float = 0.5
    bt_max: int = 8
    mu_maxiter: int = 25
    optimizer: str = "cg"
    tr_delta0: float = 0.5
    tr_cg_max: int = 20
    deflation_targets: Any = None
    deflation_sigma: float = 0.0
    deflation_length: float = 1.0
    block_sizes: tuple[int, ...] | None = None
    project_fn: Any = None
    return_Q: bool = True
    return_density: bool = True
    return_fock: bool = True
    # Mixed precision for the CG loop (no counterpart in the reference;
    # "fp64" keeps its behaviour): "mixed" uses the FP32 exchange workspace
    # until the gradient norm reaches mixed_switch, then FP64 builds.
    exchange_precision: str = "fp64"
    mixed_switch: float = 1e-6


class SolveResult(NamedTuple):
    Q: np.ndarray | None
    p: np.ndarray
    mu: np.ndarray
    density: np.ndarray | None
    fock: np.ndarray | None
    energy: np.ndarray
    n_iter: np.ndarray
    converged: np.ndarray
    history: dict


def _kernel_args_for_native(kernel) -> dict:
    """The ``kernel_args`` dict ABI of solver.py:79-134 (with the superlattice
    entries of a SuperlatticeHartreeFockKernel)."""
    refP = kernel._refP if kernel._refP is not None else kernel._empty_refP
    args = dict(
        h=np.ascontiguousarray(kernel.h, dtype=np.complex128),
        VR=np.ascontiguousarray(kernel._VR_shifted, dtype=np.complex128),
        refP=np.ascontiguousarray(refP, dtype=np.complex128),
        has_refP=bool(kernel.has_reference_density),
        w2d=np.ascontiguousarray(kernel.w2d, dtype=np.float64),
        HH=np.ascontiguousarray(kernel.HH, dtype=np.float64),
        contact_g=np.ascontiguousarray(kernel.contact_g, dtype=np.float64),
        contact_Oi=np.ascontiguousarray(kernel.contact_Oi, dtype=np.complex128),
        contact_Oj=np.ascontiguousarray(kernel.contact_Oj, dtype=np.complex128),
        weight_sum=float(kernel.weight_sum),
        T=float(kernel.T),
        include_hartree=bool(kernel.include_hartree),
        include_exchange=bool(kernel.include_exchange),
        exchange_hcp=bool(kernel.exchange_hermitian_channel_packing),
    )
    if getattr(kernel, "superlattice_fock_active", False) \
            or getattr(kernel, "superlattice_hartree_active", False):
        from .superlattice import superlattice_kernel_args

        args.update(superlattice_kernel_args(kernel))
    return args


_OPTIMIZER_CODES = {"cg": 0, "newton": 1}


def _optimizer_code(name: str) -> int:
    try:
        return _OPTIMIZER_CODES[str(name).lower()]
    except KeyError:
        raise ValueError(
            f"Unknown optimizer {name!r}; expected one of {sorted(_OPTIMIZER_CODES)}."
        ) from None


def _mixed_flag(precision: str) -> bool:
    p = str(precision).lower()
    if p not in ("fp64", "mixed"):
        raise ValueError(f"exchange_precision must be 'fp64' or 'mixed', got {precision!r}")
    return p == "mixed"


def solve_direct_minimization(kernel, P0, n_electrons: float, *,
                              config: SolverConfig | None = None) -> SolveResult:
    """Direct-minimisation Hartree-Fock solve on the GPU (solver.py:150-220)."""
    native_required()
    if config is None:
        config = SolverConfig()
    validate_electron_count(kernel.w2d, kernel.h.shape[-1], n_electrons, context="n_electrons")
    P0 = np.ascontiguousarray(P0, dtype=np.complex128)
    block_sizes = [int(s) for s in (config.block_sizes or ())]
    code = _optimizer_code(config.optimizer)
    nk1, nk2, nb = kernel.h.shape[0], kernel.h.shape[1], kernel.h.shape[-1]
    if config.deflation_targets is None:
        defl = np.empty((0, nk1, nk2, nb, nb), dtype=np.complex128)
    else:
        defl = np.ascontiguousarray(config.deflation_targets, dtype=np.complex128)
        if defl.ndim == 4:
            defl = defl[None, ...]
        if defl.ndim != 5 or defl.shape[1:] != (nk1, nk2, nb, nb):
            raise ValueError(
                "deflation_targets must have shape (n_solutions, nk1, nk2, nb, nb) "
                f"matching the kernel; got {tuple(defl.shape)}."
            )
    if defl.shape[0] > 0 and float(config.deflation_sigma) > 0.0 \
            and str(config.optimizer).lower() != "newton":
        raise ValueError(
            "deflation is only supported with optimizer='newton' "
            f"(got optimizer={config.optimizer!r})."
        )
    Q, p, density, fock, mu, energy, n_iter, converged, hE, hG = _native.solve_dm(
        _kernel_args_for_native(kernel), P0, float(n_electrons),
        int(config.max_iter), float(config.tol_E), float(config.tol_grad),
        float(config.max_step), float(config.bt_shrink), float(config.denom_scale),
        int(config.bt_max), int(config.cg_restart), int(config.plateau_window),
        int(config.mu_maxiter), int(code), float(config.tr_delta0), int(config.tr_cg_max),
        defl, float(config.deflation_sigma), float(config.deflation_length), block_sizes,
        config.project_fn, bool(config.return_Q), bool(config.return_density),
        bool(config.return_fock), exchange_fp32=_mixed_flag(config.exchange_precision),
        fp32_switch=float(config.mixed_switch),
    )
    return SolveResult(
        Q=Q, p=p, mu=np.asarray(mu), density=density, fock=fock, energy=np.asarray(energy),
        n_iter=np.asarray(int(n_iter), dtype=np.int32), converged=np.asarray(bool(converged)),
        history={"E": hE, "grad_norm": hG},
    )


solve = solve_direct_minimization


def _cayley_retract(d: np.ndarray, tau: float) -> np.ndarray:
    """Dense Cayley map ``(I + tau d / 2)^-1 (I - tau d / 2)`` (solver.py:226-236),
    kept as the algebraic reference of the spectral Cayley kernels."""
    n = d.shape[-1]
    eye = np.eye(n, dtype=d.dtype)
    A = 0.5 * float(tau) * d
    return np.linalg.solve(eye + A, eye - A)
