"""Many independent Hartree-Fock solves on one GPU (parameter sweeps).

The reference runs sweeps as one solve after another, or as a pool of CPU
processes
(``examples/tetralayer_quartermetal/deprecated/quartermetal_grid.py:26,396``).
On one MI355X a single small solve (36 x 36 to 120 x 120 meshes, 2 to 12
bands) leaves most of the GPU idle: its kernels are short and launch-bound.
:func:`solve_many` does two things about it:

* solves run concurrently, one HIP stream and host thread each (the native
  loops release the GIL), with ``concurrency`` in flight;
* with ``batch=True``, direct-minimisation solves that share a band count
  and temperature are advanced together (``_native.solve_dm_batch``), on
  whatever meshes: every launch of an iteration covers the whole group (the
  exchange passes one launch per mesh, the chemical potentials and energies
  one launch pair), every solve keeps its own stop test, and each result
  equals the solve run alone bit for bit.  A single small solve is a chain
  of latency-bound kernels; a group turns the chains of its members into
  one chain of wider kernels.

Reductions are fixed-order, so results equal those of :func:`solve` /
:func:`solve_scf` called one by one.
"""

from __future__ import annotations

from concurrent.futures import ThreadPoolExecutor
from typing import Sequence

import numpy as np

from ._compat import _native, native_required
from .reference_scf import SCFConfig, solve_scf
from .solver import SolveResult, SolverConfig, _kernel_args_for_native, solve
from .utils import validate_electron_count

#: solves advanced by one launch sequence at most (the native limit)
MAX_BATCH = 32


def _batchable(config: SolverConfig) -> bool:
    """Configurations the grouped solver takes: CG without projector, blocks,
    deflation or mixed precision."""
    return (str(config.optimizer).lower() == "cg" and config.project_fn is None
            and not config.block_sizes and config.deflation_targets is None
            and str(config.exchange_precision).lower() == "fp64")


def _groups(kernels: Sequence, max_batch: int) -> list[list[int]]:
    """Indices grouped by (bands, temperature), in input order, split into
    chunks of at most max_batch."""
    by: dict[tuple, list[int]] = {}
    for i, k in enumerate(kernels):
        by.setdefault((int(k.h.shape[-1]), float(k.T)), []).append(i)
    out = []
    for idx in by.values():
        for c in range(0, len(idx), max_batch):
            out.append(idx[c:c + max_batch])
    return out


def _solve_group(kernels, P0s, n_electrons, idx: list[int], config: SolverConfig) -> list:
    for i in idx:
        validate_electron_count(kernels[i].w2d, kernels[i].h.shape[-1], n_electrons[i],
                                context="n_electrons")
    res = _native.solve_dm_batch(
        [_kernel_args_for_native(kernels[i]) for i in idx],
        [np.ascontiguousarray(P0s[i], dtype=np.complex128) for i in idx],
        [float(n_electrons[i]) for i in idx],
        int(config.max_iter), float(config.tol_E), float(config.tol_grad),
        float(config.max_step), float(config.bt_shrink), float(config.denom_scale),
        int(config.bt_max), int(config.cg_restart), int(config.plateau_window),
        int(config.mu_maxiter), bool(config.return_Q), bool(config.return_density),
        bool(config.return_fock))
    out = []
    for Q, p, density, fock, mu, energy, n_iter, converged, hE, hG in res:
        out.append(SolveResult(
            Q=Q, p=p, mu=np.asarray(mu), density=density, fock=fock,
            energy=np.asarray(energy), n_iter=np.asarray(int(n_iter), dtype=np.int32),
            converged=np.asarray(bool(converged)), history={"E": hE, "grad_norm": hG}))
    return out


def solve_many(kernels: Sequence, P0s: Sequence, n_electrons: Sequence[float], *,
               config: SolverConfig | SCFConfig | None = None, method: str = "dm",
               concurrency: int = 8, batch: bool = False, max_batch: int = MAX_BATCH) -> list:
    """Solve every (kernel, P0, n_electrons) triple; results in input order.

    method "dm" runs :func:`solve` (``config``: :class:`SolverConfig`),
    "scf" runs :func:`solve_scf` (``config``: :class:`SCFConfig`).
    ``concurrency`` solves (or groups) run at once (1 runs them one after
    another).  With ``batch=True`` (direct minimisation with the CG optimizer
    and no projector, blocks, deflation or mixed precision), solves with one
    band count and temperature are advanced together in groups of up to
    ``max_batch`` (results unchanged, bit for bit).
    """
    if not (len(kernels) == len(P0s) == len(n_electrons)):
        raise ValueError("kernels, P0s and n_electrons must have the same length")
    if method not in ("dm", "scf"):
        raise ValueError("method must be 'dm' or 'scf'")
    if concurrency < 1:
        raise ValueError("concurrency must be >= 1")
    if not 1 <= max_batch <= 32:
        raise ValueError("max_batch must be in 1..32")
    if method == "dm" and config is None:
        config = SolverConfig()

    def one(i: int):
        if method == "dm":
            return solve(kernels[i], P0s[i], n_electrons[i], config=config)
        return solve_scf(kernels[i], P0s[i], n_electrons[i], config=config)

    if method == "dm" and batch and max_batch > 1 and _batchable(config):
        native_required()
        groups = _groups(kernels, max_batch)
        results: list = [None] * len(kernels)

        def run(idx):
            if len(idx) == 1:
                return idx, [one(idx[0])]
            return idx, _solve_group(kernels, P0s, n_electrons, idx, config)

        if concurrency == 1:
            done = [run(g) for g in groups]
        else:
            with ThreadPoolExecutor(max_workers=concurrency) as pool:
                done = list(pool.map(run, groups))
        for idx, rs in done:
            for i, r in zip(idx, rs):
                results[i] = r
        return results

    if concurrency == 1:
        return [one(i) for i in range(len(kernels))]
    with ThreadPoolExecutor(max_workers=concurrency) as pool:
        return list(pool.map(one, range(len(kernels))))
