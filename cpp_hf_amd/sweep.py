"""Many independent Hartree-Fock solves on one GPU (parameter sweeps).

The reference runs sweeps as one solve after another, or as a pool of CPU
processes
(``examples/tetralayer_quartermetal/deprecated/quartermetal_grid.py:26,396``).
On one MI355X a single small solve (36 x 36 to 120 x 120 meshes, 2 to 12
bands) leaves most of the GPU idle: its kernels are short and each iteration
waits on a host read-back. :func:`solve_many` overlaps the solves instead:
every solve owns its HIP stream and its device workspaces (the library's
memory pool means no solve synchronises the device to free memory), the
native solver loop releases the GIL, and a pool of host threads keeps
``concurrency`` solves in flight. Each solve's arithmetic is the same as
running it alone, with fixed-order reductions, so the results equal those of
:func:`solve` / :func:`solve_scf` called one by one.
"""

from __future__ import annotations

from concurrent.futures import ThreadPoolExecutor
from typing import Sequence

from .reference_scf import SCFConfig, solve_scf
from .solver import SolverConfig, solve


def solve_many(kernels: Sequence, P0s: Sequence, n_electrons: Sequence[float], *,
               config: SolverConfig | SCFConfig | None = None, method: str = "dm",
               concurrency: int = 8) -> list:
    """Solve every (kernel, P0, n_electrons) triple; results in input order.

    method "dm" runs :func:`solve` (``config``: :class:`SolverConfig`),
    "scf" runs :func:`solve_scf` (``config``: :class:`SCFConfig`).
    ``concurrency`` solves run at once (1 reproduces the sequential loop).
    """
    if not (len(kernels) == len(P0s) == len(n_electrons)):
        raise ValueError("kernels, P0s and n_electrons must have the same length")
    if method not in ("dm", "scf"):
        raise ValueError("method must be 'dm' or 'scf'")
    if concurrency < 1:
        raise ValueError("concurrency must be >= 1")

    def one(i: int):
        if method == "dm":
            return solve(kernels[i], P0s[i], n_electrons[i], config=config)
        return solve_scf(kernels[i], P0s[i], n_electrons[i], config=config)

    if concurrency == 1:
        return [one(i) for i in range(len(kernels))]
    with ThreadPoolExecutor(max_workers=concurrency) as pool:
        return list(pool.map(one, range(len(kernels))))
