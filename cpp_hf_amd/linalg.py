"""Hermitian eigendecomposition with optional block-diagonal dispatch.

Same interface as the reference's ``cpp_hf.linalg`` (src/cpp_hf/linalg.py):
``eigh`` runs the batched GPU Jacobi eigensolver; block specs restrict it to
the diagonal blocks after an off-block check, running each block size as
one batched GPU launch.
"""

from __future__ import annotations

from typing import Any, Sequence

import numpy as np

from ._compat import batched_eigh

BlockSpec = tuple[str, tuple[Any, ...]]


def _as_sequence(x: Any) -> Sequence[Any]:
    if isinstance(x, np.ndarray):
        return x.tolist()
    if isinstance(x, (list, tuple)):
        return x
    raise TypeError(f"Expected a sequence, got {type(x)}.")


def _normalize_one_spec(spec: Any) -> BlockSpec:
    if spec is None:
        raise TypeError("block_specs entries must be non-None.")
    if (isinstance(spec, tuple) and len(spec) == 2 and isinstance(spec[0], str)
            and spec[0].strip().lower() in ("sizes", "indices")):
        kind = spec[0].strip().lower()
        if kind == "sizes":
            return ("sizes", tuple(int(x) for x in _as_sequence(spec[1])))
        return ("indices", tuple(tuple(int(i) for i in _as_sequence(b))
                                 for b in _as_sequence(spec[1])))
    if isinstance(spec, dict):
        if "block_sizes" in spec:
            return ("sizes", tuple(int(x) for x in _as_sequence(spec["block_sizes"])))
        if "block_indices" in spec:
            return ("indices", tuple(tuple(int(i) for i in _as_sequence(b))
                                     for b in _as_sequence(spec["block_indices"])))
        raise TypeError("Spec dict must contain 'block_sizes' or 'block_indices'.")
    seq = _as_sequence(spec)
    if len(seq) == 0:
        raise TypeError("Spec sequences must be non-empty.")
    if all(isinstance(x, (int, np.integer)) for x in seq):
        return ("sizes", tuple(int(x) for x in seq))
    return ("indices", tuple(tuple(int(i) for i in _as_sequence(b)) for b in seq))


def normalize_block_specs(block_specs: Any) -> tuple[BlockSpec, ...] | None:
    """Canonical ``(("sizes"|"indices", data), ...)`` form (linalg.py:20-72)."""
    if block_specs is None:
        return None
    items = (block_specs,) if isinstance(block_specs, dict) else block_specs
    if not isinstance(items, (list, tuple)):
        raise TypeError(
            "block_specs must be a sequence of specs or a single spec dict; "
            f"got {type(block_specs)}."
        )
    return tuple(_normalize_one_spec(s) for s in items)


def _offblock_mask(n: int, kind: str, data) -> np.ndarray:
    mask = np.ones((n, n), dtype=bool)
    if kind == "sizes":
        start = 0
        for size in data:
            mask[start:start + size, start:start + size] = False
            start += size
    else:
        for idx in data:
            ix = np.asarray(idx, dtype=int)
            mask[np.ix_(ix, ix)] = False
    return mask


def _eigh_blocks(array: np.ndarray, blocks: list[np.ndarray], sort: bool):
    """Eigh of each diagonal block (batched on the GPU per block), assembled
    into full-size eigenvalues / eigenvectors, optionally sorted."""
    n = array.shape[-1]
    lead = array.shape[:-2]
    w = np.empty(lead + (n,), dtype=array.real.dtype)
    V = np.zeros(lead + (n, n), dtype=array.dtype)
    col = 0
    for ix in blocks:
        sub = array[..., ix[:, None], ix[None, :]]
        wb, vb = batched_eigh(sub)
        s = len(ix)
        w[..., col:col + s] = wb
        V[..., ix[:, None], np.arange(col, col + s)[None, :]] = vb
        col += s
    if sort:
        order = np.argsort(w, axis=-1, kind="stable")
        w = np.take_along_axis(w, order, axis=-1)
        V = np.take_along_axis(V, order[..., None, :], axis=-1)
    return w, V


def _eigh_block_specs(array, specs, *, sort, check_offdiag, offdiag_atol, offdiag_rtol):
    n = int(array.shape[-1])
    tol = None
    if check_offdiag:
        tol = float(offdiag_atol) + float(offdiag_rtol) * float(np.max(np.abs(array)))
    for kind, data in specs:
        if kind not in ("sizes", "indices"):
            raise ValueError("block_specs kind must be 'sizes' or 'indices'.")
        if kind == "sizes":
            if any(s <= 0 for s in data):
                raise ValueError("block_sizes entries must be positive integers.")
            if sum(data) != n:
                raise ValueError(f"block_sizes must sum to {n} (got {sum(data)}).")
        else:
            used: set[int] = set()
            for b in data:
                if len(b) == 0:
                    raise ValueError("block_indices entries must be non-empty.")
                for i in b:
                    if i < 0 or i >= n:
                        raise ValueError(
                            f"block_indices contains index {i} outside [0, {n - 1}].")
                    if i in used:
                        raise ValueError(
                            "block_indices must form a disjoint partition of the basis.")
                    used.add(i)
            if len(used) != n:
                raise ValueError(f"block_indices must cover all basis indices 0..{n - 1}.")
        mask = _offblock_mask(n, kind, data)
        if (not check_offdiag) or float(np.max(np.abs(array) * mask)) <= tol:
            if kind == "sizes":
                starts = np.cumsum((0,) + tuple(data))
                blocks = [np.arange(starts[i], starts[i + 1]) for i in range(len(data))]
            else:
                blocks = [np.asarray(b, dtype=np.intp) for b in data]
            return _eigh_blocks(array, blocks, sort)
    return batched_eigh(array)


def eigh(
    array: np.ndarray,
    *,
    block_specs: Any | None = None,
    block_sizes: Any | None = None,
    block_indices: Any | None = None,
    sort: bool = True,
    check_offdiag: bool | None = None,
    offdiag_atol: float = 1e-12,
    offdiag_rtol: float = 0.0,
    **kwargs,
) -> tuple[np.ndarray, np.ndarray]:
    """Batched Hermitian eigh on the GPU (linalg.py:239-297).

    Ascending eigenvalues and column eigenvectors, cast back to the input
    precision.  The input is hermitised first.
    """
    array = np.asarray(array)
    check = True if check_offdiag is None else bool(check_offdiag)
    specs = None
    if block_specs is not None:
        specs = normalize_block_specs(block_specs)
    elif block_sizes is not None:
        specs = normalize_block_specs(({"block_sizes": block_sizes},))
    elif block_indices is not None:
        specs = normalize_block_specs(({"block_indices": block_indices},))
    if specs:
        return _eigh_block_specs(array, specs, sort=bool(sort), check_offdiag=check,
                                 offdiag_atol=float(offdiag_atol),
                                 offdiag_rtol=float(offdiag_rtol))
    return batched_eigh(array)
