"""Symmetry projectors for ``project_fn`` (reference: src/cpp_hf/symmetry.py).

``make_project_fn`` returns a callable that averages a density or Fock
matrix of shape (nk1, nk2, nb, nb) over
  - a unitary group G:           A_k -> (1/|G|) sum_g g A_k g^H,
  - a spatial group S taking k to -k (joined with G, or with the identity
    when G is absent): (1/(|G|+|S|)) [sum_g g A_k g^H + sum_s s A_-k s^H],
  - time reversal with U:        A_k -> (A_k + U conj(A_-k) U^H) / 2,
applied in that order.  -k is the index map i -> -i mod n ("mod") or the
array flip ("flip") along the chosen k axes.

Called directly, the projector averages a numpy array on the host.  Passed
as ``project_fn`` to ``solve`` / ``solve_scf``, it runs on the GPU inside the
solver loops (libhfgpu's device projector, ``hf_symmetry_desc``), with no
host round trip per projection; a plain callable still goes through the host
(as the reference's ProjectFn, fock.hpp:295-298).
"""

from __future__ import annotations

from typing import Callable

import numpy as np

ProjectFn = Callable[[np.ndarray], np.ndarray]

__all__ = ["ProjectFn", "SymmetryProjector", "make_project_fn"]


def _minus_k(A: np.ndarray, convention: str, axes: tuple[int, ...]) -> np.ndarray:
    if convention == "flip":
        return np.flip(A, axis=axes)
    if convention == "mod":
        n1, n2 = A.shape[0], A.shape[1]
        i = (-np.arange(n1)) % n1 if 0 in axes else np.arange(n1)
        j = (-np.arange(n2)) % n2 if 1 in axes else np.arange(n2)
        return A[i[:, None], j[None, :], ...]
    raise ValueError(f"k_convention must be 'mod' or 'flip', got {convention!r}")


def _conj_sum(A: np.ndarray, group: np.ndarray) -> np.ndarray:
    """sum_g g A g^H over the group elements (leading axis of group)."""
    out = np.zeros_like(A)
    for g in group:
        out += (g @ A) @ np.conj(g.T)
    return out


def _axes_bits(axes: tuple[int, ...]):
    bits = 0
    for a in axes:
        if a not in (0, 1):
            return None
        bits |= 1 << a
    return bits


class SymmetryProjector:
    """The averaging of make_project_fn as a callable (numpy, on the host)
    that also describes itself to the GPU solvers."""

    def __init__(self, G, S, s_conv, s_axes, U, t_conv, t_axes):
        self.unitary_group, self.spatial_group, self.time_reversal_U = G, S, U
        self.spatial_k_convention, self.spatial_k_flip_axes = s_conv, s_axes
        self.time_reversal_k_convention, self.time_reversal_k_flip_axes = t_conv, t_axes

    def __call__(self, A: np.ndarray) -> np.ndarray:
        G, S, U = self.unitary_group, self.spatial_group, self.time_reversal_U
        out = np.asarray(A)
        if S is not None:
            same = G if G is not None else np.eye(S.shape[-1], dtype=S.dtype)[None]
            total = _conj_sum(out, same) + _conj_sum(
                _minus_k(out, self.spatial_k_convention, self.spatial_k_flip_axes), S)
            out = total / np.asarray(float(same.shape[0] + S.shape[0]), dtype=out.dtype)
        elif G is not None:
            out = _conj_sum(out, G) / np.asarray(float(G.shape[0]), dtype=out.dtype)
        if U is not None:
            tr = U @ np.conj(_minus_k(out, self.time_reversal_k_convention,
                                      self.time_reversal_k_flip_axes)) @ np.conj(U.T)
            out = 0.5 * (out + tr)
        return out

    def _hfgpu_symmetry(self):
        """(unitary, spatial, spatial_mod, spatial_axes, tr_U, tr_mod,
        tr_axes) for the device projector, or None when the k axes are not
        a subset of (0, 1).  Unknown k conventions raise ValueError, as a
        call would."""
        for conv, used in ((self.spatial_k_convention, self.spatial_group is not None),
                           (self.time_reversal_k_convention, self.time_reversal_U is not None)):
            if used and conv not in ("mod", "flip"):
                raise ValueError(f"k_convention must be 'mod' or 'flip', got {conv!r}")
        s_bits = _axes_bits(self.spatial_k_flip_axes)
        t_bits = _axes_bits(self.time_reversal_k_flip_axes)
        if s_bits is None or t_bits is None:
            return None
        c = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.complex128)
        return (c(self.unitary_group), c(self.spatial_group),
                self.spatial_k_convention == "mod", s_bits, c(self.time_reversal_U),
                self.time_reversal_k_convention == "mod", t_bits)


def make_project_fn(*, unitary_group: np.ndarray | None = None,
                    spatial_group: np.ndarray | None = None,
                    spatial_k_convention: str = "mod",
                    spatial_k_flip_axes: tuple[int, ...] = (0, 1),
                    time_reversal_U: np.ndarray | None = None,
                    time_reversal_k_convention: str = "mod",
                    time_reversal_k_flip_axes: tuple[int, ...] = (0, 1)) -> ProjectFn:
    """Symmetry-averaging projection (symmetry.py:81-117)."""
    G = None if unitary_group is None else np.asarray(unitary_group)
    S = None if spatial_group is None else np.asarray(spatial_group)
    U = None if time_reversal_U is None else np.asarray(time_reversal_U)
    if G is None and S is None and U is None:
        return lambda A: A
    s_conv, s_axes = str(spatial_k_convention), tuple(spatial_k_flip_axes)
    t_conv, t_axes = str(time_reversal_k_convention), tuple(time_reversal_k_flip_axes)
    return SymmetryProjector(G, S, s_conv, s_axes, U, t_conv, t_axes)
