"""Physics utilities of the Python layer (reference: src/cpp_hf/utils.py).

Compute goes to the GPU library: the exchange self-energy, the chemical
potential, the density from a Fock matrix and the k-grid resampling.  The
block-diagonal exchange path slices the band blocks here and runs each block
through the GPU exchange, like the reference's per-block numpy loop
(utils.py:156-251).
"""

from __future__ import annotations

from typing import Any

import numpy as np

from ._compat import _native, expit, hermitize, ifftshift2d, native_required
from .linalg import eigh, normalize_block_specs


def validate_electron_count(weights: Any, nbands: int, n_electrons: Any, *,
                            context: str = "n_electrons") -> None:
    """Reject electron counts outside [0, sum(w) * nbands] (utils.py:22-40)."""
    total_weight = float(np.real(np.asarray(weights)).sum())
    target = float(np.asarray(n_electrons))
    max_electrons = total_weight * int(nbands)
    tol = max(1e-8, 1e-6 * max(1.0, abs(max_electrons)))
    if not np.isfinite(target):
        raise ValueError(f"{context} must be finite, got {target!r}.")
    if target < -tol or target > max_electrons + tol:
        raise ValueError(
            f"{context}={target} is outside the physically reachable range "
            f"[0, {max_electrons}] for {int(nbands)} bands and weights.sum()={total_weight}."
        )


def fermidirac(x: np.ndarray, T: float) -> np.ndarray:
    return expit(-np.asarray(x) / (float(T) + 1e-12))


def electron_density(P: np.ndarray) -> np.ndarray:
    return np.real(np.trace(P, axis1=-2, axis2=-1))


def density_spectrum(bands: np.ndarray, mu: float, T: float) -> np.ndarray:
    return fermidirac(bands - mu, T).sum(axis=-1)


def _exchange_gpu(VR: np.ndarray, P: np.ndarray, hermitian_channel_packing: bool) -> np.ndarray:
    native_required()
    return _native.selfenergy_fft_full(
        np.ascontiguousarray(VR, dtype=np.complex128),
        np.ascontiguousarray(P, dtype=np.complex128),
        apply_ifftshift=False,
        hermitian_channel_packing=bool(hermitian_channel_packing),
    )


def _slice_interaction(VR: np.ndarray, idx) -> np.ndarray:
    if VR.shape[-2:] == (1, 1):
        return VR
    return VR[..., idx, idx] if isinstance(idx, slice) else VR[..., idx[:, None], idx[None, :]]


def _exchange_blocks_sizes(VR, P, sizes, hcp):
    n = int(P.shape[-1])
    if sum(int(s) for s in sizes) != n:
        raise ValueError(f"block_sizes must sum to {n}.")
    out = np.zeros(P.shape, dtype=np.complex128)
    start = 0
    for size in sizes:
        s = slice(start, start + int(size))
        out[..., s, s] = _exchange_gpu(_slice_interaction(VR, s), P[..., s, s], hcp)
        start += int(size)
    return out


def _exchange_blocks_indices(VR, P, blocks, hcp):
    out = np.zeros(P.shape, dtype=np.complex128)
    for idx in blocks:
        ix = np.asarray(idx, dtype=np.intp)
        sub = P[..., ix[:, None], ix[None, :]]
        out[..., ix[:, None], ix[None, :]] = _exchange_gpu(_slice_interaction(VR, ix), sub, hcp)
    return out


def selfenergy_fft(
    VR: np.ndarray,
    P: np.ndarray,
    *,
    block_specs: Any | None = None,
    check_offdiag: bool | None = None,
    offdiag_atol: float = 1e-12,
    offdiag_rtol: float = 0.0,
    _apply_ifftshift: bool = True,
    hermitian_channel_packing: bool = False,
) -> np.ndarray:
    """Exchange self-energy ``Sigma = -IFFT[FFT(P) * VR]`` on the GPU.

    Same arguments and block-spec dispatch as utils.py:55-108.
    """
    VR = np.asarray(VR)
    P = np.asarray(P)
    if hermitian_channel_packing and VR.shape[-2:] != (1, 1):
        raise ValueError(
            "hermitian_channel_packing requires a scalar interaction kernel "
            "with shape (..., 1, 1)."
        )
    result = None
    specs = normalize_block_specs(block_specs) if block_specs is not None else None
    if specs:
        check = bool(check_offdiag) if check_offdiag is not None else True
        n = int(P.shape[-1])
        tol = None
        if check:
            tol = float(offdiag_atol) + float(offdiag_rtol) * float(np.max(np.abs(P)))
        for kind, data in specs:
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
            if (not check) or float(np.max(np.abs(P) * mask)) <= tol:
                if kind == "sizes":
                    result = _exchange_blocks_sizes(VR, P, data, hermitian_channel_packing)
                else:
                    result = _exchange_blocks_indices(VR, P, data, hermitian_channel_packing)
                break
    if result is None:
        result = _exchange_gpu(VR, P, hermitian_channel_packing)
    if _apply_ifftshift:
        return ifftshift2d(result)
    return result


def find_chemical_potential(
    bands: np.ndarray,
    weights: np.ndarray,
    n_electrons: float,
    T: float,
    *,
    maxiter: int | None = None,
    method: str = "bisection",
) -> np.ndarray:
    """mu with sum_k w_k sum_j f(eps_kj - mu) = n_electrons (utils.py:254-274).

    ``bisection`` runs the GPU multisection kernel (30 steps for float64
    bands, 54 for float32, as the reference); ``newton`` runs the
    reference's Newton/bisection hybrid in one GPU workgroup.
    """
    validate_electron_count(weights, np.asarray(bands).shape[-1], n_electrons,
                            context="n_electrons")
    method = str(method).lower()
    bands = np.asarray(bands)
    real_dtype = bands.real.dtype if np.iscomplexobj(bands) else bands.dtype
    if method == "bisection":
        if maxiter is None:
            maxiter = 30 if real_dtype == np.float64 else 54
        native_required()
        mu = _native.find_mu_bisection(
            np.ascontiguousarray(bands, dtype=np.float64),
            np.ascontiguousarray(weights, dtype=np.float64),
            float(n_electrons), float(T), int(maxiter),
        )
        return np.asarray(mu, dtype=real_dtype)
    if method == "newton":
        return _find_mu_newton(bands, weights, n_electrons, T, maxiter=maxiter)
    raise ValueError(f"method must be 'bisection' or 'newton', got {method!r}")


def _find_mu_newton(bands, weights, n_electrons, T, *, maxiter=None) -> np.ndarray:
    """Newton with bisection safeguard on the GPU (utils.py:311-360).

    The reference's algorithm step for step: raw weights, the bracket
    [emin - span, emax + span], its midpoint as the start, ``maxiter`` steps
    (15 for float64 bands, 25 otherwise) and the bracket midpoint when the
    final count misses the target by more than 1e-12.
    """
    native_required()
    bands = np.asarray(bands)
    real_dtype = bands.real.dtype if np.iscomplexobj(bands) else bands.dtype
    if maxiter is None:
        maxiter = 15 if real_dtype == np.float64 else 25
    mu = _native.find_mu_newton(
        np.ascontiguousarray(bands, dtype=np.float64),
        np.ascontiguousarray(weights, dtype=np.float64),
        float(n_electrons), float(T), int(maxiter),
    )
    return np.asarray(mu, dtype=real_dtype)


def density_matrix_from_fock(
    F: np.ndarray,
    weights: np.ndarray,
    n_electrons: float,
    T: float,
    *,
    eigh_block_specs: object | None = None,
    eigh_check_offdiag: bool | None = None,
    eigh_offdiag_atol: float = 1e-12,
    eigh_offdiag_rtol: float = 0.0,
) -> tuple[np.ndarray, np.ndarray]:
    """Physical density P(k) from a Hermitian Fock matrix (utils.py:363-389)."""
    F = hermitize(np.asarray(F))
    validate_electron_count(weights, F.shape[-1], n_electrons, context="n_electrons")
    native_required()
    if eigh_block_specs is None:
        maxiter = 30 if F.real.dtype == np.float64 else 54
        P, mu = _native.density_from_fock(
            np.ascontiguousarray(F, dtype=np.complex128),
            np.ascontiguousarray(weights, dtype=np.float64),
            float(n_electrons), float(T), int(maxiter),
        )
        return P.astype(F.dtype, copy=False), np.asarray(mu, dtype=F.real.dtype)
    eps, U = eigh(F, block_specs=eigh_block_specs, check_offdiag=eigh_check_offdiag,
                  offdiag_atol=eigh_offdiag_atol, offdiag_rtol=eigh_offdiag_rtol)
    mu = find_chemical_potential(eps, np.asarray(weights), n_electrons=float(n_electrons),
                                 T=float(T))
    occ = fermidirac(eps - mu, float(T))
    P = np.einsum("...in,...n,...jn->...ij", U, occ, np.conj(U))
    return hermitize(P), mu


def resample_kgrid(values: np.ndarray, nk: int, *, method: str = "linear") -> np.ndarray:
    """Resample a centred periodic (nk, nk, ...) k-grid array (utils.py:392-411)."""
    nk = int(nk)
    x = np.asarray(values)
    if x.ndim < 2:
        raise ValueError(
            "resample_kgrid expects an array with at least 2 dimensions (nk,nk,...)."
        )
    if x.shape[0] == nk and x.shape[1] == nk:
        return x
    if str(method).lower() != "linear":
        raise ValueError(f"Unsupported resample method {method!r}. Only 'linear' is supported.")
    native_required()
    return _native.resample_kgrid_2d(np.ascontiguousarray(x, dtype=np.complex128), nk)
