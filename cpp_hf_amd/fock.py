"""Fock matrix and energies (reference: src/cpp_hf/fock.py).

``build_fock`` runs the whole Fock build on the GPU (exchange, Hartree
diagonal, contact terms, hermitisation) through ``_native.build_fock_apply``
and returns the reference's ``(Sigma, H, F)`` with ``H`` as the (1, 1, nb, nb)
Hartree matrix.  ``hf_energy`` and ``occupation_entropy`` are GPU reductions.
"""

from __future__ import annotations

from typing import Any

import numpy as np

from ._compat import _native, native_required
from .utils import selfenergy_fft


def _herm(X: np.ndarray) -> np.ndarray:
    return 0.5 * (X + np.conj(np.swapaxes(X, -1, -2)))


def _is_zero_broadcast(refP) -> bool:
    """The kernel's stand-in for an absent reference density."""
    if refP is None:
        return True
    a = np.asarray(refP)
    return bool(a.size and all(s == 0 for s in a.strides) and a.flat[0] == 0)


def build_fock(
    P: np.ndarray,
    *,
    h: np.ndarray,
    VR: np.ndarray,
    refP: np.ndarray,
    HH: np.ndarray,
    w2d: np.ndarray,
    include_exchange: bool,
    include_hartree: bool,
    exchange_hermitian_channel_packing: bool,
    contact_g: np.ndarray | None = None,
    contact_Oi: np.ndarray | None = None,
    contact_Oj: np.ndarray | None = None,
    exchange_block_specs: Any | None = None,
    exchange_check_offdiag: bool | None = None,
    exchange_offdiag_atol: float = 1e-12,
    exchange_offdiag_rtol: float = 0.0,
    project_fn=None,
) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """``(Sigma, H_hartree, F)`` with F = project(herm(h + Sigma[P] + H[P])).

    Signature and outputs of src/cpp_hf/fock.py:16-88.
    """
    native_required()
    h = np.ascontiguousarray(h, dtype=np.complex128)
    nb = h.shape[-1]
    has_refP = not _is_zero_broadcast(refP)
    if exchange_block_specs is not None and include_exchange:
        # block-restricted exchange: Sigma from the per-block GPU exchange,
        # everything else from the GPU Fock build with exchange off
        dP = P - np.asarray(refP) if has_refP else P
        sigma_x = selfenergy_fft(
            VR, dP, block_specs=exchange_block_specs, check_offdiag=exchange_check_offdiag,
            offdiag_atol=exchange_offdiag_atol, offdiag_rtol=exchange_offdiag_rtol,
            _apply_ifftshift=False,
            hermitian_channel_packing=exchange_hermitian_channel_packing,
        )
        if not include_hartree and contact_g is None:
            Sigma = sigma_x
            F = _herm(h + Sigma)
            if project_fn is not None:
                F = _herm(np.asarray(project_fn(F), dtype=F.dtype))
            return Sigma, np.zeros_like(h), F
        rest_S, rest_H, _ = build_fock(
            P, h=h, VR=VR, refP=refP, HH=HH, w2d=w2d, include_exchange=False,
            include_hartree=include_hartree, exchange_hermitian_channel_packing=False,
            contact_g=contact_g, contact_Oi=contact_Oi, contact_Oj=contact_Oj,
        )
        Sigma = sigma_x + rest_S
        F = _herm(h + Sigma + rest_H)
        if project_fn is not None:
            F = _herm(np.asarray(project_fn(F), dtype=F.dtype))
        return Sigma, rest_H, F

    if contact_g is None or contact_Oi is None or contact_Oj is None:
        contact_g = np.zeros((1,))
        contact_Oi = np.zeros((1, nb, nb), dtype=np.complex128)
        contact_Oj = np.zeros((1, nb, nb), dtype=np.complex128)
    w2d = np.ascontiguousarray(w2d, dtype=np.float64)
    ka = dict(
        h=h,
        VR=np.ascontiguousarray(VR, dtype=np.complex128),
        refP=(np.ascontiguousarray(refP, dtype=np.complex128) if has_refP
              else np.empty((0,), np.complex128)),
        has_refP=has_refP,
        w2d=w2d,
        HH=np.ascontiguousarray(HH, dtype=np.float64),
        contact_g=np.ascontiguousarray(contact_g, dtype=np.float64),
        contact_Oi=np.ascontiguousarray(contact_Oi, dtype=np.complex128),
        contact_Oj=np.ascontiguousarray(contact_Oj, dtype=np.complex128),
        weight_sum=float(np.sum(w2d)),
        T=0.0,
        include_hartree=bool(include_hartree),
        include_exchange=bool(include_exchange),
        exchange_hcp=bool(exchange_hermitian_channel_packing),
    )
    if not include_exchange:
        ka["VR"] = np.zeros(h.shape[:2] + (1, 1), np.complex128)
    Sigma, Hh, F, _E = _native.build_fock_apply(
        ka, np.ascontiguousarray(P, dtype=np.complex128), project_fn
    )
    if include_hartree:
        H = Hh[:1, :1].real.copy()
    else:
        H = np.zeros_like(h)
    return Sigma, H, F


def hf_energy(
    P: np.ndarray,
    *,
    h: np.ndarray,
    Sigma: np.ndarray,
    H: np.ndarray,
    weights_b: np.ndarray,
    refP: np.ndarray | None = None,
) -> np.ndarray:
    """E = sum_k w_k Re[Tr(h P) + 1/2 Tr((Sigma + H)(P - refP))] (fock.py:91-107)."""
    native_required()
    P = np.ascontiguousarray(P, dtype=np.complex128)
    nb = P.shape[-1]
    H = np.asarray(H, dtype=np.complex128)
    if H.shape != P.shape:
        H = np.broadcast_to(H, (1, 1, nb, nb)) if H.size == nb * nb else np.broadcast_to(H, P.shape)
    w = np.ascontiguousarray(np.broadcast_to(np.asarray(weights_b)[..., 0, 0], P.shape[:2]),
                             dtype=np.float64)
    ref = None if refP is None or _is_zero_broadcast(refP) else np.ascontiguousarray(
        np.broadcast_to(refP, P.shape), dtype=np.complex128)
    E = _native.hf_energy(
        P, np.ascontiguousarray(np.broadcast_to(h, P.shape), dtype=np.complex128),
        np.ascontiguousarray(np.broadcast_to(Sigma, P.shape), dtype=np.complex128),
        np.ascontiguousarray(H), w, ref,
    )
    return np.float64(E)


def occupation_entropy(p: np.ndarray, w_norm: np.ndarray) -> np.ndarray:
    """S = -sum w [p ln p + (1-p) ln(1-p)], p clipped to 1e-14 (fock.py:110-113)."""
    native_required()
    p = np.ascontiguousarray(p, dtype=np.float64)
    w = np.ascontiguousarray(np.broadcast_to(w_norm, p.shape[:-1]), dtype=np.float64)
    return np.float64(_native.occupation_entropy(p, w))


def free_energy(E: np.ndarray, p: np.ndarray, w_norm: np.ndarray, T: np.ndarray) -> np.ndarray:
    """Omega = E - max(T, 1e-14) S(p) (fock.py:116-124)."""
    T_val = max(float(T), 1e-14)
    return E - T_val * occupation_entropy(p, w_norm)
