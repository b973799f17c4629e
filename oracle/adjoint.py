"""ORACLE — test infrastructure only.

numpy statement of the vector-Jacobian product of the Fock build: the
adjoint of build_fock_compact + build_fock (cpp/include/cpp_hf/fock.hpp:
34-321) together with HartreeFockKernel's interaction transform
VR = phase * fftn(w Vq) (src/cpp_hf/problem.py:84-107).  The reference has
no derivatives, so this restatement is pinned by the mathematics alone:
central finite differences of the oracle's own Fock build (which is linear
in P, h, Vq and HH, so the differences are exact to rounding) in
tests/test_oracle_adjoint.py.  Only tests/ import it.

Convention: dL = Re sum conj(g) dX for L = Re<cot_sigma, Sigma> +
Re<cot_hartree, Hh> + Re<cot_F, F>.
"""

from __future__ import annotations

import numpy as np


def _herm(X):
    return 0.5 * (X + np.conj(np.swapaxes(X, -1, -2)))


def _phase(n0: int, n1: int) -> np.ndarray:
    s0, s1 = n0 // 2, n1 // 2
    ph0 = np.exp(2j * np.pi * np.arange(n0) * s0 / n0)
    ph1 = np.exp(2j * np.pi * np.arange(n1) * s1 / n1)
    return ph0[:, None] * ph1[None, :]


def build_fock_vjp(ka: dict, P, cot_sigma=None, cot_hartree=None, cot_F=None):
    """(gP, gVq, gh, gHH) for the kernel_args dict ka at density P."""
    h = np.asarray(ka["h"])
    n0, n1, nb, _ = h.shape
    N = n0 * n1
    w = np.asarray(ka["w2d"], dtype=np.float64).reshape(n0, n1)
    VR = np.asarray(ka["VR"])
    P = np.asarray(P, dtype=np.complex128)
    zero = np.zeros_like(P)
    GFh = _herm(np.asarray(cot_F)) if cot_F is not None else zero
    G = (np.asarray(cot_sigma) if cot_sigma is not None else zero) + GFh
    Gd = GFh + (np.asarray(cot_hartree) if cot_hartree is not None else zero)
    dP = P - np.asarray(ka["refP"]) if ka["has_refP"] else P

    gP = np.zeros_like(P)
    gVq = np.zeros(VR.shape, dtype=np.complex128)
    if ka["include_exchange"]:
        X = np.fft.fft2(G, axes=(0, 1))
        Y = np.fft.fft2(dP, axes=(0, 1))
        gam = -(X * np.conj(Y)) / N
        if VR.shape[-1] == 1:
            gam = gam.sum(axis=(-2, -1), keepdims=True)
        gVq = w[..., None, None] * (N * np.fft.ifft2(np.conj(_phase(n0, n1))[..., None, None] * gam,
                                                     axes=(0, 1)))
        # Sigma_x = -ifft2(VR fft2(dP)) -> adjoint -ifft2(conj(VR) fft2(G))
        gP = -np.fft.ifft2(np.conj(VR) * X, axes=(0, 1))

    r = np.real(np.einsum("...aa->...a", Gd)).sum(axis=(0, 1))
    n = np.real(np.einsum("...aa->...a", w[..., None, None] * dP)).sum(axis=(0, 1))
    M = np.zeros((nb, nb), dtype=np.complex128)
    gHH = np.zeros((nb, nb))
    if ka["include_hartree"]:
        HH = np.asarray(ka["HH"], dtype=np.float64)
        M += np.diag(HH.T @ r)
        gHH = np.outer(r, n)
    cg = np.asarray(ka["contact_g"], dtype=np.float64)
    if np.any(cg != 0.0) or cg.shape[0] > 1:
        S = G.sum(axis=(0, 1))
        for g_t, Oi, Oj in zip(cg, np.asarray(ka["contact_Oi"]), np.asarray(ka["contact_Oj"])):
            if g_t == 0.0:
                continue
            c = np.sum(np.conj(S) * Oi)
            M += g_t * (np.conj(c) * Oj.conj().T - Oi.conj().T @ S @ Oj.conj().T)
    gP = gP + w[..., None, None] * M
    return gP, gVq, GFh, gHH
