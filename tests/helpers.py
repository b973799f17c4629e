"""Shared builders for the tests (seeded, small)."""

from __future__ import annotations

import numpy as np


def herm(X):
    return 0.5 * (X + np.conj(np.swapaxes(X, -1, -2)))


def rand_herm(rng, shape):
    X = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    return herm(X)


def rand_c(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def kernel_args_dict(kernel):
    from cpp_hf_amd.solver import _kernel_args_for_native

    return _kernel_args_for_native(kernel)


def two_band_kernel(nk=1, T=0.2, exchange=0.25):
    """The reference tests' two-band model (tests/test_solver.py:24-35)."""
    from cpp_hf_amd import HartreeFockKernel

    w = np.ones((nk, nk), dtype=np.float32)
    h = np.zeros((nk, nk, 2, 2), dtype=np.complex64)
    h[..., 0, 0] = -0.5
    h[..., 1, 1] = 0.5
    v = np.full((nk, nk, 1, 1), exchange, dtype=np.complex64)
    return HartreeFockKernel(weights=w, hamiltonian=h, coulomb_q=v, T=T)


def contact_kernel(g, nk=2, T=0.05):
    """tests/test_solver.py:180-200."""
    from cpp_hf_amd import HartreeFockKernel

    w = np.ones((nk, nk), dtype=np.float32)
    h = np.zeros((nk, nk, 2, 2), dtype=np.complex64)
    h[..., 0, 0] = -1.0
    h[..., 1, 1] = 1.0
    for i in range(nk):
        for j in range(nk):
            h[i, j, 0, 1] = 0.4 + 0.1 * (i + j)
            h[i, j, 1, 0] = 0.4 + 0.1 * (i + j)
    v = np.full((nk, nk, 1, 1), 0.1, dtype=np.complex64)
    sz = np.array([[1.0, 0.0], [0.0, -1.0]], dtype=np.complex64)
    return HartreeFockKernel(weights=w, hamiltonian=h, coulomb_q=v, T=T,
                             contact_terms=[(np.float32(g), sz, sz)])


def hartree_kernel(nk=3, nb=3, seed=4, dense_v=False):
    """Random Hermitian h, refP, Hartree matrix (tests/test_main.py:61-85)."""
    from cpp_hf_amd import HartreeFockKernel

    rng = np.random.default_rng(seed)
    w = np.ones((nk, nk)) / (nk * nk)
    h = rand_herm(rng, (nk, nk, nb, nb))
    P = rand_herm(rng, (nk, nk, nb, nb))
    refP = 0.1 * rand_herm(rng, (nk, nk, nb, nb))
    HH = rng.standard_normal((nb, nb))
    HH = 0.5 * (HH + HH.T) + nb * np.eye(nb)
    if dense_v:
        # physical interaction: real, symmetric in the band pair, even in q
        # (index i <-> (-i) mod nk on the centred mesh)
        A = rng.standard_normal((nk, nk, nb, nb))
        A = A + np.swapaxes(A, -1, -2)
        refl = (-np.arange(nk)) % nk
        Vq = 0.1 * (A + A[refl][:, refl])
    else:
        Vq = np.full((nk, nk, 1, 1), 0.15)
    k = HartreeFockKernel(w, h, Vq, T=0.1, include_hartree=True, include_exchange=True,
                          reference_density=refP, hartree_matrix=HH)
    return k, P


def brute_force_exchange(VR, P):
    """Sigma_ab(k) = -sum_q W_ab(k - q) P_ab(q) with W the inverse DFT of VR,
    both by explicit DFT sums (no FFT)."""
    n1, n2 = P.shape[0], P.shape[1]
    i1 = np.arange(n1)
    i2 = np.arange(n2)
    E1 = np.exp(2j * np.pi * np.outer(i1, i1) / n1)
    E2 = np.exp(2j * np.pi * np.outer(i2, i2) / n2)
    W = np.einsum("qm,rn,qr...->mn...", E1, E2, VR) / (n1 * n2)
    out = np.zeros(P.shape, dtype=np.complex128)
    for a in range(n1):
        for b in range(n2):
            for q1 in range(n1):
                for q2 in range(n2):
                    out[a, b] -= W[(a - q1) % n1, (b - q2) % n2] * P[q1, q2]
    return out


def block_kernel(nk=6, T=0.1, couple=0.0):
    """Two 2-band QWZ copies on one mesh (masses 0.6 and -0.9, the second
    shifted by 0.2), scalar V, Hartree with refP = I/2: block-diagonal in the
    band blocks (2, 2) unless `couple` adds an off-block hopping."""
    from cpp_hf_amd import HartreeFockKernel

    k = 2.0 * np.pi * np.arange(nk) / nk - np.pi
    KX, KY = np.meshgrid(k, k, indexing="ij")
    sx = np.array([[0, 1], [1, 0]], dtype=complex)
    sy = np.array([[0, -1j], [1j, 0]])
    sz = np.diag([1.0, -1.0]).astype(complex)

    def qwz(m):
        return (np.sin(KX)[..., None, None] * sx + np.sin(KY)[..., None, None] * sy
                + (m - np.cos(KX) - np.cos(KY))[..., None, None] * sz)

    h = np.zeros((nk, nk, 4, 4), dtype=complex)
    h[..., :2, :2] = qwz(0.6)
    h[..., 2:, 2:] = qwz(-0.9) + 0.2 * np.eye(2)
    h[..., 0, 2] = couple
    h[..., 2, 0] = couple
    w = np.full((nk, nk), 1.0 / (nk * nk))
    V = np.full((nk, nk, 1, 1), 0.3)
    HH = 0.2 * np.ones((4, 4)) + 0.3 * np.eye(4)
    refP = np.broadcast_to(0.5 * np.eye(4, dtype=complex), h.shape).copy()
    return HartreeFockKernel(w, h, V, T=T, include_hartree=True, reference_density=refP,
                             hartree_matrix=HH)
