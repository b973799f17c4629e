"""Parameter gradients of the converged free energy through the oracle's own
re-solves (CPU).

At a converged SCF state the free energy Omega = E - T S is stationary in
the density at fixed electron count, so dOmega/dtheta for a model parameter
theta is the partial derivative at fixed density (Hellmann-Feynman): the
oracle's Fock-build adjoint (oracle/adjoint.py) contracted with the model
tangent.  Here it is checked against central differences of complete oracle
SCF re-solves at theta +- h for the layered model's u, scale, screening
length and interlayer hopping.  The GPU adjoint at the same state is checked
against the same differences in tests/test_gpu_adjoint.py.
"""

from __future__ import annotations

import numpy as np
import pytest

from helpers import kernel_args_dict

from oracle import adjoint as adj

N_MESH, N_LAYERS, T, N_E = 12, 2, 0.1, 2.1
BASE = dict(u=0.1, gamma1=0.4, scale=0.8, gate=4.0)
STEP = 1e-5


def _kernel(**kw):
    from cpp_hf_amd.workloads import layered_kernel

    p = dict(BASE)
    p.update(kw)
    return layered_kernel(N_MESH, N_LAYERS, T=T, filling=N_E, **p)


def oracle_converged(oracle, **kw):
    """(kernel, P, mu) of the oracle's DIIS SCF to a tight commutator
    residual; the density residual floors at the resolution of the 30-level
    chemical-potential bisection."""
    wl = _kernel(**kw)
    r = oracle.solve_scf(kernel_args_dict(wl.kernel), np.ascontiguousarray(wl.P0), wl.n_electrons,
                         400, 1e-7, 1e-9, 0.5, 0.0, None, [], "diis", 6, 2, 1.0, 0.0, True, True)
    P, _, _, mu, n, conv = r[:6]
    assert conv, (kw, n)
    return wl.kernel, np.asarray(P), float(mu)


def oracle_omega(oracle, k, P, mu):
    """E - T S at P from the oracle's build, moved to the electron count N_E
    to first order (the solve's count carries the bisection's resolution and
    dOmega/dN = mu)."""
    *_, E = oracle.build_fock_apply(kernel_args_dict(k), P)
    f = np.clip(np.linalg.eigvalsh(0.5 * (P + np.conj(np.swapaxes(P, -1, -2)))), 1e-14,
                1 - 1e-14)
    wn = np.asarray(k.w2d) / np.sum(k.w2d)
    S = -np.sum(wn[..., None] * (f * np.log(f) + (1 - f) * np.log1p(-f)))
    N = float(np.sum(np.asarray(k.w2d)[..., None] * np.real(np.einsum("...aa->...a", P))))
    return float(E) - max(k.T, 1e-14) * S - mu * (N - N_E)


def oracle_derivative(k, P, dh, dVq, dHH):
    """dOmega/dtheta from the oracle adjoint: dE/dh = w P^H, and the
    interaction and Hartree-matrix parts are the Fock-build VJP with
    cotangents (w/2)(P - refP)^H on Sigma and the Hartree matrix."""
    ka = kernel_args_dict(k)
    w = np.asarray(k.w2d, dtype=np.float64)[..., None, None]
    dP = P - k._refP if k._refP is not None else P
    half = 0.5 * w * np.conj(np.swapaxes(dP, -1, -2))
    _, gVq, _, gHH = adj.build_fock_vjp(ka, P, cot_sigma=half, cot_hartree=half)
    gh = w * np.conj(np.swapaxes(P, -1, -2))
    return float(np.real(np.vdot(gh, dh)) + np.real(np.vdot(gVq, dVq)) + np.sum(gHH * dHH))


def central_differences(oracle, name):
    kp, Pp, mup = oracle_converged(oracle, **{name: BASE[name] + STEP})
    km, Pm, mum = oracle_converged(oracle, **{name: BASE[name] - STEP})
    return (oracle_omega(oracle, kp, Pp, mup) - oracle_omega(oracle, km, Pm, mum)) / (2 * STEP)


@pytest.fixture(scope="module")
def base_state(oracle):
    return oracle_converged(oracle)


@pytest.mark.parametrize("name", ["u", "scale", "gate", "gamma1"])
def test_oracle_adjoint_matches_oracle_resolves(oracle, base_state, name):
    from cpp_hf_amd.workloads import layered_tangents

    k0, P0, _ = base_state
    dh, dVq, dHH = layered_tangents(N_MESH, N_LAYERS, **BASE)[name]
    ad = oracle_derivative(k0, P0, dh, dVq, dHH)
    fd = central_differences(oracle, name)
    assert abs(ad - fd) <= 1e-6 * max(abs(fd), 1e-3), (name, ad, fd)


def test_fd_detects_a_dropped_term(oracle, base_state):
    """The check has teeth: the adjoint without its interaction part misses
    the screening-length difference by far more than the tolerance."""
    from cpp_hf_amd.workloads import layered_tangents

    k0, P0, _ = base_state
    dh, dVq, dHH = layered_tangents(N_MESH, N_LAYERS, **BASE)["gate"]
    full = oracle_derivative(k0, P0, dh, dVq, dHH)
    dropped = oracle_derivative(k0, P0, dh, np.zeros_like(dVq), dHH)
    fd = central_differences(oracle, "gate")
    assert abs(full - fd) <= 1e-6 * max(abs(fd), 1e-3)
    assert abs(dropped - fd) > 1e-3 * max(abs(fd), 1e-3)
