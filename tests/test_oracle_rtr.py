"""Pins of the oracle's trust-region Newton solver and deflation bias (CPU).

The oracle restates solver_rtr.hpp (joint (Q, p) Hessian-vector product,
Steihaug truncated CG, trust-radius control, deflation bias).  Each part is
checked against something independent of its own formulas, following the
reference's tests/test_solver_rtr.py and tests/test_deflation.py: the
Hessian-vector product against finite differences of the gradient, the bias
against its closed form and finite differences of Phi, and the solver against
the oracle's CG and SCF fixed points, stationarity and particle number.
"""

from __future__ import annotations

import numpy as np
import pytest

from helpers import contact_kernel, herm, kernel_args_dict, rand_herm, two_band_kernel


def _comm_rms(F, P, w):
    c = F @ P - P @ F
    return float(np.sqrt(np.sum(w * np.sum(np.abs(c) ** 2, axis=(-2, -1))) / np.sum(w)))


def multiband_kernel(nk, nb, T, ex, seed):
    """tests/test_solver_rtr.py:57-71."""
    from cpp_hf_amd import HartreeFockKernel

    rng = np.random.default_rng(seed)
    w = np.ones((nk, nk), dtype=np.float32)
    diag = np.linspace(-1.0, 1.0, nb)
    A = (rng.standard_normal((nb, nb)) + 1j * rng.standard_normal((nb, nb))) * 0.35
    A = 0.5 * (A + A.conj().T)
    h = np.zeros((nk, nk, nb, nb), dtype=np.complex128)
    kx = np.arange(nk) / nk * 2 * np.pi
    for i in range(nk):
        for j in range(nk):
            hk = np.diag(diag).astype(np.complex128) + A * (np.cos(kx[i]) + np.cos(kx[j]))
            h[i, j] = 0.5 * (hk + hk.conj().T)
    return HartreeFockKernel(weights=w, hamiltonian=h.astype(np.complex64),
                             coulomb_q=np.full((nk, nk, 1, 1), ex, np.complex64), T=T)


def hybridized_kernel(nk=2, T=0.05, ex=0.25):
    """tests/test_deflation.py:26-39."""
    from cpp_hf_amd import HartreeFockKernel

    w = np.ones((nk, nk), np.float32)
    h = np.zeros((nk, nk, 2, 2), np.complex64)
    h[..., 0, 0] = -0.5
    h[..., 1, 1] = 0.5
    h[..., 0, 1] = 0.3
    h[..., 1, 0] = 0.3
    return HartreeFockKernel(weights=w, hamiltonian=h,
                             coulomb_q=np.full((nk, nk, 1, 1), ex, np.complex64), T=T)


def bistable_kernel(nk=1, T=0.02, ex=1.0, delta=0.1):
    """tests/test_deflation.py:42-57: two self-consistent states at one k point,
    diag(1, 0) with E = -delta - ex/2 and diag(0, 1) with E = delta - ex/2."""
    from cpp_hf_amd import HartreeFockKernel

    w = np.ones((nk, nk), np.float32)
    h = np.zeros((nk, nk, 2, 2), np.complex64)
    h[..., 0, 0] = -delta
    h[..., 1, 1] = +delta
    return HartreeFockKernel(weights=w, hamiltonian=h,
                             coulomb_q=np.full((nk, nk, 1, 1), ex, np.complex64), T=T)


def solve_dm(oracle, k, n_e, *, P0=None, max_iter=200, tol_E=1e-7, tol_grad=0.0,
             plateau_window=5, optimizer=0, targets=None, sigma=0.0, length=1.0,
             project_fn=None, tr_delta0=0.5, tr_cg_max=20):
    """The oracle's solve_dm with SolverConfig's defaults (solver.py:19-64)."""
    P0 = np.zeros_like(k.h) if P0 is None else P0
    defl = np.empty((0,) + k.h.shape) if targets is None else targets
    Q, p, P, F, mu, E, n, conv, hE, hG = oracle.solve_dm(
        kernel_args_dict(k), P0, n_e, max_iter, tol_E, tol_grad, 0.6, 0.5, 1e-3, 8, 10,
        plateau_window, 25, optimizer, tr_delta0, tr_cg_max, defl, sigma, length, [],
        project_fn, True, True, True)
    return dict(Q=Q, p=p, density=P, fock=F, mu=mu, energy=E, n_iter=n, converged=conv,
                hist_E=hE, hist_grad=hG)


def newton(oracle, k, n_e, **kw):
    return solve_dm(oracle, k, n_e, optimizer=1, **kw)


# ------------------------------------------------- Hessian-vector product ---

def test_oracle_rtr_hvp_matches_finite_difference(oracle):
    # tests/test_solver_rtr.py:80-134: the Hessian-vector product is the
    # directional derivative of the joint gradient (G_Q, projected g_p)
    from scipy.linalg import expm

    def offdiag(X):
        X = X.copy()
        i = np.arange(X.shape[-1])
        X[..., i, i] = 0
        return X

    def skew(X):
        return 0.5 * (X - np.conj(np.swapaxes(X, -1, -2)))

    def ft(Q, M):
        return np.conj(np.swapaxes(Q, -1, -2)) @ M @ Q

    def dens(Q, p):
        return (Q * p[..., None, :]) @ np.conj(np.swapaxes(Q, -1, -2))

    K = multiband_kernel(5, 5, 0.05, 0.4, 4)
    args = kernel_args_dict(K)
    T = float(K.T)
    nk1, nk2, nb = K.h.shape[:3]

    def fock(P):
        return oracle.build_fock_apply(args, P)[2]

    rng = np.random.default_rng(1)
    eps, Q = np.linalg.eigh(fock(np.zeros_like(K.h)))
    X0 = skew(offdiag(rng.standard_normal(Q.shape) + 1j * rng.standard_normal(Q.shape)))
    rot = np.stack([expm(0.2 * X0[i, j]) for i in range(nk1) for j in range(nk2)])
    Q = Q @ rot.reshape(Q.shape)
    p = np.clip(1.0 / (1 + np.exp((eps - np.median(eps)) / T)), 0.05, 0.95)
    mu = float(np.median(eps))

    def grad(Q, p):
        Ft = ft(Q, fock(dens(Q, p)))
        GQ = offdiag((p[..., None, :] - p[..., :, None]) * Ft)
        gp = np.real(np.diagonal(Ft, axis1=-2, axis2=-1)) + T * np.log(p / (1 - p)) - mu
        return GQ, gp, Ft

    GQ, gp, Ft = grad(Q, p)
    X = skew(offdiag(rng.standard_normal(Q.shape) + 1j * rng.standard_normal(Q.shape)))
    dp = rng.standard_normal(p.shape) * 0.1
    HX, Hp = oracle._rtr_joint_hvp(args, X, dp, Q, p, Ft)
    w2 = np.asarray(K.w2d)
    errs_Q, errs_p = [], []
    for e in (1e-5, 1e-6):
        rot = np.stack([expm(e * X[i, j]) for i in range(nk1) for j in range(nk2)])
        GQe, gpe, _ = grad(Q @ rot.reshape(Q.shape), p + e * dp)
        fdQ = (GQe - GQ) / e
        fdp = (gpe - gp) / e
        fdp = fdp - np.sum(w2[..., None] * fdp) / (np.sum(w2) * nb)
        errs_Q.append(np.max(np.abs(fdQ - HX)) / np.max(np.abs(HX)))
        errs_p.append(np.max(np.abs(fdp - Hp)) / np.max(np.abs(Hp)))
    assert errs_Q[1] < 1e-4 and errs_p[1] < 1e-4
    assert errs_Q[1] < 0.2 * errs_Q[0] + 1e-12
    assert errs_p[1] < 0.2 * errs_p[0] + 1e-12
    # HX is skew-Hermitian with a zero diagonal; Hp conserves particles
    np.testing.assert_allclose(HX, -np.conj(np.swapaxes(HX, -1, -2)), atol=1e-14)
    assert np.max(np.abs(np.diagonal(HX, axis1=-2, axis2=-1))) == 0.0
    assert abs(np.sum(w2[..., None] * Hp)) < 1e-12 * np.sum(np.abs(Hp))


def test_oracle_rtr_hvp_is_linear(oracle):
    # H is a linear operator of the tangent (X, dp): H(2 v1 - v2) = 2 H v1 - H v2
    K = multiband_kernel(3, 4, 0.1, 0.3, 2)
    args = kernel_args_dict(K)
    rng = np.random.default_rng(5)
    F0 = oracle.build_fock_apply(args, np.zeros_like(K.h))[2]
    eps, Q = np.linalg.eigh(F0)
    p = np.clip(1.0 / (1 + np.exp(eps / K.T)), 0.05, 0.95)
    Ft = np.conj(np.swapaxes(Q, -1, -2)) @ F0 @ Q
    w = np.asarray(K.w2d, np.float64)
    nb = K.h.shape[-1]

    def tangent():
        X = rng.standard_normal(K.h.shape) + 1j * rng.standard_normal(K.h.shape)
        X = 0.5 * (X - np.conj(np.swapaxes(X, -1, -2)))
        i = np.arange(nb)
        X[..., i, i] = 0
        d = rng.standard_normal(p.shape)
        return X, d - np.sum(w[..., None] * d) / (np.sum(w) * nb)

    (X1, d1), (X2, d2) = tangent(), tangent()
    H1 = oracle._rtr_joint_hvp(args, X1, d1, Q, p, Ft)
    H2 = oracle._rtr_joint_hvp(args, X2, d2, Q, p, Ft)
    H12 = oracle._rtr_joint_hvp(args, 2.0 * X1 - X2, 2.0 * d1 - d2, Q, p, Ft)
    np.testing.assert_allclose(H12[0], 2.0 * H1[0] - H2[0], atol=1e-12)
    np.testing.assert_allclose(H12[1], 2.0 * H1[1] - H2[1], atol=1e-12)
    assert np.max(np.abs(H1[0])) > 1e-3 and np.max(np.abs(H1[1])) > 1e-3


# ---------------------------------------------------------- deflation bias ---

def test_oracle_deflation_bias_gradient_phi_and_diag(oracle):
    # tests/test_deflation.py:111-152
    rng = np.random.default_rng(0)
    nk, nb = 2, 2
    w2d = np.arange(nk * nk).reshape(nk, nk).astype(np.float64) + 1.0
    sigma, length = 0.7, 0.5
    P = rand_herm(rng, (nk, nk, nb, nb))
    targets = np.stack([rand_herm(rng, (nk, nk, nb, nb)) for _ in range(2)])

    def bias(Pin):
        return oracle._deflation_bias(Pin, targets, w2d, sigma, length)

    Phi, S_pen, diag = bias(P)
    dP = rand_herm(rng, (nk, nk, nb, nb))
    analytic = float(np.sum(w2d[..., None, None] * np.real(np.conj(S_pen) * dP)))
    fd = [(bias(P + e * dP)[0] - bias(P - e * dP)[0]) / (2 * e) for e in (1e-4, 1e-5)]
    assert abs(fd[1] - analytic) < 1e-4 * (abs(analytic) + 1.0)
    assert abs(fd[1] - analytic) < 0.2 * abs(fd[0] - analytic) + 1e-9

    inv2L2 = 1.0 / (2.0 * length * length)
    d2 = np.array([np.sum(w2d[..., None, None] * np.abs(P - t) ** 2) for t in targets])
    phi = sigma * np.exp(-d2 * inv2L2)
    np.testing.assert_allclose(Phi, np.sum(phi), rtol=1e-13)
    np.testing.assert_allclose(diag, 2.0 * np.sum(-phi * inv2L2), rtol=1e-13)


def test_oracle_deflation_bias_empty_is_zero(oracle):
    # tests/test_deflation.py:154-162
    nk, nb = 2, 2
    P = np.zeros((nk, nk, nb, nb), np.complex128)
    empty = np.empty((0, nk, nk, nb, nb), np.complex128)
    Phi, S_pen, diag = oracle._deflation_bias(P, empty, np.ones((nk, nk)), 1.0, 1.0)
    assert Phi == 0.0 and diag == 0.0 and np.max(np.abs(S_pen)) == 0.0


# ------------------------------------------------------------------ solver ---

def test_oracle_rtr_matches_cg_and_scf_unique_minimum(oracle):
    # tests/test_solver_rtr.py:142-153
    k = contact_kernel(0.3, nk=2)
    cg = solve_dm(oracle, k, 4.0, max_iter=400, tol_E=1e-10)
    nw = newton(oracle, k, 4.0, tol_grad=1e-7)
    scf = oracle.solve_scf(kernel_args_dict(k), np.zeros_like(k.h), 4.0, 600, 1e-9, 1e-8, 0.3,
                           0.0, None, [], "linear", 6, 2, 1.0, 0.0, True, True)
    assert nw["converged"]
    np.testing.assert_allclose(nw["energy"], cg["energy"], atol=1e-6, rtol=1e-7)
    np.testing.assert_allclose(nw["energy"], scf[2], atol=1e-6, rtol=1e-7)
    np.testing.assert_allclose(nw["density"], scf[0], atol=1e-5)


def test_oracle_rtr_stationary_and_conserves_particles(oracle):
    # tests/test_solver_rtr.py:155-166
    k = multiband_kernel(4, 4, 0.05, 0.3, 7)
    ne = float(4 * 4 * 4 * 0.5)
    r = newton(oracle, k, ne, tol_grad=1e-6)
    assert r["converged"]
    assert _comm_rms(r["fock"], r["density"], k.w2d) < 1e-4
    np.testing.assert_allclose(np.sum(k.w2d[..., None] * r["p"]), ne, atol=1e-4)
    np.testing.assert_allclose(r["density"], np.conj(np.swapaxes(r["density"], -1, -2)),
                               atol=1e-12)
    # the trust region accepts only energy decreases (up to float noise)
    assert np.all(np.diff(r["hist_E"]) <= 1e-12 * np.maximum(1.0, np.abs(r["hist_E"][1:])))


def test_oracle_rtr_superlinear_far_fewer_outer_iterations_than_cg(oracle):
    # tests/test_solver_rtr.py:168-182
    k = multiband_kernel(6, 6, 0.015, 0.4, 21)
    ne = float(6 * 6 * 6 * 0.5)
    cg = solve_dm(oracle, k, ne, max_iter=2000, tol_E=0.0, tol_grad=1e-6, plateau_window=0)
    nw = newton(oracle, k, ne, tol_grad=1e-6)
    assert nw["converged"]
    assert nw["n_iter"] < 40
    assert nw["n_iter"] * 4 < cg["n_iter"]
    np.testing.assert_allclose(nw["energy"], cg["energy"], atol=1e-4, rtol=1e-6)
    # superlinear tail: the last gradient drops far below the one before it
    g = nw["hist_grad"]
    assert g[-1] < 1e-6 and g[-1] < 0.1 * g[-2]


def test_oracle_rtr_two_band_stationary(oracle):
    # tests/test_solver_rtr.py:188-194
    k = two_band_kernel(nk=2, T=0.1)
    r = newton(oracle, k, 4.0, max_iter=100, tol_grad=1e-6)
    assert np.isfinite(r["energy"])
    assert _comm_rms(r["fock"], r["density"], k.w2d) < 1e-4


def test_oracle_rtr_projector_enforced(oracle):
    # tests/test_deflation.py:62-84
    k = hybridized_kernel(nk=2, ex=0.25)

    def proj_diag(P):
        P = np.array(P, copy=True)
        P[..., 0, 1] = 0.0
        P[..., 1, 0] = 0.0
        return P

    r = newton(oracle, k, 4.0, tol_grad=1e-6, project_fn=proj_diag)
    assert np.max(np.abs(r["density"][..., 0, 1])) < 1e-8
    r0 = newton(oracle, k, 4.0, tol_grad=1e-6)
    assert np.max(np.abs(r0["density"][..., 0, 1])) > 1e-3


def test_oracle_rtr_deflation_off_is_strict_noop(oracle):
    # tests/test_deflation.py:88-103
    k = hybridized_kernel(nk=2, ex=0.3)
    r1 = newton(oracle, k, 4.0, tol_grad=1e-7)
    r2 = newton(oracle, k, 4.0, tol_grad=1e-7, sigma=0.0)
    r3 = newton(oracle, k, 4.0, tol_grad=1e-7, targets=r1["density"][None], sigma=0.0)
    for r in (r2, r3):
        assert np.array_equal(r["density"], r1["density"])
        assert r["energy"] == r1["energy"]


def _oracle_solve(oracle):
    """cpp_hf_amd.solve's signature and SolveResult, computed by the oracle."""
    from cpp_hf_amd.solver import SolveResult

    def run(kernel, P0, n_electrons, *, config):
        code = {"cg": 0, "newton": 1}[config.optimizer]
        targets = config.deflation_targets
        r = solve_dm(oracle, kernel, n_electrons, P0=P0, max_iter=config.max_iter,
                     tol_E=config.tol_E, tol_grad=config.tol_grad,
                     plateau_window=config.plateau_window, optimizer=code,
                     targets=None if targets is None else np.asarray(targets),
                     sigma=config.deflation_sigma, length=config.deflation_length,
                     project_fn=config.project_fn, tr_delta0=config.tr_delta0,
                     tr_cg_max=config.tr_cg_max)
        return SolveResult(Q=r["Q"], p=r["p"], mu=np.asarray(r["mu"]), density=r["density"],
                           fock=r["fock"], energy=np.asarray(r["energy"]),
                           n_iter=np.asarray(r["n_iter"]), converged=np.asarray(r["converged"]),
                           history={"E": r["hist_E"], "grad_norm": r["hist_grad"]})

    return run


def test_solve_deflated_finds_distinct_second_minimum(oracle, monkeypatch):
    # tests/test_deflation.py:166-187, with the driver's solves computed by the
    # oracle: the host-side search (auto-scaled bias, escalation, re-polish,
    # distinctness, ordering) is what is under test here
    import cpp_hf_amd.deflation as dfl
    from cpp_hf_amd import SolverConfig

    monkeypatch.setattr(dfl, "solve", _oracle_solve(oracle))
    k = bistable_kernel()
    P0 = np.zeros_like(k.h)
    base = SolverConfig(optimizer="newton", tol_grad=1e-6, max_iter=300)
    res = dfl.solve_deflated(k, P0, 1.0, base_config=base, n_solutions=2, seed=0)
    assert res.n_found >= 2
    for s in res.solutions:
        assert _comm_rms(np.asarray(s.fock), np.asarray(s.density), k.w2d) < 1e-3
    d = np.asarray(res.solutions[0].density) - np.asarray(res.solutions[1].density)
    assert float(np.max(np.abs(d))) > 1e-2
    cold = dfl.solve(k, P0, 1.0, config=base)
    assert float(res.best.energy) <= float(cold.energy) + 1e-7
    np.testing.assert_allclose(float(res.best.energy), -0.6, atol=2e-2)
    np.testing.assert_allclose(sorted(res.energies)[:2], [-0.6, -0.4], atol=3e-2)
    assert list(res.energies) == sorted(res.energies)


def test_solve_deflated_stops_when_no_new_basin(oracle, monkeypatch):
    # a unique-minimum model: every biased attempt re-polishes back to the
    # first solution, so the search gives up after max_retries attempts
    import cpp_hf_amd.deflation as dfl
    from cpp_hf_amd import SolverConfig

    calls = []
    run = _oracle_solve(oracle)

    def counted(*a, **kw):
        calls.append(kw["config"].deflation_sigma)
        return run(*a, **kw)

    monkeypatch.setattr(dfl, "solve", counted)
    k = contact_kernel(0.3, nk=2)
    res = dfl.solve_deflated(k, np.zeros_like(k.h), 4.0, n_solutions=3, max_retries=2,
                             base_config=SolverConfig(tol_grad=1e-7))
    assert res.n_found == 1
    # first solve, then (biased, re-polish) per attempt with sigma growing x3
    assert len(calls) == 1 + 2 * 2
    assert calls[0] == 0.0 and calls[2] == 0.0 and calls[4] == 0.0
    assert calls[3] == pytest.approx(3.0 * calls[1])
