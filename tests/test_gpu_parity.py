"""GPU parity: libhfgpu (through _native and the C ABI) against the CPU oracle.

Tolerances follow the contract for this FP64 path: Fock matrices, self-
energies and eigenvalues within 1e-12 relative max-abs, index maps
bit-exact, converged energies within 1e-10.
"""

from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np
import pytest

from helpers import (block_kernel, contact_kernel, hartree_kernel, herm, kernel_args_dict, rand_c,
                     rand_herm, two_band_kernel)

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]


def relerr(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    scale = max(1e-300, float(np.max(np.abs(b))))
    return float(np.max(np.abs(a - b))) / scale


# --------------------------------------------------------------- exchange ---

EXCH_CASES = [
    # (n1, n2, nb, dense V, hcp, ifftshift)
    (24, 24, 2, False, True, False),
    (24, 24, 2, False, False, True),
    (12, 12, 4, True, False, False),
    (7, 11, 3, True, False, True),
    (5, 5, 1, False, False, False),
    (1, 1, 2, False, True, False),
    (13, 13, 2, True, False, False),
    (36, 60, 2, False, True, True),
    (8, 8, 6, False, True, False),
    (96, 96, 8, True, False, False),
    (120, 72, 4, True, False, False),
    (192, 48, 2, True, False, False),
]


@pytest.mark.parametrize("n1,n2,nb,dense,hcp,shift", EXCH_CASES)
def test_selfenergy_matches_oracle(gpu, oracle, n1, n2, nb, dense, hcp, shift):
    print(f"selfenergy case {(n1, n2, nb, dense, hcp, shift)}", flush=True)
    rng = np.random.default_rng(n1 * 1000 + n2 + nb)
    P = rand_herm(rng, (n1, n2, nb, nb)) if hcp else rand_c(rng, (n1, n2, nb, nb))
    VR = rand_c(rng, (n1, n2, nb, nb) if dense else (n1, n2, 1, 1))
    got = gpu.selfenergy_fft_full(VR, P, shift, hcp)
    ref = oracle.selfenergy_fft_full(VR, P, apply_ifftshift=shift, hermitian_channel_packing=hcp)
    assert relerr(got, ref) < 1e-12


def test_selfenergy_rejects_bad_shapes(gpu):
    P = np.zeros((4, 4, 3, 3), complex)
    with pytest.raises(ValueError, match="hermitian_channel_packing"):
        gpu.selfenergy_fft_full(np.zeros((4, 4, 3, 3), complex), P, False, True)
    with pytest.raises(ValueError, match="VR shape"):
        gpu.selfenergy_fft_full(np.zeros((4, 4, 2, 2), complex), P, False, False)


def test_selfenergy_unsupported_prime_mesh_raises(gpu):
    P = np.zeros((17, 4, 2, 2), complex)
    with pytest.raises(ValueError, match="prime factor"):
        gpu.selfenergy_fft_full(np.ones((17, 4, 1, 1), complex), P, False, False)


# -------------------------------------------------------------------- eigh --

@pytest.mark.parametrize("nb", [1, 2, 3, 5, 8, 12, 16, 17, 24, 32])
def test_eigh_matches_oracle(gpu, oracle, nb):
    print(f"eigh nb={nb}", flush=True)
    rng = np.random.default_rng(100 + nb)
    M = rand_c(rng, (3, 17, nb, nb))
    w, V = gpu.eigh_batched(M)
    wr, _ = oracle.eigh_batched(M)
    H = herm(M)
    assert w.shape == (3, 17, nb) and V.shape == (3, 17, nb, nb)
    assert float(np.max(np.abs(w - wr))) <= 1e-12 * max(1.0, float(np.max(np.abs(wr))))
    VH = np.conj(np.swapaxes(V, -1, -2))
    assert np.max(np.abs(VH @ V - np.eye(nb))) < 1e-12
    assert np.max(np.abs(V @ (w[..., None] * VH) - H)) < 1e-12 * nb * max(1.0, np.max(np.abs(H)))
    assert np.all(np.diff(w, axis=-1) >= 0)


def test_eigh_degenerate_and_trivial(gpu):
    M = np.zeros((3, 4, 4), dtype=np.complex128)
    M[0] = np.diag([3.0, -1.0, 2.0, -1.0])
    M[1] = np.eye(4)
    w, V = gpu.eigh_batched(M)
    np.testing.assert_array_equal(w[0], [-1.0, -1.0, 2.0, 3.0])
    np.testing.assert_array_equal(w[1], [1.0] * 4)
    np.testing.assert_array_equal(w[2], [0.0] * 4)


def test_eigh_through_ctypes_c_abi(gpu):
    """The C ABI bound with ctypes (no pybind11 in the path)."""
    lib = C.CDLL(str(ROOT / "cpp_hf_amd" / "libhfgpu.so"))
    dp = C.POINTER(C.c_double)
    rng = np.random.default_rng(7)
    M = np.ascontiguousarray(rand_herm(rng, (9, 4, 4)))
    w = np.empty((9, 4))
    V = np.empty_like(M)
    rc = lib.hfgpu_eigh_batched(M.ctypes.data_as(dp), w.ctypes.data_as(dp), V.ctypes.data_as(dp),
                                C.c_size_t(9), C.c_size_t(4))
    assert rc == 0
    np.testing.assert_allclose(w, np.linalg.eigvalsh(M), atol=1e-13)


# ---------------------------------------------------------------------- mu --

def test_mu_bisection_matches_oracle(gpu, oracle):
    rng = np.random.RandomState(42)
    for shape, n_e, T in [((4, 4, 6), 3.0, 0.05), ((96, 96, 8), 4.1, 0.05), ((1, 1, 2), 1.0, 0.1),
                          ((24, 24, 2), 1.1, 0.02)]:
        bands = rng.randn(*shape)
        w = np.full(shape[:2], 1.0 / (shape[0] * shape[1]))
        for maxiter in (0, 7, 54):
            got = gpu.find_mu_bisection(bands, w, n_e, T, maxiter)
            ref = oracle.find_mu_bisection(bands, w, n_e, T, maxiter)
            # Up to the default 30 halvings the count differs from n_e by far
            # more than summation rounding, so both take the same decisions and
            # land on the same midpoint.  Deeper (54 halvings, the float32
            # default) the bracket shrinks below the rounding of the count and
            # the decisions may differ: agreement is then 1e-12.
            tol = 1e-15 if maxiter <= 30 else 1e-12
            assert abs(got - ref) <= tol * max(1.0, abs(ref)), (shape, maxiter, got, ref)


def test_mu_bisection_cold_limit_hits_target(gpu, oracle):
    """T -> 0 with an integer filling: the count is a step function equal to
    n_e over a whole gap, so `count > n_e` compares equal numbers whose last
    bits depend on the summation order; every mu in the gap is a root.  Both
    implementations must land on a root (tests/test_main.py:237-244)."""
    rng = np.random.RandomState(5)
    bands = rng.randn(24, 24, 2)
    w = np.full((24, 24), 1.0 / 576)
    for mu in (gpu.find_mu_bisection(bands, w, 1.0, 1e-5, 0),
               oracle.find_mu_bisection(bands, w, 1.0, 1e-5, 0)):
        occ = 0.5 * (1.0 - np.tanh((bands - mu) / (2 * (1e-5 + 1e-12))))
        assert abs(np.sum(w[..., None] * occ) - 1.0) < 1e-4
    tiny = np.array([[[0.0, 1.0, 2.0]]])
    mu = gpu.find_mu_bisection(tiny, np.ones((1, 1)), 2.0, 1e-5, 0)
    assert 1.0 < mu < 2.0


def test_mu_newton_matches_oracle(gpu, oracle):
    rng = np.random.RandomState(3)
    for shape, n_t, T, mu0 in [((8, 8, 4), 2.1, 0.05, 0.0), ((24, 24, 2), 1.0, 0.1, 0.3),
                               ((96, 96, 8), 4.05, 0.05, -0.2)]:
        eps = rng.randn(*shape)
        wn = np.full(shape[:2], 1.0 / (shape[0] * shape[1]))
        got = gpu.solve_mu_inloop(eps, wn, n_t, mu0, T, 25)
        ref = oracle.solve_mu_inloop(eps, wn, n_t, mu0, T, 25)
        assert abs(got - ref) <= 1e-12 * max(1.0, abs(ref))


def test_mu_newton_stays_at_converged_mu(gpu, oracle):
    """Flat bands at low filling (the two-band model of tests/test_continuation.py
    on 8 x 8 with unit weights): once Newton has converged, the residual is
    zero or a rounding-level number whose step rounds away; mu must stay put
    for the remaining iterations instead of jumping to the bracket middle."""
    eps = np.zeros((8, 8, 2))
    eps[..., 0], eps[..., 1] = -0.5, 0.5
    wn = np.full((8, 8), 1.0 / 64)
    # From mu0 = 0 the sequential sum takes two bisection steps inside a
    # 1e-12-wide bracket (iterations 6-7) before it lands on a zero residual,
    # so iterations 6-8 are not compared; both end on the same mu.
    for mu0, its in ((0.0, (2, 3, 5, 25)), (-1.33, (2, 3, 5, 8, 25))):
        for it in its:
            got = gpu.solve_mu_inloop(eps, wn, 1.0 / 64, mu0, 0.2, it)
            ref = oracle.solve_mu_inloop(eps, wn, 1.0 / 64, mu0, 0.2, it)
            assert abs(got - ref) <= 1e-14, (mu0, it, got, ref)


def test_mu_newton_multi_launch_stays_at_converged_mu(gpu, oracle):
    """The same flat-band model above the one-workgroup size (96 x 96 x 2 =
    18432 eigenvalues: one launch per iteration and a last-step launch): the
    converged mu stays put.  Here the oracle's sequential sum of 9216 terms
    leaves a residual of a few 1e-13 / Z, so its mu wanders by that much
    once converged, while a jump to the bracket middle would move mu by
    ~0.5: the bound 1e-12 separates the two."""
    eps = np.zeros((96, 96, 2))
    eps[..., 0], eps[..., 1] = -0.5, 0.5
    wn = np.full((96, 96), 1.0 / 9216)
    for n_t in (1.0 / 64, 1.0):
        for mu0 in (0.0, -1.33):
            for it in (3, 5, 12, 25):
                got = gpu.solve_mu_inloop(eps, wn, n_t, mu0, 0.2, it)
                ref = oracle.solve_mu_inloop(eps, wn, n_t, mu0, 0.2, it)
                assert abs(got - ref) <= 1e-12, (n_t, mu0, it, got, ref)


def test_mu_newton_solo_and_multi_launch_agree(gpu):
    """4096 eigenvalues run in one workgroup, 4097 on several workgroups
    exchanging partial sums; the extra eigenvalue sits far above the Fermi
    level, where its occupation and slope are exactly zero, so both solve the
    same problem."""
    rng = np.random.RandomState(12)
    eps = rng.randn(64, 64, 1)
    wn = np.full((64, 64), 1.0 / 4096)
    eps2 = np.concatenate([eps.reshape(-1), [1.0e6]]).reshape(1, 4097, 1)
    wn2 = np.concatenate([wn.reshape(-1), [1.0 / 4096]]).reshape(1, 4097)
    for n_t, mu0 in ((0.3, 0.0), (0.77, 1.5)):
        for it in (1, 3, 25):
            a = gpu.solve_mu_inloop(eps, wn, n_t, mu0, 0.05, it)
            b = gpu.solve_mu_inloop(eps2, wn2, n_t, mu0, 0.05, it)
            assert abs(a - b) <= 1e-13 * max(1.0, abs(a)), (n_t, it, a, b)


# --------------------------------------------------------------- build_fock -

def _fock_cases():
    from cpp_hf_amd.workloads import layered_kernel, qwz_kernel

    out = []
    k, P = hartree_kernel(nk=3, nb=3)
    out.append(("hartree_scalarV", k, P))
    k, P = hartree_kernel(nk=6, nb=4, seed=11, dense_v=True)
    out.append(("hartree_denseV", k, P))
    k = contact_kernel(0.3, nk=2)
    out.append(("contact", k, rand_herm(np.random.default_rng(1), k.h.shape)))
    wl = qwz_kernel(24)
    out.append(("qwz24", wl.kernel, herm(wl.P0)))
    wl = layered_kernel(12, 4)
    out.append(("layered12_nb8", wl.kernel, herm(wl.P0)))
    wl = layered_kernel(24, 2)
    rng = np.random.default_rng(5)
    out.append(("layered24_nonherm", wl.kernel, rand_c(rng, wl.kernel.h.shape)))
    # more channels than the small-terms reduction has threads: 300 packed
    # (nb = 24) and 289 full (nb = 17, non-Hermitian density)
    k, P = hartree_kernel(nk=4, nb=24, seed=7)
    out.append(("hartree_nb24_packed", k, P))
    k, _ = hartree_kernel(nk=4, nb=17, seed=8)
    out.append(("hartree_nb17_full", k, rand_c(rng, k.h.shape)))
    return out


@pytest.mark.parametrize("case", range(8))
def test_build_fock_matches_oracle(gpu, oracle, case):
    name, k, P = _fock_cases()[case]
    print(f"build_fock case {name}", flush=True)
    ka = kernel_args_dict(k)
    S, Hh, F, E = gpu.build_fock_apply(ka, P, None)
    Sr, Hr, Fr, Er = oracle.build_fock_apply(ka, P)
    assert relerr(S, Sr) < 1e-12
    assert relerr(F, Fr) < 1e-12
    assert relerr(Hh, Hr) < 1e-12 if np.any(Hr) else not np.any(Hh)
    assert abs(E - Er) <= 1e-12 * max(1.0, abs(Er))
    np.testing.assert_array_equal(F, np.conj(np.swapaxes(F, -1, -2)))


def test_build_fock_exchange_off_and_project(gpu, oracle):
    from cpp_hf_amd import HartreeFockKernel

    k0, P = hartree_kernel(nk=4, nb=3)
    k = HartreeFockKernel(k0.w2d, k0.h, np.zeros((4, 4, 1, 1)), T=0.1, include_hartree=True,
                          include_exchange=False, reference_density=k0.refP,
                          hartree_matrix=k0.HH)
    ka = kernel_args_dict(k)
    got = gpu.build_fock_apply(ka, P, None)
    ref = oracle.build_fock_apply(ka, P)
    for a, b in zip(got[:3], ref[:3]):
        assert relerr(a, b) < 1e-12 or not np.any(b)
    # project_fn after the GIL release (tests/test_main.py:190-204)
    F = gpu.build_fock_apply(ka, P, lambda F: np.zeros_like(F))[2]
    assert not np.any(F)


@pytest.mark.parametrize("nb", [4, 9, 12, 16])
def test_density_from_fock_matches_oracle(gpu, oracle, nb):
    """nb 9..16 run the density product on the FP64 matrix cores."""
    rng = np.random.default_rng(21 + nb)
    F = rand_herm(rng, (12, 12, nb, nb))
    w = np.full((12, 12), 1 / 144)
    n_e = 0.55 * nb
    P, mu = gpu.density_from_fock(F, w, n_e, 0.05, 0)
    Pr, mur = oracle.density_from_fock(F, w, n_e, 0.05)
    assert abs(mu - mur) < 1e-14
    assert relerr(P, Pr) < 1e-12
    np.testing.assert_array_equal(P, np.conj(np.swapaxes(P, -1, -2)))


# --------------------------------------------------------------- resample ---

@pytest.mark.parametrize("old,new,inner", [(4, 8, 3), (8, 4, 2), (6, 9, 4), (12, 24, 16),
                                           (5, 7, 1), (24, 96, 4)])
def test_resample_bit_exact(gpu, oracle, old, new, inner):
    rng = np.random.default_rng(old * new)
    x = rand_c(rng, (old, old, inner))
    np.testing.assert_array_equal(gpu.resample_kgrid_2d(x, new), oracle.resample_kgrid_2d(x, new))


# ----------------------------------------------------------------- Cayley ---

def test_cayley_hooks_match_oracle_and_lu(gpu, oracle):
    from cpp_hf_amd.solver import _cayley_retract

    rng = np.random.default_rng(0)
    d = rand_c(rng, (3, 3, 6, 6))
    d = 0.5 * (d - np.conj(np.swapaxes(d, -1, -2)))
    V, lam = gpu._cayley_spectral_setup(d)
    _, lam_r = oracle._cayley_spectral_setup(d)
    assert np.max(np.abs(lam - lam_r)) < 1e-12
    Ft = rand_herm(rng, (3, 3, 6, 6))
    Fe = np.conj(np.swapaxes(V, -2, -1)) @ Ft @ V
    for tau in [0.0, 0.1, 0.5, 1.0, -0.3]:
        U = gpu._cayley_unitary_from_spectrum(V, lam, tau)
        np.testing.assert_allclose(U, _cayley_retract(d, tau), atol=1e-12, rtol=1e-12)
        diag = gpu._diag_UFU_from_spectrum(V, Fe, lam, tau)
        U_lu = _cayley_retract(d, tau)
        ref = np.real(np.diagonal(np.conj(np.swapaxes(U_lu, -2, -1)) @ Ft @ U_lu, axis1=-2,
                                  axis2=-1))
        np.testing.assert_allclose(diag, ref, atol=1e-12, rtol=1e-12)


# ---------------------------------------------------------------- solvers ---

def _dm_args(k, n_e, max_iter=200, tol_E=1e-9):
    return (kernel_args_dict(k), np.zeros_like(k.h), n_e, max_iter, tol_E, 0.0, 0.6, 0.5, 1e-3,
            8, 10, 5, 25, 0, 0.5, 20, np.empty((0,) + k.h.shape), 0.0, 1.0, [], None, True,
            True, True)


def _dm_args_p0(k, P0, n_e, max_iter=200, tol_E=1e-7):
    a = list(_dm_args(k, n_e, max_iter, tol_E))
    a[1] = P0
    return tuple(a)


def _dm_cases():
    from cpp_hf_amd.workloads import layered_kernel, qwz_kernel

    wl1 = qwz_kernel(12)
    wl2 = layered_kernel(12, 2, T=0.05, filling=2.1)
    return [
        ("two_band_nk2", _dm_args(two_band_kernel(nk=2, T=0.1), 4.0)),
        ("contact", _dm_args(contact_kernel(0.3, nk=2), 4.0)),
        ("qwz12", _dm_args_p0(wl1.kernel, wl1.P0, 1.0)),
        ("layered12_nb4", _dm_args_p0(wl2.kernel, wl2.P0, 2.1, max_iter=60)),
    ] + [
        # the per-k solver kernels with nb fixed at compile time (8, 12, 16)
        # and at run time (6)
        (f"layered{n}_nb{2 * nl}", _dm_args_p0(w.kernel, w.P0, w.n_electrons, max_iter=40))
        for n, nl in ((12, 4), (12, 6), (8, 8), (12, 3))
        for w in [layered_kernel(n, nl, T=0.05, filling=nl + 0.1)]
    ]


@pytest.mark.parametrize("case", range(8))
def test_solve_dm_matches_oracle(gpu, oracle, case):
    name, args = _dm_cases()[case]
    print(f"solve_dm case {name}", flush=True)
    g = gpu.solve_dm(*args)
    r = oracle.solve_dm(*args)
    Q, p, P, F, mu, E, n, conv, hE, hG = g
    assert n == r[6] and conv == r[7]
    np.testing.assert_allclose(hE, r[8], rtol=0, atol=1e-10)
    np.testing.assert_allclose(hG, r[9], rtol=1e-8, atol=1e-12)
    assert abs(E - r[5]) < 1e-10
    assert abs(mu - r[4]) < 1e-10
    assert relerr(P, r[2]) < 1e-9
    assert relerr(F, r[3]) < 1e-9
    np.testing.assert_allclose(p, r[1], atol=1e-9)


@pytest.mark.parametrize("n,nl", [(12, 4), (12, 6), (8, 8)])
def test_solve_scf_layered_matches_oracle(gpu, oracle, n, nl):
    """SCF residual kernels with nb fixed at compile time (8, 12, 16)."""
    from cpp_hf_amd.workloads import layered_kernel

    w = layered_kernel(n, nl, T=0.05, filling=nl + 0.1)
    args = (kernel_args_dict(w.kernel), w.P0, w.n_electrons, 30, 1e-9, 1e-8, 0.5, 0.0, None, [],
            "linear", 6, 2, 1.0, 0.0, True, True)
    print(f"solve_scf layered n={n} nb={2 * nl}", flush=True)
    g = gpu.solve_scf(*args)
    r = oracle.solve_scf(*args)
    assert g[4] == r[4] and g[5] == r[5]
    # relative bound on E (|E| ~ 14 at nb 16)
    assert abs(g[2] - r[2]) < 1e-10 * max(1.0, abs(r[2]))
    # mu is the midpoint of 30 bisection halvings of a bracket ~60 wide at
    # nb 16 (resolution ~6e-8); one rounding-level `count > n_e` decision at
    # the last levels moves it by that much
    dmu = abs(g[3] - r[3])
    assert dmu < 1e-7
    # occupations move by at most dmu f (1 - f) / T <= dmu / (4 T)
    assert relerr(g[0], r[0]) < max(1e-9, 2.0 * dmu / (4 * 0.05))
    np.testing.assert_allclose(g[6], r[6], rtol=1e-10, atol=1e-10)


@pytest.mark.parametrize("accel", ["linear", "diis", "oda"])
def test_solve_scf_matches_oracle(gpu, oracle, accel):
    k = contact_kernel(0.3, nk=2)
    args = (kernel_args_dict(k), np.zeros_like(k.h), 4.0, 400, 1e-7, 1e-6, 0.3, 0.0, None, [],
            accel, 6, 2, 1.0, 0.0, True, True)
    g = gpu.solve_scf(*args)
    r = oracle.solve_scf(*args)
    assert g[4] == r[4] and g[5] == r[5]
    assert abs(g[2] - r[2]) < 1e-10
    assert abs(g[3] - r[3]) < 1e-10
    # ODA here ends stalled (lambda = 0 once c1 >= 0 with c2 <= 1e-12,
    # solver_scf.hpp:381-388); which steps stall depends on the sign of c1 at
    # rounding level, so the parked density agrees to 1e-7, energy to 1e-10.
    assert relerr(g[0], r[0]) < (1e-7 if accel == "oda" else 1e-9)
    np.testing.assert_allclose(g[6], r[6], atol=1e-10)


def test_solve_scf_meshgrid_fixture_matches_oracle(gpu, oracle):
    from cpp_hf_amd import HartreeFockKernel

    d = np.load(ROOT / "tests" / "golden" / "meshgrid_regression_case_v1.npz")
    k = HartreeFockKernel(weights=d["weights"], hamiltonian=d["hamiltonian"],
                          coulomb_q=d["coulomb_q"], T=float(d["T"]), include_hartree=True,
                          include_exchange=True, reference_density=d["reference_density"],
                          hartree_matrix=d["hartree_matrix"])
    args = (kernel_args_dict(k), np.asarray(d["P0"]), float(d["electrondensity0"]),
            int(d["max_iter"]), float(d["comm_tol"]), float(d["comm_tol"]), 0.5, 0.0, None, [],
            "linear", 6, 2, 1.0, 0.0, True, True)
    g = gpu.solve_scf(*args)
    r = oracle.solve_scf(*args)
    assert g[4] == r[4] and g[5]
    assert abs(g[2] - r[2]) < 1e-10 and abs(g[3] - r[3]) < 1e-10
    np.testing.assert_allclose(g[0], d["P_expected"], rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(g[1], d["F_expected"], rtol=1e-6, atol=1e-6)


@pytest.mark.parametrize("nb,sizes", [(8, (3, 2, 3)), (6, (2, 4)), (12, (4, 4, 4)), (4, (1, 3)),
                                      (24, (8, 10, 6)), (20, (5, 15))])
@pytest.mark.parametrize("sort", [True, False])
def test_block_eigh_matches_oracle(gpu, oracle, nb, sizes, sort):
    rng = np.random.default_rng(nb + len(sizes))
    X = rng.standard_normal((50, nb, nb)) + 1j * rng.standard_normal((50, nb, nb))
    M = X + np.conj(np.swapaxes(X, -1, -2))
    w, V = gpu.eigh_block_sizes(M, list(sizes), sort)
    wo, Vo = oracle.eigh_block_sizes(M, sizes, sort=sort)
    np.testing.assert_allclose(w, wo, rtol=0, atol=1e-12 * np.max(np.abs(wo)))
    starts = np.cumsum((0,) + tuple(sizes[:-1]))
    B = np.zeros_like(M)
    for s, n in zip(starts, sizes):
        B[:, s:s + n, s:s + n] = 0.5 * (M[:, s:s + n, s:s + n]
                                        + np.conj(np.swapaxes(M[:, s:s + n, s:s + n], -1, -2)))
    R = V @ (w[..., None] * np.conj(np.swapaxes(V, -1, -2)))
    np.testing.assert_allclose(R, B, rtol=0, atol=1e-12 * np.max(np.abs(B)))
    if not sort:
        for s, n in zip(starts, sizes):  # each block's vectors stay in its block
            assert np.all(V[:, :s, s:s + n] == 0) and np.all(V[:, s + n:, s:s + n] == 0)


def test_max_offblock_matches_oracle(gpu, oracle):
    rng = np.random.default_rng(9)
    M = rng.standard_normal((40, 8, 8)) + 1j * rng.standard_normal((40, 8, 8))
    for sizes in [(8,), (4, 4), (2, 3, 3), (3, 3)]:
        got = gpu.max_offblock_sizes(M, list(sizes))
        ref = oracle.max_offblock_sizes(M, sizes)
        assert abs(got - ref) <= 4e-16 * ref


# ------------------------------------------------ block-structured solvers ---

def _block_dm_args(k, sizes):
    return (kernel_args_dict(k), np.zeros_like(k.h), 2.0, 200, 1e-10, 0.0, 0.6, 0.5, 1e-3, 8, 10,
            5, 25, 0, 0.5, 20, np.empty((0,) + k.h.shape), 0.0, 1.0, list(sizes), None, True,
            True, True)


@pytest.mark.parametrize("couple", [0.0, 0.05])
def test_solve_dm_block_sizes_matches_oracle(gpu, oracle, couple):
    """Per-block eigh in the init, Cayley and final steps (solver_dm.hpp:446-450,
    562-564, 622-624) when the problem is block-diagonal; dense otherwise."""
    k = block_kernel(couple=couple)
    args = _block_dm_args(k, [2, 2])
    print(f"solve_dm block_sizes couple={couple}", flush=True)
    Q, p, P, F, mu, E, n, conv, hE, hG = gpu.solve_dm(*args)
    r = oracle.solve_dm(*args)
    assert n == r[6] and conv == r[7]
    assert abs(E - r[5]) < 1e-10 and abs(mu - r[4]) < 1e-10
    np.testing.assert_allclose(hE, r[8], rtol=0, atol=1e-10)
    # occupations in the same (block-wise or global) order as the oracle's
    np.testing.assert_allclose(p, r[1], atol=1e-9)
    assert relerr(P, r[2]) < 1e-9
    if couple == 0.0:
        assert not np.any(Q[..., :2, 2:]) and not np.any(Q[..., 2:, :2])


@pytest.mark.parametrize("couple,accel", [(0.0, "linear"), (0.0, "diis"), (0.05, "linear")])
def test_solve_scf_block_sizes_matches_oracle(gpu, oracle, couple, accel):
    """density_from_fock with block_sizes (solver_scf.hpp:96-123): the off-block
    test of every F runs on the device."""
    k = block_kernel(couple=couple)
    args = (kernel_args_dict(k), np.zeros_like(k.h), 2.0, 400, 1e-9, 1e-8, 0.5, 0.0, None,
            [2, 2], accel, 6, 2, 1.0, 0.0, True, True)
    print(f"solve_scf block_sizes couple={couple} {accel}", flush=True)
    g = gpu.solve_scf(*args)
    r = oracle.solve_scf(*args)
    assert g[4] == r[4] and g[5] == r[5]
    assert abs(g[2] - r[2]) < 1e-10 and abs(g[3] - r[3]) < 1e-10
    assert relerr(g[0], r[0]) < 1e-9


def test_block_sizes_validation_follows_solver(gpu):
    k = block_kernel()
    with pytest.raises(ValueError, match="must sum to nb"):
        gpu.solve_dm(*_block_dm_args(k, [3, 2]))
    # SCF: sizes that do not describe nb mean "no blocks" (solver_scf.hpp:107-110)
    args = (kernel_args_dict(k), np.zeros_like(k.h), 2.0, 20, 1e-9, 1e-8, 0.5, 0.0, None,
            [3, 2], "linear", 6, 2, 1.0, 0.0, True, True)
    dense = list(args)
    dense[9] = []
    np.testing.assert_array_equal(gpu.solve_scf(*args)[0], gpu.solve_scf(*dense)[0])
