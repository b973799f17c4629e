"""GPU tests of the public Python surface, in the manner of the reference's own
tests (tests/test_main.py, tests/test_linalg.py, tests/test_solver.py).

Each public function runs on the GPU and is compared with the oracle or with
an independent library routine (LAPACK eigh, scipy's entr).  The reference's
tests hold float32 inputs at 1e-6; everything here is FP64, so the bounds are
the path's own: 1e-12 relative for Fock matrices, self-energies and
eigenvalues, 1e-12 absolute for chemical potentials and entropies.
"""

from __future__ import annotations

import threading

import numpy as np
import pytest

from helpers import contact_kernel, hartree_kernel, herm, kernel_args_dict, rand_c, rand_herm

pytestmark = pytest.mark.gpu


def relerr(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.max(np.abs(a - b))) / max(1e-300, float(np.max(np.abs(b))))


# ------------------------------------------------- find_chemical_potential --

def test_find_chemical_potential_hits_target_density(gpu):
    """tests/test_main.py:18-24."""
    from cpp_hf_amd.utils import fermidirac, find_chemical_potential

    bands = np.array([[[0.0, 1.0]]], dtype=np.float32)
    weights = np.ones((1, 1), dtype=np.float32)
    mu = find_chemical_potential(bands, weights, n_electrons=1.0, T=0.1)
    assert mu.dtype == np.float32
    assert abs(float(np.sum(weights[..., None] * fermidirac(bands - mu, 0.1))) - 1.0) < 1e-4


def test_find_chemical_potential_rejects_unreachable_density(gpu):
    """tests/test_main.py:27-33."""
    from cpp_hf_amd.utils import density_matrix_from_fock, find_chemical_potential

    with pytest.raises(ValueError, match="physically reachable range"):
        find_chemical_potential(np.array([[[0.0]]]), np.ones((1, 1)), n_electrons=2.0, T=0.1)
    with pytest.raises(ValueError, match="physically reachable range"):
        density_matrix_from_fock(np.zeros((1, 1, 1, 1), complex), np.ones((1, 1)),
                                 n_electrons=2.0, T=0.1)


@pytest.mark.parametrize("method", ["bisection", "newton"])
def test_find_chemical_potential_multiband_matches_oracle(gpu, oracle, method):
    """tests/test_main.py:197-209, in float64 against the oracle's restatement
    of the same algorithm (utils.py:277-360)."""
    from cpp_hf_amd.utils import fermidirac, find_chemical_potential

    rng = np.random.RandomState(42)
    for shape, n_t, T in [((4, 4, 6), 3.0, 0.05), ((24, 24, 2), 1.1, 0.02),
                          ((96, 96, 8), 4.1, 0.05), ((3, 5, 7), 0.2, 0.3)]:
        bands = rng.randn(*shape)
        w = np.ones(shape[:2]) / (shape[0] * shape[1])
        mu = float(find_chemical_potential(bands, w, n_t, T, method=method))
        if method == "bisection":
            ref = oracle.find_mu_bisection(bands, w, n_t, T, 30)
            assert abs(mu - ref) <= 1e-15 * max(1.0, abs(ref)), (shape, mu, ref)
            tol = 1e-6
        else:
            ref = oracle.find_mu_newton(bands, w, n_t, T, 15)
            assert abs(mu - ref) <= 1e-12 * max(1.0, abs(ref)), (shape, mu, ref)
            tol = 1e-10
        total = float(np.sum(w[..., None] * fermidirac(bands - mu, T)))
        assert abs(total - n_t) < tol


def test_find_mu_newton_unconverged_and_float32(gpu, oracle):
    """Too few steps: the final count check (utils.py:357-359) returns the
    bracket midpoint; float32 bands default to 25 steps (utils.py:323-324)."""
    from cpp_hf_amd.utils import find_chemical_potential

    rng = np.random.RandomState(7)
    bands = rng.randn(8, 8, 4)
    w = np.ones((8, 8)) / 64
    for it in (0, 1, 2, 3):
        got = float(find_chemical_potential(bands, w, 1.3, 0.05, method="newton", maxiter=it))
        ref = oracle.find_mu_newton(bands, w, 1.3, 0.05, it)
        assert abs(got - ref) <= 1e-13 * max(1.0, abs(ref)), (it, got, ref)
    b32 = bands.astype(np.float32)
    got = find_chemical_potential(b32, w.astype(np.float32), 1.3, 0.05, method="newton")
    assert got.dtype == np.float32
    ref = oracle.find_mu_newton(b32.astype(np.float64), w.astype(np.float32).astype(np.float64),
                                1.3, 0.05, 25)
    assert abs(float(got) - np.float32(ref)) == 0.0


# ------------------------------------------------ entropy and free energy ---

def test_occupation_entropy_and_free_energy(gpu, oracle):
    """fock.py:110-124 / fock.hpp:385-407 against the oracle and closed forms."""
    from scipy.special import entr

    from cpp_hf_amd import free_energy, occupation_entropy

    rng = np.random.default_rng(11)
    for shape in [(5, 7, 6), (96, 96, 8), (1, 1, 1)]:
        w = rng.uniform(0.1, 1.0, shape[:2])
        p = rng.uniform(0.0, 1.0, shape)
        p.flat[0] = 0.0
        p.flat[-1] = 1.0
        S = float(occupation_entropy(p, w))
        assert abs(S - oracle.occupation_entropy(p, w)) <= 1e-12 * max(1.0, abs(S))
        c = 1e-14
        pc = np.clip(p, c, 1 - c)
        ref = float(np.sum(w[..., None] * (entr(pc) + entr(1.0 - pc))))
        assert abs(S - ref) <= 1e-12 * max(1.0, abs(ref))
    # p = 1/2: ln 2 per band
    w = np.full((6, 6), 1.0 / 36)
    S = float(occupation_entropy(np.full((6, 6, 3), 0.5), w))
    assert abs(S - 3 * np.log(2.0)) < 1e-14
    # Omega = E - max(T, 1e-14) S
    p = rng.uniform(0.0, 1.0, (6, 6, 3))
    S = float(occupation_entropy(p, w))
    assert float(free_energy(-1.25, p, w, 0.05)) == pytest.approx(-1.25 - 0.05 * S, abs=1e-14)
    assert float(free_energy(-1.25, p, w, 0.0)) == pytest.approx(-1.25 - 1e-14 * S, abs=1e-15)
    assert float(free_energy(-1.25, p, w, -3.0)) == pytest.approx(-1.25 - 1e-14 * S, abs=1e-15)


# ------------------------------------------------- build_fock / hf_energy ---

def _public_build(k, P, **kw):
    from cpp_hf_amd import build_fock

    return build_fock(P, h=k.h, VR=k._VR_shifted, refP=k.refP, HH=k.HH, w2d=k.w2d,
                      include_exchange=k.include_exchange, include_hartree=k.include_hartree,
                      exchange_hermitian_channel_packing=k.exchange_hermitian_channel_packing,
                      contact_g=k.contact_g, contact_Oi=k.contact_Oi, contact_Oj=k.contact_Oj, **kw)


@pytest.mark.parametrize("case", ["hartree_scalarV", "hartree_denseV", "contact", "layered12"])
def test_public_build_fock_and_energy_match_oracle(gpu, oracle, case):
    """tests/test_main.py:51-109 through cpp_hf_amd.build_fock / hf_energy."""
    from cpp_hf_amd import hf_energy
    from cpp_hf_amd.workloads import layered_kernel

    if case == "hartree_scalarV":
        k, P = hartree_kernel(nk=3, nb=3)
    elif case == "hartree_denseV":
        k, P = hartree_kernel(nk=6, nb=4, seed=11, dense_v=True)
    elif case == "contact":
        k = contact_kernel(0.3, nk=2)
        P = rand_herm(np.random.default_rng(1), k.h.shape)
    else:
        wl = layered_kernel(12, 4)
        k, P = wl.kernel, herm(wl.P0)
    Sigma, H, F = _public_build(k, P)
    Sr, Hr, Fr, Er = oracle.build_fock_apply(kernel_args_dict(k), P)
    nb = k.h.shape[-1]
    assert relerr(Sigma, Sr) < 1e-12
    assert relerr(F, Fr) < 1e-12
    if k.include_hartree:
        # fock.py:68-69: the Hartree term comes back as a (1, 1, nb, nb) matrix
        assert H.shape == (1, 1, nb, nb)
        assert relerr(np.broadcast_to(H, Hr.shape), Hr) < 1e-12
    else:
        # fock.py:71
        assert H.shape == k.h.shape and not np.any(H)
    E = hf_energy(P, h=k.h, Sigma=Sigma, H=H, weights_b=k.weights_b, refP=k.refP)
    assert abs(float(E) - Er) <= 1e-12 * max(1.0, abs(Er))


def test_public_energy_has_fock_derivative(gpu):
    """tests/test_main.py:112-177: dE/dP = w F by central differences."""
    from cpp_hf_amd import HartreeFockKernel, hf_energy

    h = np.array([[-0.3, 0.04 - 0.02j], [0.04 + 0.02j, 0.5]]).reshape(1, 1, 2, 2)
    P = np.array([[0.8, 0.07 + 0.03j], [0.07 - 0.03j, 0.2]]).reshape(1, 1, 2, 2)
    refP = np.array([[0.35, -0.02 + 0.01j], [-0.02 - 0.01j, 0.15]]).reshape(1, 1, 2, 2)
    dP = np.array([[0.11, 0.05 - 0.04j], [0.05 + 0.04j, -0.07]]).reshape(1, 1, 2, 2)
    HH = np.array([[1.2, 0.1], [0.1, 0.8]])
    k = HartreeFockKernel(np.ones((1, 1)), h, np.full((1, 1, 1, 1), 0.7), T=0.1,
                          include_hartree=True, include_exchange=True, reference_density=refP,
                          hartree_matrix=HH)

    def energy_and_fock(Pe):
        S, H, F = _public_build(k, Pe)
        return float(hf_energy(Pe, h=k.h, Sigma=S, H=H, weights_b=k.weights_b, refP=k.refP)), F

    eps = 1e-6
    fd = (energy_and_fock(P + eps * dP)[0] - energy_and_fock(P - eps * dP)[0]) / (2 * eps)
    F = energy_and_fock(P)[1]
    an = float(np.sum(np.real(np.einsum("...ij,...ji->...", k.weights_b * F, dP))))
    np.testing.assert_allclose(fd, an, rtol=1e-7, atol=1e-8)


def test_public_build_fock_project_fn(gpu, oracle):
    """fock.py:80-87: F is projected and hermitised after assembly."""
    k, P = hartree_kernel(nk=4, nb=3)
    flip = np.diag([1.0, -1.0, 1.0])

    def project(F):
        return 0.5 * (F + flip @ F @ flip)

    _, _, F = _public_build(k, P, project_fn=project)
    _, _, Fr, _ = oracle.build_fock_apply(kernel_args_dict(k), P)
    assert relerr(F, herm(project(Fr))) < 1e-12


@pytest.mark.parametrize("hcp", [False, True])
def test_public_build_fock_exchange_block_specs(gpu, oracle, hcp):
    """fock.py:36-58: the block-restricted exchange.  With a block-diagonal
    density, reference density and interaction the block route equals the full
    build; with a coupled density and check_offdiag it falls back to it; forced
    blocks drop the off-block exchange."""
    from cpp_hf_amd import HartreeFockKernel

    rng = np.random.default_rng(17)
    nk, sizes = 6, (2, 3)
    nb = sum(sizes)
    mask = np.zeros((nb, nb), bool)
    mask[:2, :2] = mask[2:, 2:] = True
    h = rand_herm(rng, (nk, nk, nb, nb)) * mask
    refP = 0.1 * rand_herm(rng, (nk, nk, nb, nb)) * mask
    P = rand_herm(rng, (nk, nk, nb, nb)) * mask
    HH = rng.uniform(0.2, 1.0, (nb, nb))
    HH = 0.5 * (HH + HH.T)
    q = np.fft.fftfreq(nk) * 2 * np.pi
    Vq = (1.0 / (1.0 + q[:, None] ** 2 + q[None, :] ** 2))[..., None, None]
    k = HartreeFockKernel(np.full((nk, nk), 1.0 / nk ** 2), h, Vq, T=0.1, include_hartree=True,
                          include_exchange=True, reference_density=refP, hartree_matrix=HH)
    k.exchange_hermitian_channel_packing = hcp
    ka = kernel_args_dict(k)
    spec = [{"block_sizes": list(sizes)}]
    S, H, F = _public_build(k, P, exchange_block_specs=spec, exchange_check_offdiag=True)
    Sr, Hr, Fr, _ = oracle.build_fock_apply(ka, P)
    assert relerr(S, Sr) < 1e-12 and relerr(F, Fr) < 1e-12
    assert relerr(np.broadcast_to(H, Hr.shape), Hr) < 1e-12
    # coupled density: the off-block check sends it to the full exchange
    Pc = P + 0.05 * herm(rand_c(rng, P.shape)) * ~mask
    S, H, F = _public_build(k, Pc, exchange_block_specs=spec, exchange_check_offdiag=True)
    Sr, Hr, Fr, _ = oracle.build_fock_apply(ka, Pc)
    assert relerr(S, Sr) < 1e-12 and relerr(F, Fr) < 1e-12
    # forced blocks: the off-block exchange is absent, and with a scalar
    # interaction every channel is its own convolution, so the blocks'
    # entries are those of the full build
    S, H, F = _public_build(k, Pc, exchange_block_specs=spec, exchange_check_offdiag=False)
    assert np.max(np.abs((S - Sr) * ~mask)) > 1e-3
    assert relerr(S * mask, Sr * mask) < 1e-12


# ----------------------------------------------------------- linalg.eigh ----

def _lapack_blocks(M, blocks, sort):
    n = M.shape[-1]
    w = np.empty(M.shape[:-1])
    col = 0
    for ix in blocks:
        ix = np.asarray(ix)
        wb = np.linalg.eigvalsh(herm(M[..., ix[:, None], ix[None, :]]))
        w[..., col:col + len(ix)] = wb
        col += len(ix)
    return np.sort(w, axis=-1) if sort else w


def test_eigh_block_falls_back_to_full_when_coupled(gpu):
    """tests/test_linalg.py:9-19."""
    from cpp_hf_amd.linalg import eigh

    h = np.diag([-1.0, -0.5, 0.5, 1.0]).astype(complex)
    h[0, 2] = h[2, 0] = 1e-2
    w, v = eigh(h, block_sizes=(2, 2), check_offdiag=True, offdiag_atol=1e-6)
    np.testing.assert_allclose(w, np.linalg.eigvalsh(h), rtol=0, atol=1e-14)
    np.testing.assert_allclose(v @ np.diag(w) @ np.conj(v.T), h, rtol=0, atol=1e-14)


def test_eigh_block_forced_matches_full_when_block_diagonal(gpu):
    """tests/test_linalg.py:22-46."""
    from cpp_hf_amd.linalg import eigh

    h = np.zeros((4, 4), complex)
    h[:2, :2] = [[-1.0, 0.2 + 0.1j], [0.2 - 0.1j, -0.3]]
    h[2:, 2:] = [[0.4, -0.15j], [0.15j, 1.1]]
    w, v = eigh(h, block_sizes=(2, 2), check_offdiag=False)
    np.testing.assert_allclose(w, np.linalg.eigvalsh(h), rtol=0, atol=1e-14)
    np.testing.assert_allclose(v @ np.diag(w) @ np.conj(v.T), h, rtol=0, atol=1e-14)


@pytest.mark.parametrize("sort", [True, False])
def test_eigh_block_specs_sizes_and_indices(gpu, sort):
    """linalg.py:150-297: sizes, indices and spec lists, sorted or block-ordered,
    each block against LAPACK."""
    from cpp_hf_amd.linalg import eigh

    rng = np.random.default_rng(23)
    nb = 7
    idx = ([0, 3, 5], [1, 2], [4, 6])
    mask = np.zeros((nb, nb), bool)
    for b in idx:
        mask[np.ix_(b, b)] = True
    M = rand_herm(rng, (5, 4, nb, nb)) * mask
    w, V = eigh(M, block_indices=idx, sort=sort)
    np.testing.assert_allclose(w, _lapack_blocks(M, idx, sort), rtol=0, atol=1e-12)
    R = V @ (w[..., None] * np.conj(np.swapaxes(V, -1, -2)))
    np.testing.assert_allclose(R, M, rtol=0, atol=1e-12)
    # a first spec that does not fit is skipped; the second one is used
    w2, _ = eigh(M, block_specs=[{"block_sizes": [3, 4]}, ("indices", idx)], sort=sort)
    np.testing.assert_allclose(w2, w, rtol=0, atol=1e-14)
    sizes = (2, 5)
    Ms = rand_herm(rng, (9, 7, 7))
    Ms[:, :2, 2:] = 0
    Ms[:, 2:, :2] = 0
    blocks = [np.arange(0, 2), np.arange(2, 7)]
    w, V = eigh(Ms, block_sizes=sizes, sort=sort)
    np.testing.assert_allclose(w, _lapack_blocks(Ms, blocks, sort), rtol=0, atol=1e-12)
    if not sort:
        assert not np.any(V[:, 2:, :2]) and not np.any(V[:, :2, 2:])
    with pytest.raises(ValueError, match="must sum to"):
        eigh(Ms, block_sizes=(3, 3))
    with pytest.raises(ValueError, match="disjoint"):
        eigh(M, block_indices=([0, 1], [1, 2, 3, 4, 5, 6]))


# ------------------------------------------------ utils.selfenergy_fft ------

def test_selfenergy_fft_single_point_matches_definition(gpu):
    """tests/test_main.py:43-48."""
    from cpp_hf_amd.utils import selfenergy_fft

    VR = np.array([[[[0.3]]]], complex)
    P = np.array([[[[1.0 + 0.5j, 0.2], [0.2, 0.8 - 0.1j]]]])
    np.testing.assert_allclose(selfenergy_fft(VR, P), -VR * P, rtol=0, atol=1e-15)


@pytest.mark.parametrize("hcp", [False, True])
def test_selfenergy_fft_block_specs(gpu, oracle, hcp):
    """tests/test_linalg.py:49-162 in FP64: block route == full route for a
    block-diagonal density (with or without Hermitian channel packing),
    fallback to full when coupled, forced blocks differ."""
    from cpp_hf_amd.utils import selfenergy_fft

    rng = np.random.default_rng(7)
    nk, sizes = 8, (2, 2, 2)
    nb = sum(sizes)
    P = np.zeros((nk, nk, nb, nb), complex)
    for s in range(0, nb, 2):
        P[..., s:s + 2, s:s + 2] = rand_herm(rng, (nk, nk, 2, 2))
    VR = np.fft.fftn(rng.normal(size=(nk, nk, 1, 1)).astype(complex), axes=(0, 1))
    full = oracle.selfenergy_fft_full(VR, P, apply_ifftshift=True, hermitian_channel_packing=False)
    blk = selfenergy_fft(VR, P, block_specs=[{"block_sizes": list(sizes)}], check_offdiag=True,
                         hermitian_channel_packing=hcp)
    assert relerr(blk, full) < 1e-12
    Pc = P.copy()
    Pc[..., 0, 2] = 0.1
    Pc[..., 2, 0] = 0.1
    full_c = oracle.selfenergy_fft_full(VR, Pc, apply_ifftshift=True,
                                        hermitian_channel_packing=False)
    auto = selfenergy_fft(VR, Pc, block_specs=[{"block_sizes": list(sizes)}],
                          check_offdiag=True, offdiag_atol=1e-6, hermitian_channel_packing=hcp)
    assert relerr(auto, full_c) < 1e-12
    forced = selfenergy_fft(VR, Pc, block_specs=[{"block_sizes": list(sizes)}],
                            check_offdiag=False, hermitian_channel_packing=hcp)
    assert float(np.max(np.abs(forced - full_c))) > 1e-3


def test_selfenergy_plan_cache_concurrent_with_clear(gpu, oracle):
    """The geometry cache (the counterpart of the reference's 4-entry FFT plan
    cache, selfenergy.hpp:20-55) under concurrent calls on more meshes than it
    holds, with the cache cleared from another thread meanwhile: every call's
    geometry stays alive until the call has finished, so each result equals
    the single-threaded one."""
    rng = np.random.default_rng(31)
    cases = []
    for n1, n2 in [(12, 12), (16, 10), (9, 15), (20, 20), (24, 8), (10, 30)]:
        P = rand_herm(rng, (n1, n2, 3, 3))
        VR = rand_c(rng, (n1, n2, 1, 1))
        cases.append((VR, P, oracle.selfenergy_fft_full(VR, P, apply_ifftshift=False,
                                                        hermitian_channel_packing=True)))
    errors = []
    stop = threading.Event()

    def worker(i):
        try:
            for r in range(12):
                VR, P, ref = cases[(i + r) % len(cases)]
                got = gpu.selfenergy_fft_full(VR, P, False, True)
                if relerr(got, ref) >= 1e-12:
                    errors.append((i, r, relerr(got, ref)))
        except Exception as e:  # noqa: BLE001 - reported below
            errors.append(repr(e))

    def clearer():
        while not stop.is_set():
            gpu.clear_fft_plan_cache()
            stop.wait(0.001)

    threads = [threading.Thread(target=worker, args=(i,)) for i in range(6)]
    cl = threading.Thread(target=clearer)
    cl.start()
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    stop.set()
    cl.join(timeout=10)
    assert not any(t.is_alive() for t in threads)
    assert not errors, errors[:3]


# ------------------------------------------------------ SCF trust radius ----

@pytest.mark.parametrize("accel,radius", [("linear", 0.02), ("diis", 0.05), ("diis", 0.02)])
def test_solve_scf_trust_radius_matches_oracle(gpu, oracle, accel, radius):
    """solver_scf.hpp:405-420: the density step is rescaled to the trust
    radius when it is longer.  Run to convergence (density_tol 1e-7,
    comm_tol 1e-6): 159 (linear), 29 and 60 (DIIS) iterations.

    Linear mixing is contractive: the whole trajectory matches the oracle.
    DIIS solves a least-squares system whose condition number grows as the
    residuals shrink, so rounding-level differences of its Gram matrix grow
    to the 1e-8 level near convergence, where the iteration count may differ
    by a few steps; the clipped phase matches step by step and both runs end
    on the same fixed point within the convergence tolerance."""
    from cpp_hf_amd.workloads import layered_kernel

    w = layered_kernel(12, 2, T=0.05, filling=2.1)
    args = (kernel_args_dict(w.kernel), w.P0, w.n_electrons, 200, 1e-7, 1e-6, 0.5, 0.0, None,
            [], accel, 6, 2, 1.0, radius, True, True)
    g = gpu.solve_scf(*args)
    r = oracle.solve_scf(*args)
    assert r[5] and g[5]
    if accel == "linear":
        assert g[4] == r[4]
        assert abs(g[2] - r[2]) < 1e-10 * max(1.0, abs(r[2]))
        assert relerr(g[0], r[0]) < 1e-9
        np.testing.assert_allclose(g[7], r[7], rtol=1e-8, atol=1e-12)
    else:
        assert abs(g[4] - r[4]) <= 5
        assert abs(g[2] - r[2]) < 1e-9 * max(1.0, abs(r[2]))
        # both densities satisfy comm_tol = 1e-6 and density_tol = 1e-7
        assert relerr(g[0], r[0]) < 1e-5
        np.testing.assert_allclose(g[7][:10], r[7][:10], rtol=1e-8, atol=1e-12)
    # the clip is active: each step moves the density by the radius, where
    # the unclipped run's second step halves the residual
    assert abs(r[7][1] - r[7][0]) <= 1.05 * radius
    free = list(args)
    free[14] = 0.0
    assert oracle.solve_scf(*free)[7][1] < 0.6 * r[7][1]
