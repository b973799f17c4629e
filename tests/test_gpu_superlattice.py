"""GPU superlattice exchange and Hartree shift (libhfgpu SlPlan / SlHartree)
against the oracle (pinned by a brute-force momentum sum in
test_oracle_superlattice.py): the streamed primitive with the reference
test's tolerance (rel 1e-12, test_superlattice.py:116-131), its zero-density
and shape checks (test_superlattice.py:134-153), channel batches, the Fock
build of a SuperlatticeHartreeFockKernel, and SCF / direct-minimisation
solves with that kernel.
"""

from __future__ import annotations

import numpy as np
import pytest

from helpers import kernel_args_dict, rand_c, rand_herm
from test_oracle_superlattice import call, layout, sl_kernel

pytestmark = pytest.mark.gpu


def relerr(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.max(np.abs(a - b))) / max(1e-300, float(np.max(np.abs(b))))


# (nk, g_cut, orb, orbital V, Hermitian fill); the last two run the
# channels in 2 and 5 batches (at most 1024 channels per batch; 22 and 43
# transformed dG of 64 and 100 channels each)
PRIM = [(3, 1, 1, False, False), (3, 1, 2, False, True), (4, 2, 2, True, False),
        (4, 2, 2, True, True), (2, 1, 3, True, True), (3, 2, 8, False, True),
        (2, 2, 10, True, False)]


@pytest.mark.parametrize("case", range(len(PRIM)))
def test_streamed_matches_oracle(gpu, oracle, case):
    nk, g_cut, orb, orbital, herm_fill = PRIM[case]
    print(f"superlattice primitive nk={nk} g_cut={g_cut} orb={orb} orbital={orbital} "
          f"hermitian={herm_fill}", flush=True)
    L, g, _ = layout(nk=nk, g_cut=g_cut)
    n_G = g.shape[0]
    D = n_G * orb
    rng = np.random.default_rng(case)
    rho = (rand_herm(rng, (nk, nk, D, D)) if herm_fill else rand_c(rng, (nk, nk, D, D)))
    rho6 = rho.reshape(nk, nk, n_G, orb, n_G, orb)
    V_orb = None
    if orbital:
        V_orb = L.V_lag_fft[:, :, None, None] * (1.0 + 0.2 * rng.standard_normal((orb, orb)))
    got = call(gpu, rho6, L, nk, n_G, orb, hermitian=herm_fill, V_orb=V_orb)
    ref = call(oracle, rho6, L, nk, n_G, orb, hermitian=herm_fill, V_orb=V_orb)
    assert got.shape == rho6.shape
    assert relerr(got, ref) < 1e-12


def test_streamed_zero_density_and_validation(gpu):
    # tests/test_superlattice.py:134-153, plus indices outside the grid
    nk, orb = 3, 2
    L, g, _ = layout(nk=nk, g_cut=1)
    n_G = g.shape[0]
    z = np.zeros((nk, nk, n_G, orb, n_G, orb), dtype=complex)
    for herm_fill in (False, True):
        assert not np.any(call(gpu, z, L, nk, n_G, orb, hermitian=herm_fill))
    with pytest.raises(ValueError):
        call(gpu, np.zeros((nk, nk, n_G * orb, n_G * orb), dtype=complex), L, nk, n_G, orb)
    bad_off = L.g_a_off.copy()
    bad_off[0, 0] = L.N_ext_x // nk
    with pytest.raises(ValueError, match="outside"):
        gpu.selfenergy_superlattice_streamed(
            z, L.V_lag_fft, bad_off, L.delta_pair_i, L.delta_pair_j, L.delta_pair_start, nk, nk,
            n_G, orb, L.N_ext_x, L.N_ext_y)
    bad_i = L.delta_pair_i.copy()
    bad_i[-1] = n_G
    with pytest.raises(ValueError, match="n_G"):
        gpu.selfenergy_superlattice_streamed(
            z, L.V_lag_fft, L.g_a_off, bad_i, L.delta_pair_j, L.delta_pair_start, nk, nk, n_G,
            orb, L.N_ext_x, L.N_ext_y)


def test_superlattice_fock_wrapper(gpu, oracle):
    from cpp_hf_amd.superlattice import superlattice_fock

    nk, orb = 4, 2
    L, g, _ = layout(nk=nk, g_cut=1)
    n_G = g.shape[0]
    D = n_G * orb
    rho = rand_herm(np.random.default_rng(5), (nk, nk, D, D))
    got = superlattice_fock(rho, L, n_G, orb, nk, nk, hermitian_rho=True)
    ref = call(oracle, rho.reshape(nk, nk, n_G, orb, n_G, orb), L, nk, n_G, orb,
               hermitian=True).reshape(nk, nk, D, D)
    assert got.shape == (nk, nk, D, D)
    assert relerr(got, ref) < 1e-12


# (orbital Hartree and V, reference density, exchange on)
FOCK = [(False, True, True), (True, True, True), (False, False, True), (True, False, False)]


@pytest.mark.parametrize("case", range(len(FOCK)))
def test_build_fock_superlattice_matches_oracle(gpu, oracle, case):
    orbital, ref, exch = FOCK[case]
    print(f"superlattice build_fock orbital={orbital} refP={ref} exchange={exch}", flush=True)
    k = sl_kernel(nk=4, g_cut=1, orb=2, orbital=orbital, refP=ref)
    if not exch:
        k.include_exchange = False
        k.superlattice_fock_active = False
    ka = kernel_args_dict(k)
    P = rand_herm(np.random.default_rng(11 + case), k.h.shape)
    S, Hh, F, E = gpu.build_fock_apply(ka, P, None)
    Sr, Hr, Fr, Er = oracle.build_fock_apply(ka, P)
    assert relerr(S, Sr) < 1e-12
    assert relerr(F, Fr) < 1e-12
    assert not np.any(Hh) and not np.any(Hr)
    assert abs(E - Er) <= 1e-12 * max(1.0, abs(Er))
    np.testing.assert_array_equal(F, np.conj(np.swapaxes(F, -1, -2)))


def test_superlattice_kernel_rejects_mesh_exchange_beside_shift(gpu):
    k = sl_kernel(nk=3, g_cut=1, orb=1)
    ka = kernel_args_dict(k)
    ka["superlattice_fock_active"] = False
    with pytest.raises(ValueError, match="superlattice_fock_active"):
        gpu.build_fock_apply(ka, rand_herm(np.random.default_rng(0), k.h.shape), None)


def test_solve_scf_superlattice_matches_oracle(gpu, oracle):
    k = sl_kernel(nk=4, g_cut=1, orb=2, T=0.05)
    args = (kernel_args_dict(k), k._refP.copy(), 7.0, 60, 1e-9, 1e-8, 0.5, 0.0, None, [],
            "diis", 6, 2, 1.0, 0.0, True, True)
    print("superlattice solve_scf diis nk=4 D=14", flush=True)
    g = gpu.solve_scf(*args)
    r = oracle.solve_scf(*args)
    assert g[5] and r[5] and g[4] == r[4]
    assert abs(g[2] - r[2]) < 1e-10 * max(1.0, abs(r[2]))
    assert abs(g[3] - r[3]) < 1e-8
    assert relerr(g[0], r[0]) < 1e-8
    assert relerr(g[1], r[1]) < 1e-8


@pytest.mark.parametrize("orbital", [False, True])
def test_solve_dm_superlattice_matches_oracle(gpu, oracle, orbital):
    k = sl_kernel(nk=4, g_cut=1, orb=2, T=0.05, orbital=orbital)
    args = (kernel_args_dict(k), k._refP.copy(), 7.0, 60, 1e-9, 0.0, 0.6, 0.5, 1e-3, 8, 10, 5,
            25, 0, 0.5, 20, np.empty((0,) + k.h.shape), 0.0, 1.0, [], None, True, True, True)
    print(f"superlattice solve_dm orbital={orbital}", flush=True)
    g = gpu.solve_dm(*args)
    r = oracle.solve_dm(*args)
    Q, p, P, F, mu, E, n, conv, hE, hG = g
    assert conv and r[7] and n == r[6]
    np.testing.assert_allclose(hE, r[8], rtol=0, atol=1e-10)
    assert abs(E - r[5]) < 1e-10
    assert abs(mu - r[4]) < 1e-8
    assert relerr(P, r[2]) < 1e-8
    assert relerr(F, r[3]) < 1e-8


def test_build_fock_superlattice_k_dependent_reference(gpu, oracle):
    # refP that varies with k: the scatter subtracts refP_k entry by entry
    # (the constant I/2 case goes through the one-matrix path)
    from cpp_hf_amd.superlattice import SuperlatticeHartreeFockKernel

    k0 = sl_kernel(nk=4, g_cut=1, orb=2, refP=False)
    rng = np.random.default_rng(21)
    ref = rand_herm(rng, k0.h.shape) * 0.1 + 0.5 * np.eye(k0.h.shape[-1])
    k = SuperlatticeHartreeFockKernel(k0.w2d, k0.h, k0.layout, HH_GG=k0.HH_GG, T=0.1,
                                      reference_density=ref, hartree_degeneracy=2.0)
    ka = kernel_args_dict(k)
    P = rand_herm(rng, k.h.shape)
    S, Hh, F, E = gpu.build_fock_apply(ka, P, None)
    Sr, Hr, Fr, Er = oracle.build_fock_apply(ka, P)
    assert relerr(S, Sr) < 1e-12
    assert relerr(F, Fr) < 1e-12
    assert abs(E - Er) <= 1e-12 * max(1.0, abs(Er))


def test_newton_solve_superlattice_matches_oracle(gpu, oracle):
    # the Newton solver's Hessian-vector products build the Fock matrix of a
    # density variation without the reference density
    k = sl_kernel(nk=4, g_cut=1, orb=2, T=0.05)
    args = (kernel_args_dict(k), k._refP.copy(), 7.0, 40, 1e-10, 0.0, 0.6, 0.5, 1e-3, 8, 10,
            5, 25, 1, 0.5, 20, np.empty((0,) + k.h.shape), 0.0, 1.0, [], None, True, True, True)
    print("superlattice newton solve", flush=True)
    g = gpu.solve_dm(*args)
    r = oracle.solve_dm(*args)
    assert g[6] == r[6] and g[7] == r[7]
    np.testing.assert_allclose(g[8], r[8], rtol=0, atol=1e-9)
    assert abs(g[5] - r[5]) < 1e-9
    assert relerr(g[2], r[2]) < 1e-7


@pytest.mark.parametrize("herm_fill", [False, True])
def test_streamed_incomplete_channel_table(gpu, oracle, herm_fill):
    # a channel table without one dG: its blocks (and, with the partner fill,
    # nothing else) have no source and stay zero, as in the reference's
    # channel loop; the GPU then stores through the separate gather pass
    nk, orb = 3, 2
    L, g, _ = layout(nk=nk, g_cut=1)
    n_G = g.shape[0]
    D = n_G * orb
    drop = 3
    ps = L.delta_pair_start
    keep = np.r_[0:ps[drop], ps[drop + 1]:ps[-1]]
    pi, pj = L.delta_pair_i[keep], L.delta_pair_j[keep]
    cnt = np.diff(ps).copy()
    cnt = np.delete(cnt, drop)
    start = np.zeros(len(cnt) + 1, dtype=np.int64)
    start[1:] = np.cumsum(cnt)
    rng = np.random.default_rng(4)
    rho = rand_herm(rng, (nk, nk, D, D)).reshape(nk, nk, n_G, orb, n_G, orb)
    out = []
    for mod in (gpu, oracle):
        out.append(mod.selfenergy_superlattice_streamed(
            rho, L.V_lag_fft, L.g_a_off, pi, pj, start, nk, nk, n_G, orb, L.N_ext_x, L.N_ext_y,
            VR_fft_orbital=None, hermitian_rho=herm_fill))
    got, ref = out
    assert relerr(got, ref) < 1e-12
    dropped = [(int(L.delta_pair_i[p]), int(L.delta_pair_j[p])) for p in range(ps[drop], ps[drop + 1])]
    G = got.reshape(nk, nk, n_G, orb, n_G, orb)
    assert all(not np.any(G[:, :, i, :, j, :]) for i, j in dropped)
