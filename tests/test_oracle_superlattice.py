"""Pins of the superlattice layout (cpp_hf_amd.superlattice) and the oracle's
superlattice exchange and Hartree shift (superlattice.hpp, fock.hpp:66-173),
CPU only.

The exchange is checked against the physics it encodes: a brute-force sum
Sigma[k, a G_i, b G_j] = -w sum_{q, i'} V(|(k - G_i) - (q - G_i')|)
rho[q, a G_i', b G_j'] with G_j' = G_i' - (G_i - G_j), independent of the
extended grid, the channel tables and the FFT; and against a numpy
streaming restatement (the reference's own test, test_superlattice.py:62-
90).  The Hermitian partner fill must equal the direct computation.
"""

from __future__ import annotations

import numpy as np
import pytest

from helpers import herm, kernel_args_dict, rand_c, rand_herm


def hex_basis(g_cut):
    """tests/test_superlattice.py:26-59: gamma plus g_cut shells."""
    b0 = 4 * np.pi / np.sqrt(3)
    B = np.stack([np.array([np.sqrt(3) / 2, -0.5]) * b0, np.array([0.0, 1.0]) * b0], axis=1)
    fr, nr = [], []
    for n1 in range(-g_cut - 1, g_cut + 2):
        for n2 in range(-g_cut - 1, g_cut + 2):
            fr.append((n1, n2))
            nr.append(np.linalg.norm(B @ np.array([n1, n2])))
    nr = np.array(nr)
    shells = np.sort(np.unique(np.round(nr, 6)))
    cut = shells[min(g_cut, len(shells) - 1)] + 1e-3
    return np.array(fr, dtype=np.int64)[nr <= cut], B


def V_of(q):
    return 1.0 / np.sqrt(q ** 2 + 0.5 ** 2)


def layout(nk=3, g_cut=1):
    from cpp_hf_amd.superlattice import build_extended_layout

    g, B = hex_basis(g_cut)
    L = build_extended_layout(g, B, nk, nk, V_of, 1.0 / (nk * nk))
    return L, g, B


def call(oracle, rho6, L, nk, n_G, orb, hermitian=False, V_orb=None):
    return oracle.selfenergy_superlattice_streamed(
        rho6, L.V_lag_fft, L.g_a_off, L.delta_pair_i, L.delta_pair_j, L.delta_pair_start, nk, nk,
        n_G, orb, L.N_ext_x, L.N_ext_y, VR_fft_orbital=V_orb, hermitian_rho=hermitian)


def test_layout_tables():
    # tests/test_superlattice.py:97-108 plus the dG and offset definitions
    L, g, _ = layout(nk=4, g_cut=2)
    n_G = g.shape[0]
    seen = set()
    for d in range(L.n_delta):
        for p in range(L.delta_pair_start[d], L.delta_pair_start[d + 1]):
            i, j = int(L.delta_pair_i[p]), int(L.delta_pair_j[p])
            assert L.pair_to_delta[i, j] == d
            seen.add((i, j))
    assert len(seen) == n_G * n_G
    dg = {tuple(g[i] - g[j]): L.pair_to_delta[i, j] for i in range(n_G) for j in range(n_G)}
    assert len(dg) == L.n_delta  # one channel per distinct dG
    np.testing.assert_array_equal(L.g_a_off, -g - (-g).min(axis=0))
    assert (L.N_ext_x, L.N_ext_y) == (2 * L.n_supp_x, 2 * L.n_supp_y)
    assert L.n_supp_x == 4 * (np.ptp(g[:, 0]) + 1)


@pytest.mark.parametrize("orb", [1, 2])
def test_oracle_superlattice_matches_brute_force_sum(oracle, orb):
    nk = 3
    L, g, B = layout(nk=nk, g_cut=1)
    n_G = g.shape[0]
    D = n_G * orb
    rng = np.random.default_rng(orb)
    rho = rand_c(rng, (nk, nk, D, D))
    got = call(oracle, rho.reshape(nk, nk, n_G, orb, n_G, orb), L, nk, n_G, orb)
    got = got.reshape(nk, nk, D, D)
    w = 1.0 / (nk * nk)
    idx = {tuple(v): n for n, v in enumerate(g)}
    kf = np.stack(np.meshgrid(np.arange(nk) / nk, np.arange(nk) / nk, indexing="ij"), -1)
    ref = np.zeros_like(rho).reshape(nk, nk, n_G, orb, n_G, orb)
    r6 = rho.reshape(nk, nk, n_G, orb, n_G, orb)
    for i in range(n_G):
        for j in range(n_G):
            for ip in range(n_G):
                jp = idx.get(tuple(g[ip] - (g[i] - g[j])))
                if jp is None:
                    continue
                # lag (k - G_i) - (q - G_i') for every (k, q)
                lag = (kf[:, :, None, None, :] - g[i]) - (kf[None, None, :, :, :] - g[ip])
                q = np.einsum("xy,...y->...x", B, lag)
                Vk = V_of(np.linalg.norm(q, axis=-1))  # (k1, k2, q1, q2)
                ref[:, :, i, :, j, :] -= w * np.einsum("klmn,mnab->klab", Vk, r6[:, :, ip, :, jp, :])
    ref = ref.reshape(nk, nk, D, D)
    assert np.max(np.abs(got - ref)) < 1e-12 * np.max(np.abs(ref))


def _np_streaming(rho, L, n_G, orb, nk, V_orb=None):
    """tests/test_superlattice.py:62-90 (scatter, numpy FFT convolution, gather)."""
    r = rho.reshape(nk, nk, n_G, orb, n_G, orb)
    out = np.zeros_like(r)
    for d in range(L.n_delta):
        buf = np.zeros((L.N_ext_x, L.N_ext_y, orb, orb), dtype=complex)
        ps = range(L.delta_pair_start[d], L.delta_pair_start[d + 1])
        for p in ps:
            i, j = L.delta_pair_i[p], L.delta_pair_j[p]
            ox, oy = nk * L.g_a_off[i, 0], nk * L.g_a_off[i, 1]
            buf[ox:ox + nk, oy:oy + nk] = r[:, :, i, :, j, :]
        V = L.V_lag_fft[:, :, None, None] if V_orb is None else V_orb
        s = -np.fft.ifftn(np.fft.fftn(buf, axes=(0, 1)) * V, axes=(0, 1))
        for p in ps:
            i, j = L.delta_pair_i[p], L.delta_pair_j[p]
            ox, oy = nk * L.g_a_off[i, 0], nk * L.g_a_off[i, 1]
            out[:, :, i, :, j, :] = s[ox:ox + nk, oy:oy + nk]
    return out


@pytest.mark.parametrize("orbital_v", [False, True])
def test_oracle_superlattice_matches_numpy_streaming(oracle, orbital_v):
    nk, orb = 4, 3
    L, g, _ = layout(nk=nk, g_cut=1)
    n_G = g.shape[0]
    D = n_G * orb
    rng = np.random.default_rng(0)
    rho = (rng.standard_normal((nk, nk, D, D)) + 1j * rng.standard_normal((nk, nk, D, D))) * 1e-3
    V_orb = None
    if orbital_v:
        V_orb = L.V_lag_fft[:, :, None, None] * (1.0 + 0.1 * rng.standard_normal((orb, orb)))
    got = call(oracle, rho.reshape(nk, nk, n_G, orb, n_G, orb), L, nk, n_G, orb, V_orb=V_orb)
    ref = _np_streaming(rho, L, n_G, orb, nk, V_orb)
    assert np.max(np.abs(got - ref)) < 1e-12 * np.max(np.abs(ref))


def test_oracle_superlattice_hermitian_partner_fill(oracle):
    nk, orb = 4, 2
    L, g, _ = layout(nk=nk, g_cut=2)
    n_G = g.shape[0]
    D = n_G * orb
    rho = rand_herm(np.random.default_rng(3), (nk, nk, D, D)).reshape(nk, nk, n_G, orb, n_G, orb)
    a = call(oracle, rho, L, nk, n_G, orb, hermitian=False).reshape(nk, nk, D, D)
    b = call(oracle, rho, L, nk, n_G, orb, hermitian=True).reshape(nk, nk, D, D)
    assert np.max(np.abs(a - b)) < 1e-13 * np.max(np.abs(a))
    np.testing.assert_allclose(b, np.conj(np.swapaxes(b, -1, -2)), atol=1e-14)
    z = call(oracle, np.zeros_like(rho), L, nk, n_G, orb, hermitian=True)
    assert np.max(np.abs(z)) == 0.0


def sl_kernel(nk=4, g_cut=1, orb=2, T=0.1, orbital=False, seed=0, refP=True):
    from cpp_hf_amd.superlattice import SuperlatticeHartreeFockKernel

    L, g, B = layout(nk=nk, g_cut=g_cut)
    n_G = g.shape[0]
    D = n_G * orb
    rng = np.random.default_rng(seed)
    gc = (B @ g.T).T
    kf = np.stack(np.meshgrid(np.arange(nk) / nk, np.arange(nk) / nk, indexing="ij"), -1)
    h = np.zeros((nk, nk, D, D), dtype=complex)
    for i in range(n_G):
        kin = np.sum((np.einsum("xy,...y->...x", B, kf - g[i])) ** 2, axis=-1)
        for a in range(orb):
            h[:, :, i * orb + a, i * orb + a] = 0.02 * kin + 0.3 * a
    hop = 0.2 * rand_herm(rng, (D, D))
    h = h + hop[None, None]
    extra = {}
    if orbital:
        HHo = np.abs(rng.standard_normal((n_G, n_G, orb, orb)))
        HHo[np.arange(n_G), np.arange(n_G)] = 0.0
        extra["HH_GG_orbital"] = HHo
        extra["V_lag_fft_orbital"] = L.V_lag_fft[:, :, None, None] * (
            1.0 + 0.1 * np.abs(rng.standard_normal((orb, orb))))
    ref = np.broadcast_to(0.5 * np.eye(D, dtype=complex), h.shape).copy() if refP else None
    return SuperlatticeHartreeFockKernel(np.full((nk, nk), 1.0 / nk ** 2), h, L, coulomb_V=V_of,
                                         g_cart=gc, T=T, reference_density=ref,
                                         hartree_degeneracy=2.0, **extra)


@pytest.mark.parametrize("orbital", [False, True])
def test_oracle_superlattice_fock_build_composition(oracle, orbital):
    # Sigma of the Fock build = exchange(P - refP) + the Hartree shift, with
    # the orbital Hartree reducing to the scalar one for HH_orb = HH
    from cpp_hf_amd.superlattice import _hartree_shift

    k = sl_kernel(orbital=orbital)
    ka = kernel_args_dict(k)
    nk, D = k.h.shape[0], k.h.shape[-1]
    P = rand_herm(np.random.default_rng(7), k.h.shape)
    S, H, F, E = oracle.build_fock_apply(ka, P)
    dP = P - k._refP
    x = call(oracle, dP.reshape(nk, nk, k.n_G, k.dim_orb, k.n_G, k.dim_orb), k.layout, nk,
             k.n_G, k.dim_orb, hermitian=True, V_orb=k.V_lag_fft_orbital).reshape(S.shape)
    sh = _hartree_shift(dP, k.w2d, k.layout, k.n_G, k.dim_orb, HH_GG=k.HH_GG,
                        degeneracy=k.hartree_degeneracy, HH_GG_orbital=k.HH_GG_orbital)
    np.testing.assert_allclose(S, x + sh[None, None], atol=1e-13)
    assert not np.any(H)  # no mesh Hartree in superlattice mode
    np.testing.assert_allclose(F, herm(k.h + S), atol=1e-14)
    if not orbital:
        HHo = np.broadcast_to(k.HH_GG[:, :, None, None], k.HH_GG.shape + (k.dim_orb,) * 2)
        sh2 = _hartree_shift(dP, k.w2d, k.layout, k.n_G, k.dim_orb, HH_GG=k.HH_GG,
                             degeneracy=k.hartree_degeneracy, HH_GG_orbital=HHo)
        np.testing.assert_allclose(sh2, sh, atol=1e-14)
