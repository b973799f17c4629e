"""Symmetry projectors (reference: src/cpp_hf/symmetry.py, tests/test_symmetry.py).

Each check is a property the averaging must have (invariance under the
group it averages over, idempotence on a closed group, Hermiticity kept) or
a closed form (the Z2 group {1, sigma_z} keeps exactly the block-diagonal
part; "mod" and "flip" differ by one cyclic shift).
"""

from __future__ import annotations

import numpy as np
import pytest

from cpp_hf_amd.symmetry import make_project_fn

SX = np.array([[0, 1], [1, 0]], dtype=np.complex128)
SY = np.array([[0, -1j], [1j, 0]], dtype=np.complex128)
SZ = np.diag([1.0, -1.0]).astype(np.complex128)


def _rand(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def _minus_k_mod(A):
    n1, n2 = A.shape[:2]
    return A[((-np.arange(n1)) % n1)[:, None], ((-np.arange(n2)) % n2)[None, :]]


def test_no_symmetry_is_identity():
    A = _rand(np.random.default_rng(0), (2, 3, 3, 3))
    out = make_project_fn()(A)
    np.testing.assert_array_equal(out, A)


def test_z2_group_keeps_block_diagonal_part():
    g = np.diag([1, 1, -1, -1]).astype(np.complex128)
    proj = make_project_fn(unitary_group=np.stack([np.eye(4), g]))
    A = _rand(np.random.default_rng(1), (3, 3, 4, 4))
    want = A.copy()
    want[..., :2, 2:] = 0
    want[..., 2:, :2] = 0
    np.testing.assert_allclose(proj(A), want, atol=1e-15)


def test_cyclic_group_invariant_idempotent_hermitian():
    perm = np.roll(np.eye(3), 1, axis=0).astype(np.complex128)
    G = np.stack([np.eye(3), perm, perm @ perm])
    proj = make_project_fn(unitary_group=G)
    A = _rand(np.random.default_rng(2), (2, 2, 3, 3))
    A = 0.5 * (A + np.conj(np.swapaxes(A, -1, -2)))
    out = proj(A)
    for g in G:
        np.testing.assert_allclose(g @ out @ g.conj().T, out, atol=1e-14)
    np.testing.assert_allclose(proj(out), out, atol=1e-14)
    np.testing.assert_allclose(out, np.conj(np.swapaxes(out, -1, -2)), atol=1e-14)


@pytest.mark.parametrize("conv", ["mod", "flip"])
def test_time_reversal_invariance(conv):
    proj = make_project_fn(time_reversal_U=SY, time_reversal_k_convention=conv)
    A = _rand(np.random.default_rng(3), (5, 4, 2, 2))
    out = proj(A)
    neg = _minus_k_mod(out) if conv == "mod" else np.flip(out, axis=(0, 1))
    np.testing.assert_allclose(SY @ np.conj(neg) @ SY.conj().T, out, atol=1e-14)
    np.testing.assert_allclose(proj(out), out, atol=1e-14)


def test_mod_is_flip_shifted_by_one():
    """-k mod n is the array flip rolled by one site along each flipped axis."""
    A = _rand(np.random.default_rng(4), (6, 5, 2, 2))
    pm = make_project_fn(time_reversal_U=np.eye(2), time_reversal_k_convention="mod")
    pf = make_project_fn(time_reversal_U=np.eye(2), time_reversal_k_convention="flip")
    np.testing.assert_allclose(pm(A), 0.5 * (A + np.conj(_minus_k_mod(A))), atol=1e-15)
    flipped = np.roll(np.flip(A, axis=(0, 1)), 1, axis=(0, 1))
    np.testing.assert_allclose(pm(A), 0.5 * (A + np.conj(flipped)), atol=1e-15)
    assert not np.allclose(pm(A), pf(A))


def test_combined_unitary_and_spatial_group_formula():
    G = np.stack([np.eye(2), SZ])
    S = SX[None]
    proj = make_project_fn(unitary_group=G, spatial_group=S, spatial_k_convention="flip")
    A = _rand(np.random.default_rng(5), (4, 4, 2, 2))
    neg = np.flip(A, axis=(0, 1))
    want = (A + SZ @ A @ SZ + SX @ neg @ SX) / 3.0
    np.testing.assert_allclose(proj(A), want, atol=1e-15)


def test_spatial_group_alone_joins_identity():
    proj = make_project_fn(spatial_group=SX[None], spatial_k_convention="flip")
    A = _rand(np.random.default_rng(6), (4, 4, 2, 2))
    out = proj(A)
    np.testing.assert_allclose(out, 0.5 * (A + SX @ np.flip(A, axis=(0, 1)) @ SX), atol=1e-15)
    np.testing.assert_allclose(SX @ np.flip(out, axis=(0, 1)) @ SX, out, atol=1e-15)


def test_single_axis_flip_idempotent():
    proj = make_project_fn(unitary_group=np.eye(2)[None], spatial_group=SX[None],
                           spatial_k_convention="flip", spatial_k_flip_axes=(0,))
    A = _rand(np.random.default_rng(7), (6, 4, 2, 2))
    once = proj(A)
    np.testing.assert_allclose(proj(once), once, atol=1e-14)
    np.testing.assert_allclose(once, 0.5 * (A + SX @ np.flip(A, axis=0) @ SX), atol=1e-15)


def test_unknown_k_convention_raises():
    proj = make_project_fn(time_reversal_U=np.eye(2), time_reversal_k_convention="mirror")
    with pytest.raises(ValueError, match="k_convention"):
        proj(np.zeros((2, 2, 2, 2), dtype=np.complex128))


def test_device_description_of_the_projector():
    # what the GPU solvers receive: groups as complex arrays, conventions as
    # flags, k axes as bits (k1 = 1, k2 = 2); axes outside (0, 1) have no
    # device form, an unknown convention raises as a call would
    G = np.stack([np.eye(2), SZ]).real
    p = make_project_fn(unitary_group=G, spatial_group=SX[None], spatial_k_convention="flip",
                        spatial_k_flip_axes=(1,), time_reversal_U=SY)
    U, S, smod, sax, T, tmod, tax = p._hfgpu_symmetry()
    assert U.dtype == np.complex128 and U.shape == (2, 2, 2)
    np.testing.assert_array_equal(U, G)
    np.testing.assert_array_equal(S, SX[None])
    assert (smod, sax, tmod, tax) == (False, 2, True, 3)
    np.testing.assert_array_equal(T, SY)
    assert make_project_fn(unitary_group=G, spatial_group=SX[None],
                           spatial_k_flip_axes=(0, 2))._hfgpu_symmetry() is None
    with pytest.raises(ValueError):
        make_project_fn(time_reversal_U=SY, time_reversal_k_convention="mirror")._hfgpu_symmetry()
