"""Pins of the oracle's Fock-build adjoint (oracle/adjoint.py), CPU only.

The Fock build is affine in the density P and linear in h, the interaction
coulomb_q and the Hartree matrix, so central differences of the oracle's
own build_fock_apply along random directions are exact up to rounding: each
gradient of the adjoint must reproduce them to ~1e-12.  Kernels cover the
scalar (Hermitian-packed) and dense interactions, a k-dependent and a
constant reference density, the Hartree term and contact terms.
"""

from __future__ import annotations

import numpy as np
import pytest

from helpers import herm, kernel_args_dict, rand_c, rand_herm

from oracle import adjoint as adj


def _model(kind, seed=3):
    """Constructor arguments of a HartreeFockKernel, so Vq, h and HH can be perturbed."""
    rng = np.random.default_rng(seed)
    if kind == "scalar":
        n, nb = 6, 2
        kw = dict(T=0.1)
        Vq = np.abs(rng.standard_normal((n, n, 1, 1))) + 0.5
    elif kind == "dense_hartree":
        n, nb = 5, 3
        A = rng.standard_normal((n, n, nb, nb))
        A = A + np.swapaxes(A, -1, -2)
        refl = (-np.arange(n)) % n
        Vq = 0.1 * (A + A[refl][:, refl])
        HH = rng.standard_normal((nb, nb))
        kw = dict(T=0.1, include_hartree=True, reference_density=0.1 * rand_herm(rng, (n, n, nb, nb)),
                  hartree_matrix=0.5 * (HH + HH.T))
    elif kind == "contact":
        n, nb = 4, 2
        Vq = np.full((n, n, 1, 1), 0.1)
        Oi = rand_herm(rng, (nb, nb))
        Oj = rand_c(rng, (nb, nb))
        kw = dict(T=0.05, contact_terms=[(0.3, Oi, Oj), (-0.2, Oj, Oi)])
    elif kind == "layered":
        from cpp_hf_amd.workloads import layered_kernel

        wl = layered_kernel(6, 2, T=0.05)
        k = wl.kernel
        return dict(weights=k.w2d, hamiltonian=k.h,
                    coulomb_q=_layered_vq(6, 2), T=0.05, include_hartree=True,
                    reference_density=k._refP, hartree_matrix=k.HH)
    else:
        raise ValueError(kind)
    w = rng.uniform(0.5, 1.5, (n, n)) / (n * n)
    h = rand_herm(rng, (n, n, nb, nb))
    return dict(weights=w, hamiltonian=h, coulomb_q=Vq, **kw)


def _layered_vq(n, nl):
    from cpp_hf_amd.workloads import kgrid, _v_of_q

    nb = 2 * nl
    k = kgrid(n)
    KX, KY = np.meshgrid(k, k, indexing="ij")
    layer = np.arange(nb) // 2
    dl = np.abs(layer[:, None] - layer[None, :]).astype(np.float64)
    absq = np.sqrt(KX ** 2 + KY ** 2)
    base, nz = _v_of_q(absq, 0.8, 4.0)
    Vq = base[..., None, None] * np.exp(-0.3 * absq[..., None, None] * dl[None, None])
    Vq[~nz] = 2.0 * np.pi * 0.8 * 4.0
    return Vq


def _kernel(args):
    from cpp_hf_amd import HartreeFockKernel

    return HartreeFockKernel(**args)


KINDS = ["scalar", "dense_hartree", "contact", "layered"]


def _loss(oracle, ka, P, cots):
    S, H, F, _ = oracle.build_fock_apply(ka, P)
    Gs, Gh, GF = cots
    return float(np.real(np.vdot(Gs, S) + np.vdot(Gh, H) + np.vdot(GF, F)))


def _setup(kind, seed=5):
    args = _model(kind)
    k = _kernel(args)
    rng = np.random.default_rng(seed)
    shape = k.h.shape
    P = rand_herm(rng, shape)
    cots = (rand_c(rng, shape), rand_c(rng, shape), rand_c(rng, shape))
    return args, k, P, cots, rng


def _close(fd, an):
    assert abs(fd - an) <= 1e-10 * max(1.0, abs(an)), (fd, an)


@pytest.mark.parametrize("kind", KINDS)
def test_oracle_vjp_density(oracle, kind):
    args, k, P, cots, rng = _setup(kind)
    ka = kernel_args_dict(k)
    gP, _, _, _ = adj.build_fock_vjp(ka, P, *cots)
    for _ in range(3):
        dP = rand_herm(rng, P.shape)
        fd = (_loss(oracle, ka, P + dP, cots) - _loss(oracle, ka, P - dP, cots)) / 2.0
        _close(fd, float(np.real(np.vdot(gP, dP))))


@pytest.mark.parametrize("kind", KINDS)
def test_oracle_vjp_interaction_h_and_hartree_matrix(oracle, kind):
    args, k, P, cots, rng = _setup(kind)
    _, gVq, gh, gHH = adj.build_fock_vjp(kernel_args_dict(k), P, *cots)

    def loss_with(**over):
        a = dict(args)
        a.update(over)
        return _loss(oracle, kernel_args_dict(_kernel(a)), P, cots)

    Vq = np.asarray(args["coulomb_q"], dtype=np.float64)
    dV = rng.standard_normal(Vq.shape)
    fd = (loss_with(coulomb_q=Vq + dV) - loss_with(coulomb_q=Vq - dV)) / 2.0
    _close(fd, float(np.sum(np.real(gVq) * dV)))

    h = np.asarray(args["hamiltonian"])
    dh = rand_herm(rng, h.shape)
    fd = (loss_with(hamiltonian=h + dh) - loss_with(hamiltonian=h - dh)) / 2.0
    _close(fd, float(np.real(np.vdot(gh, dh))))

    if args.get("include_hartree"):
        HH = np.asarray(args["hartree_matrix"])
        dH = rng.standard_normal(HH.shape)
        fd = (loss_with(hartree_matrix=HH + dH) - loss_with(hartree_matrix=HH - dH)) / 2.0
        _close(fd, float(np.sum(gHH * dH)))
    else:
        assert not np.any(gHH)


def test_oracle_vjp_is_the_transpose_of_the_exchange(oracle):
    # <G, Sigma_x[X]> = <A^H G, X> for non-Hermitian X and G on the full
    # (dense) path: the exchange adjoint is the convolution with conj(VR)
    rng = np.random.default_rng(9)
    k = _kernel(_model("dense_hartree"))
    ka = dict(kernel_args_dict(k))
    ka.update(include_hartree=False, has_refP=False)
    X = rand_c(rng, k.h.shape)
    G = rand_c(rng, k.h.shape)
    sx = oracle.selfenergy_fft_full(ka["VR"], X, apply_ifftshift=False,
                                    hermitian_channel_packing=False)
    gP, _, _, _ = adj.build_fock_vjp(ka, np.zeros_like(X), cot_sigma=G)
    lhs = np.vdot(G, sx)
    rhs = np.vdot(gP, X)
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)


def test_layered_tangents_match_builder_differences():
    # the closed-form model tangents used for the free-energy parameter
    # gradients against central differences of the model builder (through
    # VR for the interaction, whose transform is linear)
    from cpp_hf_amd.problem import _shifted_interaction
    from cpp_hf_amd.workloads import layered_kernel, layered_tangents

    base = dict(u=0.1, gamma1=0.4, scale=0.8, gate=4.0)
    tan = layered_tangents(12, 3, **base)
    e = 1e-6
    for name, (dh, dV, dHH) in tan.items():
        kp = layered_kernel(12, 3, **{**base, name: base[name] + e}).kernel
        km = layered_kernel(12, 3, **{**base, name: base[name] - e}).kernel
        np.testing.assert_allclose((kp.h - km.h) / (2 * e), dh, atol=1e-8)
        np.testing.assert_allclose((kp.HH - km.HH) / (2 * e), dHH, rtol=1e-8)
        dVR = (kp._VR_shifted - km._VR_shifted) / (2 * e)
        pred = _shifted_interaction(kp.weights_b, dV, np.complex128)
        assert np.max(np.abs(dVR - pred)) <= 1e-7 * max(1.0, np.max(np.abs(pred)))
