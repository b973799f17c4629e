"""Device symmetry projector (libhfgpu SymPlan, hf_symmetry_desc) against
make_project_fn's numpy averaging (whose properties test_symmetry.py pins:
invariance, idempotence, Hermiticity, the Z2 closed form, mod vs flip), and
solves that project on the device against the oracle's solves calling the
same projector as a host callback.
"""

from __future__ import annotations

import numpy as np
import pytest

from helpers import kernel_args_dict, rand_c

from cpp_hf_amd.symmetry import make_project_fn

pytestmark = pytest.mark.gpu

SZ = np.diag([1.0, -1.0]).astype(np.complex128)
SX = np.array([[0, 1], [1, 0]], dtype=np.complex128)


def relerr(a, b):
    return float(np.max(np.abs(a - b))) / max(1e-300, float(np.max(np.abs(b))))


def _unitaries(rng, n, nb):
    out = []
    for _ in range(n):
        q, _ = np.linalg.qr(rand_c(rng, (nb, nb)))
        out.append(q)
    return np.stack(out)


# (nk1, nk2, nb, unitary, spatial, spatial conv/axes, TR conv/axes)
CASES = [
    (4, 4, 2, 2, 0, None, None),
    (5, 3, 4, 3, 0, None, ("mod", (0, 1))),
    (4, 6, 4, 0, 2, ("flip", (0, 1)), None),
    (6, 4, 8, 4, 2, ("mod", (0,)), ("flip", (1,))),
    (3, 5, 6, 0, 0, None, ("mod", (1,))),
    (4, 4, 17, 2, 1, ("mod", (0, 1)), ("mod", (0, 1))),
    (3, 3, 32, 2, 2, ("flip", (1,)), ("mod", (0, 1))),
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_device_projector_matches_numpy(gpu, case):
    n1, n2, nb, nu, ns, sp, tr = CASES[case]
    print(f"symmetry projector case {CASES[case]}", flush=True)
    rng = np.random.default_rng(case)
    kw = {}
    if nu:
        kw["unitary_group"] = _unitaries(rng, nu, nb)
    if ns:
        kw["spatial_group"] = _unitaries(rng, ns, nb)
        kw["spatial_k_convention"], kw["spatial_k_flip_axes"] = sp
    if tr:
        kw["time_reversal_U"] = _unitaries(rng, 1, nb)[0]
        kw["time_reversal_k_convention"], kw["time_reversal_k_flip_axes"] = tr
    proj = make_project_fn(**kw)
    A = rand_c(rng, (n1, n2, nb, nb))
    U, S, smod, sax, T, tmod, tax = proj._hfgpu_symmetry()
    got = gpu.project_symmetry(A, U, S, smod, sax, T, tmod, tax)
    want = proj(A)
    assert relerr(got, want) < 1e-14


def _layered():
    from cpp_hf_amd.workloads import layered_kernel

    return layered_kernel(12, 2, T=0.05, filling=2.1)


def _projectors(nb):
    G = np.stack([np.eye(nb), np.kron(np.eye(nb // 2), SZ)]).astype(np.complex128)
    return {
        "unitary": make_project_fn(unitary_group=G),
        "tr_mod": make_project_fn(time_reversal_U=np.eye(nb)),
        "spatial_flip_tr": make_project_fn(unitary_group=G, spatial_group=G,
                                           spatial_k_convention="flip",
                                           time_reversal_U=np.eye(nb),
                                           time_reversal_k_convention="flip"),
    }


@pytest.mark.parametrize("name,accel,iters", [("tr_mod", "diis", 80), ("unitary", "linear", 20),
                                              ("spatial_flip_tr", "linear", 20)])
def test_solve_scf_device_projector_matches_oracle(gpu, oracle, name, accel, iters):
    w = _layered()
    pf = _projectors(w.kernel.h.shape[-1])[name]
    args = (kernel_args_dict(w.kernel), w.P0, w.n_electrons, iters, 1e-9, 1e-8, 0.5, 0.0, pf, [],
            accel, 6, 2, 1.0, 0.0, True, True)
    print(f"symmetry solve_scf {name} {accel}", flush=True)
    g = gpu.solve_scf(*args)
    r = oracle.solve_scf(*args)
    assert g[4] == r[4] and g[5] == r[5]
    np.testing.assert_allclose(g[6], r[6], rtol=0, atol=1e-9)
    assert abs(g[2] - r[2]) < 1e-9
    assert relerr(g[0], r[0]) < 1e-8
    assert relerr(g[1], r[1]) < 1e-8
    # the same solve through the host callback (a plain function hides the
    # device form) agrees to rounding
    h = gpu.solve_scf(*(args[:8] + (lambda A: pf(A),) + args[9:]))
    assert h[4] == g[4]
    assert abs(h[2] - g[2]) < 1e-10
    assert relerr(h[0], g[0]) < 1e-9


@pytest.mark.parametrize("name", ["unitary", "tr_mod"])
def test_solve_dm_device_projector_matches_oracle(gpu, oracle, name):
    w = _layered()
    k = w.kernel
    pf = _projectors(k.h.shape[-1])[name]
    args = (kernel_args_dict(k), w.P0, w.n_electrons, 60, 1e-9, 0.0, 0.6, 0.5, 1e-3, 8, 10, 5,
            25, 0, 0.5, 20, np.empty((0,) + k.h.shape), 0.0, 1.0, [], pf, True, True, True)
    print(f"symmetry solve_dm {name}", flush=True)
    g = gpu.solve_dm(*args)
    r = oracle.solve_dm(*args)
    assert g[7] and r[7] and g[6] == r[6]
    np.testing.assert_allclose(g[8], r[8], rtol=0, atol=1e-10)
    assert abs(g[5] - r[5]) < 1e-10
    assert relerr(g[2], r[2]) < 1e-8


def test_device_projector_shape_mismatch_raises(gpu):
    w = _layered()
    bad = make_project_fn(time_reversal_U=np.eye(3))
    args = (kernel_args_dict(w.kernel), w.P0, w.n_electrons, 5, 1e-9, 1e-8, 0.5, 0.0, bad, [],
            "linear", 6, 2, 1.0, 0.0, True, True)
    with pytest.raises(ValueError, match="nb"):
        gpu.solve_scf(*args)
