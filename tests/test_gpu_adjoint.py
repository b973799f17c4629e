"""GPU adjoint of the Fock build (libhfgpu fock_vjp) against the oracle's
numpy adjoint (pinned by finite differences in test_oracle_adjoint.py), the
C ABI through ctypes, the torch autograd wrapper, and free-energy parameter
gradients at a converged 144 x 144, 8-band state against central finite
differences of full GPU re-solves.
"""

from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np
import pytest

from helpers import kernel_args_dict, rand_c, rand_herm
from test_oracle_adjoint import KINDS, _kernel, _model

from oracle import adjoint as adj

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]


def relerr(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.max(np.abs(a - b))) / max(1e-300, float(np.max(np.abs(b))))


def _workload_kernels():
    from cpp_hf_amd.workloads import layered_kernel, qwz_kernel

    return {
        "qwz24": qwz_kernel(24).kernel,
        "layered24_nb8": layered_kernel(24, 4).kernel,
        "layered12_nb12": layered_kernel(12, 6).kernel,
        "layered20_nb6_generic_mesh": layered_kernel(20, 3).kernel,
    }


CASES = KINDS + ["qwz24", "layered24_nb8", "layered12_nb12", "layered20_nb6_generic_mesh"]


def _get_kernel(name):
    return _kernel(_model(name)) if name in KINDS else _workload_kernels()[name]


@pytest.mark.parametrize("name", CASES)
def test_fock_vjp_matches_oracle(gpu, name):
    from cpp_hf_amd import build_fock_vjp

    print(f"fock vjp {name}", flush=True)
    k = _get_kernel(name)
    rng = np.random.default_rng(2)
    P = rand_herm(rng, k.h.shape)
    cots = [rand_c(rng, k.h.shape) for _ in range(3)]
    g = build_fock_vjp(k, P, *cots)
    r = adj.build_fock_vjp(kernel_args_dict(k), P, *cots)
    assert relerr(g.density, r[0]) < 1e-12
    rv = np.real(r[1]) if k.coulomb_q_is_real else r[1]
    assert relerr(g.coulomb_q, rv) < 1e-12
    assert relerr(g.hamiltonian, r[2]) < 1e-14
    if np.any(r[3]):
        assert relerr(g.hartree_matrix, r[3]) < 1e-12
    else:
        assert not np.any(g.hartree_matrix)
    # single cotangents: the NULL paths of the C ABI
    for i in range(3):
        one = [None, None, None]
        one[i] = cots[i]
        g1 = build_fock_vjp(k, P, *one)
        r1 = adj.build_fock_vjp(kernel_args_dict(k), P, *one)
        assert relerr(g1.density, r1[0]) < 1e-12


def test_fock_vjp_through_ctypes_c_abi(gpu, oracle):
    """hfgpu_kernel_create + hfgpu_build_fock_vjp bound with ctypes."""
    lib = C.CDLL(str(ROOT / "cpp_hf_amd" / "libhfgpu.so"))
    dp = C.POINTER(C.c_double)
    k = _get_kernel("dense_hartree")
    ka = kernel_args_dict(k)
    desc = oracle._Desc(ka)
    handle = C.c_void_p()
    assert lib.hfgpu_kernel_create(C.byref(desc.desc), C.byref(handle)) == 0
    try:
        rng = np.random.default_rng(4)
        P = np.ascontiguousarray(rand_herm(rng, k.h.shape))
        G = np.ascontiguousarray(rand_c(rng, k.h.shape))
        gP = np.empty_like(P)
        gV = np.empty(ka["VR"].shape, np.complex128)
        gh = np.empty_like(P)
        gHH = np.empty(k.h.shape[-2:])
        rc = lib.hfgpu_build_fock_vjp(handle, P.ctypes.data_as(dp), None, None,
                                      G.ctypes.data_as(dp), gP.ctypes.data_as(dp),
                                      gV.ctypes.data_as(dp), gh.ctypes.data_as(dp),
                                      gHH.ctypes.data_as(dp))
        assert rc == 0
    finally:
        lib.hfgpu_kernel_destroy(handle)
    r = adj.build_fock_vjp(ka, P, cot_F=G)
    assert relerr(gP, r[0]) < 1e-12
    assert relerr(gV, r[1]) < 1e-12
    assert relerr(gHH, r[3]) < 1e-12


_AUTOGRAD_CHILD = r"""
import sys
import numpy as np
import torch  # first: torch and libhfgpu share libamdhip64.so.7, the first loaded serves both
sys.path.insert(0, sys.argv[1])
sys.path.insert(0, sys.argv[1] + "/tests")
from cpp_hf_amd.adjoint import fock_autograd
from cpp_hf_amd.workloads import layered_kernel
from helpers import rand_c, rand_herm

k = layered_kernel(24, 4).kernel
rng = np.random.default_rng(6)
P = rand_herm(rng, k.h.shape)
G = rand_c(rng, k.h.shape)
FockBuild = fock_autograd()
out = {}
for dev in ("cuda", "cpu"):
    Pt = torch.tensor(P, dtype=torch.complex128, device=dev, requires_grad=True)
    Gt = torch.tensor(G, dtype=torch.complex128, device=dev)
    F = FockBuild.apply(Pt, k)
    torch.real(torch.sum(torch.conj(Gt) * F)).backward()
    out[dev] = Pt.grad.detach().cpu().numpy()
np.savez(sys.argv[2], P=P, G=G, cuda=out["cuda"], cpu=out["cpu"])
print("autograd child done", flush=True)
"""


def test_fock_autograd_matches_oracle(gpu, tmp_path):
    """The torch.autograd.Function over the Fock build, on a GPU tensor (device
    pointers, torch's stream) and a CPU tensor, in a child process that
    imports torch first."""
    import subprocess
    import sys

    print("torch autograd through the Fock build (child process)", flush=True)
    out = tmp_path / "grads.npz"
    r = subprocess.run([sys.executable, "-c", _AUTOGRAD_CHILD, str(ROOT), str(out)],
                       capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    d = np.load(out)
    k = _get_kernel("layered24_nb8")
    ref = adj.build_fock_vjp(kernel_args_dict(k), d["P"], cot_F=d["G"])
    assert relerr(d["cuda"], ref[0]) < 1e-12
    assert relerr(d["cpu"], ref[0]) < 1e-12


def _omega(k, P, mu, n_e):
    """Free energy E - T S at a density P (occupations = eigenvalues of P),
    moved to the electron count n_e to first order: the solver's P carries
    the count n_e + dN of its chemical-potential resolution, and
    dOmega/dN = mu."""
    from cpp_hf_amd import build_fock, hf_energy

    S, H, F = build_fock(P, h=k.h, VR=k._VR_shifted,
                         refP=k._refP if k._refP is not None else k._zero_refP, HH=k.HH,
                         w2d=k.w2d, include_exchange=k.include_exchange,
                         include_hartree=k.include_hartree,
                         exchange_hermitian_channel_packing=k.exchange_hermitian_channel_packing)
    E = float(hf_energy(P, h=k.h, Sigma=S, H=H, weights_b=k.weights_b, refP=k._refP))
    f = np.clip(np.linalg.eigvalsh(0.5 * (P + np.conj(np.swapaxes(P, -1, -2)))), 1e-14,
                1 - 1e-14)
    wn = np.asarray(k.w2d) / np.sum(k.w2d)
    Sent = -np.sum(wn[..., None] * (f * np.log(f) + (1 - f) * np.log1p(-f)))
    N = float(np.sum(np.asarray(k.w2d)[..., None] * np.real(np.einsum("...aa->...a", P))))
    return E - max(k.T, 1e-14) * Sent - mu * (N - n_e)


def test_free_energy_parameter_gradients_144(gpu):
    """dOmega/d(u, scale, screening length, gamma1) at the converged 144 x 144,
    8-band state by the adjoint, against central differences (step 1e-5) of
    full GPU re-solves."""
    from cpp_hf_amd import SCFConfig, free_energy_derivative, solve_scf
    from cpp_hf_amd.workloads import layered_kernel, layered_tangents

    base = dict(u=0.1, gamma1=0.4, scale=0.8, gate=4.0)
    # the density residual floors at ~1e-8, the resolution of the 30-level
    # chemical-potential bisection over a ~60-wide bracket; the commutator
    # residual reaches ~5e-10.  The free energy is stationary in the density
    # at fixed count, and the count error is corrected to first order
    cfg = SCFConfig(max_iter=300, mixing=0.5, acceleration="diis", density_tol=1e-7,
                    comm_tol=1e-8)
    n_e = 4.1

    def converged(**kw):
        p = dict(base)
        p.update(kw)
        wl = layered_kernel(144, 4, T=0.05, filling=n_e, **p)
        r = solve_scf(wl.kernel, wl.P0, wl.n_electrons, config=cfg)
        assert r.converged, (p, r.message)
        return wl.kernel, np.asarray(r.density_matrix), float(r.chemical_potential)

    print("converged base state 144 x 144 nb 8", flush=True)
    k0, P0, _ = converged()
    tangents = layered_tangents(144, 4, **base)
    h = 1e-5
    for name, (dh, dV, dHH) in tangents.items():
        adjoint = free_energy_derivative(k0, P0, dh=dh, dVq=dV, dHH=dHH)
        print(f"finite difference {name}", flush=True)
        kp, Pp, mup = converged(**{name: base[name] + h})
        km, Pm, mum = converged(**{name: base[name] - h})
        fd = (_omega(kp, Pp, mup, n_e) - _omega(km, Pm, mum, n_e)) / (2 * h)
        print(f"  {name}: adjoint {adjoint:.12e} fd {fd:.12e}", flush=True)
        assert abs(adjoint - fd) <= 1e-6 * max(abs(fd), 1e-3), (name, adjoint, fd)


def test_fock_vjp_144_seed2_cotangent(gpu):
    """VJPs of build_fock with respect to P and coulomb_q at 144 x 144, 8
    bands, for a seed-2 Gaussian Hermitian cotangent on F, against central
    differences (step 1e-5) of GPU Fock builds and against the oracle."""
    from cpp_hf_amd import HartreeFockKernel, build_fock_vjp
    from cpp_hf_amd.solver import _kernel_args_for_native
    from cpp_hf_amd._compat import _native
    from cpp_hf_amd.workloads import layered_kernel
    from test_oracle_adjoint import _layered_vq

    wl = layered_kernel(144, 4, T=0.05, filling=4.1)
    k = wl.kernel
    rng = np.random.default_rng(2)
    G = rand_herm(rng, k.h.shape)
    P = rand_herm(np.random.default_rng(0), k.h.shape) * 0.05 + 0.5 * np.eye(8)
    print("vjp 144 x 144 nb 8", flush=True)
    g = build_fock_vjp(k, P, cot_F=G)
    r = adj.build_fock_vjp(kernel_args_dict(k), P, cot_F=G)
    assert relerr(g.density, r[0]) < 1e-12
    assert relerr(g.coulomb_q, np.real(r[1])) < 1e-12

    def loss(kern, Pm):
        _, _, F, _ = _native.build_fock_apply(_kernel_args_for_native(kern), Pm, None)
        return float(np.real(np.vdot(G, F)))

    # the build is linear in P and in coulomb_q, so the differences are exact
    # up to rounding; at step 1e-5 the rounding of the 1.3M-term loss
    # (~1e-12) sets a floor of ~1e-7 relative: the bar is 1e-6
    step = 1e-5
    dP = rand_herm(np.random.default_rng(11), k.h.shape)
    fd = (loss(k, P + step * dP) - loss(k, P - step * dP)) / (2 * step)
    an = float(np.real(np.vdot(g.density, dP)))
    assert abs(fd - an) <= 1e-6 * abs(an), (fd, an)

    Vq = _layered_vq(144, 4)
    dV = np.random.default_rng(12).standard_normal(Vq.shape)

    def kern_with(V):
        return HartreeFockKernel(k.w2d, k.h, V, T=k.T, include_hartree=True,
                                 reference_density=k._refP, hartree_matrix=k.HH)

    fd = (loss(kern_with(Vq + step * dV), P) - loss(kern_with(Vq - step * dV), P)) / (2 * step)
    an = float(np.sum(g.coulomb_q * dV))
    assert abs(fd - an) <= 1e-6 * abs(an), (fd, an)


@pytest.mark.parametrize("name", ["u", "scale", "gate", "gamma1"])
def test_gpu_adjoint_matches_oracle_resolves(gpu, oracle, name):
    """The GPU adjoint's dOmega/dtheta at the oracle's converged 12 x 12,
    4-band state against central differences of the oracle's own SCF
    re-solves (the differences and the oracle adjoint are pinned against
    each other in test_oracle_fd_resolve.py) and against the oracle adjoint
    at the same state."""
    import test_oracle_fd_resolve as fdr
    from cpp_hf_amd import free_energy_derivative
    from cpp_hf_amd.workloads import layered_tangents

    k0, P0, _ = fdr.oracle_converged(oracle)
    dh, dVq, dHH = layered_tangents(fdr.N_MESH, fdr.N_LAYERS, **fdr.BASE)[name]
    print(f"gpu adjoint vs oracle re-solves: {name}", flush=True)
    g = free_energy_derivative(k0, P0, dh=dh, dVq=dVq, dHH=dHH)
    o = fdr.oracle_derivative(k0, P0, dh, dVq, dHH)
    fd = fdr.central_differences(oracle, name)
    assert abs(g - o) <= 1e-11 * max(abs(o), 1e-3), (name, g, o)
    assert abs(g - fd) <= 1e-6 * max(abs(fd), 1e-3), (name, g, fd)
