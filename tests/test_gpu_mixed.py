"""Mixed-precision Fock build (FP32 exchange workspace and interaction,
FP64 arithmetic) and the FP64 correction phase of the solvers.

A single mixed build carries the rounding of three complex-float
crossings (two workspace crossings and the interaction): Sigma within
~1e-7 relative of the FP64 oracle.  The solvers switch to FP64 builds below
the switch residual and count convergence only on FP64 builds, so their
converged states agree with the FP64 solves to the convergence tolerance.
"""

from __future__ import annotations

import numpy as np
import pytest

from helpers import kernel_args_dict, rand_herm

pytestmark = pytest.mark.gpu


def relerr(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.max(np.abs(a - b))) / max(1e-300, float(np.max(np.abs(b))))


def _kernels():
    from cpp_hf_amd.workloads import layered_kernel, qwz_kernel

    return {
        "layered96_nb8": layered_kernel(96, 4).kernel,
        "qwz48_scalar_nb2": qwz_kernel(48).kernel,
        "layered24_nb16": layered_kernel(24, 8).kernel,
        "layered20_nb4_no_two_step": layered_kernel(20, 2).kernel,
    }


@pytest.mark.parametrize("name", ["layered96_nb8", "qwz48_scalar_nb2", "layered24_nb16",
                                  "layered20_nb4_no_two_step"])
def test_mixed_build_error_bound(gpu, oracle, name):
    k = _kernels()[name]
    ka = kernel_args_dict(k)
    P = rand_herm(np.random.default_rng(1), k.h.shape) * 0.1 + 0.5 * np.eye(k.h.shape[-1])
    print(f"mixed build {name}", flush=True)
    S64, H64, F64, E64 = gpu.build_fock_apply(ka, P, None)
    S32, H32, F32, E32 = gpu.build_fock_apply(ka, P, None, exchange_fp32=True)
    So, Ho, Fo, Eo = oracle.build_fock_apply(ka, P)
    assert relerr(S64, So) < 1e-12
    np.testing.assert_array_equal(H32, H64)  # Hartree terms stay FP64
    capable = gpu.GpuKernel(ka).mixed_capable(True)
    # the two-step transforms (and so the FP32 workspace) need tiles of 8
    # channels and a mesh of the two-step table
    assert capable == (name in ("layered96_nb8", "layered24_nb16"))
    if not capable:
        np.testing.assert_array_equal(S32, S64)  # FP64 workspace here
        return
    err = relerr(S32, So)
    assert 1e-10 < err < 1e-6, err  # the FP32 path ran, within its bound
    assert relerr(F32, Fo) < 1e-6
    np.testing.assert_allclose(F32, np.conj(np.swapaxes(F32, -1, -2)), atol=0)


@pytest.mark.parametrize("accel", ["diis"])
def test_mixed_scf_matches_fp64(gpu, accel):
    from cpp_hf_amd import SCFConfig, solve_scf
    from cpp_hf_amd.workloads import layered_kernel

    wl = layered_kernel(48, 4, T=0.05, filling=4.1)
    base = dict(max_iter=400, mixing=0.5, acceleration=accel, density_tol=1e-7, comm_tol=1e-8)
    print(f"mixed scf {accel}", flush=True)
    r64 = solve_scf(wl.kernel, wl.P0, wl.n_electrons, config=SCFConfig(**base))
    r32 = solve_scf(wl.kernel, wl.P0, wl.n_electrons,
                    config=SCFConfig(exchange_precision="mixed", **base))
    assert r64.converged and r32.converged, (r64.message, r32.message)
    dE = abs(float(r32.energy) - float(r64.energy))
    dmu = abs(float(r32.chemical_potential) - float(r64.chemical_potential))
    dF = np.sqrt(np.sum(np.abs(np.asarray(r32.fock_matrix) - np.asarray(r64.fock_matrix)) ** 2,
                        axis=(-2, -1)))
    print(f"  dE {dE:.2e} dmu {dmu:.2e} max_k dF {np.max(dF):.2e} iterations "
          f"{r64.iterations} / {r32.iterations}", flush=True)
    # both states satisfy density_tol = 1e-7 (E is first order in that
    # residual), and mu is resolved to ~6e-8 by the bisection
    assert dE < 1e-7
    assert dmu < 1e-6
    assert float(np.max(dF)) < 1e-6
    # the FP32 phase ran: the early iterations differ from the FP64 run
    assert r32.history["E"][1] != r64.history["E"][1]


def test_mixed_dm_matches_fp64(gpu):
    from cpp_hf_amd import SolverConfig, solve
    from cpp_hf_amd.workloads import layered_kernel

    # the reference's CG loop stalls near |G| ~ 1e-6 on the layered models
    # (the oracle's too), so both runs take a fixed 300 iterations; the
    # mixed run switches to FP64 builds once |G| <= 1e-3
    wl = layered_kernel(24, 4, T=0.05, filling=4.1)
    base = dict(max_iter=300, tol_E=0.0, tol_grad=0.0, plateau_window=0)
    print("mixed dm", flush=True)
    r64 = solve(wl.kernel, wl.P0, wl.n_electrons, config=SolverConfig(**base))
    r32 = solve(wl.kernel, wl.P0, wl.n_electrons,
                config=SolverConfig(exchange_precision="mixed", mixed_switch=1e-3, **base))
    g32 = np.asarray(r32.history["grad_norm"])
    dE = abs(float(r32.energy) - float(r64.energy))
    print(f"  dm dE {dE:.2e} |G| {r64.history['grad_norm'][-1]:.1e} / {g32[-1]:.1e}", flush=True)
    assert np.any(g32 <= 1e-3)
    assert r32.history["E"][1] != r64.history["E"][1]  # the FP32 phase ran
    assert dE < 1e-6
    with pytest.raises(ValueError):
        solve(wl.kernel, wl.P0, wl.n_electrons, config=SolverConfig(exchange_precision="fp16"))


def test_mixed_dm_192x192_nb16_fock_within_1e8_of_fp64(gpu):
    """16 bands on 192 x 192: the mixed-precision DM solve (FP32 workspace
    until |G| <= 1e-4, then FP64) ends within 1e-8 per k point of the FP64
    solve's Fock matrix (the DM's Newton mu is resolved to rounding; the
    SCF's 30-step bisection mu is not, which leaves two FP64 SCF runs from
    different starts ~3e-7 apart)."""
    from cpp_hf_amd import SolverConfig, solve
    from cpp_hf_amd.workloads import layered_kernel

    wl = layered_kernel(192, 8, T=0.05, filling=8.1)
    base = dict(max_iter=300, tol_E=0.0, tol_grad=0.0, plateau_window=0, mixed_switch=1e-4)
    print("mixed dm 192x192 nb16", flush=True)
    r64 = solve(wl.kernel, wl.P0, wl.n_electrons, config=SolverConfig(**base))
    r32 = solve(wl.kernel, wl.P0, wl.n_electrons,
                config=SolverConfig(exchange_precision="mixed", **base))
    assert r32.history["E"][1] != r64.history["E"][1]  # the FP32 phase ran
    dF = np.sqrt(np.sum(np.abs(np.asarray(r32.fock) - np.asarray(r64.fock)) ** 2, axis=(-2, -1)))
    print(f"  max_k |dF| {np.max(dF):.2e}", flush=True)
    assert float(np.max(dF)) <= 1e-8
    assert abs(float(r32.mu) - float(r64.mu)) <= 1e-9
