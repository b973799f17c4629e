"""Coarse-to-fine continuation (reference: src/cpp_hf/continuation.py,
tests/test_continuation.py).

The validation cases run on the CPU (they fail before any solve).  The GPU
cases check the driver against the same stages run by hand, the fine solve
against a cold start on the fine mesh, and both stages against the oracle.
"""

from __future__ import annotations

import numpy as np
import pytest

from helpers import two_band_kernel, kernel_args_dict


def test_rejects_mismatched_orbital_dims():
    from cpp_hf_amd import HartreeFockKernel, solve_continuation

    kc = two_band_kernel(2)
    kf = HartreeFockKernel(np.ones((4, 4)), np.zeros((4, 4, 3, 3), np.complex128),
                           np.full((4, 4, 1, 1), 0.1), T=0.1)
    with pytest.raises(ValueError, match="orbital dimensions"):
        solve_continuation(kc, kf, np.zeros_like(kc.h), 1.0, 1.0)


def test_rejects_mismatched_hartree_flag():
    from cpp_hf_amd import HartreeFockKernel, solve_continuation

    kc = two_band_kernel(2)
    h = np.zeros((4, 4, 2, 2), np.complex128)
    kf = HartreeFockKernel(np.ones((4, 4)), h, np.full((4, 4, 1, 1), 0.1), T=0.1,
                           include_hartree=True, reference_density=np.zeros_like(h),
                           hartree_matrix=np.eye(2))
    with pytest.raises(ValueError, match="include_hartree"):
        solve_continuation(kc, kf, np.zeros_like(kc.h), 1.0, 1.0)


def test_rejects_non_square_mesh_and_bad_seed():
    from cpp_hf_amd import HartreeFockKernel, solve_continuation

    kc = two_band_kernel(2)
    kf = HartreeFockKernel(np.ones((4, 6)), np.zeros((4, 6, 2, 2), np.complex128),
                           np.full((4, 6, 1, 1), 0.1), T=0.2)
    with pytest.raises(ValueError, match="fine_kernel must have a square k-mesh"):
        solve_continuation(kc, kf, np.zeros_like(kc.h), 1.0, 1.0)
    with pytest.raises(ValueError, match="P0_coarse shape"):
        solve_continuation(kc, two_band_kernel(4), np.zeros((3, 3, 2, 2)), 1.0, 1.0)


def test_rejects_unknown_config_type():
    from cpp_hf_amd import solve_continuation

    kc, kf = two_band_kernel(2), two_band_kernel(4)
    with pytest.raises(TypeError, match="coarse_config"):
        solve_continuation(kc, kf, np.zeros_like(kc.h), 1.0, 1.0, coarse_config={"max_iter": 5})


@pytest.mark.gpu
def test_dm_stages_equal_manual_driver(gpu, oracle):
    from cpp_hf_amd import SolverConfig, resample_kgrid, solve, solve_continuation

    kc, kf = two_band_kernel(4, exchange=0.6), two_band_kernel(8, exchange=0.6)
    cfg = SolverConfig(max_iter=50, tol_E=1e-8)
    got = solve_continuation(kc, kf, np.zeros_like(kc.h), 1.0, 1.0,
                             coarse_config=cfg, fine_config=cfg)
    coarse = solve(kc, np.zeros_like(kc.h), 1.0, config=cfg)
    P = resample_kgrid(coarse.density, 8, method="linear")
    P0f = 0.5 * (P + np.conj(np.swapaxes(P, -1, -2)))
    fine = solve(kf, P0f, 1.0, config=cfg)
    np.testing.assert_array_equal(got.coarse.density, coarse.density)
    np.testing.assert_array_equal(got.P0_fine, P0f)
    np.testing.assert_array_equal(got.fine.density, fine.density)
    # the resampled seed is the oracle's index map applied to the coarse density
    Pr = oracle.resample_kgrid_2d(coarse.density.reshape(4, 4, -1), 8).reshape(8, 8, 2, 2)
    np.testing.assert_array_equal(P, Pr)


@pytest.mark.gpu
def test_fine_stage_reaches_cold_start_fixed_point(gpu, oracle):
    from cpp_hf_amd import SolverConfig, solve, solve_continuation

    kc, kf = two_band_kernel(4, exchange=0.6), two_band_kernel(8, exchange=0.6)
    cfg = SolverConfig(max_iter=200, tol_E=1e-10)
    cold = solve(kf, np.zeros_like(kf.h), 1.0, config=cfg)
    cont = solve_continuation(kc, kf, np.zeros_like(kc.h), 1.0, 1.0,
                              coarse_config=cfg, fine_config=cfg)
    assert bool(cold.converged) and bool(cont.fine.converged)
    assert abs(float(cont.fine.energy) - float(cold.energy)) < 1e-6
    np.testing.assert_allclose(cont.fine.density, cold.density, atol=1e-5)
    # the fine stage from the resampled seed is the oracle's solve from that seed
    r = oracle.solve_dm(kernel_args_dict(kf), cont.P0_fine, 1.0, 200, 1e-10, 0.0, 0.6, 0.5,
                        1e-3, 8, 10, 5, 25, 0, 0.5, 20, np.empty((0, 8, 8, 2, 2)), 0.0, 1.0,
                        [], None, True, True, True)
    assert int(cont.fine.n_iter) == r[6]
    assert abs(float(cont.fine.energy) - r[5]) < 1e-10
    np.testing.assert_allclose(cont.fine.density, r[2], atol=1e-9)


@pytest.mark.gpu
def test_scf_and_mixed_stages(gpu):
    from cpp_hf_amd import (SCFConfig, SCFResult, SolveResult, SolverConfig,
                            solve_continuation)

    kc, kf = two_band_kernel(2), two_band_kernel(4)
    scf = SCFConfig(max_iter=40, density_tol=1e-6, comm_tol=1e-5, mixing=0.5)
    both = solve_continuation(kc, kf, np.zeros_like(kc.h), 1.0, 1.0,
                              coarse_config=scf, fine_config=scf)
    assert isinstance(both.coarse, SCFResult) and isinstance(both.fine, SCFResult)
    assert both.fine.converged and both.P0_fine.shape == kf.h.shape
    mixed = solve_continuation(kc, kf, np.zeros_like(kc.h), 1.0, 1.0, coarse_config=scf,
                               fine_config=SolverConfig(max_iter=30, tol_E=1e-8),
                               retain_P0_fine=False)
    assert isinstance(mixed.coarse, SCFResult) and isinstance(mixed.fine, SolveResult)
    assert mixed.P0_fine is None
    default = solve_continuation(kc, kf, np.zeros_like(kc.h), 1.0, 1.0)
    assert isinstance(default.coarse, SolveResult) and isinstance(default.fine, SolveResult)


@pytest.mark.gpu
def test_solve_with_time_reversal_projector_matches_oracle(gpu, oracle):
    """A host project_fn from make_project_fn inside the native CG loop."""
    from cpp_hf_amd import SolverConfig, make_project_fn, solve
    from cpp_hf_amd.workloads import qwz_kernel

    wl = qwz_kernel(12)
    nb = wl.kernel.h.shape[-1]
    proj = make_project_fn(time_reversal_U=np.eye(nb), time_reversal_k_convention="mod")
    cfg = SolverConfig(max_iter=60, tol_E=1e-9, project_fn=proj)
    got = solve(wl.kernel, wl.P0, 1.0, config=cfg)
    r = oracle.solve_dm(kernel_args_dict(wl.kernel), wl.P0, 1.0, 60, 1e-9, 0.0, 0.6, 0.5, 1e-3,
                        8, 10, 5, 25, 0, 0.5, 20, np.empty((0,) + wl.kernel.h.shape), 0.0, 1.0,
                        [], proj, True, True, True)
    assert int(got.n_iter) == r[6]
    assert abs(float(got.energy) - r[5]) < 1e-10
    np.testing.assert_allclose(got.density, r[2], atol=1e-9)
