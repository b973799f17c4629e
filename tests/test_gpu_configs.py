"""GPU parity at the production configurations.

- 24 x 24 two-band QWZ, scalar V (the CPU-runnable case): the default
  ``solve`` and ``solve_scf(mixing=0.5)`` run to the end.
- 96 x 96, 8 bands, layer-resolved V, Hartree and reference density:
  ``solve`` and ``solve_scf(mixing=0.5)`` run to the end.
- the 48-kernel parameter sweep (meshes 36..120, 2..12 bands, seed-1 draws):
  a fixed subset solved against the oracle, and all 48 solved concurrently
  against the same solves one by one.
- 192 x 192, 16 bands: the FP64 Fock build against the oracle and the bound
  of the FP32-workspace build.

Each GPU solve is compared with the oracle's solve from the same inputs:
same iteration count and convergence flag, energies within 1e-10, densities
within 1e-9 relative; Fock matrices within 1e-12 relative.
"""

from __future__ import annotations

import numpy as np
import pytest

from helpers import herm, kernel_args_dict

pytestmark = pytest.mark.gpu


def relerr(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.max(np.abs(a - b))) / max(1e-300, float(np.max(np.abs(b))))


def _dm_args(kern, P0, n_e):
    """_native.solve_dm arguments of the default SolverConfig (solver.py:150-220)."""
    from cpp_hf_amd import SolverConfig

    c = SolverConfig()
    return (kernel_args_dict(kern), np.ascontiguousarray(P0, dtype=np.complex128), float(n_e),
            int(c.max_iter), float(c.tol_E), float(c.tol_grad), float(c.max_step),
            float(c.bt_shrink), float(c.denom_scale), int(c.bt_max), int(c.cg_restart),
            int(c.plateau_window), int(c.mu_maxiter), 0, float(c.tr_delta0), int(c.tr_cg_max),
            np.empty((0,) + kern.h.shape, dtype=np.complex128), 0.0, 1.0, [], None, True, True,
            True)


def _scf_args(kern, P0, n_e, mixing=0.5):
    """_native.solve_scf arguments of SCFConfig(mixing=mixing) (reference_scf.py:114-165)."""
    from cpp_hf_amd import SCFConfig

    c = SCFConfig(mixing=mixing)
    return (kernel_args_dict(kern), np.ascontiguousarray(P0, dtype=np.complex128), float(n_e),
            int(c.max_iter), float(c.density_tol), float(c.comm_tol), float(c.mixing),
            float(c.level_shift), None, [], c.acceleration, int(c.diis_size), int(c.diis_start),
            float(c.diis_damping), float(c.trust_radius), True, True)


def _check_dm(g, r):
    Q, p, P, F, mu, E, n, conv, hE, hG = g
    assert n == r[6] and conv == r[7], (n, r[6], conv, r[7])
    assert abs(E - r[5]) < 1e-10 * max(1.0, abs(r[5])), (E, r[5])
    np.testing.assert_allclose(hE, r[8], rtol=1e-10, atol=1e-10)
    assert abs(mu - r[4]) < 1e-9
    assert relerr(P, r[2]) < 1e-9
    assert relerr(F, r[3]) < 1e-9


def _check_scf(g, r, T):
    """The SCF's mu is the midpoint of 30 bisection halvings of a bracket
    about 20-30 wide (resolution ~2e-8); a rounding-level difference of the
    density moves the root, and when the last midpoints fall between the two
    roots one `count > n_e` decision differs and mu moves by the resolution.
    The occupations then move by at most dmu f (1 - f) / T <= dmu / (4 T)."""
    assert g[4] == r[4] and g[5] == r[5], (g[4], r[4], g[5], r[5])
    assert abs(g[2] - r[2]) < 1e-10 * max(1.0, abs(r[2])), (g[2], r[2])
    dmu = abs(g[3] - r[3])
    assert dmu < 1e-7
    assert relerr(g[0], r[0]) < max(1e-9, 2.0 * dmu / (4 * T))
    np.testing.assert_allclose(g[6], r[6], rtol=1e-10, atol=1e-10)


# ---------------------------------------------- 24 x 24 two-band QWZ -------

def test_qwz24_default_solve_matches_oracle(gpu, oracle):
    """The public solve() with the default SolverConfig, to convergence."""
    from cpp_hf_amd import SolverConfig, solve
    from cpp_hf_amd.workloads import qwz_kernel

    wl = qwz_kernel(24)
    print("qwz24 default solve", flush=True)
    res = solve(wl.kernel, wl.P0, wl.n_electrons, config=SolverConfig())
    r = oracle.solve_dm(*_dm_args(wl.kernel, wl.P0, wl.n_electrons))
    assert bool(res.converged) and r[7]
    assert int(res.n_iter) == r[6]
    assert abs(float(res.energy) - r[5]) < 1e-10
    assert abs(float(res.mu) - r[4]) < 1e-9
    assert relerr(res.density, r[2]) < 1e-9
    np.testing.assert_allclose(res.history["E"], r[8], rtol=0, atol=1e-10)


def test_qwz24_scf_matches_oracle(gpu, oracle):
    from cpp_hf_amd.workloads import qwz_kernel

    wl = qwz_kernel(24)
    args = _scf_args(wl.kernel, wl.P0, wl.n_electrons)
    print("qwz24 scf", flush=True)
    _check_scf(gpu.solve_scf(*args), oracle.solve_scf(*args), 0.1)


# ------------------------------------------- 96 x 96, 8 layered bands ------

@pytest.fixture(scope="module")
def c2():
    from cpp_hf_amd.workloads import layered_nb8

    return layered_nb8(96)


def test_layered96_dm_solve_matches_oracle(gpu, oracle, c2):
    args = _dm_args(c2.kernel, c2.P0, c2.n_electrons)
    print("layered96 dm solve", flush=True)
    g = gpu.solve_dm(*args)
    r = oracle.solve_dm(*args)
    _check_dm(g, r)


def test_layered96_scf_solve_matches_oracle(gpu, oracle, c2):
    """solve_scf(mixing=0.5): 200 iterations, unconverged on both sides."""
    args = _scf_args(c2.kernel, c2.P0, c2.n_electrons)
    print("layered96 scf solve", flush=True)
    g = gpu.solve_scf(*args)
    r = oracle.solve_scf(*args)
    _check_scf(g, r, 0.05)


# --------------------------------------------------- the parameter sweep ----

def _sweep():
    from cpp_hf_amd.workloads import sweep_params, sweep_workload

    return [sweep_workload(p) for p in sweep_params(48, seed=1)]


# one kernel of each band count, on meshes the oracle solves in seconds
SWEEP_SUBSET = (3, 4, 5, 11, 17, 21, 29)


@pytest.mark.parametrize("idx", SWEEP_SUBSET)
def test_sweep_kernel_solve_matches_oracle(gpu, oracle, idx):
    w = _sweep()[idx]
    n, nb = w.kernel.h.shape[0], w.kernel.h.shape[-1]
    print(f"sweep kernel {idx}: {n}x{n} nb {nb}", flush=True)
    args = _dm_args(w.kernel, w.P0, w.n_electrons)
    _check_dm(gpu.solve_dm(*args), oracle.solve_dm(*args))


def test_sweep_subset_covers_every_band_count():
    from cpp_hf_amd.workloads import sweep_params

    ps = sweep_params(48, seed=1)
    assert {ps[i]["nb"] for i in SWEEP_SUBSET} == {2, 4, 8, 12}
    assert max(ps[i]["n"] for i in SWEEP_SUBSET) <= 72


def test_sweep_48_grouped_equals_sequential(gpu):
    """All 48 sweep kernels: grouped by band count and run concurrently, and
    concurrent one solve per stream, against one solve after another."""
    from cpp_hf_amd import SolverConfig, solve_many

    wls = _sweep()
    ks, P0s, nes = [w.kernel for w in wls], [w.P0 for w in wls], [w.n_electrons for w in wls]
    print("sweep 48 sequential", flush=True)
    seq = solve_many(ks, P0s, nes, config=SolverConfig(), concurrency=1)
    print("sweep 48 grouped and concurrent", flush=True)
    grouped = solve_many(ks, P0s, nes, config=SolverConfig(), concurrency=16, batch=True)
    print("sweep 48 concurrent streams", flush=True)
    par = solve_many(ks, P0s, nes, config=SolverConfig(), concurrency=16)
    for a, b, c in zip(seq, grouped, par):
        for x in (b, c):
            assert int(a.n_iter) == int(x.n_iter)
            assert float(a.energy) == float(x.energy)
            np.testing.assert_array_equal(a.density, x.density)
            np.testing.assert_array_equal(a.Q, x.Q)
            np.testing.assert_array_equal(a.history["E"], x.history["E"])


# -------------------------------------------- 192 x 192, 16 layered bands --

def test_layered192_nb16_fock_build_matches_oracle(gpu, oracle):
    from cpp_hf_amd.workloads import layered_kernel

    wl = layered_kernel(192, 8, T=0.05, filling=8.1)
    ka = kernel_args_dict(wl.kernel)
    P = herm(wl.P0)
    print("layered192 nb16 build (FP64 and FP32 workspace)", flush=True)
    S, Hh, F, E = gpu.build_fock_apply(ka, P, None)
    Sr, Hr, Fr, Er = oracle.build_fock_apply(ka, P)
    assert relerr(S, Sr) < 1e-12
    assert relerr(F, Fr) < 1e-12
    # the Hartree term HH n with n = sum_k w_k Re(P - refP)_ii: 36864 terms
    # of ~1e-6 with cancellation, so its summation order sets ~4e-13 of a
    # 1.9e-3 value; bounded on the scale of F, as the matrix it enters
    assert float(np.max(np.abs(Hh - Hr))) < 1e-12 * float(np.max(np.abs(Fr)))
    assert abs(E - Er) <= 1e-12 * max(1.0, abs(Er))
    np.testing.assert_array_equal(F, np.conj(np.swapaxes(F, -1, -2)))
    S32, H32, F32, E32 = gpu.build_fock_apply(ka, P, None, exchange_fp32=True)
    err = relerr(S32, Sr)
    assert 1e-10 < err < 1e-6, err  # the FP32 workspace ran, within its bound
    dF = np.sqrt(np.sum(np.abs(F32 - Fr) ** 2, axis=(-2, -1)))
    assert float(np.max(dF)) < 1e-6
