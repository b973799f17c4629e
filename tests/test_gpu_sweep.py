"""Concurrent solves (solve_many) equal the same solves run one by one."""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("method", ["dm", "scf"])
def test_solve_many_matches_sequential(gpu, method):
    from cpp_hf_amd import SCFConfig, SolverConfig, solve_many
    from cpp_hf_amd.workloads import sweep_workload

    params = [dict(n=n, nb=nb, u=0.1 * i, scale=1.0, gate=3.0 + i, offset=0.05 * (i - 2))
              for i, (n, nb) in enumerate([(12, 2), (16, 4), (24, 2), (12, 4), (20, 6), (16, 2)])]
    wls = [sweep_workload(p) for p in params]
    kernels = [w.kernel for w in wls]
    P0s = [w.P0 for w in wls]
    n_es = [w.n_electrons for w in wls]
    cfg = SolverConfig(max_iter=40) if method == "dm" else SCFConfig(max_iter=40, mixing=0.5)
    seq = solve_many(kernels, P0s, n_es, config=cfg, method=method, concurrency=1)
    par = solve_many(kernels, P0s, n_es, config=cfg, method=method, concurrency=4)
    for a, b in zip(seq, par):
        if method == "dm":
            assert a.n_iter == b.n_iter
            assert a.energy == b.energy
            np.testing.assert_array_equal(a.density, b.density)
            np.testing.assert_array_equal(a.history["E"], b.history["E"])
        else:
            assert a.iterations == b.iterations
            assert a.energy == b.energy
            np.testing.assert_array_equal(a.density_matrix, b.density_matrix)


def _assert_same_dm(a, b):
    assert int(a.n_iter) == int(b.n_iter) and bool(a.converged) == bool(b.converged)
    assert float(a.energy) == float(b.energy) and float(a.mu) == float(b.mu)
    for x, y in ((a.Q, b.Q), (a.p, b.p), (a.density, b.density), (a.fock, b.fock),
                 (a.history["E"], b.history["E"]),
                 (a.history["grad_norm"], b.history["grad_norm"])):
        np.testing.assert_array_equal(np.asarray(x), np.asarray(y))


@pytest.mark.parametrize("max_batch", [16, 2])
def test_grouped_solves_equal_single_solves(gpu, max_batch):
    """Direct-minimisation solves advanced together (one launch sequence per
    iteration for solves of one band count and temperature, meshes mixed)
    equal the same solves run alone, bit for bit, including solves that stop
    many iterations before the rest of their group and groups split into
    chunks.  The kernels are 13 of the 48-kernel sweep's (36 x 36 and
    72 x 72 with 4 bands, 36 x 36 with 8 bands: 9 to 200 iterations within
    a group)."""
    from cpp_hf_amd import SolverConfig, solve, solve_many
    from cpp_hf_amd.workloads import sweep_params, sweep_workload

    ps = sweep_params(48, seed=1)
    pick = [i for i, p in enumerate(ps) if (p["n"], p["nb"]) in ((36, 4), (36, 8), (72, 4))]
    assert len(pick) == 13
    wls = [sweep_workload(ps[i]) for i in pick]
    ks, P0s, nes = [w.kernel for w in wls], [w.P0 for w in wls], [w.n_electrons for w in wls]
    cfg = SolverConfig()
    print(f"grouped solves, max_batch={max_batch}", flush=True)
    grouped = solve_many(ks, P0s, nes, config=cfg, concurrency=3, batch=True,
                         max_batch=max_batch)
    print("the same solves one by one", flush=True)
    single = [solve(k, P, n, config=cfg) for k, P, n in zip(ks, P0s, nes)]
    iters = sorted(int(r.n_iter) for r in single)
    assert iters[0] < 20 and iters[-1] >= 100  # stops far apart inside the groups
    for a, b in zip(grouped, single):
        _assert_same_dm(a, b)


def test_grouped_solve_argument_checks(gpu):
    from cpp_hf_amd._compat import _native
    from cpp_hf_amd.solver import _kernel_args_for_native
    from cpp_hf_amd.workloads import layered_kernel

    a = layered_kernel(12, 1, T=0.05)
    b = layered_kernel(12, 2, T=0.05)
    c = layered_kernel(12, 1, T=0.1)
    args = (50, 1e-7, 0.0, 0.6, 0.5, 1e-3, 8, 10, 5, 25, False, False, False)
    for other in (b, c):  # another band count, another temperature
        with pytest.raises(ValueError):
            _native.solve_dm_batch([_kernel_args_for_native(a.kernel),
                                    _kernel_args_for_native(other.kernel)],
                                   [a.P0, other.P0], [a.n_electrons, other.n_electrons], *args)
    with pytest.raises(ValueError):
        _native.solve_dm_batch([], [], [], *args)


def test_grouped_mixed_paths_equal_single_solves(gpu):
    """One group whose solves take every per-solve path of the grouped
    iteration: two-step meshes sharing the exchange launches (12 x 12 twice,
    48 x 48), a Stockham mesh built on its own (20 x 20), the one-workgroup
    Newton mu (up to 4096 eigenvalues: 12 x 12 x 8, 20 x 20 x 8) and the
    multi-workgroup one (48 x 48 x 8), with solves that stop at different
    iterations;
    each result equals the solve run alone, bit for bit."""
    from cpp_hf_amd import SolverConfig, solve, solve_many
    from cpp_hf_amd.workloads import sweep_workload

    params = [dict(n=12, nb=8, u=0.10, scale=1.0, gate=3.0, offset=0.00),
              dict(n=20, nb=8, u=0.20, scale=0.8, gate=5.0, offset=0.10),
              dict(n=48, nb=8, u=0.05, scale=1.2, gate=4.0, offset=-0.10),
              dict(n=12, nb=8, u=0.25, scale=0.6, gate=7.0, offset=0.20)]
    wls = [sweep_workload(p) for p in params]
    ks, P0s, nes = [w.kernel for w in wls], [w.P0 for w in wls], [w.n_electrons for w in wls]
    cfg = SolverConfig(max_iter=60)
    print("mixed-path group", flush=True)
    grouped = solve_many(ks, P0s, nes, config=cfg, concurrency=1, batch=True)
    print("the same solves one by one", flush=True)
    single = [solve(k, P, n, config=cfg) for k, P, n in zip(ks, P0s, nes)]
    assert len({int(r.n_iter) for r in single}) > 1
    for a, b in zip(grouped, single):
        _assert_same_dm(a, b)


def _exchange_off(wl):
    """The workload's kernel with the exchange off (Hartree only)."""
    from cpp_hf_amd import HartreeFockKernel

    k = wl.kernel
    nb = k.h.shape[-1]
    refP = np.broadcast_to(0.5 * np.eye(nb, dtype=np.complex128), k.h.shape).copy()
    return HartreeFockKernel(k.w2d, k.h, np.zeros(k.h.shape[:2] + (1, 1)), T=k.T,
                             include_hartree=True, include_exchange=False,
                             reference_density=refP, hartree_matrix=k.HH)


@pytest.mark.parametrize("case", ["nb16", "nb20", "kinds"])
def test_grouped_band_counts_and_build_kinds_equal_single_solves(gpu, case):
    """Groups on the less common paths, each against its solves run alone,
    bit for bit: 16 bands (matrix-core density, 16-lane eigh) on two-step and
    Stockham meshes; 20 bands (one-wavefront loop kernels, 32-lane eigh with
    its rows in LDS); and one mesh with three kinds of Fock build side by
    side (Hartree + exchange, exchange alone, Hartree alone), which the
    grouped build runs as separate exchange launches or on its own."""
    from cpp_hf_amd import SolverConfig, solve, solve_many
    from cpp_hf_amd.workloads import layered_kernel

    if case == "nb16":
        wls = [layered_kernel(n, 8, T=0.05, u=0.1 + 0.02 * i, seed=i)
               for i, n in enumerate((12, 16, 20))]
        ks = [w.kernel for w in wls]
        cfg = SolverConfig(max_iter=25)
    elif case == "nb20":
        wls = [layered_kernel(n, 10, T=0.05, u=0.1 + 0.03 * i, seed=i)
               for i, n in enumerate((8, 12))]
        ks = [w.kernel for w in wls]
        cfg = SolverConfig(max_iter=15)
    else:
        wls = [layered_kernel(12, 4, T=0.05, u=0.1, seed=1),
               layered_kernel(12, 4, T=0.05, u=0.2, seed=2, hartree=False),
               layered_kernel(12, 4, T=0.05, u=0.15, seed=3)]
        ks = [wls[0].kernel, wls[1].kernel, _exchange_off(wls[2])]
        cfg = SolverConfig(max_iter=40)
    P0s, nes = [w.P0 for w in wls], [w.n_electrons for w in wls]
    print(f"grouped {case}", flush=True)
    grouped = solve_many(ks, P0s, nes, config=cfg, concurrency=1, batch=True)
    print(f"single {case}", flush=True)
    single = [solve(k, P, n, config=cfg) for k, P, n in zip(ks, P0s, nes)]
    for a, b in zip(grouped, single):
        _assert_same_dm(a, b)


def test_grouped_solve_rejects_a_repeated_handle(gpu, oracle):
    """Through the C ABI (ctypes): one kernel handle passed twice in a batch
    is refused before any work, since both solves would build in the one
    handle's workspaces."""
    import ctypes as C
    from pathlib import Path

    from cpp_hf_amd.solver import _kernel_args_for_native
    from cpp_hf_amd.workloads import layered_kernel

    lib = C.CDLL(str(Path(__file__).resolve().parents[1] / "cpp_hf_amd" / "libhfgpu.so"))
    lib.hfgpu_last_error.restype = C.c_char_p
    wl = layered_kernel(12, 1, T=0.05)
    desc = oracle._Desc(_kernel_args_for_native(wl.kernel))
    handle = C.c_void_p()
    assert lib.hfgpu_kernel_create(C.byref(desc.desc), C.byref(handle)) == 0
    try:
        P0 = np.ascontiguousarray(wl.P0, dtype=np.complex128)
        ks = (C.c_void_p * 2)(handle, handle)
        p0s = (C.c_void_p * 2)(P0.ctypes.data, P0.ctypes.data)
        nes = (C.c_double * 2)(wl.n_electrons, wl.n_electrons)
        cfg = C.create_string_buffer(1024)
        outs = C.create_string_buffer(2048)
        rc = lib.hfgpu_solve_dm_batch(ks, C.c_size_t(2), p0s, nes, cfg, outs)
        assert rc != 0
        assert b"twice" in lib.hfgpu_last_error()
    finally:
        lib.hfgpu_kernel_destroy(handle)
