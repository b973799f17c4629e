"""Solver iterations replayed from captured hipGraphs against the same
iterations launched one kernel at a time, and the SCF loop that enqueues
each iteration before the previous one's convergence test against the loop
that tests first.

After its first iteration each solver captures its per-iteration launch
sequence once per variant (buffer parity, CG restart, precision) and replays
it; the replays run the same kernels on the same buffers in the same order,
so every iterate, and with it the final state, must equal the eager run's bit
for bit.  Each solve replays its graphs many times, with host read-backs and
eager launches in between.
"""

from __future__ import annotations

import os

import numpy as np
import pytest

from helpers import kernel_args_dict

pytestmark = pytest.mark.gpu


def _dm_args(k, P0, n_e, max_iter):
    # tol_E = -1: the plateau test never stops the run before max_iter
    return (kernel_args_dict(k), P0, n_e, max_iter, -1.0, 0.0, 0.6, 0.5, 1e-3, 8, 10, 5, 25, 0,
            0.5, 20, np.empty((0,) + k.h.shape), 0.0, 1.0, [], None, True, True, True)


def _with_env(name, on, fn):
    old = os.environ.get(name)
    os.environ[name] = "1" if on else "0"
    try:
        return fn()
    finally:
        if old is None:
            del os.environ[name]
        else:
            os.environ[name] = old


def _with_graphs(on, fn):
    return _with_env("HFGPU_SOLVER_GRAPH", on, fn)


@pytest.mark.parametrize("n,nl", [(96, 4), (24, 6), (36, 1)])
def test_dm_graph_replay_equals_eager(gpu, n, nl):
    from cpp_hf_amd.workloads import layered_kernel

    w = layered_kernel(n, nl, T=0.05, filling=nl + 0.1)
    args = _dm_args(w.kernel, w.P0, w.n_electrons, 40)
    print(f"dm graph replay {n}x{n} nb {2 * nl}", flush=True)
    g = _with_graphs(True, lambda: gpu.solve_dm(*args))
    e = _with_graphs(False, lambda: gpu.solve_dm(*args))
    assert g[6] == e[6] == 40  # both buffer parities and CG restarts replayed
    for a, b in zip(g[:6], e[:6]):
        np.testing.assert_array_equal(np.asarray(a), np.asarray(b))
    np.testing.assert_array_equal(g[8], e[8])
    np.testing.assert_array_equal(g[9], e[9])


@pytest.mark.parametrize("accel", ["linear", "diis"])
def test_scf_graph_replay_equals_eager(gpu, accel):
    from cpp_hf_amd.workloads import layered_kernel

    w = layered_kernel(48, 4, T=0.05, filling=4.1)
    args = (kernel_args_dict(w.kernel), w.P0, w.n_electrons, 30, 1e-9, 1e-8, 0.5, 0.0, None, [],
            accel, 6, 2, 1.0, 0.0, True, True)
    print(f"scf graph replay {accel}", flush=True)
    g = _with_graphs(True, lambda: gpu.solve_scf(*args))
    e = _with_graphs(False, lambda: gpu.solve_scf(*args))
    assert g[4] == e[4] and g[4] > 5
    for a, b in zip(g, e):
        if a is not None:
            np.testing.assert_array_equal(np.asarray(a), np.asarray(b))


@pytest.mark.parametrize("graphs", [True, False])
@pytest.mark.parametrize("case", ["converges", "max_iter", "one_iteration"])
def test_scf_pipelined_equals_synchronous(gpu, case, graphs):
    """Linear mixing runs iteration j + 1 ahead of iteration j's test; a
    converged test discards it.  Converging before max_iter (a speculative
    iteration is thrown away), stopping at max_iter (none is enqueued past
    it) and max_iter = 1 must all equal the synchronous loop bit for bit."""
    from cpp_hf_amd.workloads import layered_kernel, qwz_kernel

    if case == "converges":
        w = qwz_kernel(24)
        max_iter, tol = 200, (1e-7, 1e-6)
    else:
        w = layered_kernel(48, 4, T=0.05, filling=4.1)
        max_iter, tol = (30 if case == "max_iter" else 1), (1e-12, 1e-12)
    args = (kernel_args_dict(w.kernel), w.P0, w.n_electrons, max_iter, tol[0], tol[1], 0.5, 0.0,
            None, [], "linear", 6, 2, 1.0, 0.0, True, True)
    print(f"scf pipelined {case} graphs={graphs}", flush=True)
    p = _with_graphs(graphs, lambda: _with_env("HFGPU_SCF_PIPELINE", True,
                                                lambda: gpu.solve_scf(*args)))
    s = _with_graphs(graphs, lambda: _with_env("HFGPU_SCF_PIPELINE", False,
                                                lambda: gpu.solve_scf(*args)))
    assert p[4] == s[4] and p[5] == s[5]
    if case == "converges":
        assert p[5] and 1 < p[4] < max_iter
    else:
        assert not p[5] and p[4] == max_iter
    for a, b in zip(p, s):
        if a is not None:
            np.testing.assert_array_equal(np.asarray(a), np.asarray(b))


@pytest.mark.parametrize("graphs", [True, False])
@pytest.mark.parametrize("case", ["converges", "max_iter", "one_iteration", "layered_converges"])
def test_dm_pipelined_equals_synchronous(gpu, case, graphs):
    """Direct minimisation enqueues iteration k + 1 before iteration k's stop
    test; a stop discards it.  Stopping on the energy test before max_iter,
    at max_iter and after one iteration must all equal the test-first loop
    bit for bit (orbitals, occupations, density, Fock matrix, histories)."""
    from cpp_hf_amd import SolverConfig
    from cpp_hf_amd.workloads import layered_kernel, qwz_kernel

    if case == "converges":
        w = qwz_kernel(24)
        c = SolverConfig()
        args = (kernel_args_dict(w.kernel), w.P0, w.n_electrons, int(c.max_iter), float(c.tol_E),
                float(c.tol_grad), float(c.max_step), float(c.bt_shrink), float(c.denom_scale),
                int(c.bt_max), int(c.cg_restart), int(c.plateau_window), int(c.mu_maxiter), 0,
                float(c.tr_delta0), int(c.tr_cg_max), np.empty((0,) + w.kernel.h.shape), 0.0,
                1.0, [], None, True, True, True)
    elif case == "layered_converges":
        w = layered_kernel(48, 4, T=0.05, filling=4.1)
        args = _dm_args(w.kernel, w.P0, w.n_electrons, 300)
        args = args[:4] + (1e-9,) + args[5:]
    else:
        w = layered_kernel(36, 2, T=0.05, filling=2.1)
        args = _dm_args(w.kernel, w.P0, w.n_electrons, 40 if case == "max_iter" else 1)
    print(f"dm pipelined {case} graphs={graphs}", flush=True)
    p = _with_graphs(graphs, lambda: _with_env("HFGPU_DM_PIPELINE", True,
                                                lambda: gpu.solve_dm(*args)))
    s = _with_graphs(graphs, lambda: _with_env("HFGPU_DM_PIPELINE", False,
                                                lambda: gpu.solve_dm(*args)))
    assert p[6] == s[6] and p[7] == s[7]
    if case in ("converges", "layered_converges"):
        assert p[7] and 1 < p[6] < args[3]
    else:
        assert not p[7] and p[6] == args[3]
    for a, b in zip(p, s):
        if a is not None:
            np.testing.assert_array_equal(np.asarray(a), np.asarray(b))
