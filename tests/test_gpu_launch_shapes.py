"""The two-step exchange passes under every launch shape they accept: row
passes with 1 to 5 channel tiles per workgroup (the parts of one k1 row on
one XCD) or the whole row, and 1, 2 or 4 wavefronts (adjacent columns of one
tile) per column workgroup.  Each shape must give the oracle's Fock build."""

from __future__ import annotations

import numpy as np
import pytest

from helpers import herm, kernel_args_dict, rand_c

pytestmark = pytest.mark.gpu

# (tiles per row workgroup, 0 = whole row; wavefronts per column workgroup)
SHAPES = [("0", "4"), ("1", "1"), ("2", "2"), ("3", "4"), ("5", "2"), ("4", "1")]


def _cases():
    from cpp_hf_amd.workloads import layered_kernel, qwz_kernel

    wl = layered_kernel(96, 4)
    yield "layered96_nb8", wl.kernel, herm(wl.P0)
    wl = qwz_kernel(24)
    yield "qwz24", wl.kernel, herm(wl.P0)
    wl = layered_kernel(24, 2)
    yield "layered24_nonherm", wl.kernel, rand_c(np.random.default_rng(9), wl.kernel.h.shape)


@pytest.mark.parametrize("shape", SHAPES, ids=["-".join(s) for s in SHAPES])
def test_launch_shapes_match_oracle(gpu, oracle, shape, monkeypatch):
    # read when a kernel's exchange geometry is built
    monkeypatch.setenv("HFGPU_ROW_WAVES", shape[0])
    monkeypatch.setenv("HFGPU_COL_WAVES", shape[1])
    for name, k, P in _cases():
        print(f"launch shape {shape} case {name}", flush=True)
        ka = kernel_args_dict(k)
        S, H, F, E = gpu.build_fock_apply(ka, P, None)
        Sr, Hr, Fr, Er = oracle.build_fock_apply(ka, P)
        scale = float(np.max(np.abs(Fr)))
        assert float(np.max(np.abs(F - Fr))) / scale < 1e-12, name
        assert float(np.max(np.abs(S - Sr))) / max(1e-300, float(np.max(np.abs(Sr)))) < 1e-12
        assert abs(E - Er) <= 1e-12 * max(1.0, abs(Er))


@pytest.mark.parametrize("split", ["1", "2"])
def test_wavefronts_per_tile_match_oracle(gpu, oracle, split, monkeypatch):
    # one or two wavefronts per channel tile in all three passes, with the
    # row and column workgroup shapes they combine with
    monkeypatch.setenv("HFGPU_SPLIT", split)
    for rw, cw in (("1", "2"), ("0", "4"), ("2", "1")):
        monkeypatch.setenv("HFGPU_ROW_WAVES", rw)
        monkeypatch.setenv("HFGPU_COL_WAVES", cw)
        for name, k, P in _cases():
            print(f"split {split} shape {rw}/{cw} case {name}", flush=True)
            ka = kernel_args_dict(k)
            S, H, F, E = gpu.build_fock_apply(ka, P, None)
            Sr, Hr, Fr, Er = oracle.build_fock_apply(ka, P)
            assert float(np.max(np.abs(F - Fr))) / float(np.max(np.abs(Fr))) < 1e-12, name
            assert abs(E - Er) <= 1e-12 * max(1.0, abs(Er))


def test_autotune_keeps_results(gpu, oracle, monkeypatch):
    from cpp_hf_amd.workloads import layered_kernel

    wl = layered_kernel(48, 4)
    ka = kernel_args_dict(wl.kernel)
    P = herm(wl.P0)
    before = gpu.build_fock_apply(ka, P, None)
    gk = gpu.GpuKernel(ka)
    t = gk.autotune(True)
    print(f"autotune 48x48 nb8: {t}", flush=True)
    assert t["us_per_build"] > 0.0
    assert t["split_fwd"] in (1, 2) and t["split_col"] in (1, 2) and t["split_inv"] in (1, 2)
    assert t["col_waves"] * t["split_col"] <= 4 and t["row_waves"] >= 0
    # HFGPU_AUTOTUNE=1: new geometries take the tuned (cached) shape; launch
    # shapes move tasks between lanes only, so every output is bit-identical
    monkeypatch.setenv("HFGPU_AUTOTUNE", "1")
    after = gpu.build_fock_apply(ka, P, None)
    for a, b in zip(before[:3], after[:3]):
        np.testing.assert_array_equal(a, b)
    assert before[3] == after[3]
    Fr = oracle.build_fock_apply(ka, P)[2]
    assert float(np.max(np.abs(after[2] - Fr))) / float(np.max(np.abs(Fr))) < 1e-12


_EIGH_CHILD = r"""
import hashlib, sys
import numpy as np
sys.path.insert(0, sys.argv[1])
from cpp_hf_amd._compat import _native
rng = np.random.default_rng(5)
for nb, nmat in ((4, 700), (8, 500), (12, 300), (16, 200)):
    X = rng.standard_normal((nmat, nb, nb)) + 1j * rng.standard_normal((nmat, nb, nb))
    M = 0.5 * (X + np.conj(np.swapaxes(X, 1, 2)))
    w, V = _native.eigh_batched(M)
    print(nb, hashlib.sha1(w.tobytes() + V.tobytes()).hexdigest(), flush=True)
"""


def test_eigh_matrices_per_wavefront_keep_results(gpu):
    """The eigh with 4, 2 or 1 matrices per wavefront (HFGPU_EIGH_MPW: the
    autotuner's candidates) gives the full-wavefront results bit for bit, at
    4, 8, 12 and 16 bands (one child process per count, the variable being
    read once per process)."""
    import os
    import subprocess
    import sys
    from pathlib import Path

    root = str(Path(__file__).resolve().parents[1])
    outs = []
    for mpw in (None, "4", "2", "1"):
        env = dict(os.environ)
        env.pop("HFGPU_EIGH_MPW", None)
        if mpw:
            env["HFGPU_EIGH_MPW"] = mpw
        print(f"eigh child HFGPU_EIGH_MPW={mpw}", flush=True)
        r = subprocess.run([sys.executable, "-c", _EIGH_CHILD, root], env=env,
                           capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr[-2000:]
        outs.append(r.stdout)
    assert len(outs[0].splitlines()) == 4
    assert all(o == outs[0] for o in outs[1:])


@pytest.mark.parametrize("nb,nmat", [(4, 3000), (12, 1200), (20, 64)])
def test_eigh_autotune(gpu, nb, nmat):
    """The matrices-per-wavefront autotuner returns one of its candidates and
    later calls of that shape give the same results."""
    rng = np.random.default_rng(nb)
    M = herm(rand_c(rng, (nmat, nb, nb)))
    w0, V0 = gpu.eigh_batched(M)
    r = gpu.eigh_autotune(nmat, nb)
    print(f"eigh autotune nb={nb} nmat={nmat}: {r}", flush=True)
    if nb > 16:
        assert r["matrices_per_wavefront"] == 2 and r["us_per_call"] == 0.0
    else:
        g = 64 // (4 if nb <= 4 else 16)
        assert r["matrices_per_wavefront"] in (g, g // 2, g // 4, g // 8)
        assert r["us_per_call"] > 0.0
    w1, V1 = gpu.eigh_batched(M)
    np.testing.assert_array_equal(w0, w1)
    np.testing.assert_array_equal(V0, V1)
    with pytest.raises(ValueError):
        gpu.eigh_autotune(0, nb)
    with pytest.raises(ValueError):
        gpu.eigh_autotune(nmat, 33)


def test_dataflow_floor(gpu):
    """The build's dataflow as three plain streaming launches: the bytes of
    the two-step passes (P and the packed tiles in, Sigma and F out) and a
    time below the build's own."""
    from cpp_hf_amd.workloads import layered_kernel, qwz_kernel

    wl = layered_kernel(96, 4)
    gk = gpu.GpuKernel(kernel_args_dict(wl.kernel))
    r = gk.dataflow_floor(True, 50)
    print(f"dataflow floor 96x96 nb8: {r}", flush=True)
    nd, nt = 96 * 96 * 64, 5 * 8 * 96 * 96  # 5 tiles of 8 packed channels
    assert r["bytes_per_build"] == 16.0 * (3 * nd + 6 * nt)
    t = gk.autotune(True)
    assert 0.0 < r["us_per_build"] < t["us_per_build"]
    # a mesh without the two-step transform has no such floor
    gk13 = gpu.GpuKernel(kernel_args_dict(qwz_kernel(13).kernel))
    with pytest.raises(ValueError):
        gk13.dataflow_floor(True, 5)
