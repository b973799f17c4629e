"""CPU-side checks of the boundary and the host logic (no GPU needed)."""

from __future__ import annotations

import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from helpers import two_band_kernel

ROOT = Path(__file__).resolve().parents[1]
LIB = ROOT / "cpp_hf_amd" / "libhfgpu.so"


def _declared_symbols():
    names = []
    for hdr in (ROOT / "include").glob("*.h"):
        text = hdr.read_text()
        names += re.findall(r"^\s*(?:const\s+)?\w+\**\s+\**(hfgpu_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(str(LIB))
    names = _declared_symbols()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_library_reports_errors_without_a_gpu_call():
    lib = C.CDLL(str(LIB))
    lib.hfgpu_last_error.restype = C.c_char_p
    assert lib.hfgpu_device_count() >= 0
    rc = lib.hfgpu_kernel_create(None, None)
    assert rc == 1  # HF_EINVAL
    assert b"null" in lib.hfgpu_last_error()


def test_extension_entry_points_check_arguments_before_any_gpu_call():
    """The autotuner and dataflow-floor entry points refuse null pointers,
    an empty batch and a repeated kernel handle before touching the GPU."""
    lib = C.CDLL(str(LIB))
    lib.hfgpu_last_error.restype = C.c_char_p
    mpw, us, nbytes = C.c_int(), C.c_double(), C.c_double()
    assert lib.hfgpu_eigh_autotune(C.c_size_t(10), C.c_size_t(8), None, C.byref(us)) == 1
    assert b"null" in lib.hfgpu_last_error()
    assert lib.hfgpu_eigh_autotune(C.c_size_t(0), C.c_size_t(8), C.byref(mpw), C.byref(us)) == 1
    assert b"nmat" in lib.hfgpu_last_error()
    assert lib.hfgpu_kernel_dataflow_floor(None, 1, 10, C.byref(us), C.byref(nbytes)) == 1
    assert b"null" in lib.hfgpu_last_error()
    assert lib.hfgpu_kernel_autotune(None, 1, None, None) == 1
    # a handle passed twice to the grouped solve (never dereferenced)
    fake = C.c_void_p(0x1000)
    ks = (C.c_void_p * 2)(fake, fake)
    p0 = (C.c_void_p * 2)(C.c_void_p(0x2000), C.c_void_p(0x2000))
    ne = (C.c_double * 2)(1.0, 1.0)
    cfg = C.create_string_buffer(1024)
    outs = C.create_string_buffer(2048)
    assert lib.hfgpu_solve_dm_batch(ks, C.c_size_t(2), p0, ne, cfg, outs) == 1
    assert b"twice" in lib.hfgpu_last_error()


def test_native_module_exports_reference_names():
    import cpp_hf_amd._native as nat

    for name in ["selfenergy_fft_full", "clear_fft_plan_cache", "eigh_batched",
                 "find_mu_bisection", "build_fock_apply", "solve_scf", "solve_dm",
                 "_cayley_spectral_setup", "_cayley_unitary_from_spectrum",
                 "_diag_UFU_from_spectrum", "resample_kgrid_2d", "eigh_block_sizes",
                 "max_offblock_sizes"]:
        assert hasattr(nat, name), name


def test_package_surface():
    import cpp_hf_amd

    for name in ["HartreeFockKernel", "SolverConfig", "SolveResult", "solve",
                 "solve_direct_minimization", "SCFConfig", "SCFResult", "solve_scf",
                 "build_fock", "hf_energy", "free_energy", "occupation_entropy",
                 "resample_kgrid"]:
        assert hasattr(cpp_hf_amd, name), name
    from cpp_hf_amd.linalg import eigh  # noqa: F401


def test_compute_without_gpu_fails_loudly():
    import cpp_hf_amd._native as nat
    from cpp_hf_amd import solve

    if nat.device_count() > 0:
        pytest.skip("a GPU is visible")
    k = two_band_kernel()
    with pytest.raises(RuntimeError, match="GPU"):
        solve(k, np.zeros_like(k.h), 1.0)


def test_interaction_transform_absorbs_ifftshift():
    """IFFT(VR) must be w * Vq rolled so that q = 0 sits at index 0
    (problem.py:95-104): the phase exp(2 pi i m (n//2)/n) is that shift."""
    from cpp_hf_amd import HartreeFockKernel

    rng = np.random.default_rng(0)
    for n0, n1 in [(6, 6), (5, 8), (7, 3)]:
        w = rng.random((n0, n1))
        Vq = rng.random((n0, n1, 2, 2))
        h = np.zeros((n0, n1, 2, 2), complex)
        k = HartreeFockKernel(w, h, Vq, T=0.1)
        back = np.fft.ifftn(k._VR_shifted, axes=(0, 1))
        ref = np.fft.ifftshift(w[..., None, None] * Vq, axes=(0, 1))
        np.testing.assert_allclose(back, ref, atol=1e-14)
        assert not k.exchange_hermitian_channel_packing
    k = HartreeFockKernel(np.ones((4, 4)), np.zeros((4, 4, 2, 2)), np.ones((4, 4, 1, 1)), T=0.1)
    assert k.exchange_hermitian_channel_packing
    kc = HartreeFockKernel(np.ones((4, 4)), np.zeros((4, 4, 2, 2)),
                           np.ones((4, 4, 1, 1)) * (1 + 1e-3j), T=0.1)
    assert not kc.exchange_hermitian_channel_packing


def test_kernel_validation_errors():
    from cpp_hf_amd import HartreeFockKernel

    w = np.ones((2, 2))
    h = np.zeros((2, 2, 2, 2))
    V = np.ones((2, 2, 1, 1))
    with pytest.raises(ValueError, match="at least one"):
        HartreeFockKernel(w, h, V, T=0.1, include_exchange=False)
    with pytest.raises(ValueError, match="hartree_matrix"):
        HartreeFockKernel(w, h, V, T=0.1, include_hartree=True)
    with pytest.raises(ValueError, match="reference_density"):
        HartreeFockKernel(w, h, V, T=0.1, include_hartree=True, hartree_matrix=np.eye(2))
    with pytest.raises(ValueError, match="reference_density must have shape"):
        HartreeFockKernel(w, h, V, T=0.1, reference_density=np.zeros((3, 2, 2, 2)))
    with pytest.raises(ValueError, match="must be"):
        HartreeFockKernel(w, h, V, T=0.1, contact_terms=[(1.0, np.eye(2))])
    k = HartreeFockKernel(w, h, V, T=0.1)
    assert k.refP.strides == (0, 0, 0, 0) and not k.refP.flags.writeable


def test_electron_count_validation_message():
    from cpp_hf_amd.utils import validate_electron_count

    with pytest.raises(ValueError, match="physically reachable range"):
        validate_electron_count(np.ones((1, 1)), 2, 3.0)
    validate_electron_count(np.ones((1, 1)), 2, 2.0 + 1e-9)


def test_solve_many_validates_arguments():
    from cpp_hf_amd import solve_many

    with pytest.raises(ValueError, match="same length"):
        solve_many([1, 2], [1], [1.0, 2.0])
    with pytest.raises(ValueError, match="method"):
        solve_many([], [], [], method="newton")
    with pytest.raises(ValueError, match="concurrency"):
        solve_many([], [], [], concurrency=0)
