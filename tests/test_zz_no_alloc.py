"""The allocation-free density_from_fock entry point, through the C ABI.

It must give the same P and mu as hfgpu_density_from_fock (the same kernels
on the same inputs: bit for bit), on two calls in a row with the second one
smaller (the shared scratch array keeps the first call's contents), and
reject an input larger than its scratch array.
"""

from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np
import pytest

from helpers import rand_herm
from hipmem import DeviceArray, hip

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]
dp = C.POINTER(C.c_double)


def _lib():
    lib = C.CDLL(str(ROOT / "cpp_hf_amd" / "libhfgpu.so"))
    lib.hfgpu_last_error.restype = C.c_char_p
    lib.hfgpu_density_from_fock_noalloc.argtypes = [
        C.c_void_p, C.c_void_p, C.c_size_t, C.c_size_t, C.c_double, C.c_double, C.c_int,
        C.c_void_p, C.c_void_p, C.c_void_p]
    return lib


def _host(lib, F, w, n_e, T):
    nk, nb = F.shape[0] * F.shape[1], F.shape[-1]
    P = np.empty_like(F)
    mu = C.c_double()
    rc = lib.hfgpu_density_from_fock(F.ctypes.data_as(dp), w.ctypes.data_as(dp), C.c_size_t(nk),
                                     C.c_size_t(nb), C.c_double(n_e), C.c_double(T), C.c_int(0),
                                     P.ctypes.data_as(dp), C.byref(mu))
    assert rc == 0, lib.hfgpu_last_error()
    return P, mu.value


def _noalloc(lib, F, w, n_e, T):
    nk, nb = F.shape[0] * F.shape[1], F.shape[-1]
    dF, dw = DeviceArray(F.shape, np.complex128), DeviceArray(w.shape, np.float64)
    dP, dmu = DeviceArray(F.shape, np.complex128), DeviceArray((1,), np.float64)
    dF.upload(F)
    dw.upload(w)
    rc = lib.hfgpu_density_from_fock_noalloc(dF.addr, dw.addr, nk, nb, n_e, T, 0, dP.addr,
                                             dmu.addr, None)
    assert rc == 0, lib.hfgpu_last_error()
    assert hip().hipDeviceSynchronize() == 0
    return dP.download(), float(dmu.download()[0])


def test_noalloc_density_from_fock_matches_allocating_entry_point(gpu):
    lib = _lib()
    rng = np.random.default_rng(8)
    # the second call is smaller, and 96 x 96 with 8 bands fills most of the array
    for (n1, n2, nb, n_e, T) in [(96, 96, 8, 4.1, 0.05), (12, 10, 4, 1.7, 0.1),
                                 (24, 24, 2, 1.0, 0.02)]:
        print(f"noalloc density_from_fock {n1}x{n2} nb {nb}", flush=True)
        F = rand_herm(rng, (n1, n2, nb, nb))
        w = np.full((n1, n2), 1.0 / (n1 * n2))
        P_ref, mu_ref = _host(lib, F, w, n_e, T)
        P, mu = _noalloc(lib, F, w, n_e, T)
        np.testing.assert_array_equal(P, P_ref)
        assert mu == mu_ref


def test_noalloc_rejects_input_beyond_scratch(gpu):
    lib = _lib()
    nk, nb = 192 * 192, 16  # 151 MB of eigenvectors: more than the 16 MiB array
    rc = lib.hfgpu_density_from_fock_noalloc(C.c_void_p(8), C.c_void_p(8), nk, nb, 8.0, 0.05, 0,
                                             C.c_void_p(8), C.c_void_p(8), None)
    assert rc == 1  # HF_EINVAL, before any device work
    assert b"scratch" in lib.hfgpu_last_error()
