"""With HFGPU_GRAPH=1 the device-pointer Fock build replays a captured
hipGraph; every replay must equal the eager result, with eager work and
synchronisations in between and new data written into the captured input
buffer."""

from __future__ import annotations

import numpy as np
import pytest

from helpers import herm, kernel_args_dict

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("which", ["layered", "qwz", "full_channels"])
def test_graph_replays_match_eager(gpu, which, monkeypatch):
    # read when the kernel is created
    monkeypatch.setenv("HFGPU_GRAPH", "1")
    from hipmem import DeviceArray, synchronize

    from cpp_hf_amd.workloads import layered_kernel, qwz_kernel

    wl = layered_kernel(24, 2) if which != "qwz" else qwz_kernel(24)
    ka = kernel_args_dict(wl.kernel)
    shape = wl.kernel.h.shape
    nb = shape[-1]
    gk = gpu.GpuKernel(ka)
    dP = DeviceArray(shape, np.complex128)
    dS = DeviceArray(shape, np.complex128)
    dF = DeviceArray(shape, np.complex128)
    dH = DeviceArray((nb,), np.float64)
    hermitian = which != "full_channels"
    rng = np.random.default_rng(3)
    for rep in range(4):
        P = herm(wl.P0 + 0.01 * rep * rng.standard_normal(shape))
        dP.upload(P)
        gk.build_fock_device(dP.addr, dS.addr, dF.addr, dH.addr, hermitian, 0)
        synchronize()
        S_g, F_g = dS.download(), dF.download()
        # eager reference through the host-pointer entry point
        S_e, Hh_e, F_e, _ = gpu.build_fock_apply(ka, P, None)
        if hermitian:
            np.testing.assert_array_equal(F_g, F_e)
            np.testing.assert_array_equal(S_g, S_e)
        else:
            # full nb^2 channels versus the eager (packed) path: same values
            # to rounding
            assert np.max(np.abs(F_g - F_e)) <= 1e-13 * np.max(np.abs(F_e))
        H_e = np.real(np.diagonal(Hh_e[0, 0]))
        if hermitian:
            np.testing.assert_array_equal(dH.download(), H_e)
        else:
            # the full-channel row sums group the density differently
            np.testing.assert_allclose(dH.download(), H_e, rtol=0, atol=1e-14 * np.max(np.abs(H_e)))
