"""Entry points called from many host threads at once (each call on a stream
from the library's pool or on its kernel handle's own stream, buffers from
the device pool, transforms from the shared geometry cache) give the results
of the same calls made one after another, bit for bit."""

from __future__ import annotations

from concurrent.futures import ThreadPoolExecutor

import numpy as np

from helpers import herm, kernel_args_dict, rand_c

import pytest

pytestmark = pytest.mark.gpu


def test_entry_points_from_many_threads(gpu):
    from cpp_hf_amd.workloads import layered_kernel, qwz_kernel

    rng = np.random.default_rng(3)
    kernels = [layered_kernel(24, 2).kernel, qwz_kernel(18).kernel, layered_kernel(12, 8).kernel]
    kas = [kernel_args_dict(k) for k in kernels]
    jobs = []
    for i in range(12):
        k = kernels[i % 3]
        nb = k.h.shape[-1]
        jobs.append((i % 3, herm(rand_c(rng, k.h.shape)), herm(rand_c(rng, (300, nb + i % 5, nb + i % 5)))))

    def run(j):
        ki, P, M = jobs[j]
        ka = kas[ki]
        w, V = gpu.eigh_batched(M)
        S, H, F, E = gpu.build_fock_apply(ka, P, None)
        sig = gpu.selfenergy_fft_full(ka["VR"], P, True, bool(ka["exchange_hcp"]))
        return w, V, S, F, E, sig

    serial = [run(j) for j in range(len(jobs))]
    for rep in range(3):
        print(f"threads pass {rep}", flush=True)
        with ThreadPoolExecutor(max_workers=8) as ex:
            par = list(ex.map(run, range(len(jobs))))
        for a, b in zip(serial, par):
            for x, y in zip(a, b):
                np.testing.assert_array_equal(np.asarray(x), np.asarray(y))
