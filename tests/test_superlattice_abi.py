"""Host-side checks of the superlattice exchange's C entry point
(hfgpu_selfenergy_superlattice_streamed): the layout is validated before any
device call, so malformed tables raise ValueError on any host.  The reference
binding checks shapes only (cpp_hf_native.cpp:731-770); these checks keep
every scatter and gather index inside the extended grid."""

from __future__ import annotations

import numpy as np
import pytest

from test_oracle_superlattice import layout


def _call(native, L, nk, n_G, orb, **over):
    args = dict(rho=np.zeros((nk, nk, n_G, orb, n_G, orb), dtype=complex), VR_fft=L.V_lag_fft,
                g_a_off=L.g_a_off, pair_i=L.delta_pair_i, pair_j=L.delta_pair_j,
                pair_start=L.delta_pair_start, nkx=nk, nky=nk, n_G=n_G, dim_orb=orb,
                N_ext_x=L.N_ext_x, N_ext_y=L.N_ext_y)
    args.update(over)
    return native.selfenergy_superlattice_streamed(**args)


@pytest.fixture(scope="module")
def native():
    from cpp_hf_amd._compat import _native

    if _native is None:
        pytest.skip("native module not built")
    return _native


def test_layout_validation_before_any_device_call(native):
    nk, orb = 3, 2
    L, g, _ = layout(nk=nk, g_cut=1)
    n_G = g.shape[0]
    off = L.g_a_off.copy()
    off[1, 1] = L.N_ext_y // nk  # tile past the grid edge
    with pytest.raises(ValueError, match="outside"):
        _call(native, L, nk, n_G, orb, g_a_off=off)
    off = L.g_a_off.copy()
    off[0, 0] = -1
    with pytest.raises(ValueError, match="outside"):
        _call(native, L, nk, n_G, orb, g_a_off=off)
    pj = L.delta_pair_j.copy()
    pj[3] = -1
    with pytest.raises(ValueError, match="n_G"):
        _call(native, L, nk, n_G, orb, pair_j=pj)
    ps = L.delta_pair_start.copy()
    ps[-1] = ps[-1] + 5  # past the pair arrays
    with pytest.raises(ValueError, match="past the end"):
        _call(native, L, nk, n_G, orb, pair_start=ps)
    ps = L.delta_pair_start.copy()
    ps[1], ps[2] = ps[2], ps[1]
    with pytest.raises(ValueError, match="non-decreasing"):
        _call(native, L, nk, n_G, orb, pair_start=ps)
    with pytest.raises(ValueError, match="rho shape"):
        _call(native, L, nk, n_G, orb, rho=np.zeros((nk, nk, n_G, orb, n_G, orb + 1), complex))
