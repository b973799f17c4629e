"""Superlattice (plane-wave basis) Fock self-energy on the GPU.

Reference: src/cpp_hf/superlattice.py and cpp/include/cpp_hf/superlattice.hpp.
For a Hamiltonian in a basis (G, orbital) over a moire reciprocal-lattice
basis ``g_basis``, the Fock self-energy

    Sigma[a G_a, b G_b](k) = -sum_{q, G0} V(|k - q + G0|) rho[a (G_a + G0), b (G_b + G0)](q)

is one 2-D convolution in the absolute momentum p = k - G for each channel
(a, b, dG = G_a - G_b).  :func:`build_extended_layout` lays the (k, G)
support out on an extended grid (twice the support per axis, so the
circular transform is an aperiodic convolution) and lists the (G_a, G_b)
pairs of each dG channel; the GPU scatters the density onto the grid,
convolves every channel at once with the lag Coulomb kernel and gathers the
result (``_native.selfenergy_superlattice_streamed``).  With a Hermitian
density only one channel of each (+dG, -dG) pair is transformed.

:class:`SuperlatticeHartreeFockKernel` carries the layout, the G-basis
Hartree coupling and the bare Hamiltonian into the solvers (``solve``,
``solve_scf``), whose Fock builds then use the superlattice exchange and the
superlattice Hartree shift in place of the mesh exchange.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Callable

import numpy as np

from ._compat import _native, native_required


@dataclass(frozen=True)
class ExtendedGridLayout:
    """Index tables and the transformed lag kernel of the extended grid
    (fields of src/cpp_hf/superlattice.py:41-88)."""

    N_ext_x: int
    N_ext_y: int
    n_supp_x: int
    n_supp_y: int
    g_a_off: np.ndarray          # (n_G, 2) tile offsets in units of (nkx, nky)
    n_delta: int                 # distinct dG = G_a - G_b
    pair_to_delta: np.ndarray    # (n_G, n_G) channel of each pair
    V_lag_fft: np.ndarray        # (N_ext_x, N_ext_y) FFT of w * V on the lag grid
    delta_pair_count: np.ndarray  # (n_delta,)
    delta_pair_i: np.ndarray     # (n_pairs,) G_a of each pair, grouped by channel
    delta_pair_j: np.ndarray     # (n_pairs,) G_b
    delta_pair_start: np.ndarray  # (n_delta + 1,) CSR row pointer


def build_extended_layout(g_basis_fractional, g_basis_B, nkx: int, nky: int,
                          coulomb_V: Callable[[np.ndarray], np.ndarray],
                          w_scalar: float) -> ExtendedGridLayout:
    """Extended-grid layout for a G-basis (superlattice.py:91-200).

    ``g_basis_fractional`` (n_G, 2) integer coordinates in the reciprocal
    basis ``g_basis_B`` (2, 2) (columns are the basis vectors); ``coulomb_V``
    maps momentum magnitudes to V; ``w_scalar`` is the uniform BZ weight.
    The plane wave of G sits at momentum p = k - G, so G's tile is placed
    at the offset of -G from the most negative one.  Channels are numbered
    in the order their dG first appears in the row-major (G_a, G_b) scan.
    """
    g = np.asarray(g_basis_fractional, dtype=np.int64)
    n_G = int(g.shape[0])
    p = -g
    lo = p.min(axis=0)
    span = p.max(axis=0) - lo + 1
    n_supp_x, n_supp_y = int(nkx * span[0]), int(nky * span[1])
    N_ext_x, N_ext_y = 2 * n_supp_x, 2 * n_supp_y
    g_a_off = np.ascontiguousarray(p - lo[None, :], dtype=np.int64)

    index: dict[tuple[int, int], int] = {}
    pair_to_delta = np.empty((n_G, n_G), dtype=np.int64)
    for i in range(n_G):
        for j in range(n_G):
            key = (int(g[i, 0] - g[j, 0]), int(g[i, 1] - g[j, 1]))
            pair_to_delta[i, j] = index.setdefault(key, len(index))
    n_delta = len(index)
    flat = pair_to_delta.ravel()
    count = np.bincount(flat, minlength=n_delta).astype(np.int64)
    start = np.zeros(n_delta + 1, dtype=np.int64)
    start[1:] = np.cumsum(count)
    # stable grouping keeps the row-major (i, j) order inside each channel
    order = np.argsort(flat, kind="stable")
    pair_i = (order // n_G).astype(np.int64)
    pair_j = (order % n_G).astype(np.int64)

    # lag grid in the FFT's natural order (fractions of the moire lattice)
    fx = np.arange(N_ext_x)
    fy = np.arange(N_ext_y)
    fx = np.where(fx < N_ext_x // 2, fx, fx - N_ext_x).astype(float) / float(nkx)
    fy = np.where(fy < N_ext_y // 2, fy, fy - N_ext_y).astype(float) / float(nky)
    FX, FY = np.meshgrid(fx, fy, indexing="ij")
    B = np.asarray(g_basis_B, dtype=float)
    qx = B[0, 0] * FX + B[0, 1] * FY
    qy = B[1, 0] * FX + B[1, 1] * FY
    V = np.asarray(coulomb_V(np.sqrt(qx * qx + qy * qy)), dtype=float)
    V_lag_fft = np.fft.fftn(float(w_scalar) * V, axes=(0, 1))
    return ExtendedGridLayout(
        N_ext_x=N_ext_x, N_ext_y=N_ext_y, n_supp_x=n_supp_x, n_supp_y=n_supp_y,
        g_a_off=g_a_off, n_delta=n_delta, pair_to_delta=pair_to_delta,
        V_lag_fft=np.ascontiguousarray(V_lag_fft, dtype=np.complex128),
        delta_pair_count=count, delta_pair_i=pair_i, delta_pair_j=pair_j,
        delta_pair_start=start)


def superlattice_fock(rho, layout: ExtendedGridLayout, n_G: int, dim_orb: int, nkx: int,
                      nky: int, *, V_lag_fft_orbital=None, hermitian_rho: bool = False):
    """Sigma of a density (nkx, nky, n_G dim_orb, n_G dim_orb) on the GPU
    (superlattice.py:203-260).  ``V_lag_fft_orbital`` (N_ext_x, N_ext_y,
    dim_orb, dim_orb) replaces the scalar kernel per orbital pair;
    ``hermitian_rho`` lets one channel of each (+dG, -dG) pair stand for
    both."""
    native_required()
    D = n_G * dim_orb
    r = np.ascontiguousarray(np.asarray(rho), dtype=np.complex128)
    if r.shape != (nkx, nky, D, D):
        raise ValueError(f"rho must have shape ({nkx}, {nky}, {D}, {D}); got {r.shape}")
    V_orb = None
    if V_lag_fft_orbital is not None:
        V_orb = np.ascontiguousarray(np.asarray(V_lag_fft_orbital), dtype=np.complex128)
        expected = (layout.N_ext_x, layout.N_ext_y, dim_orb, dim_orb)
        if V_orb.shape != expected:
            raise ValueError(
                f"V_lag_fft_orbital must have shape {expected}; got {V_orb.shape}")
    sigma = _native.selfenergy_superlattice_streamed(
        rho=r.reshape(nkx, nky, n_G, dim_orb, n_G, dim_orb), VR_fft=layout.V_lag_fft,
        g_a_off=layout.g_a_off, pair_i=layout.delta_pair_i, pair_j=layout.delta_pair_j,
        pair_start=layout.delta_pair_start, nkx=int(nkx), nky=int(nky), n_G=int(n_G),
        dim_orb=int(dim_orb), N_ext_x=int(layout.N_ext_x), N_ext_y=int(layout.N_ext_y),
        VR_fft_orbital=V_orb, hermitian_rho=bool(hermitian_rho))
    return np.asarray(sigma).reshape(nkx, nky, D, D)


def _hartree_shift(rho, weights, layout, n_G, dim_orb, *, HH_GG, degeneracy,
                   HH_GG_orbital=None):
    """k-independent (D, D) superlattice Hartree shift of a density
    (superlattice.py:525-573, fock.hpp:66-173)."""
    D = n_G * dim_orb
    rho_bar = np.einsum("ij,ijab->ab", np.asarray(weights), np.asarray(rho))
    rho_bar = rho_bar.reshape(n_G, dim_orb, n_G, dim_orb)
    sigma = np.zeros((n_G, dim_orb, n_G, dim_orb), dtype=complex)
    p2d = layout.pair_to_delta
    if HH_GG_orbital is None:
        per_pair = np.einsum("AxBx->AB", rho_bar)
        per_delta = np.zeros(layout.n_delta, dtype=complex)
        np.add.at(per_delta, p2d.ravel(), per_pair.ravel())
        s = degeneracy * np.asarray(HH_GG) * per_delta[p2d]
        for x in range(dim_orb):
            sigma[:, x, :, x] = s
    else:
        per_pair = np.einsum("AxBx->ABx", rho_bar)
        per_delta = np.zeros((layout.n_delta, dim_orb), dtype=complex)
        np.add.at(per_delta, p2d.ravel(), per_pair.reshape(n_G * n_G, dim_orb))
        s = degeneracy * np.einsum("ABag,ABg->ABa", np.asarray(HH_GG_orbital, dtype=complex),
                                   per_delta[p2d])
        for a in range(dim_orb):
            sigma[:, a, :, a] = s[:, :, a]
    return sigma.reshape(D, D)


class SuperlatticeHartreeFockKernel:
    """A superlattice HF problem for the solvers (superlattice.py:263-522):
    same constructor, validation and attributes as the reference's."""

    def __init__(self, weights, hamiltonian, layout: ExtendedGridLayout, *, HH_GG=None,
                 HH_GG_orbital=None, V_lag_fft_orbital=None, coulomb_V=None, g_cart=None,
                 T: float, include_hartree: bool = True, include_exchange: bool = True,
                 reference_density=None, hartree_degeneracy: float = 1.0,
                 embed_reference_potential: bool = False,
                 center_embedded_hartree: bool = True) -> None:
        h = np.ascontiguousarray(np.asarray(hamiltonian), dtype=np.complex128)
        if h.ndim != 4 or h.shape[-1] != h.shape[-2]:
            raise ValueError(f"hamiltonian must be (nkx, nky, D, D); got {h.shape}")
        nkx, nky, D, _ = h.shape
        n_G = int(layout.g_a_off.shape[0])
        if D % n_G != 0:
            raise ValueError(f"Hamiltonian dim D={D} not divisible by n_G={n_G} from layout")
        dim_orb = D // n_G
        w2d = np.ascontiguousarray(np.asarray(weights), dtype=np.float64)
        if w2d.shape != (nkx, nky):
            raise ValueError(f"weights must have shape ({nkx}, {nky}); got {w2d.shape}")
        self.T = float(T)
        self.include_hartree = bool(include_hartree)
        self.include_exchange = bool(include_exchange)
        if not (self.include_hartree or self.include_exchange):
            raise ValueError("SuperlatticeHartreeFockKernel must include at least one channel.")

        if HH_GG is None:
            if not self.include_hartree:
                HH = np.zeros((n_G, n_G))
            else:
                if coulomb_V is None or g_cart is None:
                    raise ValueError(
                        "include_hartree=True without HH_GG requires coulomb_V + g_cart")
                gc = np.asarray(g_cart, dtype=float)
                HH = np.asarray(coulomb_V(np.linalg.norm(gc[:, None, :] - gc[None, :, :],
                                                         axis=-1)), dtype=np.float64)
                np.fill_diagonal(HH, 0.0)
        else:
            HH = np.asarray(HH_GG, dtype=np.float64)
            if HH.shape != (n_G, n_G):
                raise ValueError(f"HH_GG must have shape ({n_G}, {n_G}); got {HH.shape}")
        self._HH_GG = np.ascontiguousarray(HH, dtype=np.float64)

        if HH_GG_orbital is None:
            self._HH_GG_orbital = None
        else:
            HO = np.asarray(HH_GG_orbital, dtype=np.float64)
            if HO.shape != (n_G, n_G, dim_orb, dim_orb):
                raise ValueError(f"HH_GG_orbital must have shape {(n_G, n_G, dim_orb, dim_orb)}; "
                                 f"got {HO.shape}")
            if float(np.max(np.abs(HO[np.arange(n_G), np.arange(n_G)]))) > 1e-12:
                import warnings

                warnings.warn(
                    "HH_GG_orbital has nonzero diagonal-G blocks; the q=0 (uniform/layer) "
                    "Hartree will be INCLUDED.  This is intended for layer-resolved "
                    "gate-screened systems; pass a version with HH_GG_orbital[i, i, :, :] = 0 "
                    "to drop it (canonical moiré convention).", stacklevel=2)
            self._HH_GG_orbital = np.ascontiguousarray(HO)

        if V_lag_fft_orbital is None:
            self._V_lag_fft_orbital = None
        else:
            VO = np.asarray(V_lag_fft_orbital, dtype=np.complex128)
            expected = (layout.N_ext_x, layout.N_ext_y, dim_orb, dim_orb)
            if VO.shape != expected:
                raise ValueError(f"V_lag_fft_orbital must have shape {expected}; got {VO.shape}")
            self._V_lag_fft_orbital = np.ascontiguousarray(VO)

        if reference_density is not None:
            ref = np.ascontiguousarray(np.asarray(reference_density), dtype=h.dtype)
            if ref.shape != h.shape:
                raise ValueError(f"reference_density must have shape {h.shape}; got {ref.shape}")
            self._refP = ref
        else:
            self._refP = None
        self._empty_refP = np.empty((0,), dtype=h.dtype)
        self._zero_refP = np.broadcast_to(np.array(0, dtype=h.dtype), h.shape)

        # the mesh-exchange fields of the kernel_args contract (unused in
        # superlattice mode, shape-correct for the validators)
        self.h = h
        self.weights_b = w2d[..., None, None]
        self.w2d = w2d
        self.weight_sum = float(np.sum(w2d))
        self._VR_shifted = np.zeros((nkx, nky, 1, 1), dtype=h.dtype)
        self.exchange_hermitian_channel_packing = False
        self.coulomb_q_is_real = True
        self.HH = np.zeros((D, D), dtype=np.float64)
        self.contact_g = np.zeros((1,), dtype=np.float64)
        self.contact_Oi = np.zeros((1, D, D), dtype=h.dtype)
        self.contact_Oj = np.zeros((1, D, D), dtype=h.dtype)

        self._layout = layout
        self.n_G = int(n_G)
        self.dim_orb = int(dim_orb)
        self.hartree_degeneracy = float(hartree_degeneracy)
        self.superlattice_fock_active = bool(self.include_exchange)
        self.superlattice_hartree_active = bool(self.include_hartree)

        # embed V_HF[refP] into h once (superlattice.py:455-489)
        self.embed_reference_potential = bool(embed_reference_potential)
        self.center_embedded_hartree = bool(center_embedded_hartree)
        self.embedded_energy_offset = 0.0
        if self.embed_reference_potential:
            if self._refP is None:
                raise ValueError("embed_reference_potential=True requires reference_density.")
            offset = 0.0
            V_ref = np.zeros_like(h)
            if self.include_exchange:
                sF = superlattice_fock(self._refP, layout, n_G, dim_orb, nkx, nky,
                                       V_lag_fft_orbital=self._V_lag_fft_orbital)
                V_ref = V_ref + sF
                offset += float(np.einsum("ij,ijab,ijba->", w2d, sF, self._refP).real)
            if self.include_hartree:
                sH = _hartree_shift(self._refP, w2d, layout, n_G, dim_orb, HH_GG=self._HH_GG,
                                    degeneracy=self.hartree_degeneracy,
                                    HH_GG_orbital=self._HH_GG_orbital)
                if self.center_embedded_hartree:
                    sH = sH - float(np.mean(np.real(np.diag(sH)))) * np.eye(D, dtype=h.dtype)
                V_ref = V_ref + sH[None, None, :, :]
                offset += float(np.einsum("ij,ab,ijba->", w2d, sH, self._refP).real)
            self.h = np.ascontiguousarray(self.h + V_ref, dtype=np.complex128)
            self.embedded_energy_offset = 0.5 * offset

    @property
    def has_reference_density(self) -> bool:
        return self._refP is not None

    @property
    def refP(self) -> np.ndarray:
        return self._zero_refP if self._refP is None else self._refP

    @property
    def layout(self) -> ExtendedGridLayout:
        return self._layout

    @property
    def HH_GG(self) -> np.ndarray:
        return self._HH_GG

    @property
    def HH_GG_orbital(self):
        return self._HH_GG_orbital

    @property
    def V_lag_fft_orbital(self):
        return self._V_lag_fft_orbital


def superlattice_kernel_args(kernel) -> dict:
    """The superlattice entries of the kernel_args dict (solver.py:104-134)."""
    L = kernel.layout
    args = dict(
        superlattice_fock_active=bool(kernel.superlattice_fock_active),
        superlattice_hartree_active=bool(kernel.superlattice_hartree_active),
        hartree_degeneracy=float(kernel.hartree_degeneracy),
        n_G=int(kernel.n_G), dim_orb=int(kernel.dim_orb), n_delta=int(L.n_delta),
        N_ext_x=int(L.N_ext_x), N_ext_y=int(L.N_ext_y),
        V_lag_fft=np.ascontiguousarray(L.V_lag_fft, dtype=np.complex128),
        g_a_off=np.ascontiguousarray(L.g_a_off, dtype=np.int64),
        pair_i=np.ascontiguousarray(L.delta_pair_i, dtype=np.int64),
        pair_j=np.ascontiguousarray(L.delta_pair_j, dtype=np.int64),
        pair_start=np.ascontiguousarray(L.delta_pair_start, dtype=np.int64),
        pair_to_delta=np.ascontiguousarray(L.pair_to_delta, dtype=np.int64),
        HH_GG=np.ascontiguousarray(kernel.HH_GG, dtype=np.float64),
    )
    if kernel.HH_GG_orbital is not None:
        args["HH_GG_orbital"] = np.ascontiguousarray(kernel.HH_GG_orbital, dtype=np.float64)
    if kernel.V_lag_fft_orbital is not None:
        args["V_lag_fft_orbital"] = np.ascontiguousarray(kernel.V_lag_fft_orbital,
                                                         dtype=np.complex128)
    return args


__all__ = ["ExtendedGridLayout", "build_extended_layout", "superlattice_fock",
           "SuperlatticeHartreeFockKernel", "superlattice_kernel_args"]
