"""Hartree-Fock kernel: the arrays the solvers consume, precomputed once.

Same constructor, attributes and validation as the reference's
``HartreeFockKernel`` (src/cpp_hf/problem.py:16-258).  The interaction
transform ``VR = fftn(w * Vq) * phase`` is computed here with numpy, with the
reference's operation order, so ``VR`` is bit-identical to the reference's;
the phase ``exp(2 pi i m (n//2) / n)`` per axis absorbs the ifftshift of the
exchange convolution (problem.py:95-104).  The GPU library receives these
arrays through ``_kernel_args_for_native``.
"""

from __future__ import annotations

import numpy as np

from ._compat import fftn2d
from .utils import selfenergy_fft


def _shifted_interaction(weights_b: np.ndarray, Vq: np.ndarray, dtype) -> np.ndarray:
    """fftn over the k axes of w * Vq, times the ifftshift phase."""
    VR = fftn2d(weights_b * Vq.astype(dtype, copy=False))
    n0, n1 = VR.shape[0], VR.shape[1]
    s0, s1 = n0 // 2, n1 // 2
    ph0 = np.exp(2j * np.pi * np.arange(n0) * s0 / n0)
    ph1 = np.exp(2j * np.pi * np.arange(n1) * s1 / n1)
    VR *= (ph0[:, None] * ph1[None, :])[..., None, None].astype(VR.dtype)
    return np.ascontiguousarray(VR)


class HartreeFockKernel:
    """Precomputed arrays for the Hartree-Fock solver loops.

    Parameters follow src/cpp_hf/problem.py:24-37: ``weights`` (nk1, nk2),
    ``hamiltonian`` (nk1, nk2, nb, nb), ``coulomb_q`` (nk1, nk2, 1, 1) or
    (nk1, nk2, nb, nb), temperature ``T``; optional Hartree term (needs
    ``hartree_matrix`` and ``reference_density``), contact terms
    ``[(g, O_i, O_j), ...]`` and the embedded reference potential.
    """

    def __init__(
        self,
        weights,
        hamiltonian,
        coulomb_q,
        T: float,
        include_hartree: bool = False,
        include_exchange: bool = True,
        reference_density=None,
        hartree_matrix=None,
        contact_terms=None,
        embed_reference_potential: bool = False,
        center_embedded_hartree: bool = True,
    ):
        h = np.ascontiguousarray(np.asarray(hamiltonian), dtype=np.complex128)
        self.h = h
        w2d = np.ascontiguousarray(np.asarray(weights), dtype=np.float64)
        self.w2d = w2d
        self.weights_b = w2d[..., None, None]
        self.weight_sum = np.sum(w2d)
        self.T = float(T)
        self.include_hartree = bool(include_hartree)
        self.include_exchange = bool(include_exchange)
        if not (self.include_hartree or self.include_exchange):
            raise ValueError(
                "HartreeFockKernel must include at least one of Hartree or exchange."
            )
        n0, n1 = int(h.shape[0]), int(h.shape[1])
        nb = int(h.shape[-1])

        # exchange interaction (problem.py:84-107)
        if self.include_exchange:
            Vq = np.asarray(coulomb_q)
            if np.iscomplexobj(Vq):
                if float(np.max(np.abs(np.imag(Vq)))) <= 1e-8:
                    Vq = np.real(Vq)
            else:
                Vq = Vq.astype(np.float64)
            self.exchange_hermitian_channel_packing = bool(
                Vq.shape[-2:] == (1, 1) and not np.iscomplexobj(Vq)
            )
            self.coulomb_q_is_real = not np.iscomplexobj(Vq)
            self._VR_shifted = _shifted_interaction(self.weights_b, Vq, h.dtype)
        else:
            self.exchange_hermitian_channel_packing = False
            self.coulomb_q_is_real = True
            self._VR_shifted = np.zeros((n0, n1, 1, 1), dtype=h.dtype)

        # reference density (problem.py:109-121)
        if reference_density is not None:
            ref = np.ascontiguousarray(np.asarray(reference_density), dtype=h.dtype)
            if ref.shape != h.shape:
                raise ValueError(
                    f"reference_density must have shape {h.shape}, got {ref.shape}"
                )
            self._refP = ref
        else:
            self._refP = None
        self._empty_refP = np.empty((0,), dtype=h.dtype)
        self._zero_refP = np.broadcast_to(np.array(0, dtype=h.dtype), h.shape)

        # Hartree matrix (problem.py:123-139)
        if self.include_hartree:
            if hartree_matrix is None:
                raise ValueError("include_hartree=True requires hartree_matrix to be provided")
            if reference_density is None:
                raise ValueError(
                    "include_hartree=True requires reference_density to be provided"
                )
            HH = np.asarray(hartree_matrix, dtype=np.float64)
            if HH.shape != h.shape[-2:]:
                raise ValueError(
                    f"hartree_matrix must have shape {h.shape[-2:]}, got {HH.shape}"
                )
            self.HH = HH
        else:
            self.HH = np.zeros(h.shape[-2:], dtype=np.float64)

        # contact terms (problem.py:141-172)
        if contact_terms is None or len(contact_terms) == 0:
            self.contact_g = np.zeros((1,), dtype=np.float64)
            self.contact_Oi = np.zeros((1, nb, nb), dtype=h.dtype)
            self.contact_Oj = np.zeros((1, nb, nb), dtype=h.dtype)
        else:
            gs, ois, ojs = [], [], []
            for idx, term in enumerate(contact_terms):
                if len(term) != 3:
                    raise ValueError(
                        f"contact_terms[{idx}] must be (g, O_i, O_j); got len={len(term)}"
                    )
                g_t, oi_t, oj_t = term
                oi = np.asarray(oi_t, dtype=h.dtype)
                oj = np.asarray(oj_t, dtype=h.dtype)
                for name, arr in (("O_i", oi), ("O_j", oj)):
                    if arr.shape != (nb, nb):
                        raise ValueError(
                            f"contact_terms[{idx}].{name} must have shape {(nb, nb)}, "
                            f"got {arr.shape}"
                        )
                gs.append(np.asarray(g_t, dtype=np.float64))
                ois.append(oi)
                ojs.append(oj)
            self.contact_g = np.stack(gs)
            self.contact_Oi = np.stack(ois)
            self.contact_Oj = np.stack(ojs)

        # embedded reference potential (problem.py:174-228)
        self.embed_reference_potential = bool(embed_reference_potential)
        self.center_embedded_hartree = bool(center_embedded_hartree)
        self.embedded_energy_offset = 0.0
        if self.embed_reference_potential:
            self._embed_reference(w2d, nb)

    def _embed_reference(self, w2d: np.ndarray, nb: int) -> None:
        """Fold V_HF[refP] = J[refP] + Sigma[refP] into h once."""
        if self._refP is None:
            raise ValueError("embed_reference_potential=True requires reference_density.")
        ref = self._refP
        j_diag = np.zeros(nb, dtype=np.float64)
        e_hartree = 0.0
        if self.include_hartree:
            occ_ref = np.einsum("ij,ijaa->a", w2d, ref).real
            j_diag = self.HH @ occ_ref
            if self.center_embedded_hartree:
                j_diag = j_diag - float(np.mean(j_diag))
            j_diag = j_diag.astype(np.float64, copy=False)
            e_hartree = float(np.sum(j_diag * occ_ref))
        sigma_ref = None
        e_exchange = 0.0
        if self.include_exchange:
            sigma_ref = selfenergy_fft(
                self._VR_shifted,
                ref,
                _apply_ifftshift=False,
                hermitian_channel_packing=self.exchange_hermitian_channel_packing,
            )
            sigma_ref = np.ascontiguousarray(sigma_ref, dtype=self.h.dtype)
            e_exchange = float(np.einsum("ij,ijab,ijba->", w2d, sigma_ref, ref).real)
        h_eff = self.h.copy()
        if sigma_ref is not None:
            h_eff += sigma_ref
        for a in range(nb):
            h_eff[..., a, a] += j_diag[a]
        self.h = np.ascontiguousarray(h_eff, dtype=np.complex128)
        self.embedded_energy_offset = 0.5 * (e_hartree + e_exchange)

    @property
    def has_reference_density(self) -> bool:
        return self._refP is not None

    @property
    def refP(self) -> np.ndarray:
        """Reference density; a read-only zero broadcast view when absent."""
        return self._zero_refP if self._refP is None else self._refP

    def as_args(self) -> dict:
        """Dynamic inputs of the solver functions (problem.py:241-258)."""
        return dict(
            h=self.h,
            weights_b=self.weights_b,
            weight_sum=self.weight_sum,
            VR=self._VR_shifted,
            T=self.T,
            refP=self.refP,
            has_refP=self.has_reference_density,
            HH=self.HH,
            include_hartree=self.include_hartree,
            include_exchange=self.include_exchange,
            exchange_hermitian_channel_packing=self.exchange_hermitian_channel_packing,
            contact_g=self.contact_g,
            contact_Oi=self.contact_Oi,
            contact_Oj=self.contact_Oj,
        )
