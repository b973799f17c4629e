"""Deterministic model problems: the QWZ and layered-QWZ lattice models.

These build ``HartreeFockKernel`` inputs for the benchmark and test
workloads:

* ``qwz_kernel``      2-band Qi-Wu-Zhang model, d = (sin kx, sin ky,
  m - cos kx - cos ky), scalar interaction V(q) = 2 pi s tanh(d_g |q|)/|q|
  with V(0) = 0; exchange only.
* ``layered_kernel``  nl layers x 2 sublattices (orbital 2 l + s), QWZ per
  layer plus u (l - (nl - 1)/2) on site, interlayer hopping gamma1 between
  B_l and A_{l+1}; layer-resolved V_ll'(q) = 2 pi s exp(-a |q| |l - l'|)
  tanh(d_g |q|)/|q| with V(0) = 2 pi s d_g, Hartree on with reference
  density I/2 and Hartree matrix V_ll'(0).

* ``moire_kernel``    continuum moire model in a plane-wave basis (G,
  orbital) for the superlattice exchange: kinetic c |k - G|^2 per orbital,
  a moire potential between first-shell neighbours G - G', an on-site
  orbital hybridisation; screened V(q) = 2 pi / sqrt(q^2 + kappa^2);
  superlattice Hartree from V(|G - G'|), reference density I/2.

Meshes are k_i = 2 pi i / n - pi (index n // 2 holds k = 0).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .problem import HartreeFockKernel


@dataclass
class Workload:
    kernel: HartreeFockKernel
    P0: np.ndarray
    n_electrons: float
    name: str


def kgrid(n: int) -> np.ndarray:
    return 2.0 * np.pi * np.arange(n) / n - np.pi


def _qwz_block(kx, ky, mass):
    dx, dy, dz = np.sin(kx), np.sin(ky), mass - np.cos(kx) - np.cos(ky)
    h = np.zeros(kx.shape + (2, 2), dtype=np.complex128)
    h[..., 0, 0] = dz
    h[..., 1, 1] = -dz
    h[..., 0, 1] = dx - 1j * dy
    h[..., 1, 0] = dx + 1j * dy
    return h


def _v_of_q(absq, scale, gate):
    out = np.empty_like(absq)
    nz = absq > 0
    out[nz] = 2.0 * np.pi * scale * np.tanh(gate * absq[nz]) / absq[nz]
    return out, nz


def initial_density(n: int, nb: int, seed: int = 0, amp: float = 0.05) -> np.ndarray:
    """P0 = I/2 + amp * H with H a seeded Gaussian Hermitian array."""
    rng = np.random.default_rng(seed)
    s = (n, n, nb, nb)
    X = rng.standard_normal(s) + 1j * rng.standard_normal(s)
    H = 0.5 * (X + np.conj(np.swapaxes(X, -1, -2)))
    return 0.5 * np.eye(nb)[None, None] + amp * H


def qwz_kernel(n: int = 24, *, T: float = 0.1, mass: float = 0.5, scale: float = 0.8,
               gate: float = 4.0, seed: int = 0) -> Workload:
    """2-band QWZ, exchange only, scalar V (the Hermitian-packed path)."""
    k = kgrid(n)
    KX, KY = np.meshgrid(k, k, indexing="ij")
    h = _qwz_block(KX, KY, mass)
    w = np.full((n, n), 1.0 / (n * n))
    absq = np.sqrt(KX ** 2 + KY ** 2)
    V, nz = _v_of_q(absq, scale, gate)
    V[~nz] = 0.0
    kern = HartreeFockKernel(w, h, V[..., None, None], T=T)
    return Workload(kern, initial_density(n, 2, seed), 1.0, f"qwz_n{n}_nb2")


def layered_kernel(n: int = 96, n_layers: int = 4, *, T: float = 0.05, mass: float = 0.5,
                   u: float = 0.1, gamma1: float = 0.4, scale: float = 0.8,
                   gate: float = 4.0, alpha: float = 0.3, filling: float | None = None,
                   hartree: bool = True, seed: int = 0) -> Workload:
    """Layered QWZ with layer-resolved V (nb = 2 n_layers)."""
    nb = 2 * n_layers
    k = kgrid(n)
    KX, KY = np.meshgrid(k, k, indexing="ij")
    h = np.zeros((n, n, nb, nb), dtype=np.complex128)
    blk = _qwz_block(KX, KY, mass)
    for l in range(n_layers):
        s = slice(2 * l, 2 * l + 2)
        h[..., s, s] = blk
        onsite = u * (l - 0.5 * (n_layers - 1))
        h[..., 2 * l, 2 * l] += onsite
        h[..., 2 * l + 1, 2 * l + 1] += onsite
        if l + 1 < n_layers:
            h[..., 2 * l + 1, 2 * (l + 1)] += gamma1
            h[..., 2 * (l + 1), 2 * l + 1] += gamma1
    layer = np.arange(nb) // 2
    dl = np.abs(layer[:, None] - layer[None, :]).astype(np.float64)
    absq = np.sqrt(KX ** 2 + KY ** 2)
    base, nz = _v_of_q(absq, scale, gate)
    Vq = base[..., None, None] * np.exp(-alpha * absq[..., None, None] * dl[None, None])
    Vq[~nz] = 2.0 * np.pi * scale * gate
    w = np.full((n, n), 1.0 / (n * n))
    HH = np.full((nb, nb), 2.0 * np.pi * scale * gate)
    refP = np.broadcast_to(0.5 * np.eye(nb, dtype=np.complex128), (n, n, nb, nb)).copy()
    kern = HartreeFockKernel(
        w, h, Vq, T=T, include_hartree=hartree, include_exchange=True,
        reference_density=refP if hartree else None, hartree_matrix=HH if hartree else None,
    )
    n_e = (0.5 * nb + 0.1) if filling is None else filling
    return Workload(kern, initial_density(n, nb, seed), float(n_e), f"layered_n{n}_nb{nb}")


def layered_tangents(n: int = 96, n_layers: int = 4, *, mass: float = 0.5, u: float = 0.1,
                     gamma1: float = 0.4, scale: float = 0.8, gate: float = 4.0,
                     alpha: float = 0.3) -> dict:
    """Derivatives (dh, dVq, dHH) of layered_kernel's inputs with respect to
    u, scale, gate (the screening length) and gamma1, in closed form."""
    nb = 2 * n_layers
    k = kgrid(n)
    KX, KY = np.meshgrid(k, k, indexing="ij")
    layer = np.arange(nb) // 2
    dl = np.abs(layer[:, None] - layer[None, :]).astype(np.float64)
    absq = np.sqrt(KX ** 2 + KY ** 2)
    nz = absq > 0
    screen = np.exp(-alpha * absq[..., None, None] * dl[None, None])
    zero_h = np.zeros((n, n, nb, nb), dtype=np.complex128)

    dh_u = zero_h.copy()
    for l in range(n_layers):
        c = l - 0.5 * (n_layers - 1)
        dh_u[..., 2 * l, 2 * l] = c
        dh_u[..., 2 * l + 1, 2 * l + 1] = c
    dh_g1 = zero_h.copy()
    for l in range(n_layers - 1):
        dh_g1[..., 2 * l + 1, 2 * (l + 1)] = 1.0
        dh_g1[..., 2 * (l + 1), 2 * l + 1] = 1.0

    # V = 2 pi s tanh(g q)/q exp(-a q dl), V(0) = 2 pi s g; HH = 2 pi s g
    base_s = np.zeros_like(absq)
    base_s[nz] = 2.0 * np.pi * np.tanh(gate * absq[nz]) / absq[nz]
    dV_s = base_s[..., None, None] * screen
    dV_s[~nz] = 2.0 * np.pi * gate
    base_g = np.zeros_like(absq)
    base_g[nz] = 2.0 * np.pi * scale / np.cosh(gate * absq[nz]) ** 2
    dV_g = base_g[..., None, None] * screen
    dV_g[~nz] = 2.0 * np.pi * scale
    zero_v = np.zeros((n, n, nb, nb))
    return {
        "u": (dh_u, zero_v, np.zeros((nb, nb))),
        "scale": (zero_h, dV_s, np.full((nb, nb), 2.0 * np.pi * gate)),
        "gate": (zero_h, dV_g, np.full((nb, nb), 2.0 * np.pi * scale)),
        "gamma1": (dh_g1, zero_v, np.zeros((nb, nb))),
    }


def layered_nb8(n: int = 96) -> Workload:
    """The layered model with 8 bands (4 layers), T = 0.05, 4.1 electrons per
    k, Hartree on (the 96 x 96 Fock-build and solve workload)."""
    return layered_kernel(n, 4, T=0.05, filling=4.1)


def sweep_params(count: int = 48, seed: int = 1):
    """Parameter draws of the batched sweep: mesh, bands and model knobs."""
    rng = np.random.default_rng(seed)
    meshes = np.array([36, 48, 60, 72, 96, 120])
    bands = np.array([2, 4, 8, 12])
    out = []
    for _ in range(count):
        n = int(rng.choice(meshes))
        nb = int(rng.choice(bands))
        u = float(rng.uniform(0.0, 0.3))
        scale = float(rng.uniform(0.5, 1.5))
        gate = float(rng.uniform(2.0, 8.0))
        offset = float(rng.uniform(-0.25, 0.25))
        out.append(dict(n=n, nb=nb, u=u, scale=scale, gate=gate, offset=offset))
    return out


def sweep_workload(p: dict, *, T: float = 0.05) -> Workload:
    nl = p["nb"] // 2
    return layered_kernel(p["n"], nl, T=T, u=p["u"], scale=0.8 * p["scale"], gate=p["gate"],
                          filling=0.5 * p["nb"] + p["offset"])


def hexagonal_g_basis(n_shells: int):
    """Gamma plus the n_shells nearest shells of the hexagonal reciprocal
    lattice: integer coordinates (n_G, 2) and the basis B (columns b1, b2,
    |b| = 4 pi / sqrt(3))."""
    b0 = 4.0 * np.pi / np.sqrt(3.0)
    B = np.stack([np.array([np.sqrt(3.0) / 2.0, -0.5]) * b0, np.array([0.0, 1.0]) * b0], axis=1)
    r = n_shells + 1
    cand = np.array([(i, j) for i in range(-r, r + 1) for j in range(-r, r + 1)], dtype=np.int64)
    norm = np.linalg.norm(cand @ B.T, axis=1)
    shells = np.unique(np.round(norm, 6))
    keep = norm <= shells[min(n_shells, len(shells) - 1)] + 1e-3
    return cand[keep], B


def moire_kernel(n: int = 24, n_shells: int = 2, dim_orb: int = 2, *, T: float = 0.05,
                 kinetic: float = 0.02, potential: float = 0.15, hybrid: float = 0.1,
                 kappa: float = 0.5, filling: float | None = None):
    """Continuum moire model on an n x n mesh of the moire Brillouin zone.

    Returns (SuperlatticeHartreeFockKernel, P0, n_electrons): P0 is the
    reference density I/2 plus a small deterministic Hermitian perturbation;
    the default filling is half of D = n_G dim_orb bands per k.
    """
    from .superlattice import SuperlatticeHartreeFockKernel, build_extended_layout

    g, B = hexagonal_g_basis(n_shells)
    n_G = g.shape[0]
    D = n_G * dim_orb

    def V(q):
        return 2.0 * np.pi / np.sqrt(np.asarray(q) ** 2 + kappa ** 2)

    layout = build_extended_layout(g, B, n, n, V, 1.0 / (n * n))
    kf = np.stack(np.meshgrid(np.arange(n) / n, np.arange(n) / n, indexing="ij"), -1)
    h = np.zeros((n, n, D, D), dtype=np.complex128)
    for i in range(n_G):
        kin = kinetic * np.sum(np.einsum("xy,...y->...x", B, kf - g[i]) ** 2, axis=-1)
        for a in range(dim_orb):
            h[:, :, i * dim_orb + a, i * dim_orb + a] = kin + 0.2 * a
            if a + 1 < dim_orb:
                h[:, :, i * dim_orb + a, i * dim_orb + a + 1] = hybrid
                h[:, :, i * dim_orb + a + 1, i * dim_orb + a] = hybrid
    b0 = np.linalg.norm(B[:, 0])
    for i in range(n_G):
        for j in range(n_G):
            if abs(np.linalg.norm(B @ (g[i] - g[j])) - b0) < 1e-6:
                for a in range(dim_orb):
                    h[:, :, i * dim_orb + a, j * dim_orb + a] = potential
    ref = np.broadcast_to(0.5 * np.eye(D, dtype=np.complex128), h.shape).copy()
    kern = SuperlatticeHartreeFockKernel(np.full((n, n), 1.0 / (n * n)), h, layout,
                                         coulomb_V=V, g_cart=(B @ g.T).T, T=T,
                                         reference_density=ref, hartree_degeneracy=2.0)
    x = np.cos(np.arange(D)[:, None] * 0.7 + np.arange(D)[None, :] * 1.3)
    pert = 0.02 * (x + x.T) * np.cos(2.0 * np.pi * kf[..., 0])[..., None, None]
    P0 = ref + pert
    n_e = 0.5 * D if filling is None else float(filling)
    return kern, P0, n_e
