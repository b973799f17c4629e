"""Derivatives of the Fock build and of the HF free energy, on the GPU.

The reference package has no derivatives; this module adds them around its
forward maps (``build_fock``, src/cpp_hf/fock.py:16-88, and
``hf_energy``, fock.py:91-118):

* :func:`build_fock_vjp` — vector-Jacobian product of the Fock build with
  respect to the density, the interaction ``coulomb_q``, ``h`` and the
  Hartree matrix, for cotangents on its outputs (Sigma, H_hartree, F).  The
  exchange part is the transpose of the FFT convolution (a convolution with
  conj(VR) and a cross spectrum), computed by libhfgpu's own transforms;
  nothing is stored between the forward and the backward pass.
* :func:`energy_gradients` — partial derivatives of ``hf_energy`` at a
  fixed density.  At a converged (stationary) state these are the
  derivatives of the free energy with respect to the model (the
  Hellmann-Feynman theorem for the variational free energy at fixed
  electron count): :func:`free_energy_derivative` contracts them with model
  tangents.
* :func:`fock_autograd` — a ``torch.autograd.Function`` over the Fock build
  (torch is imported only there).

Gradients follow the convention dL = Re sum conj(g) dX; for a real
``coulomb_q`` or Hartree matrix the real part is the gradient.
"""

from __future__ import annotations

from typing import NamedTuple

import numpy as np

from ._compat import _native, native_required
from .solver import _kernel_args_for_native


class FockVJP(NamedTuple):
    density: np.ndarray        # (nk1, nk2, nb, nb) complex
    coulomb_q: np.ndarray      # (nk1, nk2, dv, dv), real for a real interaction
    hamiltonian: np.ndarray    # (nk1, nk2, nb, nb) complex
    hartree_matrix: np.ndarray  # (nb, nb) real


def _cot(x, shape):
    if x is None:
        return None
    a = np.ascontiguousarray(x, dtype=np.complex128)
    if a.shape != shape:
        raise ValueError(f"cotangent must have shape {shape}, got {a.shape}")
    return a


def build_fock_vjp(kernel, P, cot_sigma=None, cot_hartree=None, cot_F=None) -> FockVJP:
    """VJP of ``build_fock(kernel, P) -> (Sigma, H_hartree, F)``.

    Any cotangent may be None (zero).  ``coulomb_q``'s gradient is taken
    through HartreeFockKernel's transform VR = phase * fftn(w Vq); it is
    real when the kernel's interaction is real.
    """
    native_required()
    shape = kernel.h.shape
    P = np.ascontiguousarray(P, dtype=np.complex128)
    if P.shape != shape:
        raise ValueError(f"P must have shape {shape}, got {P.shape}")
    gP, gV, gh, gHH = _native.build_fock_vjp(
        _kernel_args_for_native(kernel), P, _cot(cot_sigma, shape), _cot(cot_hartree, shape),
        _cot(cot_F, shape))
    if kernel.coulomb_q_is_real:
        gV = np.ascontiguousarray(np.real(gV))
    return FockVJP(gP, gV, gh, gHH)


class EnergyGradients(NamedTuple):
    hamiltonian: np.ndarray     # dE/dh
    coulomb_q: np.ndarray       # dE/dVq (complex; real part for a real interaction)
    hartree_matrix: np.ndarray  # dE/dHH


def energy_gradients(kernel, P) -> EnergyGradients:
    """Partial derivatives of hf_energy (fock.hpp:355-383) at fixed P.

    E = sum_k w_k Re[Tr(h P) + 1/2 Tr((Sigma + H)(P - refP))]: dE/dh =
    w P^H, and the interaction and Hartree-matrix parts are the VJP of the
    Fock build with cotangents (w/2) (P - refP)^H on Sigma and H_hartree.
    """
    P = np.ascontiguousarray(P, dtype=np.complex128)
    w = np.asarray(kernel.w2d, dtype=np.float64)[..., None, None]
    dP = P - kernel._refP if kernel._refP is not None else P
    half = 0.5 * w * np.conj(np.swapaxes(dP, -1, -2))
    vjp = build_fock_vjp(kernel, P, cot_sigma=half, cot_hartree=half)
    gh = w * np.conj(np.swapaxes(P, -1, -2))
    return EnergyGradients(gh, vjp.coulomb_q, vjp.hartree_matrix)


def free_energy_derivative(kernel, P, *, dh=None, dVq=None, dHH=None) -> float:
    """d(free energy)/d(theta) at a converged density P for a model tangent
    (dh, dVq, dHH) = d(h, coulomb_q, hartree_matrix)/d(theta).

    Valid at a stationary point of the free energy at fixed electron count
    (the solvers' converged states), where the density's own response drops
    out.
    """
    g = energy_gradients(kernel, P)
    total = 0.0
    if dh is not None:
        total += float(np.real(np.vdot(g.hamiltonian, np.asarray(dh, dtype=np.complex128))))
    if dVq is not None:
        total += float(np.real(np.vdot(g.coulomb_q, np.asarray(dVq, dtype=np.complex128))))
    if dHH is not None:
        total += float(np.sum(g.hartree_matrix * np.asarray(dHH, dtype=np.float64)))
    return total


def _handle(kernel):
    """The kernel's device-resident handle (made once, kept on the kernel)."""
    h = getattr(kernel, "_gpu_handle", None)
    if h is None:
        h = _native.GpuKernel(_kernel_args_for_native(kernel))
        kernel._gpu_handle = h
    return h


def fock_autograd():
    """A torch.autograd.Function ``FockBuild.apply(P, kernel) -> F``.

    P is a complex128 tensor.  On a GPU tensor the forward Fock build and
    the backward adjoint run on device pointers on torch's current stream
    (no host copies); a CPU tensor goes through the host entry points.
    torch's gradient of a real loss with respect to a complex tensor is the
    g of dL = Re sum conj(g) dz, the convention of :func:`build_fock_vjp`,
    so the cotangent passes through unchanged.  Gradients with respect to
    the model go through :func:`build_fock_vjp` and
    :func:`energy_gradients`.
    """
    import torch

    if _native is not None and _native.device_count() > 0 and not torch.cuda.is_available():
        raise RuntimeError(
            "torch sees no GPU while libhfgpu does: import torch before cpp_hf_amd (both load "
            "libamdhip64.so.7, and the copy loaded first serves the whole process)")

    class FockBuild(torch.autograd.Function):
        @staticmethod
        def forward(ctx, P, kernel):
            native_required()
            if P.dtype != torch.complex128 or tuple(P.shape) != tuple(kernel.h.shape):
                raise ValueError("P must be a complex128 tensor of the kernel's h shape")
            Pc = P.detach().contiguous()
            ctx.kernel = kernel
            ctx.save_for_backward(Pc)
            if Pc.is_cuda:
                h = _handle(kernel)
                S = torch.empty_like(Pc)
                F = torch.empty_like(Pc)
                hd = torch.empty(kernel.h.shape[-1], dtype=torch.float64, device=Pc.device)
                stream = torch.cuda.current_stream(Pc.device).cuda_stream
                h.build_fock_device(Pc.data_ptr(), S.data_ptr(), F.data_ptr(), hd.data_ptr(),
                                    False, stream)
                return F
            _, _, F, _ = _native.build_fock_apply(_kernel_args_for_native(kernel),
                                                  Pc.numpy(), None)
            return torch.as_tensor(F)

        @staticmethod
        def backward(ctx, gF):
            (Pc,) = ctx.saved_tensors
            gFc = gF.detach().contiguous().to(torch.complex128)
            if Pc.is_cuda:
                h = _handle(ctx.kernel)
                gP = torch.empty_like(Pc)
                stream = torch.cuda.current_stream(Pc.device).cuda_stream
                h.build_fock_vjp_device(Pc.data_ptr(), 0, 0, gFc.data_ptr(), gP.data_ptr(), 0, 0,
                                        0, stream)
                return gP, None
            vjp = build_fock_vjp(ctx.kernel, Pc.numpy(), cot_F=gFc.numpy())
            return torch.as_tensor(vjp.density), None

    return FockBuild


__all__ = ["FockVJP", "EnergyGradients", "build_fock_vjp", "energy_gradients",
           "free_energy_derivative", "fock_autograd"]
