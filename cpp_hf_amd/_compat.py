"""Native-module guard and small numpy helpers shared by the Python layer.

Mirrors the reference's ``cpp_hf._compat`` (src/cpp_hf/_compat.py:14-69):
``_native`` is the compiled module (here bound to the HIP library libhfgpu),
``HAVE_NATIVE`` says whether it imported, and :func:`native_required` raises
``RuntimeError`` otherwise.  Every compute entry point calls it: there is no
CPU fallback, a missing or unusable GPU library fails loudly.
"""

from __future__ import annotations

import numpy as np

try:  # the extension is built in-tree by __graft_entry__.build()
    from . import _native  # type: ignore[attr-defined]

    HAVE_NATIVE = True
    _IMPORT_ERROR: Exception | None = None
except Exception as exc:  # pragma: no cover - exercised only without a build
    _native = None  # type: ignore[assignment]
    HAVE_NATIVE = False
    _IMPORT_ERROR = exc


def native_required() -> None:
    """Raise unless the GPU extension is loaded and a GPU is visible."""
    if not HAVE_NATIVE:
        raise RuntimeError(
            "The cpp_hf_amd GPU extension (cpp_hf_amd._native) is not available: "
            f"{_IMPORT_ERROR!r}. Build it with `python -c 'import __graft_entry__ as g; g.build()'`."
        )
    if _native.device_count() < 1:
        raise RuntimeError("cpp_hf_amd needs a visible AMD GPU (none found).")


def hermitize(X: np.ndarray) -> np.ndarray:
    """(X + X^H) / 2 over the last two axes."""
    return 0.5 * (X + np.conj(np.swapaxes(X, -1, -2)))


def expit(x):
    """Numerically stable logistic function, elementwise (utils.hpp:18-25)."""
    x = np.asarray(x)
    out = np.empty_like(x, dtype=np.result_type(x.dtype, np.float64))
    pos = x >= 0
    out[pos] = 1.0 / (1.0 + np.exp(-x[pos]))
    ex = np.exp(x[~pos])
    out[~pos] = ex / (1.0 + ex)
    return out


def batched_eigh(F: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """Batched Hermitian eigh on the GPU; outputs cast back to F's precision."""
    native_required()
    F = np.asarray(F)
    w, v = _native.eigh_batched(np.ascontiguousarray(F, dtype=np.complex128))
    return w.astype(F.real.dtype, copy=False), v.astype(F.dtype, copy=False)


def fftn2d(x: np.ndarray) -> np.ndarray:
    return np.fft.fftn(x, axes=(0, 1))


def ifftshift2d(x: np.ndarray) -> np.ndarray:
    return np.fft.ifftshift(x, axes=(0, 1))
