"""ORACLE — test infrastructure only.

ctypes front-end of ``liboracle_hfcpu.so`` (the CPU FP64 restatement of the
reference hot path in ``hfcpu.cpp``).  Its functions take and return the same
arguments as the reference's pybind11 module ``cpp_hf._native``
(cpp/cpp_hf_native.cpp:242-841), so a parity test can call the oracle and the
GPU module with identical arguments.

Only ``tests/``, ``bench.py`` (its ``cpu_baseline`` leg) and
``__graft_entry__.smoke()`` import this module; the product never does.
"""

from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
LIB_PATH = _HERE / "_build" / "liboracle_hfcpu.so"

_dp = C.POINTER(C.c_double)
_sz = C.c_size_t


_i64p = C.POINTER(C.c_int64)


class SlDesc(C.Structure):
    _fields_ = [
        ("fock_active", C.c_int), ("hartree_active", C.c_int),
        ("hartree_degeneracy", C.c_double),
        ("n_G", _sz), ("dim_orb", _sz), ("n_delta", _sz), ("N_ext_x", _sz), ("N_ext_y", _sz),
        ("n_pairs", _sz),
        ("V_lag_fft", _dp), ("V_lag_fft_orbital", _dp), ("g_a_off", _i64p),
        ("pair_i", _i64p), ("pair_j", _i64p), ("pair_start", _i64p), ("pair_to_delta", _i64p),
        ("HH_GG", _dp), ("HH_GG_orbital", _dp),
    ]


class KernelDesc(C.Structure):
    _fields_ = [
        ("nk1", _sz), ("nk2", _sz), ("nb", _sz), ("dv1", _sz), ("dv2", _sz),
        ("n_contact", _sz),
        ("include_hartree", C.c_int), ("include_exchange", C.c_int),
        ("exchange_hcp", C.c_int), ("has_refP", C.c_int),
        ("T", C.c_double), ("weight_sum", C.c_double),
        ("h", _dp), ("VR", _dp), ("refP", _dp), ("w2d", _dp), ("HH", _dp),
        ("contact_g", _dp), ("contact_Oi", _dp), ("contact_Oj", _dp),
        ("superlattice", C.POINTER(SlDesc)),
    ]


class ScfConfig(C.Structure):
    _fields_ = [
        ("max_iter", _sz), ("density_tol", C.c_double), ("comm_tol", C.c_double),
        ("mixing", C.c_double), ("level_shift", C.c_double),
        ("acceleration", C.c_int), ("diis_size", _sz), ("diis_start", _sz),
        ("diis_damping", C.c_double), ("trust_radius", C.c_double),
        ("block_sizes", C.POINTER(_sz)), ("n_blocks", _sz),
        ("exchange_fp32", C.c_int), ("fp32_switch", C.c_double),
        ("symmetry", C.c_void_p),  # the device projector's slot, always NULL here
    ]


class DmConfig(C.Structure):
    _fields_ = [
        ("max_iter", _sz), ("tol_E", C.c_double), ("tol_grad", C.c_double),
        ("plateau_window", _sz), ("denom_scale", C.c_double),
        ("max_step", C.c_double), ("cg_restart", _sz), ("bt_shrink", C.c_double),
        ("bt_max", _sz), ("mu_maxiter", C.c_int),
        ("block_sizes", C.POINTER(_sz)), ("n_blocks", _sz),
        ("optimizer", C.c_int), ("tr_delta0", C.c_double), ("tr_cg_max", _sz),
        ("deflation_targets", _dp), ("n_deflation", _sz), ("deflation_sigma", C.c_double),
        ("deflation_length", C.c_double), ("exchange_fp32", C.c_int),
        ("fp32_switch", C.c_double), ("symmetry", C.c_void_p),
    ]


class ScfResult(C.Structure):
    _fields_ = [
        ("density", _dp), ("fock", _dp), ("energy", C.c_double), ("mu", C.c_double),
        ("iterations", _sz), ("converged", C.c_int),
        ("hist_E", _dp), ("hist_density", _dp), ("hist_comm", _dp),
    ]


class DmResult(C.Structure):
    _fields_ = [
        ("Q", _dp), ("p", _dp), ("density", _dp), ("fock", _dp),
        ("mu", C.c_double), ("energy", C.c_double), ("n_iter", _sz),
        ("converged", C.c_int), ("hist_E", _dp), ("hist_grad", _dp),
        ("n_ls_trials", _sz),
    ]


PROJECT_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, _dp, _sz, _sz, _sz)

_lib = None


def build() -> Path:
    """Compile the oracle with its Makefile (g++, a few seconds)."""
    subprocess.run(["make", "-s", "-C", str(_HERE)], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = C.CDLL(str(LIB_PATH))
        L.hfcpu_last_error.restype = C.c_char_p
        _lib = L
    return _lib


def _check(rc: int) -> None:
    if rc == 0:
        return
    msg = lib().hfcpu_last_error().decode()
    if rc == 1:
        raise ValueError(msg)
    raise RuntimeError(msg)


def _p(a: np.ndarray | None):
    if a is None:
        return None
    return a.ctypes.data_as(_dp)


def _c128(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.complex128)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def num_threads() -> int:
    return int(lib().hfcpu_num_threads())


def fft2d(x, inverse: bool = False) -> np.ndarray:
    """2-D FFT over axes (0, 1); inverse normalised like numpy.fft.ifftn."""
    buf = _c128(x).copy()
    nk1, nk2 = buf.shape[0], buf.shape[1]
    batch = int(np.prod(buf.shape[2:])) if buf.ndim > 2 else 1
    _check(lib().hfcpu_fft2d(_p(buf), _sz(nk1), _sz(nk2), _sz(batch), int(inverse)))
    return buf


def selfenergy_fft_full(VR, P, apply_ifftshift=True, hermitian_channel_packing=False):
    VR = _c128(VR)
    P = _c128(P)
    if P.ndim != 4:
        raise ValueError("P must be (nk1,nk2,nb,nb)")
    if VR.ndim != 4:
        raise ValueError("VR must be (nk1,nk2,dv1,dv2)")
    nk1, nk2, nb = P.shape[0], P.shape[1], P.shape[2]
    out = np.empty_like(P)
    _check(lib().hfcpu_selfenergy_fft_full(
        _p(VR), _p(P), _p(out), _sz(nk1), _sz(nk2), _sz(nb),
        _sz(VR.shape[2]), _sz(VR.shape[3]), int(bool(apply_ifftshift)),
        int(bool(hermitian_channel_packing))))
    return out


def eigh_batched(M):
    M = _c128(M)
    nb = M.shape[-1]
    nk = int(np.prod(M.shape[:-2]))
    w = np.empty(M.shape[:-1], dtype=np.float64)
    V = np.empty_like(M)
    _check(lib().hfcpu_eigh_batched(_p(M), _p(w), _p(V), _sz(nk), _sz(nb)))
    return w, V


def eigh_block_sizes(M, sizes, sort=True):
    M = _c128(M)
    nb = M.shape[-1]
    nk = M.size // (nb * nb)
    sz = (C.c_size_t * max(len(sizes), 1))(*[int(x) for x in sizes])
    w = np.empty(M.shape[:-1], dtype=np.float64)
    V = np.empty_like(M)
    _check(lib().hfcpu_eigh_block_sizes(_p(M), _p(w), _p(V), _sz(nk), _sz(nb), sz,
                                        _sz(len(sizes)), C.c_int(1 if sort else 0)))
    return w, V


def max_offblock_sizes(M, sizes):
    M = _c128(M)
    nb = M.shape[-1]
    nk = M.size // (nb * nb)
    sz = (C.c_size_t * max(len(sizes), 1))(*[int(x) for x in sizes])
    out = C.c_double(0.0)
    _check(lib().hfcpu_max_offblock_sizes(_p(M), _sz(nk), _sz(nb), sz, _sz(len(sizes)),
                                          C.byref(out)))
    return out.value


def find_mu_bisection(bands, weights, n_e, T, maxiter=0):
    bands = _f64(bands)
    weights = _f64(weights)
    nb = bands.shape[-1]
    nk = int(np.prod(bands.shape[:-1]))
    mu = C.c_double()
    _check(lib().hfcpu_find_mu_bisection(_p(bands), _p(weights), _sz(nk), _sz(nb),
                                         C.c_double(n_e), C.c_double(T), C.c_int(maxiter),
                                         C.byref(mu)))
    return mu.value


def find_mu_newton(bands, weights, n_e, T, maxiter=15):
    """The Python surface's Newton (utils.py:311-360)."""
    bands = _f64(bands)
    weights = _f64(weights)
    nb = bands.shape[-1]
    nk = int(np.prod(bands.shape[:-1]))
    mu = C.c_double()
    _check(lib().hfcpu_find_mu_newton(_p(bands), _p(weights), _sz(nk), _sz(nb), C.c_double(n_e),
                                      C.c_double(T), C.c_int(maxiter), C.byref(mu)))
    return mu.value


def solve_mu_inloop(eps, w_norm, n_target_norm, mu0, T, maxiter=25):
    eps = _f64(eps)
    w_norm = _f64(w_norm)
    nb = eps.shape[-1]
    nk = int(np.prod(eps.shape[:-1]))
    mu = C.c_double()
    _check(lib().hfcpu_solve_mu_inloop(_p(eps), _p(w_norm), _sz(nk), _sz(nb),
                                       C.c_double(n_target_norm), C.c_double(mu0),
                                       C.c_double(T), C.c_int(maxiter), C.byref(mu)))
    return mu.value


def occupation_entropy(p, w_norm):
    p = _f64(p)
    w_norm = _f64(w_norm)
    nb = p.shape[-1]
    nk = int(np.prod(p.shape[:-1]))
    S = C.c_double()
    _check(lib().hfcpu_occupation_entropy(_p(p), _p(w_norm), _sz(nk), _sz(nb), C.byref(S)))
    return S.value


class _Desc:
    """Keeps the contiguous arrays alive while the C call runs."""

    def __init__(self, ka: dict):
        self.h = _c128(ka["h"])
        self.VR = _c128(ka["VR"])
        self.refP = _c128(ka["refP"])
        self.w2d = _f64(ka["w2d"])
        self.HH = _f64(ka["HH"])
        self.cg = _f64(ka["contact_g"])
        self.Oi = _c128(ka["contact_Oi"])
        self.Oj = _c128(ka["contact_Oj"])
        h = self.h
        if h.ndim != 4:
            raise ValueError("h must be (nk1,nk2,nb,nb)")
        if self.VR.ndim != 4:
            raise ValueError("VR must be (nk1,nk2,dv1,dv2)")
        has_refP = bool(ka["has_refP"])
        if has_refP and self.refP.shape != h.shape:
            raise ValueError("refP must match h shape when has_refP is true")
        self.desc = KernelDesc(
            nk1=h.shape[0], nk2=h.shape[1], nb=h.shape[2],
            dv1=self.VR.shape[2], dv2=self.VR.shape[3], n_contact=self.cg.shape[0],
            include_hartree=int(bool(ka["include_hartree"])),
            include_exchange=int(bool(ka["include_exchange"])),
            exchange_hcp=int(bool(ka["exchange_hcp"])), has_refP=int(has_refP),
            T=float(ka["T"]), weight_sum=float(ka["weight_sum"]),
            h=_p(self.h), VR=_p(self.VR), refP=_p(self.refP) if has_refP else None,
            w2d=_p(self.w2d), HH=_p(self.HH), contact_g=_p(self.cg),
            contact_Oi=_p(self.Oi), contact_Oj=_p(self.Oj),
        )
        self.shape = h.shape
        if ka.get("superlattice_fock_active") or ka.get("superlattice_hartree_active"):
            # the superlattice entries of the kernel_args dict (solver.py:104-134)
            i64 = lambda a: np.ascontiguousarray(a, dtype=np.int64)
            self.sl_arrays = dict(
                V=_c128(ka["V_lag_fft"]),
                Vo=_c128(ka["V_lag_fft_orbital"]) if "V_lag_fft_orbital" in ka else None,
                off=i64(ka["g_a_off"]), pi=i64(ka["pair_i"]), pj=i64(ka["pair_j"]),
                ps=i64(ka["pair_start"]), p2d=i64(ka["pair_to_delta"]),
                HH=_f64(ka["HH_GG"]),
                HHo=_f64(ka["HH_GG_orbital"]) if "HH_GG_orbital" in ka else None)
            a = self.sl_arrays
            ip = lambda x: x.ctypes.data_as(_i64p)
            self.sl = SlDesc(
                fock_active=int(bool(ka["superlattice_fock_active"])),
                hartree_active=int(bool(ka["superlattice_hartree_active"])),
                hartree_degeneracy=float(ka.get("hartree_degeneracy", 1.0)),
                n_G=int(ka["n_G"]), dim_orb=int(ka["dim_orb"]), n_delta=int(ka["n_delta"]),
                N_ext_x=int(ka["N_ext_x"]), N_ext_y=int(ka["N_ext_y"]),
                n_pairs=int(a["pi"].shape[0]),
                V_lag_fft=_p(a["V"]), V_lag_fft_orbital=_p(a["Vo"]) if a["Vo"] is not None else None,
                g_a_off=ip(a["off"]), pair_i=ip(a["pi"]), pair_j=ip(a["pj"]),
                pair_start=ip(a["ps"]), pair_to_delta=ip(a["p2d"]), HH_GG=_p(a["HH"]),
                HH_GG_orbital=_p(a["HHo"]) if a["HHo"] is not None else None)
            self.desc.superlattice = C.pointer(self.sl)


def _wrap_project(project_fn):
    if project_fn is None:
        return None

    def cb(_user, M, nk1, nk2, nb):
        try:
            view = np.ctypeslib.as_array(M, shape=(nk1, nk2, nb, nb, 2))
            arr = view.view(np.complex128).reshape(nk1, nk2, nb, nb)
            out = np.ascontiguousarray(project_fn(arr.copy()), dtype=np.complex128)
            if out.shape != (nk1, nk2, nb, nb):
                return 1
            arr[...] = out
            return 0
        except Exception:  # pragma: no cover - surfaced as RuntimeError
            return 1

    return PROJECT_FN(cb)


def build_fock_apply(kernel_args, P, project_fn=None):
    d = _Desc(kernel_args)
    P = _c128(P)
    S = np.empty(d.shape, np.complex128)
    H = np.empty(d.shape, np.complex128)
    F = np.empty(d.shape, np.complex128)
    E = C.c_double()
    cb = _wrap_project(project_fn)
    _check(lib().hfcpu_build_fock_apply(C.byref(d.desc), _p(P), _p(S), _p(H), _p(F),
                                        C.byref(E), cb, None))
    return S, H, F, E.value


def density_from_fock(F, weights, n_e, T, level_shift=0.0):
    F = _c128(F)
    w = _f64(weights)
    nb = F.shape[-1]
    nk = int(np.prod(F.shape[:-2]))
    P = np.empty_like(F)
    mu = C.c_double()
    _check(lib().hfcpu_density_from_fock(_p(F), _sz(nk), _sz(nb), _p(w), C.c_double(n_e),
                                         C.c_double(T), C.c_double(level_shift), _p(P),
                                         C.byref(mu)))
    return P, mu.value


_ACCEL = {"linear": 0, "diis": 1, "oda": 2}


def _size_array(block_sizes):
    # a ctypes size_t array (None when empty); the caller's frame keeps it alive
    sizes = [int(v) for v in (block_sizes or ())]
    return (_sz * len(sizes))(*sizes) if sizes else None


def solve_scf(kernel_args, P0, n_e, max_iter, density_tol, comm_tol, mixing, level_shift,
              project_fn, block_sizes, acceleration, diis_size, diis_start, diis_damping,
              trust_radius, return_density, return_fock):
    d = _Desc(kernel_args)
    P0 = _c128(P0)
    if acceleration not in _ACCEL:
        raise ValueError("SCFConfig.acceleration must be 'linear', 'diis', or 'oda'")
    sizes = _size_array(block_sizes)
    cfg = ScfConfig(max_iter=max_iter, density_tol=density_tol, comm_tol=comm_tol,
                    mixing=mixing, level_shift=level_shift,
                    acceleration=_ACCEL[acceleration], diis_size=diis_size,
                    diis_start=diis_start, diis_damping=diis_damping,
                    trust_radius=trust_radius, block_sizes=sizes, n_blocks=len(sizes or ()))
    dens = np.empty(d.shape, np.complex128)
    fock = np.empty(d.shape, np.complex128)
    hE = np.zeros(max_iter)
    hD = np.zeros(max_iter)
    hC = np.zeros(max_iter)
    res = ScfResult(density=_p(dens), fock=_p(fock), hist_E=_p(hE), hist_density=_p(hD),
                    hist_comm=_p(hC))
    cb = _wrap_project(project_fn)
    _check(lib().hfcpu_solve_scf(C.byref(d.desc), _p(P0), C.c_double(n_e), C.byref(cfg),
                                 cb, None, C.byref(res)))
    n = int(res.iterations)
    return (dens if return_density else None, fock if return_fock else None,
            res.energy, res.mu, n, bool(res.converged), hE[:n].copy(), hD[:n].copy(),
            hC[:n].copy())


def solve_dm(kernel_args, P0, n_e, max_iter, tol_E, tol_grad, max_step, bt_shrink,
             denom_scale, bt_max, cg_restart, plateau_window, mu_maxiter, optimizer,
             tr_delta0, tr_cg_max, deflation_targets, deflation_sigma, deflation_length,
             block_sizes, project_fn, return_Q, return_density, return_fock,
             return_stats=False):
    if int(optimizer) not in (0, 1):
        raise ValueError("optimizer code must be 0 (cg) or 1 (newton)")
    d = _Desc(kernel_args)
    P0 = _c128(P0)
    sizes = _size_array(block_sizes)
    defl = _c128(deflation_targets) if deflation_targets is not None else None
    n_defl = int(defl.shape[0]) if defl is not None and defl.ndim == 5 else 0
    cfg = DmConfig(max_iter=max_iter, tol_E=tol_E, tol_grad=tol_grad,
                   plateau_window=plateau_window, denom_scale=denom_scale,
                   max_step=max_step, cg_restart=cg_restart, bt_shrink=bt_shrink,
                   bt_max=bt_max, mu_maxiter=mu_maxiter, block_sizes=sizes,
                   n_blocks=len(sizes or ()), optimizer=int(optimizer),
                   tr_delta0=float(tr_delta0), tr_cg_max=int(tr_cg_max),
                   deflation_targets=_p(defl) if n_defl else None, n_deflation=n_defl,
                   deflation_sigma=float(deflation_sigma),
                   deflation_length=float(deflation_length))
    Q = np.empty(d.shape, np.complex128)
    p = np.empty(d.shape[:-1], np.float64)
    dens = np.empty(d.shape, np.complex128)
    fock = np.empty(d.shape, np.complex128)
    hE = np.zeros(max(max_iter, 1))
    hG = np.zeros(max(max_iter, 1))
    res = DmResult(Q=_p(Q), p=_p(p), density=_p(dens), fock=_p(fock), hist_E=_p(hE),
                   hist_grad=_p(hG))
    cb = _wrap_project(project_fn)
    _check(lib().hfcpu_solve_dm(C.byref(d.desc), _p(P0), C.c_double(n_e), C.byref(cfg),
                                cb, None, C.byref(res)))
    n = int(res.n_iter)
    out = (Q if return_Q else None, p, dens if return_density else None,
           fock if return_fock else None, res.mu, res.energy, n, bool(res.converged),
           hE[:n].copy(), hG[:n].copy())
    return out + ({"n_ls_trials": int(res.n_ls_trials)},) if return_stats else out


def selfenergy_superlattice_streamed(rho, VR_fft, g_a_off, pair_i, pair_j, pair_start, nkx,
                                     nky, n_G, dim_orb, N_ext_x, N_ext_y, VR_fft_orbital=None,
                                     hermitian_rho=False):
    """The superlattice exchange primitive (superlattice.hpp:53-246) with the
    reference module's keyword signature (cpp_hf_native.cpp:718-790)."""
    rho = _c128(rho)
    V = _c128(VR_fft)
    Vo = _c128(VR_fft_orbital) if VR_fft_orbital is not None else None
    i64 = lambda a: np.ascontiguousarray(a, dtype=np.int64)
    off, pi, pj, ps = i64(g_a_off), i64(pair_i), i64(pair_j), i64(pair_start)
    out = np.empty_like(rho)
    ip = lambda x: x.ctypes.data_as(_i64p)
    _check(lib().hfcpu_selfenergy_superlattice(
        _p(rho), _p(out), _p(V), _p(Vo) if Vo is not None else None, ip(off), ip(pi), ip(pj),
        ip(ps), _sz(ps.shape[0] - 1), _sz(N_ext_x), _sz(N_ext_y), _sz(nkx), _sz(nky), _sz(n_G),
        _sz(dim_orb), C.c_int(int(bool(hermitian_rho)))))
    return out


def _rtr_joint_hvp(kernel_args, X, dp, Q, p, Ft):
    """Joint (Q, p) Hessian-vector product of the trust-region solver
    (solver_rtr.hpp:55-112), as the reference's _native._rtr_joint_hvp."""
    d = _Desc(kernel_args)
    X, Q, Ft = _c128(X), _c128(Q), _c128(Ft)
    dp, p = _f64(dp), _f64(p)
    HX = np.empty(d.shape, np.complex128)
    Hp = np.empty(d.shape[:-1], np.float64)
    _check(lib().hfcpu_rtr_joint_hvp(C.byref(d.desc), _p(X), _p(dp), _p(Q), _p(p), _p(Ft),
                                     _p(HX), _p(Hp)))
    return HX, Hp


def _deflation_bias(P, targets, w2d, sigma, length):
    """(Phi, S_pen, diag_coeff) of the deflation bias (solver_rtr.hpp:127-160)."""
    P = _c128(P)
    T = _c128(targets)
    w = _f64(w2d)
    nb = P.shape[-1]
    nk = int(np.prod(P.shape[:-2]))
    n_defl = int(T.shape[0]) if T.ndim == 5 else 0
    S = np.empty_like(P)
    Phi = C.c_double()
    dg = C.c_double()
    _check(lib().hfcpu_deflation_bias(_p(P), _p(T) if n_defl else None, _sz(n_defl), _p(w),
                                      _sz(nk), _sz(nb), C.c_double(sigma), C.c_double(length),
                                      _p(S), C.byref(Phi), C.byref(dg)))
    return Phi.value, S, dg.value


def _cayley_spectral_setup(d):
    d = _c128(d)
    nb = d.shape[-1]
    nk = int(np.prod(d.shape[:-2]))
    V = np.empty_like(d)
    lam = np.empty(d.shape[:-1])
    _check(lib().hfcpu_cayley_spectral_setup(_p(d), _p(V), _p(lam), _sz(nk), _sz(nb)))
    return V, lam


def _cayley_unitary_from_spectrum(V, lam, tau):
    V = _c128(V)
    lam = _f64(lam)
    nb = V.shape[-1]
    nk = int(np.prod(V.shape[:-2]))
    U = np.empty_like(V)
    _check(lib().hfcpu_cayley_unitary_from_spectrum(_p(V), _p(lam), C.c_double(tau), _p(U),
                                                    _sz(nk), _sz(nb)))
    return U


def _diag_UFU_from_spectrum(V, Ft_eig, lam, tau):
    V = _c128(V)
    Fe = _c128(Ft_eig)
    lam = _f64(lam)
    nb = V.shape[-1]
    nk = int(np.prod(V.shape[:-2]))
    out = np.empty(V.shape[:-1])
    _check(lib().hfcpu_diag_UFU_from_spectrum(_p(V), _p(Fe), _p(lam), C.c_double(tau),
                                              _p(out), _sz(nk), _sz(nb)))
    return out


def resample_kgrid_2d(values, nk_new):
    x = _c128(values)
    if x.ndim < 2:
        raise ValueError("need >= 2 dims")
    nk_old = x.shape[0]
    if x.shape[1] != nk_old:
        raise ValueError("First two axes must be square (nk_old, nk_old)")
    inner = int(np.prod(x.shape[2:])) if x.ndim > 2 else 1
    out = np.empty((nk_new, nk_new) + x.shape[2:], np.complex128)
    _check(lib().hfcpu_resample_kgrid_2d(_p(x), _p(out), _sz(nk_old), _sz(nk_new),
                                         _sz(inner)))
    return out


def kernel_args(kernel) -> dict:
    """Reference ``_kernel_args_for_native`` (src/cpp_hf/solver.py:79-103)."""
    refP = kernel._refP if kernel._refP is not None else kernel._empty_refP
    return dict(
        h=_c128(kernel.h), VR=_c128(kernel._VR_shifted), refP=_c128(refP),
        has_refP=bool(kernel.has_reference_density), w2d=_f64(kernel.w2d),
        HH=_f64(kernel.HH), contact_g=_f64(kernel.contact_g),
        contact_Oi=_c128(kernel.contact_Oi), contact_Oj=_c128(kernel.contact_Oj),
        weight_sum=float(kernel.weight_sum), T=float(kernel.T),
        include_hartree=bool(kernel.include_hartree),
        include_exchange=bool(kernel.include_exchange),
        exchange_hcp=bool(kernel.exchange_hermitian_channel_packing),
    )

