"""Benchmark: Fock builds/s at nb = 8 on a 96 x 96 k-mesh, FP64, one GPU.

Workload (the production-size configuration): 8-band layered QWZ model
(4 layers x 2 sublattices), 96 x 96 mesh, layer-resolved V(q) (64 band-pair
channels, Hermitian-packed to 36), Hartree with reference density I/2,
T = 0.05, 4.1 electrons per k.  One step = one device-resident Fock build
(exchange convolution + Hartree + F = herm(h + Sigma + H)) on a fixed
Hermitian density held in HBM.

The JSON line also carries:
  roofline      achieved algorithmic GB/s of the Fock build (6 A bytes per
                build, A = 16 nk nb^2: P, refP, h and VR in, Sigma and F out)
                over its HIP-event-timed duration, against 8 TB/s HBM;
                traffic = HBM bytes per build from the committed
                FETCH_SIZE/WRITE_SIZE summary of this workload.
  cpu_baseline  the CPU oracle's Fock build on this host: 50-build warm-up,
                median and interquartile range over 5 repeats (a
                restatement of the reference's FFTW/Eigen path, which cannot
                be built here), its one-thread rate, and the 48 sweep
                kernels solved one after another by the oracle.
  solves        wall time of full direct-minimisation (CG and trust-region
                Newton) and SCF (mixing 0.5) solves of the two-band QWZ model
                at 24 x 24 and the eight-band layered model at 96 x 96, on the
                GPU (Newton: GPU only) and in the
                oracle on the host, with each solve's roofline time: the
                reference dataflow's bytes per iteration (DM (30 + 2 trials) A
                with the line-search trials counted, SCF 21 A) over 8 TB/s.
  mixed_precision  the FP32-workspace Fock build against the FP64 one at
                96 x 96 (nb 8) and 192 x 192 (nb 16), and SCF solves at
                192 x 192 with and without the mixed-precision phase: times,
                |dF| per k, max |dSigma|, |dmu|, gap error.
  adjoint       time of one device VJP of the Fock build at 144 x 144, nb 8.
  superlattice  device Fock builds of the moire continuum model in its
                plane-wave basis (exchange on the extended grid + superlattice
                Hartree shift) at 24 x 24 k (13 G x 2 orbitals) and 36 x 36 k
                (19 G x 1 orbital), HIP-event timed, with the grid bytes one
                build moves, the oracle's build on this host and an SCF solve.
  autotune      the launch-shape autotuner's pick (made at set-up: the timed
                builds use it) and the default shape's time.
  symmetry      SCF solves at 96 x 96, nb 8 with a time-reversal projector
                applied on the device (SymmetryProjector) and through the
                host callback (a plain function): seconds and iterations.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
TRAFFIC_PROFILE = ROOT / "profiles" / "fock_build_96x96_nb8.md"


def measured_traffic(nk1: int, nk2: int, nb: int):
    """HBM bytes per Fock build from the committed FETCH_SIZE / WRITE_SIZE
    summary (copy-calibrated), for the workload that summary was taken on."""
    if (nk1, nk2, nb) != (96, 96, 8) or not TRAFFIC_PROFILE.exists():
        return None
    for line in TRAFFIC_PROFILE.read_text().splitlines():
        if line.startswith("traffic_bytes_per_build:"):
            return int(line.split(":", 1)[1])
    return None


def _gap_about(F: np.ndarray, mu: float) -> float:
    """Smallest band energy above mu minus the largest below (over all k)."""
    e = np.linalg.eigvalsh(F)
    return float(np.min(e[e > mu]) - np.max(e[e <= mu]))


def run_mixed_precision(torch, dev) -> dict:
    """16 bands (8 layers) on 192 x 192: the FP32-workspace Fock build against
    the FP64 one (time and error of one build), and SCF solves (DIIS) with and
    without the mixed-precision phase (errors of the converged state)."""
    from cpp_hf_amd import SCFConfig, solve_scf
    from cpp_hf_amd._compat import _native
    from cpp_hf_amd.solver import _kernel_args_for_native
    from cpp_hf_amd.workloads import initial_density, layered_kernel

    out = {}
    for n, nl in ((96, 4), (192, 8)):
        wl = layered_kernel(n, nl, T=0.05, filling=nl + 0.1)
        gk = _native.GpuKernel(_kernel_args_for_native(wl.kernel))
        nb = 2 * nl
        P0 = initial_density(n, nb, 0)
        P = torch.from_numpy(np.ascontiguousarray(0.5 * (P0 + np.conj(np.swapaxes(P0, -1, -2))))).to(dev)
        S = [torch.empty_like(P), torch.empty_like(P)]
        F = [torch.empty_like(P), torch.empty_like(P)]
        hd = torch.empty(nb, dtype=torch.float64, device=dev)
        stream = torch.cuda.Stream(device=dev)
        calls = (gk.build_fock_device, gk.build_fock_mixed_device)
        times = []
        for i, call in enumerate(calls):
            for _ in range(3):
                call(P.data_ptr(), S[i].data_ptr(), F[i].data_ptr(), hd.data_ptr(), True,
                     stream.cuda_stream)
            ev0 = torch.cuda.Event(enable_timing=True)
            ev1 = torch.cuda.Event(enable_timing=True)
            torch.cuda.synchronize()
            ev0.record(stream)
            for _ in range(20):
                call(P.data_ptr(), S[i].data_ptr(), F[i].data_ptr(), hd.data_ptr(), True,
                     stream.cuda_stream)
            ev1.record(stream)
            torch.cuda.synchronize()
            times.append(ev0.elapsed_time(ev1) * 1e3 / 20)
        dS = (S[1] - S[0]).abs().max().item()
        dF = (F[1] - F[0]).abs().max().item()
        out[f"build_{n}x{n}_nb{nb}"] = {
            "fp64_us": times[0], "mixed_us": times[1], "speedup": times[0] / times[1],
            "mixed_capable": bool(gk.mixed_capable(True)),
            "max_abs_dSigma": dS, "max_rel_dSigma": dS / S[0].abs().max().item(),
            "max_abs_dF": dF,
        }
        del S, F, P, gk
        torch.cuda.empty_cache()

    wl = layered_kernel(192, 8, T=0.05, filling=8.1)
    base = dict(max_iter=400, mixing=0.5, acceleration="diis", density_tol=1e-7, comm_tol=1e-8)
    res = {}
    for prec in ("fp64", "mixed"):
        t = time.perf_counter()
        r = solve_scf(wl.kernel, wl.P0, wl.n_electrons,
                      config=SCFConfig(exchange_precision=prec, **base))
        res[prec] = (r, time.perf_counter() - t)
    r64, t64 = res["fp64"]
    r32, t32 = res["mixed"]
    F64 = np.asarray(r64.fock_matrix)
    F32 = np.asarray(r32.fock_matrix)
    mu64, mu32 = float(r64.chemical_potential), float(r32.chemical_potential)
    out["scf_192x192_nb16"] = {
        "config": "layered QWZ, 8 layers, 192x192 k, layer-resolved V, Hartree+refP, "
                  "T=0.05, 8.1 electrons/k, SCF DIIS, density_tol 1e-7, comm_tol 1e-8",
        "fp64_s": t64, "mixed_s": t32,
        "fp64_iterations": int(r64.iterations), "mixed_iterations": int(r32.iterations),
        "converged": [bool(r64.converged), bool(r32.converged)],
        "max_k_dF": float(np.max(np.sqrt(np.sum(np.abs(F32 - F64) ** 2, axis=(-2, -1))))),
        "dmu": abs(mu32 - mu64),
        "dE": abs(float(r32.energy) - float(r64.energy)),
        "gap_fp64": _gap_about(F64, mu64),
        "gap_error": abs(_gap_about(F32, mu32) - _gap_about(F64, mu64)),
    }

    def max_k_dF(A, B):
        return float(np.max(np.sqrt(np.sum(np.abs(np.asarray(A) - np.asarray(B)) ** 2,
                                           axis=(-2, -1)))))

    # the SCF's own reproducibility: two FP64 runs from P0 and from P0 moved
    # by 1e-3 (mu from 30 bisection halvings limits how close they end)
    rng = np.random.default_rng(5)
    X = rng.standard_normal(wl.P0.shape) + 1j * rng.standard_normal(wl.P0.shape)
    P1 = wl.P0 + 5e-4 * (X + np.conj(np.swapaxes(X, -1, -2)))
    tight = dict(base, density_tol=1e-9, comm_tol=1e-9, max_iter=300)
    ra = solve_scf(wl.kernel, wl.P0, wl.n_electrons, config=SCFConfig(**tight))
    rb = solve_scf(wl.kernel, P1, wl.n_electrons, config=SCFConfig(**tight))
    out["scf_192x192_nb16"]["fp64_two_starts_max_k_dF"] = max_k_dF(ra.fock_matrix, rb.fock_matrix)

    # DM with the in-loop Newton mu (resolved to rounding): FP32-workspace
    # builds until |G| <= 1e-4, FP64 afterwards, against FP64 throughout
    from cpp_hf_amd import SolverConfig, solve

    dm = {}
    for prec in ("fp64", "mixed"):
        t = time.perf_counter()
        r = solve(wl.kernel, wl.P0, wl.n_electrons,
                  config=SolverConfig(max_iter=300, tol_E=0.0, tol_grad=0.0, plateau_window=0,
                                      exchange_precision=prec, mixed_switch=1e-4))
        dm[prec] = (r, time.perf_counter() - t)
    (d64, td64), (d32, td32) = dm["fp64"], dm["mixed"]
    out["dm_192x192_nb16"] = {
        "config": "same model, direct minimisation (CG), 300 iterations, Newton mu; mixed: "
                  "FP32 workspace until |G| <= 1e-4",
        "fp64_s": td64, "mixed_s": td32, "speedup": td64 / td32,
        "grad_norm": [float(d64.history["grad_norm"][-1]), float(d32.history["grad_norm"][-1])],
        "max_k_dF": max_k_dF(d64.fock, d32.fock),
        "dmu": abs(float(d64.mu) - float(d32.mu)),
        "dE": abs(float(d64.energy) - float(d32.energy)),
    }
    return out


def run_adjoint(torch, dev) -> dict:
    """Time of one device VJP of the Fock build at 144 x 144, 8 bands."""
    from cpp_hf_amd._compat import _native
    from cpp_hf_amd.solver import _kernel_args_for_native
    from cpp_hf_amd.workloads import initial_density, layered_kernel

    wl = layered_kernel(144, 4, T=0.05, filling=4.1)
    gk = _native.GpuKernel(_kernel_args_for_native(wl.kernel))
    P0 = initial_density(144, 8, 0)
    P = torch.from_numpy(np.ascontiguousarray(P0)).to(dev)
    G = torch.from_numpy(np.ascontiguousarray(initial_density(144, 8, 2))).to(dev)
    gP = torch.empty_like(P)
    gV = torch.empty((144, 144, 8, 8), dtype=torch.complex128, device=dev)
    gh = torch.empty_like(P)
    gHH = torch.empty((8, 8), dtype=torch.float64, device=dev)
    stream = torch.cuda.Stream(device=dev)

    def vjp():
        gk.build_fock_vjp_device(P.data_ptr(), 0, 0, G.data_ptr(), gP.data_ptr(), gV.data_ptr(),
                                 gh.data_ptr(), gHH.data_ptr(), stream.cuda_stream)

    vjp()
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(10):
        vjp()
    torch.cuda.synchronize()
    return {"workload": "layered QWZ 144x144 k, nb=8, cotangent on F",
            "ms_per_vjp": (time.perf_counter() - t) * 1e2}


def run_superlattice(torch, dev, cpu: bool) -> dict:
    """Device superlattice Fock builds (and one SCF solve) of the moire model."""
    from cpp_hf_amd import SCFConfig, solve_scf
    from cpp_hf_amd._compat import _native
    from cpp_hf_amd.solver import _kernel_args_for_native
    from cpp_hf_amd.workloads import moire_kernel

    out = {}
    for n, shells, orb in ((24, 2, 2), (36, 3, 1)):
        kern, P0, n_e = moire_kernel(n, shells, orb)
        L = kern.layout
        D = kern.h.shape[-1]
        ka = _kernel_args_for_native(kern)
        gk = _native.GpuKernel(ka)
        P = torch.from_numpy(np.ascontiguousarray(P0)).to(dev)
        S = torch.empty_like(P)
        F = torch.empty_like(P)
        hd = torch.empty(D, dtype=torch.float64, device=dev)
        stream = torch.cuda.Stream(device=dev)

        def build():
            gk.build_fock_device(P.data_ptr(), S.data_ptr(), F.data_ptr(), hd.data_ptr(), True,
                                 stream.cuda_stream)

        for _ in range(3):
            build()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        steps = 20
        e0.record(stream)
        for _ in range(steps):
            build()
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / steps
        # one channel of each (+dG, -dG) pair is transformed (the basis is
        # inversion symmetric, dG = 0 is its own partner)
        channels = (L.n_delta + 1) // 2 * orb * orb
        grid = L.N_ext_x * L.N_ext_y * channels * 16
        # the fused column pass reads and writes the whole grid (2 sweeps);
        # the row passes add only the rows that hold data, so the rate below
        # is a lower bound of the bytes moved per second
        sweeps = 2
        rec = {
            "workload": f"moire continuum {n}x{n} k, {kern.n_G} G x {orb} orbitals (D={D}), "
                        f"extended grid {L.N_ext_x}x{L.N_ext_y}, {channels} channels",
            "ms_per_build": ms,
            "grid_bytes": grid,
            "grid_sweep_GBs": sweeps * grid / (ms * 1e-3) / 1e9,
        }
        if cpu:
            from oracle import hfcpu

            t = time.perf_counter()
            Sr, Hr, Fr, Er = hfcpu.build_fock_apply(ka, P0)
            rec["cpu_oracle_ms_per_build"] = (time.perf_counter() - t) * 1e3
            rec["cpu_threads"] = hfcpu.num_threads()
            rec["parity_rel_err_F"] = float(np.max(np.abs(F.cpu().numpy() - Fr)) /
                                            np.max(np.abs(Fr)))
        t = time.perf_counter()
        r = solve_scf(kern, P0, n_e, config=SCFConfig(acceleration="diis", max_iter=100))
        rec["scf_diis"] = {"seconds": time.perf_counter() - t, "iterations": r.iterations,
                           "converged": bool(r.converged), "energy": float(r.energy)}
        out[f"moire_{n}x{n}_D{D}"] = rec
    return out


def run_symmetry() -> dict:
    """SCF of the 96 x 96 eight-band layered model with a projector: device
    averaging vs the host callback."""
    from cpp_hf_amd import SCFConfig, make_project_fn, solve_scf
    from cpp_hf_amd.workloads import layered_nb8

    wl = layered_nb8(96)
    nb = wl.kernel.h.shape[-1]
    proj = make_project_fn(time_reversal_U=np.eye(nb))
    out = {"workload": "layered QWZ 96x96 k, nb=8, SCF (DIIS, 40 iterations at most), "
                       "time reversal U = I, mod convention"}
    for name, pf in (("device", proj), ("host_callback", lambda A: proj(A))):
        cfg = SCFConfig(acceleration="diis", max_iter=40, project_fn=pf)
        solve_scf(wl.kernel, wl.P0, wl.n_electrons, config=SCFConfig(max_iter=2, project_fn=pf))
        t = time.perf_counter()
        r = solve_scf(wl.kernel, wl.P0, wl.n_electrons, config=cfg)
        out[name] = {"seconds": time.perf_counter() - t, "iterations": int(r.iterations),
                     "converged": bool(r.converged), "energy": float(r.energy)}
    out["speedup"] = out["host_callback"]["seconds"] / out["device"]["seconds"]
    return out


def _dm_args(kern, P0, n_e):
    """solve()'s positional arguments of _native.solve_dm (default SolverConfig)."""
    from cpp_hf_amd import SolverConfig
    from cpp_hf_amd.solver import _kernel_args_for_native

    c = SolverConfig()
    return (_kernel_args_for_native(kern), np.ascontiguousarray(P0, dtype=np.complex128),
            float(n_e), int(c.max_iter), float(c.tol_E), float(c.tol_grad), float(c.max_step),
            float(c.bt_shrink), float(c.denom_scale), int(c.bt_max), int(c.cg_restart),
            int(c.plateau_window), int(c.mu_maxiter), 0, float(c.tr_delta0), int(c.tr_cg_max),
            np.empty((0,) + kern.h.shape, dtype=np.complex128), 0.0, 1.0, [], None, False,
            False, False)


def _scf_args(kern, P0, n_e, mixing=0.5):
    """solve_scf()'s positional arguments of _native.solve_scf (SCFConfig(mixing))."""
    from cpp_hf_amd import SCFConfig
    from cpp_hf_amd.solver import _kernel_args_for_native

    c = SCFConfig(mixing=mixing)
    return (_kernel_args_for_native(kern), np.ascontiguousarray(P0, dtype=np.complex128),
            float(n_e), int(c.max_iter), float(c.density_tol), float(c.comm_tol),
            float(c.mixing), float(c.level_shift), None, [], c.acceleration, int(c.diis_size),
            int(c.diis_start), float(c.diis_damping), float(c.trust_radius), False, False)


def run_solves(cpu: bool) -> dict:
    """DM and SCF solves of the 24 x 24 QWZ and the 96 x 96 layered models on
    the GPU (and in the oracle), with their roofline times."""
    from cpp_hf_amd._compat import _native
    from cpp_hf_amd.workloads import layered_nb8, qwz_kernel

    out = {}
    for name, wl in (("qwz_24x24_nb2", qwz_kernel(24)), ("layered_96x96_nb8", layered_nb8(96))):
        kern = wl.kernel
        nk = kern.h.shape[0] * kern.h.shape[1]
        nb = kern.h.shape[-1]
        A = 16 * nk * nb * nb
        dm_args = _dm_args(kern, wl.P0, wl.n_electrons)
        scf_args = _scf_args(kern, wl.P0, wl.n_electrons)
        _native.solve_dm(*dm_args)  # pages in this mesh's code objects
        t = time.perf_counter()
        r = _native.solve_dm(*dm_args, return_stats=True)
        t_dm = time.perf_counter() - t
        n_it, trials = int(r[6]), int(r[10]["n_ls_trials"])
        dm_bytes = (30 * n_it + 2 * trials) * A
        t = time.perf_counter()
        s_ = _native.solve_scf(*scf_args)
        t_scf = time.perf_counter() - t
        scf_bytes = 21 * int(s_[4]) * A
        # trust-region Newton (optimizer code 1) from the same start
        nt_args = dm_args[:13] + (1,) + dm_args[14:]
        _native.solve_dm(*nt_args)
        t = time.perf_counter()
        rn = _native.solve_dm(*nt_args)
        t_nt = time.perf_counter() - t
        rec = {
            "newton": {"seconds": t_nt, "n_iter": int(rn[6]), "converged": bool(rn[7]),
                       "energy": float(rn[5])},
            "dm": {"seconds": t_dm, "n_iter": n_it, "ls_trials": trials,
                   "converged": bool(r[7]), "energy": float(r[5]),
                   "roofline_seconds": dm_bytes / (HBM_PEAK_GBS * 1e9),
                   "roofline_frac": dm_bytes / (HBM_PEAK_GBS * 1e9) / t_dm},
            "scf": {"seconds": t_scf, "iterations": int(s_[4]), "converged": bool(s_[5]),
                    "energy": float(s_[2]),
                    "roofline_seconds": scf_bytes / (HBM_PEAK_GBS * 1e9),
                    "roofline_frac": scf_bytes / (HBM_PEAK_GBS * 1e9) / t_scf},
        }
        if cpu:
            from oracle import hfcpu

            t = time.perf_counter()
            rc = hfcpu.solve_dm(*dm_args)
            rec["dm"]["cpu_oracle_seconds"] = time.perf_counter() - t
            rec["dm"]["cpu_oracle_n_iter"] = int(rc[6])
            t = time.perf_counter()
            sc = hfcpu.solve_scf(*scf_args)
            rec["scf"]["cpu_oracle_seconds"] = time.perf_counter() - t
            rec["scf"]["cpu_oracle_iterations"] = int(sc[4])
        out[name] = rec
    return out


def cpu_one_thread_builds(seconds: float) -> dict:
    """The oracle's 96 x 96 x 8 Fock builds on one host thread (a child
    process with OMP_NUM_THREADS=1: the thread count is fixed when OpenMP
    starts)."""
    import subprocess

    code = (
        "import sys, time, json, numpy as np; sys.path.insert(0, %r)\n"
        "from cpp_hf_amd.workloads import layered_nb8\n"
        "from cpp_hf_amd.solver import _kernel_args_for_native\n"
        "from oracle import hfcpu\n"
        "wl = layered_nb8(96); ka = _kernel_args_for_native(wl.kernel)\n"
        "P = 0.5 * (wl.P0 + np.conj(np.swapaxes(wl.P0, -1, -2)))\n"
        "hfcpu.build_fock_apply(ka, P); n = 0; t = time.perf_counter()\n"
        "while time.perf_counter() - t < %r:\n"
        "    hfcpu.build_fock_apply(ka, P); n += 1\n"
        "dt = time.perf_counter() - t\n"
        "print(json.dumps({'value': n / dt, 'builds': n, 'seconds': dt, "
        "'threads': hfcpu.num_threads()}))\n") % (str(ROOT), float(seconds))
    env = dict(os.environ, OMP_NUM_THREADS="1")
    r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True,
                       timeout=120)
    if r.returncode != 0:
        return {"error": r.stderr[-500:]}
    return json.loads(r.stdout.strip().splitlines()[-1])


CPU_LABEL = ("cpp_hf-semantics CPU restatement (own FFT/Jacobi, C++17 + OpenMP), "
             "not the FFTW/Eigen binary")


def cpu_baseline(ka, P_host, per_repeat: int) -> dict:
    """The oracle's Fock builds on the host: 50-call warm-up, then 5 timed
    repeats of `per_repeat` builds; median and interquartile range of the
    repeats' rates."""
    from oracle import hfcpu

    for _ in range(50):
        hfcpu.build_fock_apply(ka, P_host)
    rates = []
    for _ in range(5):
        t = time.perf_counter()
        for _ in range(per_repeat):
            hfcpu.build_fock_apply(ka, P_host)
        rates.append(per_repeat / (time.perf_counter() - t))
    q1, med, q3 = (float(x) for x in np.percentile(rates, [25, 50, 75]))
    return {
        "value": med,
        "unit": "Fock builds/s",
        "cores": hfcpu.num_threads(),
        "kind": "port",
        "label": CPU_LABEL,
        "sample": f"median of 5 repeats of {per_repeat} Fock builds (+ energy) of the same "
                  "workload after a 50-build warm-up; OpenMP over k / channels, serial "
                  "reductions",
        "iqr": [q1, q3],
        "repeats": rates,
        "one_thread": cpu_one_thread_builds(5.0),
    }


def cpu_sweep_sequential() -> dict:
    """The 48 sweep kernels solved one after another by the oracle (default
    SolverConfig, DM solver), as the reference runs a sweep without a pool."""
    from oracle import hfcpu

    from cpp_hf_amd.workloads import sweep_params, sweep_workload

    ws = [sweep_workload(p) for p in sweep_params(48, seed=1)]
    t = time.perf_counter()
    iters = 0
    for w in ws:
        iters += int(hfcpu.solve_dm(*_dm_args(w.kernel, w.P0, w.n_electrons))[6])
    dt = time.perf_counter() - t
    return {"solves_per_s": len(ws) / dt, "seconds": dt, "kernels": len(ws),
            "total_iterations": iters, "threads": hfcpu.num_threads(), "label": CPU_LABEL}


def run_sweep(concurrency: int) -> dict:
    """48-kernel parameter sweep (meshes 36..120, nb 2..12, seed-1 draws of
    u, interaction scale, screening length and filling): the DM solves one by
    one, then with `concurrency` solves in flight on separate streams, and
    with the solves of one band count grouped (one launch sequence per
    iteration per group, groups concurrent)."""
    from cpp_hf_amd import SolverConfig, solve_many
    from cpp_hf_amd.workloads import sweep_params, sweep_workload

    wls = [sweep_workload(p) for p in sweep_params(48, seed=1)]
    kernels = [w.kernel for w in wls]
    P0s = [w.P0 for w in wls]
    n_es = [w.n_electrons for w in wls]
    cfg = SolverConfig()
    # untimed passes page in the code objects and fill the memory pool
    # grouped: solves of one band count advanced together in groups of up to
    # 3, 8 groups in flight (the best of the shapes measured on MI355X)
    gkw = dict(concurrency=8, batch=True, max_batch=3)
    solve_many(kernels, P0s, n_es, config=cfg, concurrency=concurrency)
    solve_many(kernels, P0s, n_es, config=cfg, **gkw)
    # three timed passes of each mode, interleaved; medians reported (single
    # passes of the concurrent modes vary by ~10 % from run to run)
    times = {"seq": [], "par": [], "grp": []}
    for _ in range(3):
        t = time.perf_counter()
        seq = solve_many(kernels, P0s, n_es, config=cfg, concurrency=1)
        times["seq"].append(time.perf_counter() - t)
        t = time.perf_counter()
        par = solve_many(kernels, P0s, n_es, config=cfg, concurrency=concurrency)
        times["par"].append(time.perf_counter() - t)
        t = time.perf_counter()
        grp = solve_many(kernels, P0s, n_es, config=cfg, **gkw)
        times["grp"].append(time.perf_counter() - t)
    t_seq, t_par, t_grp = (float(np.median(times[k])) for k in ("seq", "par", "grp"))

    def same(xs):
        return all(int(a.n_iter) == int(b.n_iter) and float(a.energy) == float(b.energy)
                   and np.array_equal(a.density, b.density) for a, b in zip(seq, xs))

    # roofline of the sweep: every solve's reference dataflow bytes,
    # (30 n_iter + 2 line-search trials) A, over the HBM peak (the trials
    # counted by the solver itself, in an untimed pass)
    from cpp_hf_amd._compat import _native

    sweep_bytes = 0
    for w in wls:
        r = _native.solve_dm(*_dm_args(w.kernel, w.P0, w.n_electrons), return_stats=True)
        A = 16 * w.kernel.h.shape[0] * w.kernel.h.shape[1] * w.kernel.h.shape[-1] ** 2
        sweep_bytes += (30 * int(r[6]) + 2 * int(r[10]["n_ls_trials"])) * A
    roof_s = sweep_bytes / (HBM_PEAK_GBS * 1e9)

    return {
        "kernels": len(wls),
        "timing": "median of 3 timed passes per mode",
        "sequential_solves_per_s": len(wls) / t_seq,
        "batched_solves_per_s": len(wls) / t_par,
        "speedup": t_seq / t_par,
        "concurrency": concurrency,
        "method": "concurrent streams, one host thread per in-flight solve",
        "grouped_solves_per_s": len(wls) / t_grp,
        "grouped_speedup": t_seq / t_grp,
        "grouped_method": "solves of one band count (any mesh) advanced by one launch sequence "
                          "per iteration in groups of up to 3, 8 groups concurrent",
        "converged": int(sum(bool(r.converged) for r in par)),
        "total_iterations": int(sum(int(r.n_iter) for r in par)),
        "roofline_seconds": roof_s,
        "roofline_bytes": sweep_bytes,
        "roofline_frac_sequential": roof_s / t_seq,
        "roofline_frac_batched": roof_s / t_par,
        "roofline_frac_grouped": roof_s / t_grp,
        "batched_equals_sequential": bool(same(par)),
        "grouped_equals_sequential": bool(same(grp)),
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--n", type=int, default=96)
    ap.add_argument("--cpu-repeat-builds", type=int, default=200,
                    help="oracle Fock builds per timed repeat of the CPU baseline (5 repeats)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-solves", action="store_true")
    ap.add_argument("--no-sweep", action="store_true",
                    help="skip the 48-kernel sweep (sequential vs concurrent solves)")
    ap.add_argument("--sweep-concurrency", type=int, default=8,
                    help="solves in flight in the sweep (8 measured fastest on MI355X)")
    ap.add_argument("--no-mixed", action="store_true",
                    help="skip the 192 x 192 mixed-precision and the 144 x 144 adjoint sections")
    ap.add_argument("--no-superlattice", action="store_true",
                    help="skip the moire superlattice Fock builds")
    ap.add_argument("--no-symmetry", action="store_true",
                    help="skip the SCF solves with a symmetry projector")
    ap.add_argument("--full-channels", action="store_true",
                    help="time the unpacked nb^2-channel exchange instead")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a device copy of 64 MiB x 10 (scales the PMC byte counters)")
    ap.add_argument("--kernel-sample", action="store_true",
                    help="only run the solver kernels a few times each for a profiler: eigh "
                         "and density_from_fock at 96 x 96 x 8 and 192 x 192 x 16")
    ap.add_argument("--loop-sample", action="store_true",
                    help="only run one DM solve of a 96 x 96 k, 12-band sweep kernel three "
                         "times after one warm-up solve, for a profiler of the per-k loop kernels")
    ap.add_argument("--sweep-sample", choices=("sequential", "grouped"), default=None,
                    help="only run the 48-kernel sweep once after one warm-up pass, one solve "
                         "after another or as band-count groups one after another (no "
                         "concurrent host threads), for a profiler")
    args = ap.parse_args()

    if args.sweep_sample:
        from cpp_hf_amd import SolverConfig, solve_many
        from cpp_hf_amd.workloads import sweep_params, sweep_workload

        wls = [sweep_workload(p) for p in sweep_params(48, seed=1)]
        kw = dict(config=SolverConfig(), concurrency=1, batch=args.sweep_sample == "grouped",
                  max_batch=32)
        ks, P0s, nes = [w.kernel for w in wls], [w.P0 for w in wls], [w.n_electrons for w in wls]
        solve_many(ks, P0s, nes, **kw)
        t = time.perf_counter()
        rs = solve_many(ks, P0s, nes, **kw)
        dt = time.perf_counter() - t
        print(json.dumps({"sweep_sample": args.sweep_sample, "seconds": dt,
                          "total_iterations": int(sum(int(r.n_iter) for r in rs))}), flush=True)
        return 0

    if args.loop_sample:
        from cpp_hf_amd import SolverConfig, solve
        from cpp_hf_amd.workloads import sweep_params, sweep_workload

        wl = sweep_workload(sweep_params(48, seed=1)[7])  # 96 x 96 k, 12 bands
        solve(wl.kernel, wl.P0, wl.n_electrons, config=SolverConfig())
        ts, r = [], None
        for _ in range(3):
            t = time.perf_counter()
            r = solve(wl.kernel, wl.P0, wl.n_electrons, config=SolverConfig())
            ts.append(time.perf_counter() - t)
        print(json.dumps({"loop_sample": "DM solve, 96x96 k, 12 bands (sweep kernel 7 of the "
                                         "seed-1 set)", "seconds": ts, "n_iter": int(r.n_iter),
                          "energy": float(r.energy)}), flush=True)
        return 0

    if args.kernel_sample:
        from cpp_hf_amd._compat import _native

        rng = np.random.default_rng(0)
        for n, nb, n_e in ((96, 8, 4.1), (96, 12, 6.1), (120, 12, 6.1), (192, 16, 8.1)):
            X = rng.standard_normal((n, n, nb, nb)) + 1j * rng.standard_normal((n, n, nb, nb))
            F = 0.5 * (X + np.conj(np.swapaxes(X, -1, -2)))
            w = np.full((n, n), 1.0 / (n * n))
            for _ in range(5):
                _native.eigh_batched(F)
                _native.density_from_fock(F, w, n_e, 0.05, 0)
        print(json.dumps({"kernel_sample": "eigh and density_from_fock at 96x96x8, "
                                           "96x96x12, 120x120x12 and 192x192x16, "
                                           "5 calls each"}), flush=True)
        return 0

    if args.copy_calibration:
        import torch

        torch.cuda.set_device(0)
        src = torch.ones(8 << 20, dtype=torch.float64, device="cuda")  # 64 MiB
        dst = torch.empty_like(src)
        for _ in range(10):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_per_launch": src.numel() * 8}), flush=True)
        return 0

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1",
                          "n_gpus": args.gpus}), flush=True)
        return 2

    import torch


    from cpp_hf_amd._compat import _native, native_required
    from cpp_hf_amd.solver import _kernel_args_for_native
    from cpp_hf_amd.workloads import layered_nb8

    native_required()
    torch.cuda.set_device(0)
    dev = torch.device("cuda", 0)

    wl = layered_nb8(args.n)
    kern = wl.kernel
    ka = _kernel_args_for_native(kern)
    nk1, nk2, nb = kern.h.shape[0], kern.h.shape[1], kern.h.shape[-1]
    nk = nk1 * nk2
    A = 16 * nk * nb * nb
    alg_bytes = 6 * A  # P, refP, h, VR in; Sigma, F out

    P_host = 0.5 * (wl.P0 + np.conj(np.swapaxes(wl.P0, -1, -2)))
    gk = _native.GpuKernel(ka)
    P = torch.from_numpy(np.ascontiguousarray(P_host)).to(dev)
    Sigma = torch.empty_like(P)
    F = torch.empty_like(P)
    hdiag = torch.empty(nb, dtype=torch.float64, device=dev)
    # an explicit stream (default priority): the HIP events below are recorded
    # on the same stream the Fock kernels run on
    stream = torch.cuda.Stream(device=dev)
    herm = not args.full_channels

    def step():
        gk.build_fock_device(P.data_ptr(), Sigma.data_ptr(), F.data_ptr(), hdiag.data_ptr(),
                             herm, stream.cuda_stream)

    torch.cuda.synchronize()
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()

    def event_us(n):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(n):
            step()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) * 1e3 / n

    # the launch-shape autotuner runs at set-up, outside the timed region, so
    # the timed builds use the tuned shape (results do not depend on it)
    default_us = event_us(200)
    tuned = gk.autotune(herm)
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()

    # parity of this exact build against the oracle (outside the timed region)
    parity = None
    try:
        from oracle import hfcpu

        Sr, Hr, Fr, Er = hfcpu.build_fock_apply(ka, P_host)
        Fg = F.cpu().numpy()
        parity = float(np.max(np.abs(Fg - Fr)) / np.max(np.abs(Fr)))
    except Exception as exc:  # the oracle is optional on a user's machine
        parity = f"unavailable: {exc!r}"

    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        step()
    ev1.record(stream)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    wall = t1 - t0
    ev_ms = ev0.elapsed_time(ev1)
    per_build_s = ev_ms / 1e3 / args.steps
    builds_per_s = args.steps / wall
    achieved = alg_bytes / per_build_s / 1e9

    out = {
        "metric": "Fock builds/s (nb=8, 96x96 k, fp64)",
        "value": builds_per_s,
        "unit": "Fock builds/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"layered QWZ fock build, {nk1}x{nk2} k, nb={nb}, layer-resolved V, "
                        "Hartree+refP, Hermitian P",
            "nk": [nk1, nk2], "nb": nb,
            "channels": "packed nb(nb+1)/2" if herm and gk.packable else "full nb^2",
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "fock_build (k_rows_fwd_fast + k_cols_fast + k_rows_inv_fast)",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(nk1, nk2, nb) if herm else None,
            "algorithmic_bytes_per_build": alg_bytes,
            "event_us_per_build": per_build_s * 1e6,
        },
        "parity_rel_err_F": parity,
    }

    out["autotune"] = {"default_us_per_build": default_us, **tuned}
    # the same passes' bytes moved by three plain streaming launches (no
    # transform): what the three-pass dataflow costs by itself on this box
    try:
        floor = gk.dataflow_floor(herm, 200)
        out["roofline"]["dataflow_floor"] = {
            "us_per_build": floor["us_per_build"],
            "bytes_per_build": floor["bytes_per_build"],
            "build_over_floor": per_build_s * 1e6 / floor["us_per_build"],
            "frac_at_floor": (alg_bytes / (floor["us_per_build"] * 1e-6)) / 1e9 / HBM_PEAK_GBS,
        }
    except ValueError:  # a mesh without the two-step transform
        out["roofline"]["dataflow_floor"] = None
    # matrices per wavefront of the solvers' eigh at this kernel's nk x nb
    # (results do not depend on it)
    out["autotune"]["eigh"] = {"nmat": nk, "nb": nb, **_native.eigh_autotune(nk, nb)}

    if not args.no_solves:
        out["solves"] = run_solves(not args.no_cpu_baseline)

    if not args.no_sweep:
        out["sweep"] = run_sweep(args.sweep_concurrency)

    if not args.no_mixed:
        out["mixed_precision"] = run_mixed_precision(torch, dev)
        out["adjoint"] = run_adjoint(torch, dev)

    if not args.no_superlattice:
        out["superlattice"] = run_superlattice(torch, dev, not args.no_cpu_baseline)

    if not args.no_symmetry:
        out["symmetry"] = run_symmetry()

    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(ka, P_host, args.cpu_repeat_builds)
        out["gpu_over_cpu"] = out["value"] / out["cpu_baseline"]["value"]
        if not args.no_sweep:
            out["cpu_baseline"]["sweep_48_sequential"] = cpu_sweep_sequential()

    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
