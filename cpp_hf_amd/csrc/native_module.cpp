// pybind11 module `cpp_hf_amd._native`: the reference module's function
// names and signatures (cpp/cpp_hf_native.cpp:242-841), bound to the C ABI of
// libhfgpu (include/hfgpu.h).  Numpy arrays cross as raw pointers; outputs
// are allocated here as numpy arrays and filled by the GPU library; the GIL
// is released around every GPU call and re-acquired only by project_fn.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <complex>
#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "hfgpu.h"

namespace py = pybind11;
using c128 = std::complex<double>;
using CArr = py::array_t<c128, py::array::c_style | py::array::forcecast>;
using RArr = py::array_t<double, py::array::c_style | py::array::forcecast>;
using IArr = py::array_t<std::int64_t, py::array::c_style | py::array::forcecast>;

namespace {

void check(int rc) {
    if (rc == HF_OK) return;
    const std::string msg = hfgpu_last_error();
    if (rc == HF_EINVAL) throw py::value_error(msg);
    throw std::runtime_error(msg);
}

double* dp(py::array& a) { return static_cast<double*>(a.mutable_data()); }
const double* cdp(const py::array& a) { return static_cast<const double*>(a.data()); }

std::vector<py::ssize_t> lead_shape(const py::array& a, int drop) {
    std::vector<py::ssize_t> s;
    for (int i = 0; i < a.ndim() - drop; ++i) s.push_back(a.shape(i));
    return s;
}

// kernel_args dict -> hf_kernel_desc with the arrays kept alive.
struct KernelArgs {
    CArr h, VR, refP, Oi, Oj;
    RArr w2d, HH, cg;
    hf_kernel_desc d{};
    // superlattice entries (solver.py:104-134), read when a superlattice
    // flag is set (inject_superlattice_into_kernel, cpp_hf_native.cpp:105-165)
    CArr sl_V, sl_Vo;
    IArr sl_off, sl_pi, sl_pj, sl_ps, sl_p2d;
    RArr sl_HH, sl_HHo;
    hf_superlattice_desc sl{};

    void read_superlattice(py::object ka) {
        auto flag = [&](const char* key) {
            return ka.contains(key) ? py::cast<bool>(ka[key]) : false;
        };
        const bool fock = flag("superlattice_fock_active");
        const bool hart = flag("superlattice_hartree_active");
        if (!fock && !hart) return;
        sl.fock_active = fock;
        sl.hartree_active = hart;
        sl.hartree_degeneracy =
            ka.contains("hartree_degeneracy") ? py::cast<double>(ka["hartree_degeneracy"]) : 1.0;
        sl.n_G = py::cast<size_t>(ka["n_G"]);
        sl.dim_orb = py::cast<size_t>(ka["dim_orb"]);
        sl.n_delta = py::cast<size_t>(ka["n_delta"]);
        sl.N_ext_x = py::cast<size_t>(ka["N_ext_x"]);
        sl.N_ext_y = py::cast<size_t>(ka["N_ext_y"]);
        if (sl.n_G * sl.dim_orb != d.nb)
            throw py::value_error("superlattice mode requires n_G × dim_orb == nb");
        sl_V = py::cast<CArr>(ka["V_lag_fft"]);
        sl_off = py::cast<IArr>(ka["g_a_off"]);
        sl_pi = py::cast<IArr>(ka["pair_i"]);
        sl_pj = py::cast<IArr>(ka["pair_j"]);
        sl_ps = py::cast<IArr>(ka["pair_start"]);
        sl_p2d = py::cast<IArr>(ka["pair_to_delta"]);
        sl_HH = py::cast<RArr>(ka["HH_GG"]);
        if (sl_V.ndim() != 2 || (size_t)sl_V.shape(0) != sl.N_ext_x ||
            (size_t)sl_V.shape(1) != sl.N_ext_y)
            throw py::value_error("V_lag_fft must have shape (N_ext_x, N_ext_y)");
        if (sl_off.ndim() != 2 || (size_t)sl_off.shape(0) != sl.n_G || sl_off.shape(1) != 2)
            throw py::value_error("g_a_off must be (n_G, 2)");
        if (sl_ps.ndim() != 1 || (size_t)sl_ps.shape(0) != sl.n_delta + 1)
            throw py::value_error("pair_start must have length n_delta + 1");
        if (sl_pi.ndim() != 1 || sl_pj.ndim() != 1 || sl_pi.shape(0) != sl_pj.shape(0))
            throw py::value_error("pair_i and pair_j must have the same length");
        if (sl_p2d.ndim() != 2 || (size_t)sl_p2d.shape(0) != sl.n_G ||
            (size_t)sl_p2d.shape(1) != sl.n_G)
            throw py::value_error("pair_to_delta must be (n_G, n_G)");
        if (sl_HH.ndim() != 2 || (size_t)sl_HH.shape(0) != sl.n_G || (size_t)sl_HH.shape(1) != sl.n_G)
            throw py::value_error("HH_GG must be (n_G, n_G)");
        sl.n_pairs = sl_pi.shape(0);
        sl.V_lag_fft = cdp(sl_V);
        sl.g_a_off = sl_off.data();
        sl.pair_i = sl_pi.data();
        sl.pair_j = sl_pj.data();
        sl.pair_start = sl_ps.data();
        sl.pair_to_delta = sl_p2d.data();
        sl.HH_GG = cdp(sl_HH);
        if (ka.contains("HH_GG_orbital")) {
            sl_HHo = py::cast<RArr>(ka["HH_GG_orbital"]);
            if (sl_HHo.ndim() != 4 || (size_t)sl_HHo.shape(0) != sl.n_G ||
                (size_t)sl_HHo.shape(1) != sl.n_G || (size_t)sl_HHo.shape(2) != sl.dim_orb ||
                (size_t)sl_HHo.shape(3) != sl.dim_orb)
                throw py::value_error("HH_GG_orbital must be (n_G, n_G, dim_orb, dim_orb)");
            sl.HH_GG_orbital = cdp(sl_HHo);
        }
        if (ka.contains("V_lag_fft_orbital")) {
            sl_Vo = py::cast<CArr>(ka["V_lag_fft_orbital"]);
            if (sl_Vo.ndim() != 4 || (size_t)sl_Vo.shape(0) != sl.N_ext_x ||
                (size_t)sl_Vo.shape(1) != sl.N_ext_y || (size_t)sl_Vo.shape(2) != sl.dim_orb ||
                (size_t)sl_Vo.shape(3) != sl.dim_orb)
                throw py::value_error(
                    "V_lag_fft_orbital must be (N_ext_x, N_ext_y, dim_orb, dim_orb)");
            sl.V_lag_fft_orbital = cdp(sl_Vo);
        }
        d.superlattice = &sl;
    }

    explicit KernelArgs(py::object ka) {
        h = py::cast<CArr>(ka["h"]);
        VR = py::cast<CArr>(ka["VR"]);
        refP = py::cast<CArr>(ka["refP"]);
        w2d = py::cast<RArr>(ka["w2d"]);
        HH = py::cast<RArr>(ka["HH"]);
        cg = py::cast<RArr>(ka["contact_g"]);
        Oi = py::cast<CArr>(ka["contact_Oi"]);
        Oj = py::cast<CArr>(ka["contact_Oj"]);
        // shape checks of make_kernel (cpp_hf_native.cpp:58-87)
        if (h.ndim() != 4) throw py::value_error("h must be (nk1,nk2,nb,nb)");
        if (h.shape(3) != h.shape(2))
            throw py::value_error("h last two axes must be square (nb,nb)");
        if (VR.ndim() != 4) throw py::value_error("VR must be (nk1,nk2,dv1,dv2)");
        const bool has_refP = py::cast<bool>(ka["has_refP"]);
        if (has_refP && (refP.ndim() != 4 || refP.shape(0) != h.shape(0) ||
                         refP.shape(1) != h.shape(1) || refP.shape(2) != h.shape(2) ||
                         refP.shape(3) != h.shape(3)))
            throw py::value_error("refP must match h shape when has_refP is true");
        if (cg.ndim() != 1) throw py::value_error("contact_g must be 1-D");
        if (VR.shape(0) != h.shape(0) || VR.shape(1) != h.shape(1))
            throw py::value_error("VR k-mesh must match h");
        if (w2d.size() != h.shape(0) * h.shape(1))
            throw py::value_error("w2d must be (nk1,nk2)");
        d.nk1 = h.shape(0);
        d.nk2 = h.shape(1);
        d.nb = h.shape(2);
        d.dv1 = VR.shape(2);
        d.dv2 = VR.shape(3);
        d.n_contact = cg.shape(0);
        d.include_hartree = py::cast<bool>(ka["include_hartree"]);
        d.include_exchange = py::cast<bool>(ka["include_exchange"]);
        d.exchange_hcp = py::cast<bool>(ka["exchange_hcp"]);
        d.has_refP = has_refP;
        d.T = py::cast<double>(ka["T"]);
        d.weight_sum = py::cast<double>(ka["weight_sum"]);
        d.h = cdp(h);
        d.VR = cdp(VR);
        d.refP = has_refP ? cdp(refP) : nullptr;
        d.w2d = cdp(w2d);
        d.HH = cdp(HH);
        d.contact_g = cdp(cg);
        d.contact_Oi = cdp(Oi);
        d.contact_Oj = cdp(Oj);
        read_superlattice(ka);
    }
    KernelArgs(const KernelArgs&) = delete;
    KernelArgs& operator=(const KernelArgs&) = delete;
};

// Device symmetry projector from a project_fn that carries one
// (cpp_hf_amd.symmetry.SymmetryProjector): its _hfgpu_symmetry() gives
// (unitary | None, spatial | None, spatial_mod, spatial_axes, tr_U | None,
// tr_mod, tr_axes), or None when the averaging has no device form.
struct SymSpec {
    CArr U, S, T;
    hf_symmetry_desc d{};
    bool active = false;

    SymSpec(py::object project_fn, size_t nb) {
        if (project_fn.is_none() || !py::hasattr(project_fn, "_hfgpu_symmetry") || nb > 32)
            return;
        py::object spec = project_fn.attr("_hfgpu_symmetry")();
        if (spec.is_none()) return;
        py::tuple t = spec;
        d.nb = nb;
        auto group = [&](py::handle h, CArr& keep, size_t& n, const double*& ptr) {
            if (h.is_none()) return;
            keep = py::cast<CArr>(h);
            if (keep.ndim() != 3 || (size_t)keep.shape(1) != nb || (size_t)keep.shape(2) != nb)
                throw py::value_error("symmetry group matrices must be (n, nb, nb)");
            n = keep.shape(0);
            ptr = cdp(keep);
        };
        group(t[0], U, d.n_unitary, d.unitary);
        group(t[1], S, d.n_spatial, d.spatial);
        d.spatial_mod = py::cast<bool>(t[2]) ? 1 : 0;
        d.spatial_axes = py::cast<int>(t[3]);
        if (!t[4].is_none()) {
            T = py::cast<CArr>(t[4]);
            if (T.ndim() != 2 || (size_t)T.shape(0) != nb || (size_t)T.shape(1) != nb)
                throw py::value_error("time_reversal_U must be (nb, nb)");
            d.tr_U = cdp(T);
        }
        d.tr_mod = py::cast<bool>(t[5]) ? 1 : 0;
        d.tr_axes = py::cast<int>(t[6]);
        active = d.n_unitary > 0 || d.n_spatial > 0 || d.tr_U != nullptr;
    }
};

// RAII kernel handle.
struct Handle {
    hfgpu_kernel* k = nullptr;
    explicit Handle(const hf_kernel_desc& d) {
        int rc;
        {
            py::gil_scoped_release rel;
            rc = hfgpu_kernel_create(&d, &k);
        }
        check(rc);
    }
    ~Handle() {
        if (k) hfgpu_kernel_destroy(k);
    }
    Handle(const Handle&) = delete;
    Handle& operator=(const Handle&) = delete;
};

// Python-owned device-resident kernel handle.
struct PyKernel {
    hfgpu_kernel* k = nullptr;
    PyKernel() = default;
    PyKernel(const PyKernel&) = delete;
    PyKernel& operator=(const PyKernel&) = delete;
    ~PyKernel() {
        if (k) hfgpu_kernel_destroy(k);
    }
};

// project_fn trampoline (cpp_hf_native.cpp:215-238): wraps the live buffer
// without copying, calls the Python function, checks the shape, copies back.
struct ProjectCtx {
    py::object fn;
    std::string error;
    bool python_error = false;
    py::object exc_type, exc_value, exc_tb;
};

int project_trampoline(void* user, double* M, size_t nk1, size_t nk2, size_t nb) {
    auto* ctx = static_cast<ProjectCtx*>(user);
    py::gil_scoped_acquire gil;
    try {
        std::vector<py::ssize_t> shape{(py::ssize_t)nk1, (py::ssize_t)nk2, (py::ssize_t)nb,
                                       (py::ssize_t)nb};
        py::array_t<c128> view(shape, reinterpret_cast<c128*>(M), py::none());
        py::object result = ctx->fn(view);
        CArr out = py::cast<CArr>(result);
        if (out.ndim() != 4 || (size_t)out.shape(0) != nk1 || (size_t)out.shape(1) != nk2 ||
            (size_t)out.shape(2) != nb || (size_t)out.shape(3) != nb) {
            ctx->error = "project_fn returned an array of wrong shape";
            return 1;
        }
        std::memmove(M, out.data(), nk1 * nk2 * nb * nb * sizeof(c128));
        return 0;
    } catch (py::error_already_set& e) {
        ctx->python_error = true;
        ctx->error = e.what();
        return 1;
    } catch (const std::exception& e) {
        ctx->error = e.what();
        return 1;
    }
}

void check_project(int rc, const ProjectCtx* ctx) {
    if (rc != HF_OK && ctx && !ctx->error.empty()) throw std::runtime_error(ctx->error);
    check(rc);
}

py::array_t<c128> new_c(const std::vector<py::ssize_t>& shape) { return py::array_t<c128>(shape); }
py::array_t<double> new_r(const std::vector<py::ssize_t>& shape) { return py::array_t<double>(shape); }

}  // namespace

PYBIND11_MODULE(_native, m) {
    m.doc() = "cpp_hf_amd native core: HIP kernels for MI355X (gfx950) behind a C ABI";

    m.def("device_count", []() { return hfgpu_device_count(); });

    // --- Self-energy / FFT (cpp_hf_native.cpp:246-284) ---
    m.def(
        "selfenergy_fft_full",
        [](CArr VR, CArr P, bool apply_ifftshift, bool hcp) {
            if (P.ndim() != 4) throw py::value_error("P must be (nk1,nk2,nb,nb)");
            if (VR.ndim() != 4) throw py::value_error("VR must be (nk1,nk2,dv1,dv2)");
            if (VR.shape(0) != P.shape(0) || VR.shape(1) != P.shape(1))
                throw py::value_error("VR k-mesh must match P");
            if (P.shape(2) != P.shape(3)) throw py::value_error("P last two dims must be square");
            const size_t nk1 = P.shape(0), nk2 = P.shape(1), nb = P.shape(2);
            auto sigma = new_c({(py::ssize_t)nk1, (py::ssize_t)nk2, (py::ssize_t)nb, (py::ssize_t)nb});
            py::array sg = sigma;
            int rc;
            {
                py::gil_scoped_release rel;
                rc = hfgpu_selfenergy_fft_full(cdp(VR), cdp(P), dp(sg), nk1, nk2, nb, VR.shape(2),
                                               VR.shape(3), apply_ifftshift, hcp);
            }
            check(rc);
            return sigma;
        },
        py::arg("VR"), py::arg("P"), py::arg("apply_ifftshift") = true,
        py::arg("hermitian_channel_packing") = false);

    // --- device symmetry projector (no counterpart in the reference) ---
    m.def(
        "project_symmetry",
        [](CArr M, py::object unitary_group, py::object spatial_group, bool spatial_mod,
           int spatial_axes, py::object time_reversal_U, bool tr_mod, int tr_axes) {
            if (M.ndim() != 4 || M.shape(2) != M.shape(3))
                throw py::value_error("M must be (nk1, nk2, nb, nb)");
            const size_t nb = M.shape(2);
            CArr U, S, T;
            hf_symmetry_desc d{};
            d.nb = nb;
            auto group = [&](py::object h, CArr& keep, size_t& n, const double*& ptr) {
                if (h.is_none()) return;
                keep = py::cast<CArr>(h);
                if (keep.ndim() != 3 || (size_t)keep.shape(1) != nb || (size_t)keep.shape(2) != nb)
                    throw py::value_error("symmetry group matrices must be (n, nb, nb)");
                n = keep.shape(0);
                ptr = cdp(keep);
            };
            group(unitary_group, U, d.n_unitary, d.unitary);
            group(spatial_group, S, d.n_spatial, d.spatial);
            d.spatial_mod = spatial_mod ? 1 : 0;
            d.spatial_axes = spatial_axes;
            if (!time_reversal_U.is_none()) {
                T = py::cast<CArr>(time_reversal_U);
                if (T.ndim() != 2 || (size_t)T.shape(0) != nb || (size_t)T.shape(1) != nb)
                    throw py::value_error("time_reversal_U must be (nb, nb)");
                d.tr_U = cdp(T);
            }
            d.tr_mod = tr_mod ? 1 : 0;
            d.tr_axes = tr_axes;
            auto out = new_c(lead_shape(M, 0));
            py::array oa = out;
            std::memcpy(dp(oa), cdp(M), (size_t)M.size() * sizeof(c128));
            int rc;
            {
                py::gil_scoped_release rel;
                rc = hfgpu_project_symmetry(&d, dp(oa), M.shape(0), M.shape(1), nb);
            }
            check(rc);
            return out;
        },
        py::arg("M"), py::arg("unitary_group") = py::none(), py::arg("spatial_group") = py::none(),
        py::arg("spatial_mod") = true, py::arg("spatial_axes") = 3,
        py::arg("time_reversal_U") = py::none(), py::arg("tr_mod") = true, py::arg("tr_axes") = 3);

    // --- superlattice exchange (cpp_hf_native.cpp:718-790) ---
    m.def(
        "selfenergy_superlattice_streamed",
        [](CArr rho, CArr VR_fft, IArr g_a_off, IArr pair_i, IArr pair_j, IArr pair_start,
           size_t nkx, size_t nky, size_t n_G, size_t dim_orb, size_t N_ext_x, size_t N_ext_y,
           py::object VR_fft_orbital, bool hermitian_rho) {
            if (rho.ndim() != 6)
                throw py::value_error("rho must be (nkx, nky, n_G, dim_orb, n_G, dim_orb)");
            if (VR_fft.ndim() != 2) throw py::value_error("VR_fft must be (N_ext_x, N_ext_y)");
            if ((size_t)VR_fft.shape(0) != N_ext_x || (size_t)VR_fft.shape(1) != N_ext_y)
                throw py::value_error("VR_fft shape does not match (N_ext_x, N_ext_y)");
            if (g_a_off.ndim() != 2 || (size_t)g_a_off.shape(0) != n_G || g_a_off.shape(1) != 2)
                throw py::value_error("g_a_off must be (n_G, 2)");
            if (pair_start.ndim() != 1 || pair_start.shape(0) < 1)
                throw py::value_error("pair_start must be 1-D, len >= 1");
            if (pair_i.ndim() != 1 || pair_j.ndim() != 1 || pair_i.shape(0) != pair_j.shape(0))
                throw py::value_error("pair_i and pair_j must be 1-D with the same length");
            const py::ssize_t want[6] = {(py::ssize_t)nkx, (py::ssize_t)nky, (py::ssize_t)n_G,
                                         (py::ssize_t)dim_orb, (py::ssize_t)n_G,
                                         (py::ssize_t)dim_orb};
            for (int a = 0; a < 6; ++a)
                if (rho.shape(a) != want[a])
                    throw py::value_error(
                        "rho shape does not match (nkx, nky, n_G, dim_orb, n_G, dim_orb)");
            CArr Vo;
            const double* vo = nullptr;
            if (!VR_fft_orbital.is_none()) {
                Vo = py::cast<CArr>(VR_fft_orbital);
                if (Vo.ndim() != 4 || (size_t)Vo.shape(0) != N_ext_x ||
                    (size_t)Vo.shape(1) != N_ext_y || (size_t)Vo.shape(2) != dim_orb ||
                    (size_t)Vo.shape(3) != dim_orb)
                    throw py::value_error("VR_fft_orbital must be (N_ext_x, N_ext_y, dim_orb, dim_orb)");
                vo = cdp(Vo);
            }
            auto sigma = new_c(std::vector<py::ssize_t>(want, want + 6));
            py::array sg = sigma;
            const size_t n_delta = pair_start.shape(0) - 1;
            int rc;
            {
                py::gil_scoped_release rel;
                rc = hfgpu_selfenergy_superlattice_streamed(
                    cdp(rho), dp(sg), cdp(VR_fft), vo, g_a_off.data(), pair_i.data(),
                    pair_j.data(), pair_start.data(), pair_i.shape(0), n_delta, N_ext_x, N_ext_y,
                    nkx, nky, n_G, dim_orb, hermitian_rho ? 1 : 0);
            }
            check(rc);
            return sigma;
        },
        py::arg("rho"), py::arg("VR_fft"), py::arg("g_a_off"), py::arg("pair_i"),
        py::arg("pair_j"), py::arg("pair_start"), py::arg("nkx"), py::arg("nky"), py::arg("n_G"),
        py::arg("dim_orb"), py::arg("N_ext_x"), py::arg("N_ext_y"),
        py::arg("VR_fft_orbital") = py::none(), py::arg("hermitian_rho") = false);

    m.def("clear_fft_plan_cache", []() {
        py::gil_scoped_release rel;
        hfgpu_clear_plan_cache();
    });

    // --- Hermitian eigh (cpp_hf_native.cpp:287-312) ---
    m.def(
        "eigh_batched",
        [](CArr M) {
            if (M.ndim() < 2) throw py::value_error("M needs at least 2 dims");
            const size_t nb = M.shape(M.ndim() - 1);
            if ((size_t)M.shape(M.ndim() - 2) != nb)
                throw py::value_error("M last two dims must be square");
            size_t nmat = 1;
            for (int i = 0; i < M.ndim() - 2; ++i) nmat *= M.shape(i);
            auto ws = lead_shape(M, 1);
            auto w = new_r(ws);
            auto V = new_c(lead_shape(M, 0));
            py::array wa = w, va = V;
            int rc;
            {
                py::gil_scoped_release rel;
                rc = hfgpu_eigh_batched(cdp(M), dp(wa), dp(va), nmat, nb);
            }
            check(rc);
            return py::make_tuple(w, V);
        },
        py::arg("M"));

    // matrices-per-wavefront autotuner of the batched eigh (an extension)
    m.def(
        "eigh_autotune",
        [](size_t nmat, size_t nb) {
            int mpw = 0;
            double us = 0.0;
            int rc;
            {
                py::gil_scoped_release rel;
                rc = hfgpu_eigh_autotune(nmat, nb, &mpw, &us);
            }
            check(rc);
            py::dict d;
            d["matrices_per_wavefront"] = mpw;
            d["us_per_call"] = us;
            return d;
        },
        py::arg("nmat"), py::arg("nb"));

    // --- block-diagonal eigh (cpp_hf_native.cpp:314-358) ---
    m.def(
        "eigh_block_sizes",
        [](CArr M, std::vector<size_t> sizes, bool sort) {
            if (M.ndim() < 2) throw py::value_error("M needs at least 2 dims");
            const size_t nb = M.shape(M.ndim() - 1);
            size_t nmat = 1;
            for (int i = 0; i < M.ndim() - 2; ++i) nmat *= M.shape(i);
            auto w = new_r(lead_shape(M, 1));
            auto V = new_c(lead_shape(M, 0));
            py::array wa = w, va = V;
            int rc;
            {
                py::gil_scoped_release rel;
                rc = hfgpu_eigh_block_sizes(cdp(M), dp(wa), dp(va), nmat, nb, sizes.data(),
                                            sizes.size(), sort ? 1 : 0);
            }
            check(rc);
            return py::make_tuple(w, V);
        },
        py::arg("M"), py::arg("sizes"), py::arg("sort") = true);
    m.def("max_offblock_sizes", [](CArr M, std::vector<size_t> sizes) {
        if (M.ndim() < 2) throw py::value_error("M needs at least 2 dims");
        const size_t nb = M.shape(M.ndim() - 1);
        size_t nmat = 1;
        for (int i = 0; i < M.ndim() - 2; ++i) nmat *= M.shape(i);
        double mx = 0.0;
        int rc;
        {
            py::gil_scoped_release rel;
            rc = hfgpu_max_offblock_sizes(cdp(M), nmat, nb, sizes.data(), sizes.size(), &mx);
        }
        check(rc);
        return mx;
    });

    // --- mu solvers (cpp_hf_native.cpp:361-377) ---
    m.def(
        "find_mu_bisection",
        [](RArr bands, RArr weights, double n_e, double T, int maxiter) {
            if (bands.ndim() < 1) throw py::value_error("bands must be at least 1-D");
            size_t nk = 1;
            for (int i = 0; i < bands.ndim() - 1; ++i) nk *= bands.shape(i);
            const size_t nb = bands.shape(bands.ndim() - 1);
            if ((size_t)weights.size() < nk) throw py::value_error("weights too short for bands");
            double mu = 0.0;
            int rc;
            {
                py::gil_scoped_release rel;
                rc = hfgpu_find_mu_bisection(cdp(bands), cdp(weights), nk, nb, n_e, T, maxiter, &mu);
            }
            check(rc);
            return mu;
        },
        py::arg("bands"), py::arg("weights"), py::arg("n_e"), py::arg("T"), py::arg("maxiter") = 0);

    m.def(
        "find_mu_newton",
        [](RArr bands, RArr weights, double n_e, double T, int maxiter) {
            if (bands.ndim() < 1) throw py::value_error("bands must be at least 1-D");
            size_t nk = 1;
            for (int i = 0; i < bands.ndim() - 1; ++i) nk *= bands.shape(i);
            const size_t nb = bands.shape(bands.ndim() - 1);
            if ((size_t)weights.size() < nk) throw py::value_error("weights too short for bands");
            double mu = 0.0;
            int rc;
            {
                py::gil_scoped_release rel;
                rc = hfgpu_find_mu_newton(cdp(bands), cdp(weights), nk, nb, n_e, T, maxiter, &mu);
            }
            check(rc);
            return mu;
        },
        py::arg("bands"), py::arg("weights"), py::arg("n_e"), py::arg("T"),
        py::arg("maxiter") = 15);

    m.def(
        "solve_mu_inloop",
        [](RArr eps, RArr w_norm, double n_target_norm, double mu0, double T, int maxiter) {
            size_t nk = 1;
            for (int i = 0; i < eps.ndim() - 1; ++i) nk *= eps.shape(i);
            const size_t nb = eps.shape(eps.ndim() - 1);
            double mu = 0.0;
            int rc;
            {
                py::gil_scoped_release rel;
                rc = hfgpu_solve_mu_inloop(cdp(eps), cdp(w_norm), nk, nb, n_target_norm, mu0, T,
                                           maxiter, &mu);
            }
            check(rc);
            return mu;
        },
        py::arg("eps"), py::arg("w_norm"), py::arg("n_target_norm"), py::arg("mu0"),
        py::arg("T"), py::arg("maxiter") = 25);

    // --- build_fock (cpp_hf_native.cpp:380-415) ---
    m.def("build_fock_apply", [](py::object kernel_args, CArr P, py::object project_fn,
                                 bool exchange_fp32) {
        KernelArgs ka(kernel_args);
        if (P.ndim() != 4 || (size_t)P.shape(0) != ka.d.nk1 || (size_t)P.shape(1) != ka.d.nk2 ||
            (size_t)P.shape(2) != ka.d.nb || (size_t)P.shape(3) != ka.d.nb)
            throw py::value_error("P must match h shape");
        Handle H(ka.d);
        std::vector<py::ssize_t> shape{(py::ssize_t)ka.d.nk1, (py::ssize_t)ka.d.nk2,
                                       (py::ssize_t)ka.d.nb, (py::ssize_t)ka.d.nb};
        auto S = new_c(shape), Hh = new_c(shape), F = new_c(shape);
        py::array sa = S, ha = Hh, fa = F;
        ProjectCtx ctx;
        ctx.fn = project_fn;
        const bool has_proj = !project_fn.is_none();
        double E = 0.0;
        int rc;
        {
            py::gil_scoped_release rel;
            rc = (exchange_fp32 ? hfgpu_build_fock_mixed : hfgpu_build_fock_apply)(
                H.k, cdp(P), dp(sa), dp(ha), dp(fa), &E, has_proj ? project_trampoline : nullptr,
                &ctx);
        }
        check_project(rc, &ctx);
        return py::make_tuple(S, Hh, F, E);
    },
          // exchange_fp32: the mixed-precision build (no reference counterpart)
          py::arg("kernel_args"), py::arg("P"), py::arg("project_fn"),
          py::arg("exchange_fp32") = false);

    // --- adjoint of build_fock (no reference counterpart): (gP, gVq, gh, gHH) ---
    m.def("build_fock_vjp", [](py::object kernel_args, CArr P, py::object cot_sigma,
                               py::object cot_hartree, py::object cot_F) {
        KernelArgs ka(kernel_args);
        if (P.ndim() != 4 || (size_t)P.shape(0) != ka.d.nk1 || (size_t)P.shape(1) != ka.d.nk2 ||
            (size_t)P.shape(2) != ka.d.nb || (size_t)P.shape(3) != ka.d.nb)
            throw py::value_error("P must match h shape");
        std::vector<py::ssize_t> shape{(py::ssize_t)ka.d.nk1, (py::ssize_t)ka.d.nk2,
                                       (py::ssize_t)ka.d.nb, (py::ssize_t)ka.d.nb};
        // cotangents: None -> NULL (zero); arrays kept alive for the call
        std::vector<CArr> keep;
        auto cot = [&](py::object o, const char* name) -> const double* {
            if (o.is_none()) return nullptr;
            CArr a = py::cast<CArr>(o);
            if (a.ndim() != 4 || a.shape(0) != shape[0] || a.shape(1) != shape[1] ||
                a.shape(2) != shape[2] || a.shape(3) != shape[3])
                throw py::value_error(std::string(name) + " must match h shape");
            keep.push_back(a);
            return cdp(keep.back());
        };
        keep.reserve(3);
        const double* Gs = cot(cot_sigma, "cot_sigma");
        const double* Gh = cot(cot_hartree, "cot_hartree");
        const double* GF = cot(cot_F, "cot_F");
        Handle H(ka.d);
        std::vector<py::ssize_t> vshape{(py::ssize_t)ka.d.nk1, (py::ssize_t)ka.d.nk2,
                                        (py::ssize_t)ka.d.dv1, (py::ssize_t)ka.d.dv2};
        auto gP = new_c(shape), gh = new_c(shape), gV = new_c(vshape);
        auto gHH = new_r({(py::ssize_t)ka.d.nb, (py::ssize_t)ka.d.nb});
        py::array pa = gP, ha = gh, va = gV, ra = gHH;
        int rc;
        {
            py::gil_scoped_release rel;
            rc = hfgpu_build_fock_vjp(H.k, cdp(P), Gs, Gh, GF, dp(pa), dp(va), dp(ha), dp(ra));
        }
        check(rc);
        return py::make_tuple(gP, gV, gh, gHH);
    }, py::arg("kernel_args"), py::arg("P"), py::arg("cot_sigma") = py::none(),
       py::arg("cot_hartree") = py::none(), py::arg("cot_F") = py::none());

    // --- SCF solver (cpp_hf_native.cpp:418-485) ---
    m.def("solve_scf", [](py::object kernel_args, CArr P0, double n_e, size_t max_iter,
                          double density_tol, double comm_tol, double mixing, double level_shift,
                          py::object project_fn, std::vector<size_t> block_sizes,
                          std::string acceleration, size_t diis_size, size_t diis_start,
                          double diis_damping, double trust_radius, bool return_density,
                          bool return_fock, bool exchange_fp32, double fp32_switch) {
        KernelArgs ka(kernel_args);
        if (P0.ndim() != 4 || (size_t)P0.shape(0) != ka.d.nk1 ||
            (size_t)P0.shape(1) != ka.d.nk2 || (size_t)P0.shape(2) != ka.d.nb ||
            (size_t)P0.shape(3) != ka.d.nb)
            throw py::value_error("P0 must match h shape");
        int accel;
        if (acceleration == "linear") accel = 0;
        else if (acceleration == "diis") accel = 1;
        else if (acceleration == "oda") accel = 2;
        else throw py::value_error("SCFConfig.acceleration must be 'linear', 'diis', or 'oda'");
        hf_scf_config cfg{};
        cfg.max_iter = max_iter;
        cfg.density_tol = density_tol;
        cfg.comm_tol = comm_tol;
        cfg.mixing = mixing;
        cfg.level_shift = level_shift;
        cfg.acceleration = accel;
        cfg.diis_size = diis_size;
        cfg.diis_start = diis_start;
        cfg.diis_damping = diis_damping;
        cfg.trust_radius = trust_radius;
        cfg.block_sizes = block_sizes.empty() ? nullptr : block_sizes.data();
        cfg.n_blocks = block_sizes.size();
        cfg.exchange_fp32 = exchange_fp32 ? 1 : 0;
        cfg.fp32_switch = fp32_switch;
        Handle H(ka.d);
        std::vector<py::ssize_t> shape{(py::ssize_t)ka.d.nk1, (py::ssize_t)ka.d.nk2,
                                       (py::ssize_t)ka.d.nb, (py::ssize_t)ka.d.nb};
        auto dens = new_c(shape), fock = new_c(shape);
        auto hE = new_r({(py::ssize_t)max_iter}), hD = new_r({(py::ssize_t)max_iter}),
             hC = new_r({(py::ssize_t)max_iter});
        py::array da = dens, fa = fock, ea = hE, ra = hD, ca = hC;
        hf_scf_result res{};
        res.density = dp(da);
        res.fock = dp(fa);
        res.hist_E = dp(ea);
        res.hist_density = dp(ra);
        res.hist_comm = dp(ca);
        SymSpec sym(project_fn, ka.d.nb);
        if (sym.active) cfg.symmetry = &sym.d;
        ProjectCtx ctx;
        ctx.fn = project_fn;
        const bool has_proj = !project_fn.is_none() && !sym.active;
        int rc;
        {
            py::gil_scoped_release rel;
            rc = hfgpu_solve_scf(H.k, cdp(P0), n_e, &cfg, has_proj ? project_trampoline : nullptr,
                                 &ctx, &res);
        }
        check_project(rc, &ctx);
        const py::ssize_t n = (py::ssize_t)res.iterations;
        py::object d_obj = py::none(), f_obj = py::none();
        if (return_density) d_obj = dens;
        if (return_fock) f_obj = fock;
        return py::make_tuple(d_obj, f_obj, res.energy, res.mu, res.iterations,
                              (bool)res.converged, hE[py::slice(0, n, 1)].attr("copy")(),
                              hD[py::slice(0, n, 1)].attr("copy")(),
                              hC[py::slice(0, n, 1)].attr("copy")());
    },
          // the reference's positional arguments, then the mixed-precision
          // switch (no counterpart in the reference; off by default)
          py::arg("kernel_args"), py::arg("P0"), py::arg("n_e"), py::arg("max_iter"),
          py::arg("density_tol"), py::arg("comm_tol"), py::arg("mixing"), py::arg("level_shift"),
          py::arg("project_fn"), py::arg("block_sizes"), py::arg("acceleration"),
          py::arg("diis_size"), py::arg("diis_start"), py::arg("diis_damping"),
          py::arg("trust_radius"), py::arg("return_density"), py::arg("return_fock"),
          py::arg("exchange_fp32") = false, py::arg("fp32_switch") = 1e-6);

    // --- Several direct-minimisation solves of one shape at once (no
    // counterpart in the reference): lists of kernel_args, P0 and n_e, one
    // configuration; a list of solve_dm's ten-entry tuples ---
    m.def("solve_dm_batch", [](py::list kernel_args, py::list P0s, std::vector<double> n_es,
                               size_t max_iter, double tol_E, double tol_grad, double max_step,
                               double bt_shrink, double denom_scale, size_t bt_max,
                               size_t cg_restart, size_t plateau_window, int mu_maxiter,
                               bool return_Q, bool return_density, bool return_fock) {
        const size_t n = kernel_args.size();
        if (n == 0 || P0s.size() != n || n_es.size() != n)
            throw py::value_error("kernel_args, P0s and n_es must have one equal, nonzero length");
        std::vector<std::unique_ptr<KernelArgs>> kas;
        std::vector<std::unique_ptr<Handle>> hs;
        std::vector<CArr> p0;
        for (size_t i = 0; i < n; ++i) {
            kas.push_back(std::make_unique<KernelArgs>(kernel_args[i]));
            const auto& d = kas.back()->d;
            CArr a = py::cast<CArr>(P0s[i]);
            if (a.ndim() != 4 || (size_t)a.shape(0) != d.nk1 || (size_t)a.shape(1) != d.nk2 ||
                (size_t)a.shape(2) != d.nb || (size_t)a.shape(3) != d.nb)
                throw py::value_error("P0 must match h shape");
            p0.push_back(a);
        }
        for (size_t i = 0; i < n; ++i) hs.push_back(std::make_unique<Handle>(kas[i]->d));
        hf_dm_config cfg{};
        cfg.max_iter = max_iter;
        cfg.tol_E = tol_E;
        cfg.tol_grad = tol_grad;
        cfg.plateau_window = plateau_window;
        cfg.denom_scale = denom_scale;
        cfg.max_step = max_step;
        cfg.cg_restart = cg_restart;
        cfg.bt_shrink = bt_shrink;
        cfg.bt_max = bt_max;
        cfg.mu_maxiter = mu_maxiter;
        const py::ssize_t mh = (py::ssize_t)std::max<size_t>(max_iter, 1);
        std::vector<py::array_t<std::complex<double>>> Qs, Ds, Fs;
        std::vector<py::array_t<double>> ps, hEs, hGs;
        std::vector<hf_dm_result> res(n);
        std::vector<hfgpu_kernel*> ks(n);
        std::vector<const double*> P0p(n);
        for (size_t i = 0; i < n; ++i) {
            const auto& d = kas[i]->d;
            std::vector<py::ssize_t> shape{(py::ssize_t)d.nk1, (py::ssize_t)d.nk2,
                                           (py::ssize_t)d.nb, (py::ssize_t)d.nb};
            std::vector<py::ssize_t> pshape{(py::ssize_t)d.nk1, (py::ssize_t)d.nk2,
                                            (py::ssize_t)d.nb};
            Qs.push_back(new_c(shape));
            Ds.push_back(new_c(shape));
            Fs.push_back(new_c(shape));
            ps.push_back(new_r(pshape));
            hEs.push_back(new_r({mh}));
            hGs.push_back(new_r({mh}));
            py::array qa = Qs.back(), da = Ds.back(), fa = Fs.back(), pa = ps.back(),
                      ea = hEs.back(), ga = hGs.back();
            res[i] = hf_dm_result{};
            res[i].Q = return_Q ? dp(qa) : nullptr;
            res[i].p = dp(pa);
            res[i].density = return_density ? dp(da) : nullptr;
            res[i].fock = return_fock ? dp(fa) : nullptr;
            res[i].hist_E = dp(ea);
            res[i].hist_grad = dp(ga);
            ks[i] = hs[i]->k;
            P0p[i] = cdp(p0[i]);
        }
        int rc;
        {
            py::gil_scoped_release rel;
            rc = hfgpu_solve_dm_batch(ks.data(), n, P0p.data(), n_es.data(), &cfg, res.data());
        }
        check(rc);
        py::list out;
        for (size_t i = 0; i < n; ++i) {
            const py::ssize_t it = (py::ssize_t)res[i].n_iter;
            py::object q_obj = py::none(), d_obj = py::none(), f_obj = py::none();
            if (return_Q) q_obj = Qs[i];
            if (return_density) d_obj = Ds[i];
            if (return_fock) f_obj = Fs[i];
            out.append(py::make_tuple(q_obj, ps[i], d_obj, f_obj, res[i].mu, res[i].energy,
                                      res[i].n_iter, (bool)res[i].converged,
                                      hEs[i][py::slice(0, it, 1)].attr("copy")(),
                                      hGs[i][py::slice(0, it, 1)].attr("copy")()));
        }
        return out;
    },
          py::arg("kernel_args"), py::arg("P0s"), py::arg("n_es"), py::arg("max_iter"),
          py::arg("tol_E"), py::arg("tol_grad"), py::arg("max_step"), py::arg("bt_shrink"),
          py::arg("denom_scale"), py::arg("bt_max"), py::arg("cg_restart"),
          py::arg("plateau_window"), py::arg("mu_maxiter"), py::arg("return_Q"),
          py::arg("return_density"), py::arg("return_fock"));

    // --- Direct minimisation (cpp_hf_native.cpp:488-582): optimizer 1 is the
    // trust-region Newton solver (with deflation), any other code the CG loop ---
    m.def("solve_dm", [](py::object kernel_args, CArr P0, double n_e, size_t max_iter,
                         double tol_E, double tol_grad, double max_step, double bt_shrink,
                         double denom_scale, size_t bt_max, size_t cg_restart,
                         size_t plateau_window, int mu_maxiter, int optimizer, double tr_delta0,
                         size_t tr_cg_max, CArr deflation_targets, double deflation_sigma,
                         double deflation_length, std::vector<size_t> block_sizes,
                         py::object project_fn, bool return_Q, bool return_density,
                         bool return_fock, bool exchange_fp32, double fp32_switch,
                         bool return_stats) {
        KernelArgs ka(kernel_args);
        if (P0.ndim() != 4 || (size_t)P0.shape(0) != ka.d.nk1 ||
            (size_t)P0.shape(1) != ka.d.nk2 || (size_t)P0.shape(2) != ka.d.nb ||
            (size_t)P0.shape(3) != ka.d.nb)
            throw py::value_error("P0 must match h shape");
        hf_dm_config cfg{};
        cfg.max_iter = max_iter;
        cfg.tol_E = tol_E;
        cfg.tol_grad = tol_grad;
        cfg.plateau_window = plateau_window;
        cfg.denom_scale = denom_scale;
        cfg.max_step = max_step;
        cfg.cg_restart = cg_restart;
        cfg.bt_shrink = bt_shrink;
        cfg.bt_max = bt_max;
        cfg.mu_maxiter = mu_maxiter;
        cfg.block_sizes = block_sizes.empty() ? nullptr : block_sizes.data();
        cfg.n_blocks = block_sizes.size();
        cfg.optimizer = optimizer;
        cfg.tr_delta0 = tr_delta0;
        cfg.tr_cg_max = tr_cg_max;
        // (n_defl, nk1, nk2, nb, nb) or empty (cpp_hf_native.cpp:533-547)
        if (deflation_targets.ndim() == 5 && deflation_targets.shape(0) > 0) {
            cfg.n_deflation = (size_t)deflation_targets.shape(0);
            cfg.deflation_targets = cdp(deflation_targets);
            const size_t n_tot = ka.d.nk1 * ka.d.nk2 * ka.d.nb * ka.d.nb;
            if (deflation_sigma > 0.0 &&
                (size_t)deflation_targets.size() < cfg.n_deflation * n_tot)
                throw py::value_error(
                    "deflation_targets buffer smaller than n_deflation * nk * nb * nb.");
        }
        cfg.deflation_sigma = deflation_sigma;
        cfg.deflation_length = deflation_length;
        cfg.exchange_fp32 = exchange_fp32 ? 1 : 0;
        cfg.fp32_switch = fp32_switch;
        Handle H(ka.d);
        std::vector<py::ssize_t> shape{(py::ssize_t)ka.d.nk1, (py::ssize_t)ka.d.nk2,
                                       (py::ssize_t)ka.d.nb, (py::ssize_t)ka.d.nb};
        std::vector<py::ssize_t> pshape{(py::ssize_t)ka.d.nk1, (py::ssize_t)ka.d.nk2,
                                        (py::ssize_t)ka.d.nb};
        auto Q = new_c(shape), dens = new_c(shape), fock = new_c(shape);
        auto p = new_r(pshape);
        const py::ssize_t mh = (py::ssize_t)std::max<size_t>(max_iter, 1);
        auto hE = new_r({mh}), hG = new_r({mh});
        py::array qa = Q, da = dens, fa = fock, pa = p, ea = hE, ga = hG;
        hf_dm_result res{};
        res.Q = dp(qa);
        res.p = dp(pa);
        res.density = dp(da);
        res.fock = dp(fa);
        res.hist_E = dp(ea);
        res.hist_grad = dp(ga);
        SymSpec sym(project_fn, ka.d.nb);
        if (sym.active) cfg.symmetry = &sym.d;
        ProjectCtx ctx;
        ctx.fn = project_fn;
        const bool has_proj = !project_fn.is_none() && !sym.active;
        int rc;
        {
            py::gil_scoped_release rel;
            rc = hfgpu_solve_dm(H.k, cdp(P0), n_e, &cfg, has_proj ? project_trampoline : nullptr,
                                &ctx, &res);
        }
        check_project(rc, &ctx);
        const py::ssize_t n = (py::ssize_t)res.n_iter;
        py::object q_obj = py::none(), d_obj = py::none(), f_obj = py::none();
        if (return_Q) q_obj = Q;
        if (return_density) d_obj = dens;
        if (return_fock) f_obj = fock;
        py::tuple out = py::make_tuple(q_obj, p, d_obj, f_obj, res.mu, res.energy, res.n_iter,
                                       (bool)res.converged,
                                       hE[py::slice(0, n, 1)].attr("copy")(),
                                       hG[py::slice(0, n, 1)].attr("copy")());
        if (!return_stats) return out;
        // an addition of this module: the reference returns the ten entries
        py::dict stats;
        stats["n_ls_trials"] = res.n_ls_trials;
        return py::tuple(out + py::make_tuple(stats));
    },
          py::arg("kernel_args"), py::arg("P0"), py::arg("n_e"), py::arg("max_iter"),
          py::arg("tol_E"), py::arg("tol_grad"), py::arg("max_step"), py::arg("bt_shrink"),
          py::arg("denom_scale"), py::arg("bt_max"), py::arg("cg_restart"),
          py::arg("plateau_window"), py::arg("mu_maxiter"), py::arg("optimizer"),
          py::arg("tr_delta0"), py::arg("tr_cg_max"), py::arg("deflation_targets"),
          py::arg("deflation_sigma"), py::arg("deflation_length"), py::arg("block_sizes"),
          py::arg("project_fn"), py::arg("return_Q"), py::arg("return_density"),
          py::arg("return_fock"), py::arg("exchange_fp32") = false,
          py::arg("fp32_switch") = 1e-6, py::arg("return_stats") = false);

    // --- building blocks of the Python layer ---
    m.def("hf_energy", [](CArr P, CArr h, CArr Sigma, CArr H, RArr w, py::object refP) {
        if (P.ndim() != 4 || h.ndim() != 4 || Sigma.ndim() != 4)
            throw py::value_error("P, h, Sigma must be (nk1,nk2,nb,nb)");
        const size_t nk = P.shape(0) * P.shape(1), nb = P.shape(2);
        const size_t nd = nk * nb * nb;
        if ((size_t)h.size() != nd || (size_t)Sigma.size() != nd)
            throw py::value_error("P, h, Sigma shapes differ");
        const size_t nH = (size_t)H.size() / (nb * nb);
        if (nH * nb * nb != (size_t)H.size() || (nH != 1 && nH != nk))
            throw py::value_error("H must be (1,1,nb,nb) or (nk1,nk2,nb,nb)");
        if ((size_t)w.size() != nk) throw py::value_error("weights must be (nk1,nk2)");
        CArr r;
        const double* rp = nullptr;
        if (!refP.is_none()) {
            r = py::cast<CArr>(refP);
            if ((size_t)r.size() != nd) throw py::value_error("refP must match P");
            rp = cdp(r);
        }
        double E = 0.0;
        int rc;
        {
            py::gil_scoped_release rel;
            rc = hfgpu_hf_energy(cdp(h), cdp(P), cdp(Sigma), cdp(H), nH, cdp(w), rp, nk, nb, &E);
        }
        check(rc);
        return E;
    });

    m.def("occupation_entropy", [](RArr p, RArr w_norm) {
        const size_t nb = p.ndim() ? p.shape(p.ndim() - 1) : 1;
        const size_t nk = nb ? (size_t)p.size() / nb : 0;
        if ((size_t)w_norm.size() != nk) throw py::value_error("w_norm must match p's k axes");
        double S = 0.0;
        int rc;
        {
            py::gil_scoped_release rel;
            rc = hfgpu_occupation_entropy(cdp(p), cdp(w_norm), nk, nb, &S);
        }
        check(rc);
        return S;
    });

    m.def(
        "density_from_fock",
        [](CArr F, RArr weights, double n_e, double T, int maxiter) {
            if (F.ndim() < 2) throw py::value_error("F needs at least 2 dims");
            const size_t nb = F.shape(F.ndim() - 1);
            size_t nk = 1;
            for (int i = 0; i < F.ndim() - 2; ++i) nk *= F.shape(i);
            if ((size_t)weights.size() != nk) throw py::value_error("weights must match F's k axes");
            auto P = new_c(lead_shape(F, 0));
            py::array pa = P;
            double mu = 0.0;
            int rc;
            {
                py::gil_scoped_release rel;
                rc = hfgpu_density_from_fock(cdp(F), cdp(weights), nk, nb, n_e, T, maxiter, dp(pa),
                                             &mu);
            }
            check(rc);
            return py::make_tuple(P, mu);
        },
        py::arg("F"), py::arg("weights"), py::arg("n_e"), py::arg("T"), py::arg("maxiter") = 0);

    // --- Cayley test hooks (cpp_hf_native.cpp:585-651) ---
    m.def("_cayley_spectral_setup", [](CArr d) {
        if (d.ndim() < 2) throw py::value_error("d needs at least 2 dims");
        const size_t nb = d.shape(d.ndim() - 1);
        if ((size_t)d.shape(d.ndim() - 2) != nb) throw py::value_error("d last two dims must match");
        size_t nk = 1;
        for (int i = 0; i < d.ndim() - 2; ++i) nk *= d.shape(i);
        auto V = new_c(lead_shape(d, 0));
        auto lam = new_r(lead_shape(d, 1));
        py::array va = V, la = lam;
        int rc;
        {
            py::gil_scoped_release rel;
            rc = hfgpu_cayley_spectral_setup(cdp(d), dp(va), dp(la), nk, nb);
        }
        check(rc);
        return py::make_tuple(V, lam);
    });

    m.def("_cayley_unitary_from_spectrum", [](CArr V, RArr lam, double tau) {
        if (V.ndim() < 2) throw py::value_error("V needs at least 2 dims");
        const size_t nb = V.shape(V.ndim() - 1);
        size_t nk = 1;
        for (int i = 0; i < V.ndim() - 2; ++i) nk *= V.shape(i);
        auto U = new_c(lead_shape(V, 0));
        py::array ua = U;
        int rc;
        {
            py::gil_scoped_release rel;
            rc = hfgpu_cayley_unitary_from_spectrum(cdp(V), cdp(lam), tau, dp(ua), nk, nb);
        }
        check(rc);
        return U;
    });

    m.def("_diag_UFU_from_spectrum", [](CArr V, CArr Ft_eig, RArr lam, double tau) {
        if (V.ndim() < 2) throw py::value_error("V needs at least 2 dims");
        const size_t nb = V.shape(V.ndim() - 1);
        size_t nk = 1;
        for (int i = 0; i < V.ndim() - 2; ++i) nk *= V.shape(i);
        auto out = new_r(lead_shape(V, 1));
        py::array oa = out;
        int rc;
        {
            py::gil_scoped_release rel;
            rc = hfgpu_diag_UFU_from_spectrum(cdp(V), cdp(Ft_eig), cdp(lam), tau, dp(oa), nk, nb);
        }
        check(rc);
        return out;
    });

    // --- resample (cpp_hf_native.cpp:819-840) ---
    m.def("resample_kgrid_2d", [](CArr values, size_t nk_new) {
        if (values.ndim() < 2) throw py::value_error("need >= 2 dims");
        const size_t nk_old = values.shape(0);
        if ((size_t)values.shape(1) != nk_old)
            throw py::value_error("First two axes must be square (nk_old, nk_old)");
        size_t inner = 1;
        std::vector<py::ssize_t> out_shape{(py::ssize_t)nk_new, (py::ssize_t)nk_new};
        for (int i = 2; i < values.ndim(); ++i) {
            inner *= values.shape(i);
            out_shape.push_back(values.shape(i));
        }
        auto out = new_c(out_shape);
        py::array oa = out;
        int rc;
        {
            py::gil_scoped_release rel;
            rc = hfgpu_resample_kgrid_2d(cdp(values), dp(oa), nk_old, nk_new, inner);
        }
        check(rc);
        return out;
    });

    // --- device-resident handle (benchmarks, torch wrapper) ---
    py::class_<PyKernel>(m, "GpuKernel")
        .def(py::init([](py::object kernel_args) {
                 KernelArgs ka(kernel_args);
                 auto pk = std::make_unique<PyKernel>();
                 int rc;
                 {
                     py::gil_scoped_release rel;
                     rc = hfgpu_kernel_create(&ka.d, &pk->k);
                 }
                 check(rc);
                 return pk;
             }),
             py::arg("kernel_args"))
        .def_property_readonly("packable",
                               [](const PyKernel& k) { return hfgpu_kernel_packable(k.k) != 0; })
        .def(
            "build_fock_device",
            [](PyKernel& k, std::uintptr_t P, std::uintptr_t Sigma, std::uintptr_t F,
               std::uintptr_t hdiag, bool p_hermitian, std::uintptr_t stream) {
                int rc;
                {
                    py::gil_scoped_release rel;
                    rc = hfgpu_build_fock_device(k.k, (const void*)P, (void*)Sigma, (void*)F,
                                                 (void*)hdiag, p_hermitian ? 1 : 0, (void*)stream);
                }
                check(rc);
            },
            py::arg("P"), py::arg("Sigma"), py::arg("F"), py::arg("hdiag"),
            py::arg("p_hermitian") = true, py::arg("stream") = 0)
        .def(
            "build_fock_mixed_device",
            [](PyKernel& k, std::uintptr_t P, std::uintptr_t Sigma, std::uintptr_t F,
               std::uintptr_t hdiag, bool p_hermitian, std::uintptr_t stream) {
                int rc;
                {
                    py::gil_scoped_release rel;
                    rc = hfgpu_build_fock_mixed_device(k.k, (const void*)P, (void*)Sigma,
                                                       (void*)F, (void*)hdiag,
                                                       p_hermitian ? 1 : 0, (void*)stream);
                }
                check(rc);
            },
            py::arg("P"), py::arg("Sigma"), py::arg("F"), py::arg("hdiag"),
            py::arg("p_hermitian") = true, py::arg("stream") = 0)
        .def("autotune",
             [](PyKernel& k, bool p_hermitian) {
                 int shape[5] = {0, 0, 0, 0, 0};
                 double us = 0.0;
                 int rc;
                 {
                     py::gil_scoped_release rel;
                     rc = hfgpu_kernel_autotune(k.k, p_hermitian ? 1 : 0, shape, &us);
                 }
                 check(rc);
                 py::dict d;
                 d["us_per_build"] = us;
                 d["row_waves"] = shape[0];
                 d["col_waves"] = shape[1];
                 d["split_fwd"] = shape[2];
                 d["split_col"] = shape[3];
                 d["split_inv"] = shape[4];
                 return d;
             },
             py::arg("p_hermitian") = true)
        .def("dataflow_floor",
             [](PyKernel& k, bool p_hermitian, int reps) {
                 double us = 0.0, bytes = 0.0;
                 int rc;
                 {
                     py::gil_scoped_release rel;
                     rc = hfgpu_kernel_dataflow_floor(k.k, p_hermitian ? 1 : 0, reps, &us, &bytes);
                 }
                 check(rc);
                 py::dict d;
                 d["us_per_build"] = us;
                 d["bytes_per_build"] = bytes;
                 return d;
             },
             py::arg("p_hermitian") = true, py::arg("reps") = 200)
        .def("mixed_capable",
             [](PyKernel& k, bool p_hermitian) {
                 return hfgpu_kernel_mixed_capable(k.k, p_hermitian ? 1 : 0) != 0;
             },
             py::arg("p_hermitian") = true)
        .def(
            "energy_device",
            [](PyKernel& k, std::uintptr_t P, std::uintptr_t Sigma, std::uintptr_t hdiag,
               std::uintptr_t E, std::uintptr_t stream) {
                int rc;
                {
                    py::gil_scoped_release rel;
                    rc = hfgpu_energy_device(k.k, (const void*)P, (const void*)Sigma,
                                             (const void*)hdiag, (void*)E, (void*)stream);
                }
                check(rc);
            },
            py::arg("P"), py::arg("Sigma"), py::arg("hdiag"), py::arg("E"), py::arg("stream") = 0)
        .def(
            "build_fock_vjp_device",
            [](PyKernel& k, std::uintptr_t P, std::uintptr_t Gs, std::uintptr_t Gh,
               std::uintptr_t GF, std::uintptr_t gP, std::uintptr_t gVq, std::uintptr_t gh,
               std::uintptr_t gHH, std::uintptr_t stream) {
                int rc;
                {
                    py::gil_scoped_release rel;
                    rc = hfgpu_build_fock_vjp_device(k.k, (const void*)P, (const void*)Gs,
                                                     (const void*)Gh, (const void*)GF, (void*)gP,
                                                     (void*)gVq, (void*)gh, (void*)gHH,
                                                     (void*)stream);
                }
                check(rc);
            },
            py::arg("P"), py::arg("cot_sigma"), py::arg("cot_hartree"), py::arg("cot_F"),
            py::arg("gP"), py::arg("gVq"), py::arg("gh"), py::arg("gHH"), py::arg("stream") = 0);
}
