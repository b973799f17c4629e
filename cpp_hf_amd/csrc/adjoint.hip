// Adjoint of the Fock build: the vector-Jacobian product of build_fock with
// respect to the density P, the interaction coulomb_q, h and the Hartree
// matrix.  No counterpart in the reference (which has no derivatives); the
// forward map it differentiates is fock.hpp:34-321:
//
//   Sigma = A_VR[P - refP] + SC[rho],  rho = sum_k w_k (P_k - refP_k)
//   Hh    = diag(HH n),                n_b = Re rho_bb
//   F     = herm(h + Sigma + Hh)
//   A_VR[X] = -(1/N) F^H (VR o F X) per band-pair channel (F the unnormalised
//            2-D DFT over k), VR = phase o F(w Vq)   (problem.py:84-107)
//
// so with G = G_Sigma + herm(G_F):
//   gP   = A_conj(VR)[G] + w_k (grho + diag(HH^T r))
//   gVR  = -(1/N) F(G) o conj(F(P - refP))   (summed over channels, scalar V)
//   gVq  = w o F^H(conj(phase) o gVR)
//   gh   = herm(G_F),  gHH_ab = r_a n_b,  r_a = Re sum_k (herm(G_F) + G_Hh)_aa
//   grho = sum_t g_t [conj(<S, Oi_t>) Oj_t^H - Oi_t^H S Oj_t^H],  S = sum_k G_k
// The transforms recompute F(P - refP) and F(G) (no stored intermediates).
#include <cmath>
#include <memory>
#include <vector>

#include "../../include/hfgpu.h"
#include "hfgpu_core.h"

namespace hfgpu {

namespace {

// Gt = Gs + herm(GF); gh = herm(GF) (nullable); Gd = herm(GF) + Gh (nullable)
__global__ void __launch_bounds__(256)
k_cotangents(int nk, int nb, const cd* __restrict__ Gs, const cd* __restrict__ Gh,
             const cd* __restrict__ GF, cd* __restrict__ Gt, cd* __restrict__ gh,
             cd* __restrict__ Gd) {
    const size_t nb2 = static_cast<size_t>(nb) * nb, n = static_cast<size_t>(nk) * nb2;
    for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<size_t>(gridDim.x) * blockDim.x) {
        const size_t k = i / nb2;
        const int e = static_cast<int>(i - k * nb2), a = e / nb, b = e - a * nb;
        cd hf{0.0, 0.0};
        if (GF) {
            const cd x = GF[i], y = GF[k * nb2 + static_cast<size_t>(b) * nb + a];
            hf = cd{0.5 * (x.x + y.x), 0.5 * (x.y - y.y)};
        }
        Gt[i] = Gs ? cadd(Gs[i], hf) : hf;
        if (gh) gh[i] = hf;
        if (Gd) Gd[i] = Gh ? cadd(hf, Gh[i]) : hf;
    }
}

// gP_k += w_k M  (M nb x nb, k-independent)
__global__ void __launch_bounds__(256)
k_add_weighted(size_t nk, int nb2, const double* __restrict__ w, const cd* __restrict__ M,
               cd* __restrict__ gP) {
    const size_t n = nk * nb2;
    for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<size_t>(gridDim.x) * blockDim.x) {
        const size_t k = i / nb2;
        gP[i] = cadd(gP[i], cscale(M[i - k * nb2], w[k]));
    }
}

// out[q][v] = w[q] X[q][v]
__global__ void __launch_bounds__(256)
k_scale_rows(size_t nq, int nv, const double* __restrict__ w, const cd* __restrict__ X,
             cd* __restrict__ out) {
    const size_t n = nq * nv;
    for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<size_t>(gridDim.x) * blockDim.x)
        out[i] = cscale(X[i], w[i / nv]);
}

unsigned blocks_for(size_t n) { return std::max(1u, std::min<unsigned>(cdiv(n, 256), 8192u)); }

ExchangeGeom& adjoint_geometry(std::unique_ptr<ExchangeGeom>& slot, int nk1, int nk2, int nch_nb) {
    if (!slot) {
        auto g = std::make_unique<ExchangeGeom>();
        g->init(nk1, nk2, nch_nb, false);
        slot = std::move(g);
    }
    return *slot;
}

}  // namespace

void fock_vjp(GpuKernel& K, const cd* P, const cd* Gs, const cd* Gh, const cd* GF, cd* gP,
              cd* gVq, cd* gh, double* gHH, hipStream_t st) {
    if (K.superlattice())
        throw std::invalid_argument("the adjoint of the superlattice Fock build is not provided");
    const int nk = static_cast<int>(K.nk), nb = K.nb, nb2 = static_cast<int>(K.nb2);
    const size_t nd = K.nd, bytes = nd * sizeof(cd);
    const double* w = K.w2d.as<double>();
    const bool small = K.hartree || K.contact || gHH;
    DevBuf Gt(bytes), dPb(K.has_refP ? bytes : 0), Gd(small ? bytes : 0);
    hipLaunchKernelGGL(k_cotangents, dim3(blocks_for(nd)), dim3(256), 0, st, nk, nb, Gs, Gh, GF,
                       Gt.as<cd>(), gh, small ? Gd.as<cd>() : nullptr);
    HFGPU_LAUNCH_CHECK();
    const cd* dP = P;
    if (K.has_refP) {
        launch_sub_c(nd, P, K.refP.as<cd>(), dPb.as<cd>(), st);
        dP = dPb.as<cd>();
    }

    // ---- exchange: gP = A_conj(VR)[G], gVq through the cross spectrum
    if (K.exchange) {
        ExchangeGeom& ga = adjoint_geometry(K.adj_geo, K.nk1, K.nk2, nb);
        const double inv_n = 1.0 / static_cast<double>(K.nk);
        const bool scalar = K.dv == 1;
        DevBuf X(bytes);
        launch_fft2_plain(ga, Gt.as<cd>(), X.as<cd>(), nb2, false, st);
        if (gVq) {
            const int nv = scalar ? 1 : nb2;
            DevBuf Y(bytes), Gam(K.nk * nv * sizeof(cd));
            launch_fft2_plain(ga, dP, Y.as<cd>(), nb2, false, st);
            launch_cross_spectrum(K.nk, nb2, X.as<cd>(), Y.as<cd>(), -inv_n, scalar, Gam.as<cd>(),
                                  st);
            launch_unphase(K.nk1, K.nk2, nv, Gam.as<cd>(), st);
            ExchangeGeom& gv =
                scalar ? adjoint_geometry(K.adj_geo_v, K.nk1, K.nk2, 1) : ga;
            launch_fft2_plain(gv, Gam.as<cd>(), Gam.as<cd>(), nv, true, st);
            hipLaunchKernelGGL(k_scale_rows, dim3(blocks_for(K.nk * nv)), dim3(256), 0, st, K.nk,
                               nv, w, Gam.as<cd>(), gVq);
            HFGPU_LAUNCH_CHECK();
        }
        launch_mul_spectrum(K.nk, nb2, X.as<cd>(), K.VR_full.as<cd>(), scalar, true, -inv_n, st);
        launch_fft2_plain(ga, X.as<cd>(), gP, nb2, true, st);
    } else {
        fill_zero(gP, bytes, st);
        if (gVq) fill_zero(gVq, K.nk * K.dv * K.dv * sizeof(cd), st);
    }

    // ---- Hartree and contact terms: k-independent nb x nb algebra on the host
    if (small) {
        DevBuf part(kKsumParts * nb2 * sizeof(cd)), sums(3 * nb2 * sizeof(cd));
        cd* dsum = sums.as<cd>();
        launch_ksum(nk, nb2, Gt.as<cd>(), nullptr, part.as<cd>(), dsum, st);           // S
        launch_ksum(nk, nb2, Gd.as<cd>(), nullptr, part.as<cd>(), dsum + nb2, st);     // diag -> r
        launch_ksum(nk, nb2, dP, w, part.as<cd>(), dsum + 2 * nb2, st);                // rho
        std::vector<cd> hs(3 * nb2);
        std::vector<double> HH(nb2, 0.0);
        d2h(hs.data(), dsum, 3 * nb2 * sizeof(cd), st);
        if (K.hartree || gHH) d2h(HH.data(), K.HH.get(), nb2 * sizeof(double), st);
        std::vector<cd> Oi, Oj;
        if (K.contact) {
            Oi.resize(static_cast<size_t>(K.n_contact) * nb2);
            Oj.resize(Oi.size());
            d2h(Oi.data(), K.Oi.get(), Oi.size() * sizeof(cd), st);
            d2h(Oj.data(), K.Oj.get(), Oj.size() * sizeof(cd), st);
        }
        HFGPU_CHECK(hipStreamSynchronize(st));
        const cd* S = hs.data();
        const cd* Dg = hs.data() + nb2;
        const cd* rho = hs.data() + 2 * nb2;
        std::vector<cd> M(nb2, cd{0.0, 0.0});
        if (K.hartree) {
            for (int b = 0; b < nb; ++b) {
                double c = 0.0;
                for (int a = 0; a < nb; ++a) c += HH[a * nb + b] * Dg[a * nb + a].x;
                M[b * nb + b].x += c;
            }
        }
        if (gHH) {
            std::vector<double> out(nb2, 0.0);
            if (K.hartree)
                for (int a = 0; a < nb; ++a)
                    for (int b = 0; b < nb; ++b) out[a * nb + b] = Dg[a * nb + a].x * rho[b * nb + b].x;
            HFGPU_CHECK(hipMemcpyAsync(gHH, out.data(), nb2 * sizeof(double),
                                       hipMemcpyHostToDevice, st));
            HFGPU_CHECK(hipStreamSynchronize(st));
        }
        if (K.contact) {
            for (int t = 0; t < K.n_contact; ++t) {
                const double g = K.cg_host[t];
                if (g == 0.0) continue;
                const cd* oi = Oi.data() + static_cast<size_t>(t) * nb2;
                const cd* oj = Oj.data() + static_cast<size_t>(t) * nb2;
                cd c{0.0, 0.0};  // <S, Oi> = sum conj(S) Oi
                for (int e = 0; e < nb2; ++e) c = cadd(c, cmul(cconj(S[e]), oi[e]));
                // T1 = Oi^H S, T2 = T1 Oj^H
                std::vector<cd> T1(nb2, cd{0.0, 0.0});
                for (int i = 0; i < nb; ++i)
                    for (int j = 0; j < nb; ++j) {
                        cd s{0.0, 0.0};
                        for (int m = 0; m < nb; ++m) s = cadd(s, cmul(cconj(oi[m * nb + i]), S[m * nb + j]));
                        T1[i * nb + j] = s;
                    }
                for (int i = 0; i < nb; ++i)
                    for (int j = 0; j < nb; ++j) {
                        cd s{0.0, 0.0};
                        for (int m = 0; m < nb; ++m) s = cadd(s, cmul(T1[i * nb + m], cconj(oj[j * nb + m])));
                        // g (conj(c) Oj^H - Oi^H S Oj^H)
                        const cd ojh = cconj(oj[j * nb + i]);
                        M[i * nb + j] = cadd(M[i * nb + j],
                                             cscale(csub(cmul(cconj(c), ojh), s), g));
                    }
            }
        }
        if (K.hartree || K.contact) {
            DevBuf dM(nb2 * sizeof(cd));
            h2d(dM.get(), M.data(), nb2 * sizeof(cd), st);
            hipLaunchKernelGGL(k_add_weighted, dim3(blocks_for(nd)), dim3(256), 0, st, K.nk, nb2,
                               w, dM.as<cd>(), gP);
            HFGPU_LAUNCH_CHECK();
        }
    }
    // the workspaces go back to the pool on return: drain their users first
    HFGPU_CHECK(hipStreamSynchronize(st));
}

}  // namespace hfgpu

using namespace hfgpu;

extern "C" {

int hfgpu_build_fock_vjp(hfgpu_kernel* k, const double* P, const double* cot_sigma,
                         const double* cot_hartree, const double* cot_F, double* gP,
                         double* gVq, double* gh, double* gHH) {
    return guarded([&] {
        if (!k || !P || !gP) throw std::invalid_argument("null argument");
        GpuKernel& K = *k->impl;
        hipStream_t st = K.stream;
        const size_t bytes = K.nd * sizeof(cd);
        const size_t vbytes = K.nk * K.dv * K.dv * sizeof(cd);
        DevBuf dP(bytes), dGs(cot_sigma ? bytes : 0), dGh(cot_hartree ? bytes : 0),
            dGF(cot_F ? bytes : 0), dgP(bytes), dgV(gVq ? vbytes : 0), dgh(gh ? bytes : 0),
            dgHH(gHH ? K.nb2 * sizeof(double) : 0);
        h2d(dP.get(), P, bytes, st);
        if (cot_sigma) h2d(dGs.get(), cot_sigma, bytes, st);
        if (cot_hartree) h2d(dGh.get(), cot_hartree, bytes, st);
        if (cot_F) h2d(dGF.get(), cot_F, bytes, st);
        fock_vjp(K, dP.as<cd>(), cot_sigma ? dGs.as<cd>() : nullptr,
                 cot_hartree ? dGh.as<cd>() : nullptr, cot_F ? dGF.as<cd>() : nullptr,
                 dgP.as<cd>(), gVq ? dgV.as<cd>() : nullptr, gh ? dgh.as<cd>() : nullptr,
                 gHH ? dgHH.as<double>() : nullptr, st);
        d2h(gP, dgP.get(), bytes, st);
        if (gVq) d2h(gVq, dgV.get(), vbytes, st);
        if (gh) d2h(gh, dgh.get(), bytes, st);
        if (gHH) d2h(gHH, dgHH.get(), K.nb2 * sizeof(double), st);
        HFGPU_CHECK(hipStreamSynchronize(st));
    });
}

int hfgpu_build_fock_vjp_device(hfgpu_kernel* k, const void* dP, const void* dGs,
                                const void* dGh, const void* dGF, void* dgP, void* dgVq,
                                void* dgh, void* dgHH, void* stream) {
    return guarded([&] {
        if (!k || !dP || !dgP) throw std::invalid_argument("null argument");
        fock_vjp(*k->impl, static_cast<const cd*>(dP), static_cast<const cd*>(dGs),
                 static_cast<const cd*>(dGh), static_cast<const cd*>(dGF), static_cast<cd*>(dgP),
                 static_cast<cd*>(dgVq), static_cast<cd*>(dgh), static_cast<double*>(dgHH),
                 static_cast<hipStream_t>(stream));
    });
}

}  // extern "C"
