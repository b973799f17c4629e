// Symmetry projector on the device: the group averages of make_project_fn
// (symmetry.py:81-117) as per-k matrix products in LDS, replacing the host
// round trip of a project_fn callback inside the solver loops.
#pragma once

#include "hfgpu_core.h"

namespace hfgpu {

struct SymPlan {
    int nk1 = 0, nk2 = 0, nb = 0;
    int n_u = 0;          // unitary set size (the identity alone when joined)
    int n_s = 0;          // spatial set size
    int s_mod = 1, s_axes = 3, t_mod = 1, t_axes = 3;
    bool has_tr = false;
    DevBuf U_set, S_set, TR, buf;

    // Validates the description against (nk1, nk2, nb) and uploads the
    // matrices; throws std::invalid_argument on bad input.
    void init(const hf_symmetry_desc& d, int nk1, int nk2, hipStream_t st);
    // M (nk1 nk2, nb, nb) on the device, in place, on stream st.
    void apply(cd* M, hipStream_t st);
};

}  // namespace hfgpu
