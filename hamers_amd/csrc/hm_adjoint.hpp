// hm_adjoint.hpp -- discrete adjoint (vector-Jacobian product) of the
// convective-flux RHS, single-species model.
//
// Forward map of one patch: Q (ghost box) -> F_a (faces, times dt).  Given
// the adjoints F_bar_a of the face fluxes, the VJP returns Q_bar = (dF/dQ)^T
// F_bar over the ghost box.  It runs as
//   1. the forward derived-data and sensor stages (prim V, c; theta, Omega),
//   2. k_adj_face_flux: per midpoint face, the adjoint of the sixth-order
//      midpoint-to-face combination, then the reverse sweep of the face's
//      Riemann flux and back-projection (hm_adjoint_face.hpp) -- from the face
//      states a recording forward pass kept, when given, else after
//      reconstructing them again -- handed over in
//      a per-face buffer H; k_adj_face_recon: the reverse reconstruction and
//      characteristic projection, one field at a time; the six stencil
//      cells' adjoints go to a per-face buffer G (two kernels, so that each
//      half fits the register file without spilling),
//   3. k_adj_cells: per cell of the 4-ghost box, a deterministic gather of the
//      G entries of every face whose stencil holds the cell, plus the node
//      flux adjoint, mapped through the derived-data Jacobian to Q_bar.
// No atomics: the result is reproducible bit for bit.
#pragma once

#include <vector>

#include "hm_reconstructor.hpp"

namespace hm {

// Stage adjoints of a level patch: stage m at L0 + m * stride, num_eqn
// components over the box grown by g (interior cells used).
struct AdjLamSource {
    const double* L0;
    int64_t stride;
    int g;
};

struct AdjPatch {
    DevPatch f;                // forward view: n, gq, dx, Q, prim, sens
    const double* Fb[3];       // face-flux adjoints per axis (layout of F)
    AdjLamSource lam;          // level runs (KParams::lam_src >= 0): the stage adjoints instead of Fb
    double* G[3];              // per midpoint face: 6*NEQ + 2 stencil adjoints
    double* H[3];              // per midpoint face: hand-over between the two face-adjoint kernels
    double* Qb[HM_MAX_EQN];    // output over the Q ghost box
};

struct AdjPlanDev {
    SegTableT<AdjPatch> face, cell, qbox;
    int64_t n_face = 0, n_cell = 0, n_qbox = 0;  // n_qbox 0: every Q box is the 4-ghost box (no zero fill)
};

hipError_t launch_adjoint_single(int dim, int scheme, const AdjPlanDev& plan, const KParams& k, hipStream_t s);

// Owner of the tables and scratch of VJPs over fixed patch lists.
class AdjointSingle {
   public:
    explicit AdjointSingle(ConvectiveFluxReconstructorGPU* rec) : rec_(rec) {}
    ~AdjointSingle();

    struct Owned {
        ConvectiveFluxReconstructorGPU::OwnedPlan fwd;
        void* scratch = nullptr;  // forward scratch (prim, sens, ...) and the G buffers
        void* tables = nullptr;   // AdjPatch array and segment tables
        AdjPlanDev plan{};
    };

    // wrec_off (optional, 3 per patch): the patches' chunks in a face-state
    // slot, for runs given one (run's wst).
    // lam (optional, 1 per patch): the level's stage adjoints, for runs given lam_src.
    int build(const hm_vjp_patch_view* views, int np, Owned& out, const int64_t* wrec_off = nullptr,
              const AdjLamSource* lam = nullptr);
    void release(Owned& o);
    // wst: the face-state slot a recording forward pass wrote for these
    // inputs (the flux half then reads W-, W+ instead of reconstructing), or
    // NULL.  lam_src >= 0: the face-flux adjoints are lam_scale times the
    // face differences of stage adjoint lam_src over dx (else read from F_bar).
    // qoff: offset added to the indices into the views' Q arrays.
    int run(const Owned& o, double dt, hipStream_t s, const double* wst = nullptr, int lam_src = -1,
            double lam_scale = 0.0, int64_t qoff = 0);

    // One-shot VJP of a patch list (the C-ABI path); keeps the last list's plan.
    int vjp(const hm_vjp_patch_view* views, int np, double dt, hipStream_t s);

   private:
    ConvectiveFluxReconstructorGPU* rec_;
    Owned cached_{};
    std::vector<hm_vjp_patch_view> cached_views_;
    StreamOrder order_;  // calls on different streams take the cached scratch in turn
};

}  // namespace hm
