// hm_rk.hpp -- Runge-Kutta stage update descriptors and launch interface.
#pragma once

#include "hm_kernels.hpp"

namespace hm {

constexpr int kMaxRK = 5;  // coefficients per stage (HM_MAX_RK_COEFFS)

// One patch of a stage update: the scratch conservative arrays (updated in the
// interior) and, per coefficient n, the intermediate-context arrays.
struct RKPatch {
    int n[3];
    int gq[3];                           // ghost width of Q
    int gi[3];                           // ghost width of the intermediate Q
    double dx[3];
    double* Q[HM_MAX_EQN];               // num_eqn (+1 for the five-equation model) arrays
    const double* Qi[kMaxRK][HM_MAX_EQN];
    const double* Fi[kMaxRK][3];
    const double* Si[kMaxRK];
    double* Facc[3];                     // optional gamma accumulation (NULL to skip)
    double* Sacc;
    const double* Fdi[kMaxRK][3];        // Navier-Stokes: intermediate diffusive fluxes (NULL: Euler form)
    double* Fdacc[3];                    // Navier-Stokes: gamma accumulation of the diffusive flux
};

struct RKParams {
    int nc;         // number of coefficients of this stage
    int neq;        // equations updated
    int zfirst;     // first volume-fraction equation (five-equation model)
    int nz;         // number of evolved volume fractions (0: no last-fraction update)
    int last_beta;  // index of the last nonzero beta, -1 if none
    int ns;         // some patch carries intermediate diffusive fluxes (Navier-Stokes form)
    double alpha[kMaxRK], beta[kMaxRK], gamma[kMaxRK];
};

// Terms of one gamma accumulation, added in order: acc += gamma[t] * (flux,
// diffusive flux or source of intermediate src[t]).  One stage's terms, or a
// whole step's (stage by stage, each stage's in intermediate order: the same
// sequence of additions as one accumulation after every stage).
constexpr int kMaxAccTerms = kMaxRK * (kMaxRK + 1) / 2;
struct AccTerms {
    int n;
    int neq;
    int src[kMaxAccTerms];
    double gamma[kMaxAccTerms];
};

// Launch the update over `cell_items` interior cells and, when `acc` is given,
// the gamma accumulation over `acc_items` items (faces per axis, then cells).
template <class W>
hipError_t launch_rk_update(int dim, const W& cells, int64_t cell_items, const W* acc, int64_t acc_items,
                            const RKParams& r, hipStream_t s);

// Launch the gamma accumulation of `terms` over `acc_items` items.
template <class W>
hipError_t launch_rk_accumulate(int dim, const W& acc, int64_t acc_items, const AccTerms& terms, hipStream_t s);

}  // namespace hm
