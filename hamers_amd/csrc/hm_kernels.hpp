// hm_kernels.hpp -- launch interface shared by the host layer and the kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "hm_device.hpp"

namespace hm {

// Scheme constants and per-call values, passed by value to every kernel.
struct KParams {
    int p, q;                       // WCNS exponents
    double C, alpha_tau;            // WCNS6-LD constants
    double gamma[HM_MAX_SPECIES];   // species ratios of specific heats
    double R[HM_MAX_SPECIES];       // species gas constants
    double cp[HM_MAX_SPECIES];      // species c_p = gamma/(gamma-1) R (four-equation mixing)
    double cv[HM_MAX_SPECIES];      // species c_v = 1/(gamma-1) R
    double igm1[HM_MAX_SPECIES];    // species 1/(gamma-1) (volume-fraction mixing)
    int num_species;
    double dt;
    // Face-state slot (single-species model): the characteristic face values
    // W-, W+ of every midpoint face, chunk DevPatch::wrec[a] per patch and
    // axis, [2 NEQ][midpoint faces].  The recording forward pass writes it,
    // the adjoint's flux half reads it instead of reconstructing again.
    // NULL: not used.
    double* wrec;
    // Adjoint runs of a level (single-species model): with lam_src >= 0 the
    // face-flux adjoints are not read from AdjPatch::Fb but formed from the
    // level's stage adjoint lam_src, F_bar_f = lam_scale (lam_f - lam_{f-1}) / dx
    // (the transposed stage update; AdjPatch::L).
    int lam_src;
    double lam_scale;
    // Offset added to every index into DevPatch::Q by the derived-data stage
    // and the adjoint's cell gather (a level adjoint reads the stage inputs
    // from a slot of its trajectory store); 0 otherwise.
    int64_t qoff;
};

// Fused midpoint and face flux of the single-species model along axis 0
// (k_flux_single): a wavefront computes kChunk + 2 consecutive midpoint faces
// and outputs kChunk face fluxes.
constexpr int kChunk = 62;

// Work items of the axis-0 segment of one patch in the fused kernel (whole
// wavefronts): n0 + 3 midpoint faces per row, rows of the other axes.
__host__ __device__ inline int64_t fused_items_axis0(const int* n, int dim) {
    int64_t mid = n[0] + 3;
    for (int b = 1; b < dim; b++) mid *= n[b];
    return int64_t(64) * ((mid - 3) / kChunk + 1);
}

// Work of one RHS evaluation: one work source per pipeline stage
// (0 derived cells, 1 sensor cells, 2 midpoint faces, 3 output faces,
// 4 interior cells for the advective source) and the number of work items
// (padded) each stage launches over.
constexpr int kStages = 5;

template <class W>
struct Plan {
    W work[kStages];
    int64_t items[kStages];
};

template <class W>
hipError_t launch_rhs_single(int dim, int scheme, const Plan<W>& plan, const KParams& k, hipStream_t s);
template <class W>
hipError_t launch_rhs_four(int dim, int scheme, const Plan<W>& plan, const KParams& k, hipStream_t s);
template <class W>
hipError_t launch_rhs_five(int dim, int scheme, const Plan<W>& plan, const KParams& k, hipStream_t s);

}  // namespace hm
