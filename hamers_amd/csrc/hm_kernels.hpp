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

// Fused midpoint and face flux of the single-species model (k_flux_single):
// a wavefront computes kChunk + 2 consecutive midpoint faces along the axis
// and outputs kChunk face fluxes.  HM_FUSED_AXES: bit a set = axis a fused
// (else midpoint fluxes go through memory to the face-flux kernel).  Only axis
// 0 is fused by default: on axes 1 and 2 the lanes of a fused wavefront lie
// along the sweep, one memory row apart, and their loads do not coalesce (the
// 2048^2 RHS's stage 2 took 0.90 ms with all axes fused against 0.42 ms with
// axis 0 alone).
constexpr int kChunk = 62;
#ifndef HM_FUSED_AXES
#define HM_FUSED_AXES 1
#endif

__host__ __device__ inline bool fused_axis(int axis) { return (HM_FUSED_AXES >> axis) & 1; }

// Work items of the fused axis-a segment of one patch (whole wavefronts).
__host__ __device__ inline int64_t fused_items(int axis, const int* n, int dim) {
    int64_t mid = 1;
    for (int b = 0; b < dim; b++) mid *= (b == axis) ? n[b] + 3 : n[b];
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
