// hm_dt.hpp -- spectral radiuses and stable time step of patches on the GPU.
//
// Euler::computeSpectralRadiusesAndStableDtOnPatch (Euler.cpp:489-888) over
// the interior cells of every patch of a list (the reference's default,
// use_ghosts_to_compute_dt = false, reads the current data, which has no
// ghost cells; RungeKuttaLevelIntegrator.cpp:3787), with the level reduction of
// RungeKuttaLevelIntegrator::getLevelDt (:1175-1346):
//
//   lambda_a = |u_a| + c                      (maximum wave speed,
//                                              FlowModelSingleSpecies.cpp:3884-4070)
//   s_a      = max over cells of lambda_a / dx_a
//   dt_patch = 1 / max over cells of sum_a lambda_a / dx_a
//   dt_level = min(min over patches of dt_patch, 1 / sum_a s_a)
//   dt       = dt_level * cfl                  (cfl = min(cfl_init, cfl))
//
// The maxima are reductions of non-negative doubles: each wavefront reduces
// its lanes with fmax (which, like the reference's fmax, passes over NaN) and
// one lane per wavefront folds the result into an accumulator with a 64-bit
// unsigned atomic max on the value's bits (non-negative doubles order as their
// bit patterns), so the result does not depend on the order of the work.  A
// one-thread kernel then forms the dt, which stays in device memory: the RHS
// kernels of a level step read it there (KParams::dtp), so a captured step
// serves every dt with no host round trip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "hm_kernels.hpp"

namespace hm {

// Result slots of a dt reduction (doubles).
enum : int {
    kDtS = 0,         // s_0, s_1, s_2 (0 on unused axes)
    kDtPatch = 3,     // min over patches of dt_patch
    kDtLevel = 4,     // dt_level
    kDtStep = 5,      // dt_level * cfl: the dt a level step uses
    kDtTime = 6,      // sum of the step dts since the last reset
    kDtSlots = 8
};
constexpr int kDtAcc = 4;  // accumulators: s_0..s_2, max of the per-cell sums

// Device buffers of one dt reduction: kDtAcc accumulators, then kDtSlots results.
struct DtBuffers {
    unsigned long long* acc = nullptr;
    double* res = nullptr;
};

// Work of the spectral-radius kernel: interior cells of every patch, one
// segment per patch (DevPatch: n, gq, dx and the conservative arrays Q only).
// Launches: clear the accumulators, the per-cell maxima, the one-thread
// finish, which also adds the step's dt to the time sum kDtTime.
hipError_t launch_stable_dt(const hm_config& cfg, const SegTableT<DevPatch>& work, int64_t items, const KParams& k,
                            double cfl, const DtBuffers& buf, hipStream_t s);

// Zero the time sum kDtTime (a fill kernel: graph replays stay exact).
hipError_t launch_dt_reset_time(const DtBuffers& buf, hipStream_t s);

}  // namespace hm
