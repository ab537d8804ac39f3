// hm_sensor.hpp -- dilatation and vorticity magnitude kernel of the multi-species
// models (the single-species model computes them in its derived-data kernel).
#pragma once

#include "hm_kernels.hpp"
#include "hm_numerics.hpp"

namespace hm {

// ----------------------------------------------------------------------------
// Dilatation and vorticity magnitude (second-order central derivatives,
// DerivativeFirstOrder.cpp:216-228).
// ----------------------------------------------------------------------------
// VEL0 is the slot of the first velocity component in the scratch array prim.
template <int D, int VEL0, class W>
__global__ __launch_bounds__(kBlock) void k_sensor(W w, KParams k) {
    (void)k;
    const int64_t gid = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    const DevPatch* P;
    int ax;
    int64_t loc;
    if (!w.find(gid, P, ax, loc)) return;
    const Boxes<D> B(*P);
    const int e0 = B.ext(0, kGs), e1 = B.ext(1, kGs);
    // 32-bit index arithmetic (a work segment has fewer than 2^31 items).
    const uint32_t r = uint32_t(loc) / uint32_t(e0);
    const int i = int(uint32_t(loc) - r * uint32_t(e0)) - kGs;
    const int j = int((D > 2) ? r % uint32_t(e1) : r) - kGs;
    const int kk = (D > 2) ? int(r / uint32_t(e1)) - kGs : 0;
    const int64_t v4 = B.vol(kGc);
    const uint32_t c = uint32_t(B.idx(kGc, i, j, kk));
    const uint32_t st[3] = {1u, uint32_t(B.ext(0, kGc)), uint32_t(B.ext(0, kGc) * B.ext(1, kGc))};
    // d u_b / d x_a; quot rounds exactly as the division by dx_a does
    double rdx[3];
#pragma unroll
    for (int a = 0; a < D; a++) rdx[a] = rcp_rn(P->dx[a]);
    double g[3][3];
#pragma unroll
    for (int b = 0; b < D; b++) {
        const double* ub = P->prim + (VEL0 + b) * v4;
#pragma unroll
        for (int a = 0; a < D; a++)
            g[b][a] = quot(double(1) / double(2) * (ldg(ub, c + st[a]) - ldg(ub, c - st[a])), P->dx[a], rdx[a]);
    }
    double theta, Omega;
    if constexpr (D == 2) {
        theta = g[0][0] + g[1][1];
        Omega = fabs(g[1][0] - g[0][1]);
    } else {
        theta = g[0][0] + g[1][1] + g[2][2];
        const double ox = g[2][1] - g[1][2];
        const double oy = g[0][2] - g[2][0];
        const double oz = g[1][0] - g[0][1];
        Omega = sqrt(ox * ox + oy * oy + oz * oz);
    }
    const int64_t vs = B.vol(kGs);
    stg(P->sens, uint32_t(loc), theta);
    stg(P->sens + vs, uint32_t(loc), Omega);
}

}  // namespace hm
