// hm_sensor.hpp -- dilatation and vorticity magnitude of one cell, computed by
// the derived-data kernels of every flow model on the cells of the 2-ghost box.
#pragma once

#include "hm_kernels.hpp"
#include "hm_numerics.hpp"

namespace hm {

// Dilatation theta and vorticity magnitude Omega of cell (i, j, kk) from
// second-order central derivatives of the velocity (DerivativeFirstOrder.cpp:
// 216-228; WCNS56-HLLC-HLL.cpp:626-731 (2D), 1470-1656 (3D)), written to the
// patch's sensor array.  vel(a, s, v) returns the velocity components v[b] of
// the face neighbour at side s (+1, -1) along axis a, computed exactly as the
// derived-data kernel computes a cell's own velocity, so the values equal
// those the kernel stores for that neighbour.  Quotients by dx_a use quot,
// which rounds as the division does.
template <int D, class Vel>
__device__ __forceinline__ void sensor_cell(const DevPatch& P, const Boxes<D>& B, int i, int j, int kk, Vel vel) {
    double g[3][3];  // d u_b / d x_a
#pragma unroll
    for (int a = 0; a < D; a++) {
        double vp[3], vm[3];
        vel(a, 1, vp);
        vel(a, -1, vm);
        const double rdx = rcp_rn(P.dx[a]);
#pragma unroll
        for (int b = 0; b < D; b++) g[b][a] = quot(double(1) / double(2) * (vp[b] - vm[b]), P.dx[a], rdx);
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
    const uint32_t ls = uint32_t(B.idx(kGs, i, j, kk));
    stg(P.sens, ls, theta);
    stg(P.sens + vs, ls, Omega);
}

// Whether cell (i, j, kk) lies in the 2-ghost box, where the sensor is needed.
template <int D>
__device__ __forceinline__ bool in_sensor_box(const Boxes<D>& B, int i, int j, int kk) {
    return i >= -kGs && i < B.n[0] + kGs && j >= -kGs && j < B.n[1] + kGs &&
           (D < 3 || (kk >= -kGs && kk < B.n[2] + kGs));
}

}  // namespace hm
