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
    // With dtp non-NULL the kernels take dt from device memory instead (a
    // level advanced with its stable dt computed on the device each step:
    // one captured graph serves every dt).
    const double* dtp;
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
    // Block numbering of the RHS kernels (Schedule::xcd): 1 hands each XCD a
    // contiguous range of block numbers (xcd_block), 0 keeps the hardware's
    // round-robin order.
    int xcd;
    // Work-queue counter of the persistent midpoint-flux kernel (one unsigned
    // in device memory, cleared in-stream before each launch); NULL otherwise.
    unsigned* queue;
};

// Block number a RHS kernel works on.
__device__ __forceinline__ uint32_t rhs_block(const KParams& k) {
    return k.xcd ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x;
}

// Launch schedule of a batched RHS (chosen by the autotuner, or the built-in
// rule): tile rows of the single-species tiled flux stage (4, 8 or 24;
// ignored by the multi-species models), the block numbering (KParams::xcd)
// and the order of the work segments within each flux stage (0 patch order,
// 1 largest segments first, so the long ones do not trail the launch).
struct Schedule {
    int tile_rows = 24;
    int xcd = 0;
    int order = 0;
    int persistent = 0;  // multi-species midpoint flux as a persistent work-queue kernel
};

// The time step a kernel scales its fluxes with.
__device__ __forceinline__ double kdt(const KParams& k) { return k.dtp ? *k.dtp : k.dt; }

// Fused midpoint and face flux of the single-species model on axis 0
// (k_flux_single): a wavefront computes kChunk + 2 consecutive midpoint faces
// along the axis and outputs kChunk face fluxes.
constexpr int kChunk = 62;

// Fused midpoint and face flux of the single-species model on axes 1 and 2
// (k_flux_tile_single): a 256-thread block owns a tile of 64 x-positions by
// R midpoint-face rows along the axis (each wavefront R/4 rows), keeps the
// tile's midpoint fluxes in LDS and outputs the R - 2 inner rows.  R is chosen
// per call from the work it holds (tile_rows_for): 24 rows recompute the
// fewest shared rows, 8 and 4 give small calls (a lone patch, a small level)
// enough blocks and short wavefront latency.
constexpr int kTileX = 64;
#ifndef HM_TILE_WAVES
#define HM_TILE_WAVES 3  // wavefronts per SIMD asked of the 2D WCNS5-JS build
#endif

// Work items of the tiled axis-a (1 or 2) segment of one patch, R rows per
// tile: 256 per tile.
__host__ __device__ inline int64_t tile_items(int axis, const int* n, int dim, int rows) {
    int64_t trans = 1;  // the other transverse axis (3D)
    for (int b = 1; b < dim; b++)
        if (b != axis) trans *= n[b];
    const int64_t xt = (n[0] + kTileX - 1) / kTileX;
    const int64_t jt = (n[axis] + 1 + (rows - 2) - 1) / (rows - 2);
    return int64_t(256) * xt * trans * jt;
}

// Tile rows of a call whose 24-row tiles number blocks24 and whose 8-row
// tiles number blocks8 (sums over its patches and axes 1, 2): 24 when that
// gives at least 8 blocks per CU of the 256, else 8 when that gives 2, else 4.
inline int tile_rows_for(int64_t blocks24, int64_t blocks8) {
    if (blocks24 >= 2048) return 24;
    if (blocks8 >= 512) return 8;
    return 4;
}

// Work items of the fused axis-a segment of one patch (whole wavefronts).
__host__ __device__ inline int64_t fused_items(int axis, const int* n, int dim) {
    int64_t mid = 1;
    for (int b = 0; b < dim; b++) mid *= (b == axis) ? n[b] + 3 : n[b];
    return int64_t(64) * ((mid - 3) / kChunk + 1);
}

// Fused midpoint and face flux of the multi-species models (k_flux_multi): a
// 256-thread block owns a tile of R consecutive midpoint faces along the axis
// by W positions across it (R W <= 256), keeps the tile's midpoint fluxes in
// LDS and outputs the R - 2 inner faces of each of its W columns.  A column of
// n_A + 3 midpoint faces (n_A + 1 output faces) takes ntA tiles of at most
// kMTileFaces output faces each, of equal length, which share two midpoint
// faces; W is as many columns as fit.
constexpr int kMTileFaces = 62;

__host__ __device__ inline void mtile_shape(int nA, int& R, int& W, int& ntA) {
    ntA = (nA + 1 + kMTileFaces - 1) / kMTileFaces;
    R = (nA + 1 + ntA - 1) / ntA + 2;
    W = 256 / R;
}

// Work items of the fused axis-a segment of one patch: 256 per tile.
__host__ __device__ inline int64_t mtile_items(int axis, const int* n, int dim) {
    int64_t trans = 1;  // columns: positions across the axis
    for (int b = 0; b < dim; b++)
        if (b != axis) trans *= n[b];
    int R, W, ntA;
    mtile_shape(n[axis], R, W, ntA);
    return int64_t(256) * ntA * ((trans + W - 1) / W);
}

// Work of one RHS evaluation: one work source per pipeline stage
// (0 derived cells, 1 sensor cells -- empty: the derived-data kernels compute
// the sensor --, 2 midpoint faces, 3 output faces, 4 interior cells for the
// advective source) and the number of work items (padded) each stage launches
// over.
constexpr int kStages = 5;

template <class W>
struct Plan {
    W work[kStages];
    int64_t items[kStages];
    int tile_rows;  // rows of the tiled flux stage (single-species model): 4, 8 or 24
    int xcd;        // KParams::xcd of the plan's launches
    int persistent; // multi-species midpoint flux from a work queue (KParams::queue)
    unsigned* queue;
};

template <class W>
hipError_t launch_rhs_single(int dim, int scheme, const Plan<W>& plan, const KParams& k, hipStream_t s);
template <class W>
hipError_t launch_rhs_four(int dim, int scheme, const Plan<W>& plan, const KParams& k, hipStream_t s);
template <class W>
hipError_t launch_rhs_five(int dim, int scheme, const Plan<W>& plan, const KParams& k, hipStream_t s);

}  // namespace hm
