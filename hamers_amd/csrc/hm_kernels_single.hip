// hm_kernels_single.hip -- convective-flux RHS kernels, single-species model.
//
// Pipeline per RHS evaluation (each stage is one launch over all work items
// of all patches handed to it):
//   k_derived_single    cells of the 4-ghost box: V = (rho, u.., p) and c
//                       (FlowModelSingleSpecies.cpp:2649-3179); on the 2-ghost
//                       box also the dilatation and vorticity magnitude
//                       (WCNS56-HLLC-HLL.cpp:626-731 (2D), 1470-1656 (3D))
//   k_flux_single       axis 0: midpoint fluxes (characteristic projection, WCNS
//                       interpolation, back-projection, positivity fallback,
//                       HLLC / HLLC-HLL and sensor select; WCNS56-HLLC-HLL.cpp:
//                       879-1194, 1746-2320) and the sixth-order midpoint-to-face
//                       flux times dt (1197-1282, 2327-2490), fused per wavefront
//   k_flux_tile_single  axes 1 and 2: the same, fused per 64 x 24 tile
#include <type_traits>

#include "hm_cell.hpp"
#include "hm_kernels.hpp"
#include "hm_face_single.hpp"
#include "hm_numerics.hpp"
#include "hm_sensor.hpp"

namespace hm {

// ----------------------------------------------------------------------------
// Derived cell data, and on the cells of the 2-ghost box the dilatation and
// vorticity magnitude (second-order central derivatives, DerivativeFirstOrder.cpp:
// 216-228) from the velocities of the four (six) face neighbours, recomputed
// from Q as this kernel computes its own cell's: no second pass over the
// velocities in memory.
// ----------------------------------------------------------------------------
template <int D, class W>
__global__ __launch_bounds__(kBlock) void k_derived_single(W w, KParams k) {
    const int64_t gid = int64_t(rhs_block(k)) * kBlock + threadIdx.x;
    const DevPatch* P;
    int ax;
    int64_t loc;
    if (!w.find(gid, P, ax, loc)) return;
    constexpr int NEQ = D + 2;
    const Boxes<D> B(*P);
    const int e0 = B.ext(0, kGc), e1 = B.ext(1, kGc);
    // 32-bit index arithmetic (a work segment has fewer than 2^31 items).
    const uint32_t r = uint32_t(loc) / uint32_t(e0);
    const int i = int(uint32_t(loc) - r * uint32_t(e0)) - kGc;
    const int j = (D > 1) ? int((D > 2) ? r % uint32_t(e1) : r) - kGc : 0;
    const int kk = (D > 2) ? int(r / uint32_t(e1)) - kGc : 0;
    // Q component e from a wave-uniform base (trajectory-slot offset included)
    // at a 32-bit offset within the patch's Q box
    const uint32_t iq = uint32_t(q_idx<D>(*P, i, j, kk));
    auto q = [&](int e, uint32_t at) { return ldg(P->Q[e] + k.qoff, at); };
    double rho, u[3], p, c;
    cell_state_single<D>(q, iq, k.gamma[0], rho, u, p, c);

    const int64_t vol = B.vol(kGc);
    double* pr = P->prim;
    const uint32_t l32 = uint32_t(loc);
    stg(pr, l32, rho);
#pragma unroll
    for (int b = 0; b < D; b++) stg(pr + (1 + b) * vol, l32, u[b]);
    stg(pr + (D + 1) * vol, l32, p);
    stg(pr + NEQ * vol, l32, c);

    if constexpr (D > 1) {
        if (!in_sensor_box<D>(B, i, j, kk)) return;
        const uint32_t qs[3] = {1u, uint32_t(P->n[0] + 2 * P->gq[0]),
                                uint32_t((P->n[0] + 2 * P->gq[0]) * (P->n[1] + 2 * P->gq[1]))};
        sensor_cell<D>(*P, B, i, j, kk, [&](int a, int side, double (&v)[3]) {
            const uint32_t at = side > 0 ? iq + qs[a] : iq - qs[a];
            const double rn = q(0, at);
            const double rrn = rcp_rn(rn);
#pragma unroll
            for (int b = 0; b < D; b++) v[b] = quot(q(1 + b, at), rn, rrn);
        });
    }
}

// ----------------------------------------------------------------------------
// Fused midpoint and face flux along axis 0 (single-species): a wavefront
// computes kChunk + 2 consecutive midpoint faces of the patch's flattened
// face list (rows of faces -1 .. n+1, unit stride in memory), kChunk apart
// from the next wavefront's, and its lanes 1..kChunk finish their output faces
// with their neighbours' midpoint fluxes (cross-lane shuffles) and the node
// fluxes of their two cells (WCNS56-HLLC-HLL.cpp:1197-1282, 2327-2490), so
// the midpoint flux of axis 0 never goes to memory.  A row is complete in the
// list, so the neighbours of an output face are in its row.  Axes 1 and 2 run
// in tiles whose lanes lie across the sweep (coalesced loads), below.
// ----------------------------------------------------------------------------

// Midpoint flux of the face whose right cell is (c0, c1, c2) along A, from its
// stencil V (cells i-3 .. i+2) and the sound speeds of cells i-1, i.
template <int D, int A, int SCHEME, int PP, int QQ, bool REC>
__device__ __forceinline__ void midpoint_flux_at(const DevPatch& P, const Boxes<D>& B, const KParams& k,
                                                 const double (&V)[6][D + 2], double cL, double cR, int c0, int c1,
                                                 int c2, double (&Fo)[D + 2]) {
    constexpr int NEQ = D + 2;
    bool use_hll = false;
    if constexpr (D > 1) {
        const int64_t vs = B.vol(kGs);
        const uint32_t tR = uint32_t(B.idx(kGs, c0, c1, c2));
        const uint32_t sts = (A == 0) ? 1u : (A == 1) ? uint32_t(B.ext(0, kGs)) : uint32_t(B.ext(0, kGs) * B.ext(1, kGs));
        const double th_avg = 0.5 * (ldg(P.sens, tR - sts) + ldg(P.sens, tR));
        const double om_avg = 0.5 * (ldg(P.sens + vs, tR - sts) + ldg(P.sens + vs, tR));
        const double sensor = div_rn(-th_avg, fabs(th_avg) + om_avg + kEps);
        use_hll = sensor > 0.65;
    }
    if constexpr (REC) {
        const uint32_t mv = uint32_t(B.mid_vol(A)), mi = uint32_t(B.mid_idx(A, c0, c1, c2));
        double* wr = k.wrec + P.wrec[A];
        face_flux_single<D, A, SCHEME, PP, QQ>(V, cL, cR, k.gamma[0], LDConst{k.p, k.q, k.C, k.alpha_tau}, use_hll,
                                               Fo, [&](const double* Wm, const double* Wp) {
#pragma unroll
                                                   for (int c = 0; c < NEQ; c++) {
                                                       stg(wr + size_t(c) * mv, mi, Wm[c]);
                                                       stg(wr + size_t(NEQ + c) * mv, mi, Wp[c]);
                                                   }
                                               });
    } else {
        face_flux_single<D, A, SCHEME, PP, QQ>(V, cL, cR, k.gamma[0], LDConst{k.p, k.q, k.C, k.alpha_tau}, use_hll,
                                               Fo);
    }
}

// Node flux of equation e at a cell (FlowModelSingleSpecies.cpp:3329-3331,
// 3388-3391, 3470-3474) from its conservative variables (index q in the Q
// box, below 2^29) and its velocity ua along A and pressure p.
template <int D, int A>
__device__ __forceinline__ double node_flux_single(const DevPatch& P, uint32_t q, double ua, double p, int e) {
    constexpr int IP = D + 1;
    if (e == 0) return ldg(P.Q[1 + A], q);
    if (e == IP) return ua * (ldg(P.Q[IP], q) + p);
    if (e == 1 + A) return ua * ldg(P.Q[e], q) + p;
    return ua * ldg(P.Q[e], q);
}

// Output flux of face f along A: dt (1/30 (m_{f+1} + m_{f-1}) - 3/10 (n_f + n_{f-1}) + 23/15 m_f).
template <int D>
__device__ __forceinline__ void store_face_flux(double* out, int64_t fv, int64_t fi, double dt, const double* mL,
                                                const double* mM, const double* mR, const double* nL,
                                                const double* nR) {
#pragma unroll
    for (int e = 0; e < D + 2; e++)
        stg(out + e * fv, uint32_t(fi),
            dt * (double(1) / double(30) * (mR[e] + mL[e]) - double(3) / double(10) * (nR[e] + nL[e]) +
                  double(23) / double(15) * mM[e]));
}

template <int D, int A, int SCHEME, int PP, int QQ, bool REC>
__device__ __forceinline__ void fused_chunk(const DevPatch& P, int64_t loc, const KParams& k) {
    constexpr int NEQ = D + 2;
    constexpr int IP = D + 1;
    const Boxes<D> B(P);
    const uint32_t lane = uint32_t(loc) & 63u;
    const uint32_t M = (uint32_t(loc) >> 6) * kChunk + lane;  // midpoint face, flattened with axis A fastest
    const uint32_t mv = uint32_t(B.mid_vol(A));
    const bool valid = M < mv;
    const uint32_t mA = uint32_t(B.n[A] + 3);
    const uint32_t Mc = valid ? M : mv - 1;  // lanes past the end compute a real face and drop it
    const uint32_t r = Mc / mA;
    const int cm = int(Mc - r * mA);  // position along A: face between cells cm-2 and cm-1
    // the other axes from r (x first)
    constexpr int T0 = (A == 0) ? 1 : 0, T1 = (A == 2) ? 1 : 2;
    const uint32_t nT0 = uint32_t(B.n[T0]);
    const int t0 = (D > 1) ? int((D > 2) ? r % nT0 : r) : 0;
    const int t1 = (D > 2) ? int(r / nT0) : 0;
    int c[3] = {0, 0, 0};
    c[A] = cm - 1;
    if constexpr (D > 1) c[T0] = t0;
    if constexpr (D > 2) c[T1] = t1;
    const int c0 = c[0], c1 = c[1], c2 = c[2];

    const int64_t v4 = B.vol(kGc);
    const uint32_t st = (A == 0) ? 1u : (A == 1) ? uint32_t(B.ext(0, kGc)) : uint32_t(B.ext(0, kGc) * B.ext(1, kGc));
    const uint32_t iR = uint32_t(B.idx(kGc, c0, c1, c2));
    const double* pr = P.prim;
    double V[6][NEQ];
#pragma unroll
    for (int m = 0; m < 6; m++)
#pragma unroll
        for (int e = 0; e < NEQ; e++) V[m][e] = ldg(pr + e * v4, iR + uint32_t(m - 3) * st);
    const double cL = ldg(pr + NEQ * v4, iR - st);
    const double cR = ldg(pr + NEQ * v4, iR);
    double mM[NEQ];
    midpoint_flux_at<D, A, SCHEME, PP, QQ, REC>(P, B, k, V, cL, cR, c0, c1, c2, mM);

    // Neighbours' midpoint fluxes (every lane takes part in the shuffles).
    double mL[NEQ], mR[NEQ];
#pragma unroll
    for (int e = 0; e < NEQ; e++) {
        mL[e] = __shfl_up(mM[e], 1);
        mR[e] = __shfl_down(mM[e], 1);
    }
    // Output face f = c[A] of the line, for faces 0..n_A held by lanes 1..62.
    if (!valid || lane == 0u || lane > uint32_t(kChunk) || cm < 1 || cm > B.n[A] + 1) return;
    const uint32_t qst = (A == 0) ? 1u : (A == 1) ? uint32_t(P.n[0] + 2 * P.gq[0]) : uint32_t((P.n[0] + 2 * P.gq[0]) * (P.n[1] + 2 * P.gq[1]));
    const uint32_t qR = uint32_t(q_idx<D>(P, c0, c1, c2)), qL = qR - qst;
    double nL[NEQ], nR[NEQ];
#pragma unroll
    for (int e = 0; e < NEQ; e++) {
        nL[e] = node_flux_single<D, A>(P, qL, V[2][1 + A], V[2][IP], e);
        nR[e] = node_flux_single<D, A>(P, qR, V[3][1 + A], V[3][IP], e);
    }
    store_face_flux<D>(P.F[A], B.face_vol(A), B.face_idx(A, c0, c1, c2), kdt(k), mL, mM, mR, nL, nR);
}


// Stage 2 of the single-species pipeline: the fused midpoint and face flux of
// axis 0 (the stage's only segments).
template <int D, int SCHEME, int PP, int QQ, class W, bool REC = false>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(SCHEME == kWCNS5_JS && D == 2 ? 4 : 1))) void k_flux_single(
    W w, KParams k) {
    const int64_t gid = int64_t(rhs_block(k)) * kBlock + threadIdx.x;
    const DevPatch* P;
    int ax;
    int64_t loc;
    if (!w.find(gid, P, ax, loc)) return;
    if (ax == 0) fused_chunk<D, 0, SCHEME, PP, QQ, REC>(*P, loc, k);
}

// Axes 1 and 2 of the fused midpoint and face flux, one tile per 256-thread
// block: 64 x-positions (threads tx) by R midpoint faces along A (wavefront ty
// takes rows ty, ty+4, ty+8, ...), coalesced along x.  The tile's midpoint
// fluxes go to LDS; after a barrier the inner R - 2 rows become output faces (1/30, 3/10, 23/15 combination with the node fluxes
// of their cells).  Adjacent tiles share two midpoint rows, computed by both.
// The stage's segments hold whole tiles (256 items each; every segment of the
// stage is a tile segment, so each starts on a multiple of 256) and every
// block lies in one tile: the barrier is reached by all four wavefronts.
template <int D, int A, int R, int SCHEME, int PP, int QQ, bool REC>
__device__ __forceinline__ void fused_tile(const DevPatch& P, int64_t loc, const KParams& k, double* lds) {
    constexpr int NEQ = D + 2;
    constexpr int IP = D + 1;
    constexpr int TB = (A == 1) ? 2 : 1;  // the other transverse axis (3D)
    const Boxes<D> B(P);
    const uint32_t tile = uint32_t(loc) >> 8, tid = uint32_t(loc) & 255u;
    const int tx = int(tid & 63u), ty = int(tid >> 6);
    const uint32_t nxt = uint32_t(B.n[0] + kTileX - 1) / uint32_t(kTileX);
    const uint32_t nt = (D > 2) ? uint32_t(B.n[TB]) : 1u;
    const uint32_t xt = tile % nxt, r = tile / nxt;
    const int t = (D > 2) ? int(r % nt) : 0;
    const int jt = int(r / nt);
    const int x = int(xt) * kTileX + tx;
    const bool vx = x < B.n[0];
    const int xc = vx ? x : B.n[0] - 1;  // lanes past the patch edge compute a real face and drop it
    const int nA = B.n[A];
    const int j0 = jt * (R - 2) - 1;  // midpoint face of tile row 0
    auto cell = [&](int j, int& c0, int& c1, int& c2) {
        c0 = xc;
        c1 = (A == 1) ? j : t;
        c2 = (A == 2) ? j : ((D > 2) ? t : 0);
    };
    const int64_t v4 = B.vol(kGc);
    const uint32_t st = (A == 1) ? uint32_t(B.ext(0, kGc)) : uint32_t(B.ext(0, kGc) * B.ext(1, kGc));
    const double* pr = P.prim;
#pragma unroll 1
    for (int q = 0; q < R / 4; q++) {
        const int row = ty + 4 * q;
        const int j = j0 + row;
        if (j > nA + 1) break;  // past the last midpoint face (uniform per wavefront: one row each)
        int c0, c1, c2;
        cell(j, c0, c1, c2);
        const uint32_t iR = uint32_t(B.idx(kGc, c0, c1, c2));
        double V[6][NEQ];
#pragma unroll
        for (int m = 0; m < 6; m++)
#pragma unroll
            for (int e = 0; e < NEQ; e++) V[m][e] = ldg(pr + e * v4, iR + uint32_t(m - 3) * st);
        const double cL = ldg(pr + NEQ * v4, iR - st), cR = ldg(pr + NEQ * v4, iR);
        double m[NEQ];
        midpoint_flux_at<D, A, SCHEME, PP, QQ, REC>(P, B, k, V, cL, cR, c0, c1, c2, m);
#pragma unroll
        for (int e = 0; e < NEQ; e++) lds[(e * R + row) * kTileX + tx] = m[e];
    }
    __syncthreads();
    const uint32_t qst = (A == 1) ? uint32_t(P.n[0] + 2 * P.gq[0]) : uint32_t((P.n[0] + 2 * P.gq[0]) * (P.n[1] + 2 * P.gq[1]));
#pragma unroll 1
    for (int q = 0; q < R / 4; q++) {
        const int row = ty + 4 * q;
        const int f = j0 + row;  // output face of this row
        if (row < 1 || row > R - 2 || !vx || f > nA) continue;
        int c0, c1, c2;
        cell(f, c0, c1, c2);
        const uint32_t pR = uint32_t(B.idx(kGc, c0, c1, c2)), pL = pR - st;
        const uint32_t qR = uint32_t(q_idx<D>(P, c0, c1, c2)), qL = qR - qst;
        const double uaL = ldg(pr + (1 + A) * v4, pL), uaR = ldg(pr + (1 + A) * v4, pR);
        const double prL = ldg(pr + IP * v4, pL), prR = ldg(pr + IP * v4, pR);
        double mL[NEQ], mM[NEQ], mR[NEQ], nL[NEQ], nR[NEQ];
#pragma unroll
        for (int e = 0; e < NEQ; e++) {
            mL[e] = lds[(e * R + row - 1) * kTileX + tx];
            mM[e] = lds[(e * R + row) * kTileX + tx];
            mR[e] = lds[(e * R + row + 1) * kTileX + tx];
            nL[e] = node_flux_single<D, A>(P, qL, uaL, prL, e);
            nR[e] = node_flux_single<D, A>(P, qR, uaR, prR, e);
        }
        store_face_flux<D>(P.F[A], B.face_vol(A), B.face_idx(A, c0, c1, c2), kdt(k), mL, mM, mR, nL, nR);
    }
}

template <int D, int R, int SCHEME, int PP, int QQ, class W, bool REC = false>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(SCHEME == kWCNS5_JS && D == 2 ? HM_TILE_WAVES : 1))) void k_flux_tile_single(
    W w, KParams k) {
    static_assert(R % 4 == 0 && R >= 4, "tile rows: a multiple of the block's four wavefronts");
    if constexpr (D > 1) {
        __shared__ double lds[R * kTileX * (D + 2)];
        const int64_t gid = int64_t(rhs_block(k)) * kBlock + threadIdx.x;
        const DevPatch* P;
        int ax;
        int64_t loc;
        if (!w.find(gid, P, ax, loc)) return;  // whole blocks only: the stage's items are whole tiles
        if (ax == 1) {
            fused_tile<D, 1, R, SCHEME, PP, QQ, REC>(*P, loc, k, lds);
        } else if constexpr (D > 2) {
            fused_tile<D, 2, R, SCHEME, PP, QQ, REC>(*P, loc, k, lds);
        }
    }
}

// One patch (the per-patch entry point): axis 0's chunks and the tiles of axes
// 1 and 2 in one launch, so that the two flux stages of a small call overlap
// instead of running one after the other.  The axis-0 segment is padded to a
// multiple of 256 items, so a block holds either chunks or one tile.
template <int D, int R, int SCHEME, int PP, int QQ, class W>
__global__ __launch_bounds__(kBlock) void k_flux_patch_single(W w, KParams k) {
    __shared__ double lds[(D > 1) ? R * kTileX * (D + 2) : 1];
    const int64_t gid = int64_t(rhs_block(k)) * kBlock + threadIdx.x;
    const DevPatch* P;
    int ax;
    int64_t loc;
    if (!w.find(gid, P, ax, loc)) return;  // holes lie in axis-0 blocks only
    if (ax == 0) {
        fused_chunk<D, 0, SCHEME, PP, QQ, false>(*P, loc, k);
    } else if constexpr (D > 1) {
        if (ax == 1) {
            fused_tile<D, 1, R, SCHEME, PP, QQ, false>(*P, loc, k, lds);
        } else if constexpr (D > 2) {
            fused_tile<D, 2, R, SCHEME, PP, QQ, false>(*P, loc, k, lds);
        }
    }
}

// ----------------------------------------------------------------------------
// Host-side launch of the whole single-species pipeline.
// ----------------------------------------------------------------------------
namespace {

template <int D, int SCHEME, int PP, int QQ, class W>
hipError_t launch_single_dsq(const Plan<W>& plan, const KParams& k, hipStream_t s) {
    if (plan.items[0] > 0)
        hipLaunchKernelGGL((k_derived_single<D, W>), dim3(unsigned((plan.items[0] + kBlock - 1) / kBlock)),
                           dim3(kBlock), 0, s, plan.work[0], k);
    // (the dilatation and vorticity are computed by k_derived_single: no sensor stage)
    // Stage 2: fused flux on axis 0; stage 3: fused flux on axes 1 and 2 (tiles).
    if constexpr (std::is_same_v<W, SegSingle>) {
        // one patch: both flux stages in one launch (stage 2's table holds all axes)
        if (plan.items[2] > 0) {
            const dim3 grid(unsigned((plan.items[2] + kBlock - 1) / kBlock));
            switch (plan.tile_rows) {
                case 24:
                    hipLaunchKernelGGL((k_flux_patch_single<D, 24, SCHEME, PP, QQ, W>), grid, dim3(kBlock), 0, s,
                                       plan.work[2], k);
                    break;
                case 8:
                    hipLaunchKernelGGL((k_flux_patch_single<D, 8, SCHEME, PP, QQ, W>), grid, dim3(kBlock), 0, s,
                                       plan.work[2], k);
                    break;
                case 4:
                    hipLaunchKernelGGL((k_flux_patch_single<D, 4, SCHEME, PP, QQ, W>), grid, dim3(kBlock), 0, s,
                                       plan.work[2], k);
                    break;
                default: return hipErrorInvalidValue;
            }
        }
        return hipGetLastError();
    } else {
        const bool rec = k.wrec != nullptr;  // recording runs (patch tables only)
        if (plan.items[2] > 0) {
            const dim3 grid(unsigned((plan.items[2] + kBlock - 1) / kBlock));
            if (rec)
                hipLaunchKernelGGL((k_flux_single<D, SCHEME, PP, QQ, W, true>), grid, dim3(kBlock), 0, s, plan.work[2], k);
            else
                hipLaunchKernelGGL((k_flux_single<D, SCHEME, PP, QQ, W>), grid, dim3(kBlock), 0, s, plan.work[2], k);
        }
        if (plan.items[3] > 0) {
            if constexpr (D > 1) {
                const dim3 grid(unsigned((plan.items[3] + kBlock - 1) / kBlock));
                auto tiles = [&](auto rows) {
                    constexpr int R = decltype(rows)::value;
                    if (rec)
                        hipLaunchKernelGGL((k_flux_tile_single<D, R, SCHEME, PP, QQ, W, true>), grid, dim3(kBlock), 0, s,
                                           plan.work[3], k);
                    else
                        hipLaunchKernelGGL((k_flux_tile_single<D, R, SCHEME, PP, QQ, W>), grid, dim3(kBlock), 0, s,
                                           plan.work[3], k);
                };
                switch (plan.tile_rows) {
                    case 24: tiles(std::integral_constant<int, 24>{}); break;
                    case 8: tiles(std::integral_constant<int, 8>{}); break;
                    case 4: tiles(std::integral_constant<int, 4>{}); break;
                    default: return hipErrorInvalidValue;
                }
            }
        }
        return hipGetLastError();
    }
}

template <int D, class W>
hipError_t launch_single_d(const Plan<W>& plan, const KParams& k, int scheme, hipStream_t s) {
    const bool dflt = (k.p == 2 && k.q == 4);
    if (scheme == kWCNS6_LD)
        return dflt ? launch_single_dsq<D, kWCNS6_LD, 2, 4>(plan, k, s) : launch_single_dsq<D, kWCNS6_LD, -1, -1>(plan, k, s);
    return (k.p == 2) ? launch_single_dsq<D, kWCNS5_JS, 2, 4>(plan, k, s) : launch_single_dsq<D, kWCNS5_JS, -1, -1>(plan, k, s);
}

}  // namespace

template <class W>
hipError_t launch_rhs_single(int dim, int scheme, const Plan<W>& plan, const KParams& k, hipStream_t s) {
    switch (dim) {
        case 1: return launch_single_d<1>(plan, k, scheme, s);
        case 2: return launch_single_d<2>(plan, k, scheme, s);
        case 3: return launch_single_d<3>(plan, k, scheme, s);
        default: return hipErrorInvalidValue;
    }
}

template hipError_t launch_rhs_single<SegSingle>(int, int, const Plan<SegSingle>&, const KParams&, hipStream_t);
template hipError_t launch_rhs_single<SegTable>(int, int, const Plan<SegTable>&, const KParams&, hipStream_t);

}  // namespace hm
