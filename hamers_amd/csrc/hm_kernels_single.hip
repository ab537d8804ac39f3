// hm_kernels_single.hip -- convective-flux RHS kernels, single-species model.
//
// Pipeline per RHS evaluation (each stage is one launch over all work items
// of all patches handed to it):
//   k_derived   cells of the 4-ghost box: V = (rho, u.., p) and c
//               (FlowModelSingleSpecies.cpp:2649-3179)
//   k_sensor    cells of the 2-ghost box: dilatation and vorticity magnitude
//               (WCNS56-HLLC-HLL.cpp:626-731 (2D), 1470-1656 (3D))
//   k_midflux   every midpoint face of every axis: characteristic projection,
//               WCNS interpolation, back-projection, positivity fallback,
//               HLLC / HLLC-HLL and sensor select, fused in registers
//               (WCNS56-HLLC-HLL.cpp:879-1194, 1746-2320)
//   k_faceflux  every face: sixth-order midpoint-to-face flux times dt
//               (WCNS56-HLLC-HLL.cpp:1197-1282, 2327-2490)
#include <type_traits>

#include "hm_kernels.hpp"
#include "hm_face_single.hpp"
#include "hm_numerics.hpp"
#include "hm_sensor.hpp"

namespace hm {

// ----------------------------------------------------------------------------
// Derived cell data.
// ----------------------------------------------------------------------------
template <int D, class W>
__global__ __launch_bounds__(kBlock) void k_derived_single(W w, KParams k) {
    const int64_t gid = int64_t(blockIdx.x) * kBlock + threadIdx.x;
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
    const int64_t iq = q_idx<D>(*P, i, j, kk) + k.qoff;
    const double gam = k.gamma[0];

    const double rho = P->Q[0][iq];
    double u[3];
#pragma unroll
    for (int b = 0; b < D; b++) u[b] = P->Q[1 + b][iq] / rho;  // FlowModelSingleSpecies.cpp:2719
    double ke;
    if constexpr (D == 1)
        ke = u[0] * u[0];
    else if constexpr (D == 2)
        ke = u[0] * u[0] + u[1] * u[1];
    else
        ke = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    const double E = P->Q[D + 1][iq];
    const double eps_int = E / rho - double(1) / double(2) * ke;  // :2934
    const double p = (gam - double(1)) * rho * eps_int;           // EquationOfStateIdealGas.cpp:5477
    const double c = sqrt(gam * p / rho);                         // EquationOfStateIdealGas.cpp:5806

    const int64_t vol = B.vol(kGc);
    double* pr = P->prim;
    pr[loc] = rho;
#pragma unroll
    for (int b = 0; b < D; b++) pr[(1 + b) * vol + loc] = u[b];
    pr[(D + 1) * vol + loc] = p;
    pr[NEQ * vol + loc] = c;
}

// ----------------------------------------------------------------------------
// Fused midpoint flux of one face (single-species).
// ----------------------------------------------------------------------------
template <int D, int A, int SCHEME, int PP, int QQ, bool REC>
__device__ __forceinline__ void midflux_single_face(const DevPatch& P, int64_t loc, const KParams& k) {
    constexpr int NEQ = D + 2;
    const Boxes<D> B(P);
    const int m0 = B.mid_ext(A, 0), m1 = B.mid_ext(A, 1);
    // 32-bit index arithmetic (a work segment has fewer than 2^31 items).
    const uint32_t r = uint32_t(loc) / uint32_t(m0);
    int c0 = int(uint32_t(loc) - r * uint32_t(m0));
    int c1 = (D > 1) ? int((D > 2) ? r % uint32_t(m1) : r) : 0;
    int c2 = (D > 2) ? int(r / uint32_t(m1)) : 0;
    if constexpr (A == 0) c0 -= 1;
    if constexpr (A == 1) c1 -= 1;
    if constexpr (A == 2) c2 -= 1;

    const int64_t v4 = B.vol(kGc);
    const int64_t st = (A == 0) ? 1 : (A == 1) ? int64_t(B.ext(0, kGc)) : int64_t(B.ext(0, kGc)) * B.ext(1, kGc);
    const int64_t iR = B.idx(kGc, c0, c1, c2);  // cell i (right of the face); cell i-1 is iR - st
    const double* pr = P.prim;

    // Stencil of primitive variables, cells i-3 .. i+2.
    const uint32_t i32 = uint32_t(iR), st32 = uint32_t(st);
    double V[6][NEQ];
#pragma unroll
    for (int m = 0; m < 6; m++)
#pragma unroll
        for (int e = 0; e < NEQ; e++) V[m][e] = ldg(pr + e * v4, i32 + uint32_t(m - 3) * st32);
    const double cL = ldg(pr + NEQ * v4, i32 - st32);
    const double cR = ldg(pr + NEQ * v4, i32);

    // Sensor (WCNS56-HLLC-HLL.cpp:1066-1113): only when it exceeds 0.65 is the
    // HLLC-HLL flux used, so HLLC-HLL is evaluated only on those faces.
    bool use_hll = false;
    if constexpr (D > 1) {
        const int64_t vs = B.vol(kGs);
        const int64_t tR = B.idx(kGs, c0, c1, c2);
        const int64_t sts = (A == 0) ? 1 : (A == 1) ? int64_t(B.ext(0, kGs)) : int64_t(B.ext(0, kGs)) * B.ext(1, kGs);
        const int64_t tL = tR - sts;
        const double th_avg = 0.5 * (P.sens[tL] + P.sens[tR]);
        const double om_avg = 0.5 * (P.sens[vs + tL] + P.sens[vs + tR]);
        const double sensor = div_rn(-th_avg, fabs(th_avg) + om_avg + kEps);
        use_hll = sensor > 0.65;
    }

    double Fo[NEQ];
    const int64_t mv = B.mid_vol(A);
    if constexpr (REC) {
        // Recording pass: keep W-, W+ in the face-state slot.
        double* wr = k.wrec + P.wrec[A] + loc;
        face_flux_single<D, A, SCHEME, PP, QQ>(V, cL, cR, k.gamma[0], LDConst{k.p, k.q, k.C, k.alpha_tau}, use_hll,
                                               Fo, [&](const double* Wm, const double* Wp) {
#pragma unroll
                                                   for (int c = 0; c < NEQ; c++) {
                                                       wr[c * mv] = Wm[c];
                                                       wr[(NEQ + c) * mv] = Wp[c];
                                                   }
                                               });
    } else {
        face_flux_single<D, A, SCHEME, PP, QQ>(V, cL, cR, k.gamma[0], LDConst{k.p, k.q, k.C, k.alpha_tau}, use_hll,
                                               Fo);
    }

    double* out = P.fmid[A];
#pragma unroll
    for (int e = 0; e < NEQ; e++) out[e * mv + loc] = Fo[e];
}


// ----------------------------------------------------------------------------
// Fused midpoint and face flux along axis 0 (single-species): a wavefront
// computes kChunk + 2 consecutive midpoint faces of the patch's flattened
// face list (rows of faces -1 .. n+1, unit stride in memory), kChunk apart
// from the next wavefront's, and its lanes 1..kChunk finish their output faces
// with their neighbours' midpoint fluxes (cross-lane shuffles) and the node
// fluxes of their two cells (WCNS56-HLLC-HLL.cpp:1197-1282, 2327-2490), so
// the midpoint flux of axis 0 never goes to memory.  A row is complete in the
// list, so the neighbours of an output face are in its row.  On axes 1 and 2
// the lanes of a wavefront lie across the sweep (coalesced loads); there the
// midpoint flux goes through memory to the face-flux kernel.
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
                                                       stg(wr, uint32_t(c) * mv + mi, Wm[c]);
                                                       stg(wr, uint32_t(NEQ + c) * mv + mi, Wp[c]);
                                                   }
                                               });
    } else {
        face_flux_single<D, A, SCHEME, PP, QQ>(V, cL, cR, k.gamma[0], LDConst{k.p, k.q, k.C, k.alpha_tau}, use_hll,
                                               Fo);
    }
}

// Node flux of equation e at a cell (FlowModelSingleSpecies.cpp:3329-3331,
// 3388-3391, 3470-3474) from its conservative variables (index q) and its
// velocity ua along A and pressure p.
template <int D, int A>
__device__ __forceinline__ double node_flux_single(const DevPatch& P, int64_t q, double ua, double p, int e) {
    constexpr int IP = D + 1;
    if (e == 0) return P.Q[1 + A][q];
    if (e == IP) return ua * (P.Q[IP][q] + p);
    if (e == 1 + A) return ua * P.Q[e][q] + p;
    return ua * P.Q[e][q];
}

// Output flux of face f along A: dt (1/30 (m_{f+1} + m_{f-1}) - 3/10 (n_f + n_{f-1}) + 23/15 m_f).
template <int D>
__device__ __forceinline__ void store_face_flux(double* out, int64_t fv, int64_t fi, double dt, const double* mL,
                                                const double* mM, const double* mR, const double* nL,
                                                const double* nR) {
#pragma unroll
    for (int e = 0; e < D + 2; e++)
        out[e * fv + fi] = dt * (double(1) / double(30) * (mR[e] + mL[e]) - double(3) / double(10) * (nR[e] + nL[e]) +
                                 double(23) / double(15) * mM[e]);
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
    const int64_t qst = (A == 0) ? 1 : (A == 1) ? (P.n[0] + 2 * P.gq[0]) : int64_t(P.n[0] + 2 * P.gq[0]) * (P.n[1] + 2 * P.gq[1]);
    const int64_t qR = q_idx<D>(P, c0, c1, c2), qL = qR - qst;
    double nL[NEQ], nR[NEQ];
#pragma unroll
    for (int e = 0; e < NEQ; e++) {
        nL[e] = node_flux_single<D, A>(P, qL, V[2][1 + A], V[2][IP], e);
        nR[e] = node_flux_single<D, A>(P, qR, V[3][1 + A], V[3][IP], e);
    }
    store_face_flux<D>(P.F[A], B.face_vol(A), B.face_idx(A, c0, c1, c2), k.dt, mL, mM, mR, nL, nR);
}


// Stage 2 of the single-species pipeline: axis 0 runs the fused midpoint and
// face flux; axes 1 and 2 write midpoint fluxes for the face-flux kernel
// (stage 3), their lanes lying across the sweep so that the loads coalesce.
template <int D, int SCHEME, int PP, int QQ, class W, bool REC = false>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(SCHEME == kWCNS5_JS && D == 2 ? 4 : 1))) void k_flux_single(
    W w, KParams k) {
    const int64_t gid = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    const DevPatch* P;
    int ax;
    int64_t loc;
    if (!w.find(gid, P, ax, loc)) return;
    if (ax == 0) {
        fused_chunk<D, 0, SCHEME, PP, QQ, REC>(*P, loc, k);
    } else if constexpr (D > 1) {
        if (ax == 1) {
            if constexpr (HM_FUSED_AXES & 2)
                fused_chunk<D, 1, SCHEME, PP, QQ, REC>(*P, loc, k);
            else
                midflux_single_face<D, 1, SCHEME, PP, QQ, REC>(*P, loc, k);
        } else if constexpr (D > 2) {
            if constexpr (HM_FUSED_AXES & 4)
                fused_chunk<D, 2, SCHEME, PP, QQ, REC>(*P, loc, k);
            else
                midflux_single_face<D, 2, SCHEME, PP, QQ, REC>(*P, loc, k);
        }
    }
}


// ----------------------------------------------------------------------------
// Midpoint-to-face flux (sixth order), times dt.
// ----------------------------------------------------------------------------
template <int D, int A>
__device__ __forceinline__ void faceflux_single_face(const DevPatch& P, int64_t loc, const KParams& k) {
    constexpr int NEQ = D + 2;
    constexpr int IP = D + 1;
    const Boxes<D> B(P);
    const int f0 = B.face_ext(A, 0), f1 = B.face_ext(A, 1);
    // 32-bit index arithmetic (a work segment has fewer than 2^31 items).
    const uint32_t r = uint32_t(loc) / uint32_t(f0);
    const int c0 = int(uint32_t(loc) - r * uint32_t(f0));
    const int c1 = (D > 1) ? int((D > 2) ? r % uint32_t(f1) : r) : 0;
    const int c2 = (D > 2) ? int(r / uint32_t(f1)) : 0;

    const int64_t mv = B.mid_vol(A);
    const int64_t mC = B.mid_idx(A, c0, c1, c2);
    const int64_t mst = (A == 0) ? 1 : (A == 1) ? int64_t(B.mid_ext(A, 0)) : int64_t(B.mid_ext(A, 0)) * B.mid_ext(A, 1);

    // Node flux at cells i-1 and i (FlowModelSingleSpecies.cpp:3329-3331, 3388-3391, 3470-3474).
    const int64_t v4 = B.vol(kGc);
    const int64_t st4 = (A == 0) ? 1 : (A == 1) ? int64_t(B.ext(0, kGc)) : int64_t(B.ext(0, kGc)) * B.ext(1, kGc);
    const int64_t pR = B.idx(kGc, c0, c1, c2), pL = pR - st4;
    const int qst = (A == 0) ? 1 : (A == 1) ? (P.n[0] + 2 * P.gq[0]) : (P.n[0] + 2 * P.gq[0]) * (P.n[1] + 2 * P.gq[1]);
    const int64_t qR = q_idx<D>(P, c0, c1, c2), qL = qR - qst;
    const double uaL = P.prim[(1 + A) * v4 + pL], uaR = P.prim[(1 + A) * v4 + pR];
    const double prL = P.prim[IP * v4 + pL], prR = P.prim[IP * v4 + pR];

    const int64_t fv = B.face_vol(A);
    const double* fm = P.fmid[A];
    double* out = P.F[A];
    const double dt = k.dt;
#pragma unroll
    for (int e = 0; e < NEQ; e++) {
        double nL, nR;
        if (e == 0) {
            nL = P.Q[1 + A][qL];
            nR = P.Q[1 + A][qR];
        } else if (e == IP) {
            nL = uaL * (P.Q[IP][qL] + prL);
            nR = uaR * (P.Q[IP][qR] + prR);
        } else if (e == 1 + A) {
            nL = uaL * P.Q[e][qL] + prL;
            nR = uaR * P.Q[e][qR] + prR;
        } else {
            nL = uaL * P.Q[e][qL];
            nR = uaR * P.Q[e][qR];
        }
        const double* f = fm + e * mv;
        const double mR = f[mC + mst], mL = f[mC - mst], mM = f[mC];
        out[e * fv + loc] = dt * (double(1) / double(30) * (mR + mL) - double(3) / double(10) * (nR + nL) +
                                  double(23) / double(15) * mM);
    }
}

template <int D, class W>
__global__ __launch_bounds__(kBlock) void k_faceflux_single(W w, KParams k) {
    const int64_t gid = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    const DevPatch* P;
    int ax;
    int64_t loc;
    if (!w.find(gid, P, ax, loc)) return;
    if (ax == 0) {
        faceflux_single_face<D, 0>(*P, loc, k);
    } else if constexpr (D > 1) {
        if (ax == 1) {
            faceflux_single_face<D, 1>(*P, loc, k);
        } else if constexpr (D > 2) {
            faceflux_single_face<D, 2>(*P, loc, k);
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
    if constexpr (D > 1) {
        if (plan.items[1] > 0)
            hipLaunchKernelGGL((k_sensor<D, 1, W>), dim3(unsigned((plan.items[1] + kBlock - 1) / kBlock)),
                               dim3(kBlock), 0, s, plan.work[1], k);
    }
    // Stage 2: fused flux on axis 0, midpoint fluxes on axes 1 and 2; stage 3:
    // the face flux of axes 1 and 2.
    bool rec = false;
    if constexpr (std::is_same_v<W, SegTable>) rec = k.wrec != nullptr;  // recording runs use patch tables
    if (plan.items[2] > 0) {
        const dim3 grid(unsigned((plan.items[2] + kBlock - 1) / kBlock));
        if (rec) {
            if constexpr (std::is_same_v<W, SegTable>)
                hipLaunchKernelGGL((k_flux_single<D, SCHEME, PP, QQ, W, true>), grid, dim3(kBlock), 0, s, plan.work[2],
                                   k);
        } else {
            hipLaunchKernelGGL((k_flux_single<D, SCHEME, PP, QQ, W>), grid, dim3(kBlock), 0, s, plan.work[2], k);
        }
    }
    if (plan.items[3] > 0)
        hipLaunchKernelGGL((k_faceflux_single<D, W>), dim3(unsigned((plan.items[3] + kBlock - 1) / kBlock)),
                           dim3(kBlock), 0, s, plan.work[3], k);
    return hipGetLastError();
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
