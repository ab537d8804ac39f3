// hm_multi.hpp -- convective-flux RHS kernels of the multi-species models,
// templated on the model (FOUR = four-equation conservative, else five-equation
// Allaire) and instantiated by hm_kernels_four.hip / hm_kernels_five.hip.
//
// Five-equation model (FlowModelFiveEqnAllaire.cpp:16-113): partial densities
// Z_i rho_i (NS), momentum (D), total energy, volume fractions Z_i (NS-1,
// ADVECTIVE).  The caller hands over NEQ+1 conservative arrays: the last one is
// the last species' volume fraction, stored but not evolved
// (WCNS56-HLLC-HLL.cpp:223-239), which the cell EOS reads.
//
// Four-equation model (FlowModelFourEqnConservative.cpp:16-113): partial
// densities rho Y_i (NS), momentum (D), total energy, all CONSERVATIVE; the
// mixture EOS mixes c_p and c_v with mass fractions
// (EquationOfStateMixingRulesIdealGas.cpp:4677-4737).
//
// Scratch prim holds V = (partial densities, u.., p[, Z_i (NS-1)]), then c,
// then the mixture density rho, over the 4-ghost box.
//
// Pipeline: k_derived_multi (with the dilatation and vorticity on the 2-ghost
// box), k_flux_multi (midpoint and face flux fused over tiles; five-equation:
// also the HLLC normal velocity u* of every midpoint face), k_source_five.
// The persistent schedule runs the midpoint flux from a work queue
// (k_midflux_multi_queue, fluxes to memory) and then k_faceflux_multi.
#pragma once

#include <algorithm>

#include "hm_kernels.hpp"
#include "hm_numerics.hpp"
#include "hm_sensor.hpp"

// Waves per SIMD the multi-species midpoint-flux kernels are compiled for, in
// 3D and in 1D/2D (measured on the uneven 3D five-equation level: 4 waves with
// 36 B of spill beat 3 waves without by 2.5 %; the 2D kernel would spill 70 B).
#ifndef HM_MIDFLUX_MULTI_WAVES_3D
#define HM_MIDFLUX_MULTI_WAVES_3D 4
#endif
#ifndef HM_MIDFLUX_MULTI_WAVES
#define HM_MIDFLUX_MULTI_WAVES 3
#endif

namespace hm {

// Equation layout of a multi-species model.
template <int D, int NS, bool FOUR>
struct Multi {
    static constexpr int NEQ = FOUR ? NS + D + 1 : D + 2 * NS;
    static constexpr int NZ = FOUR ? 0 : NS - 1;  // evolved volume fractions
    static constexpr int IP = NS + D;             // pressure slot of V
    static constexpr int IZ = NS + D + 1;         // first volume-fraction slot of V
};

namespace {

// Mixture ratio of specific heats and c_v from NS mass fractions
// (EquationOfStateMixingRulesIdealGas.cpp:4677-4737, 6946-7150):
// c_p = sum Y_i c_p,i, c_v = sum Y_i c_v,i, gamma = c_p / c_v.
template <int NS>
__device__ __forceinline__ void mix_mass(const KParams& k, const double* Y, double& gamma, double& c_v) {
    double c_p = double(0);
    c_v = double(0);
#pragma unroll
    for (int si = 0; si < NS; si++) {
        c_p += Y[si] * k.cp[si];
        c_v += Y[si] * k.cv[si];
    }
    gamma = div_rn(c_p, c_v);
}

// Four-equation sound speed squared: Gamma p / rho + sum Y_i Psi_i with
// Gamma = gamma - 1 (EquationOfStateIdealGas.cpp:8059), eps = p / ((gamma - 1) rho)
// (:6291) and Psi_i = ((c_p,i - gamma c_v,i) / c_v + gamma - 1) eps
// (EquationOfStateMixingRulesIdealGas.cpp:6367-6564); also returns eps.
template <int NS>
__device__ __forceinline__ double sound_speed_sq_four(const KParams& k, double rho, double p, const double* Y,
                                                      double& eps) {
    double gamma, c_v;
    mix_mass<NS>(k, Y, gamma, c_v);
    const double Gamma = gamma - double(1);
    eps = div_rn(p, (gamma - double(1)) * rho);
    double c2 = div_rn(Gamma * p, rho);
    const double rcv = rcp_rn(c_v);
#pragma unroll
    for (int si = 0; si < NS; si++) {
        const double Psi = (quot(k.cp[si] - gamma * k.cv[si], c_v, rcv) + gamma - double(1)) * eps;
        c2 += Y[si] * Psi;
    }
    return c2;
}

// Mixture ratio of specific heats from all NS volume fractions
// (EquationOfStateMixingRulesIdealGas.cpp:7404-7598): xi = sum Z_i * (1/(gamma_i-1)).
template <int NS>
__device__ __forceinline__ double mix_gamma_all(const KParams& k, const double* Z) {
    double xi = double(0);
#pragma unroll
    for (int si = 0; si < NS; si++) {
        const double one_over_denominator = k.igm1[si];  // double(1) / (gamma_i - 1)
        xi += Z[si] * one_over_denominator;
    }
    return div_rn(double(1), xi) + double(1);
}

// Mixture ratio from NS-1 volume fractions, the last one being 1 - sum
// (EquationOfStateMixingRulesIdealGas.cpp:7599-7800).
template <int NS>
__device__ __forceinline__ double mix_gamma_last(const KParams& k, const double* Z) {
    double xi = double(0);
    double Z_last = double(1);
#pragma unroll
    for (int si = 0; si < NS - 1; si++) {
        const double one_over_denominator = k.igm1[si];  // double(1) / (gamma_i - 1)
        xi += Z[si] * one_over_denominator;
        Z_last -= Z[si];
    }
    xi += quot(Z_last, k.gamma[NS - 1] - double(1), k.igm1[NS - 1]);
    return div_rn(double(1), xi) + double(1);
}

}  // namespace

// Primitive state of one cell and its sound speed squared from its
// conservative variables (q(e, at): variable e of the cell at index `at`),
// exactly as the derived-data kernel computes them; shared with the stable-dt
// kernel.  Zr: partial densities; Z: the NS volume fractions (five-equation).
template <int D, int NS, bool FOUR, class QF>
__device__ __forceinline__ void cell_state_multi(const KParams& k, QF&& q, uint32_t iq, double (&Zr)[NS],
                                                 double (&Z)[NS], double& rho, double (&u)[3], double& p,
                                                 double& c2) {
#pragma unroll
    for (int si = 0; si < NS; si++) Zr[si] = q(si, iq);
    if constexpr (!FOUR) {
#pragma unroll
        for (int si = 0; si < NS; si++) Z[si] = q(NS + D + 1 + si, iq);
    }
    rho = double(0);  // EquationOfStateMixingRules.cpp:735-860
#pragma unroll
    for (int si = 0; si < NS; si++) rho += Zr[si];
    u[0] = u[1] = u[2] = 0.0;
    const double rrho = rcp_rn(rho);  // quotients by rho, correctly rounded (quot)
#pragma unroll
    for (int b = 0; b < D; b++) u[b] = quot(q(NS + b, iq), rho, rrho);
    double ke;
    if constexpr (D == 1)
        ke = u[0] * u[0];
    else if constexpr (D == 2)
        ke = u[0] * u[0] + u[1] * u[1];
    else
        ke = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    const double E = q(NS + D, iq);
    const double eps_int = quot(E, rho, rrho) - double(1) / double(2) * ke;
    if constexpr (FOUR) {
        double Y[NS];
#pragma unroll
        for (int si = 0; si < NS; si++) Y[si] = quot(Zr[si], rho, rrho);
        double gmix, c_v, eps_psi;
        mix_mass<NS>(k, Y, gmix, c_v);
        p = (gmix - double(1)) * rho * eps_int;  // EquationOfStateIdealGas.cpp:5633
        c2 = sound_speed_sq_four<NS>(k, rho, p, Y, eps_psi);
    } else {
        const double gmix = mix_gamma_all<NS>(k, Z);
        p = (gmix - double(1)) * rho * eps_int;  // EquationOfStateIdealGas.cpp:5633
        const double Gamma = gmix - double(1);
        c2 = quot(Gamma * p, rho, rrho);  // FlowModelFiveEqnAllaire.cpp:4600-4640
#pragma unroll
        for (int si = 0; si < NS; si++) {
            const double Psi = quot(p, rho, rrho);  // EquationOfStateMixingRulesIdealGas.cpp:6608
            c2 += quot(Zr[si], rho, rrho) * Psi;
        }
    }
}

// ----------------------------------------------------------------------------
// Derived cell data (FlowModelFiveEqnAllaire.cpp:3734-4876,
// FlowModelFourEqnConservative.cpp:3877-5440).
// ----------------------------------------------------------------------------
template <int D, int NS, bool FOUR, class W>
__global__ __launch_bounds__(kBlock) void k_derived_multi(W w, KParams k) {
    const int64_t gid = int64_t(rhs_block(k)) * kBlock + threadIdx.x;
    const DevPatch* P;
    int ax;
    int64_t loc;
    if (!w.find(gid, P, ax, loc)) return;
    using M = Multi<D, NS, FOUR>;
    constexpr int NEQ = M::NEQ;
    const Boxes<D> B(*P);
    const int e0 = B.ext(0, kGc), e1 = B.ext(1, kGc);
    // 32-bit index arithmetic (a work segment has fewer than 2^31 items).
    const uint32_t r = uint32_t(loc) / uint32_t(e0);
    const int i = int(uint32_t(loc) - r * uint32_t(e0)) - kGc;
    const int j = (D > 1) ? int((D > 2) ? r % uint32_t(e1) : r) - kGc : 0;
    const int kk = (D > 2) ? int(r / uint32_t(e1)) - kGc : 0;
    // Q components at 32-bit offsets within the patch's Q box (below 2^29 cells)
    const uint32_t iq = uint32_t(q_idx<D>(*P, i, j, kk));
    auto q = [&](int e, uint32_t at) { return ldg(P->Q[e], at); };

    double Zr[NS], Z[NS], rho, u[3], p, c2;
    cell_state_multi<D, NS, FOUR>(k, q, iq, Zr, Z, rho, u, p, c2);

    const int64_t vol = B.vol(kGc);
    double* pr = P->prim;
    const uint32_t l32 = uint32_t(loc);
#pragma unroll
    for (int si = 0; si < NS; si++) stg(pr + si * vol, l32, Zr[si]);
#pragma unroll
    for (int b = 0; b < D; b++) stg(pr + (NS + b) * vol, l32, u[b]);
    stg(pr + (NS + D) * vol, l32, p);
#pragma unroll
    for (int si = 0; si < M::NZ; si++) stg(pr + (NS + D + 1 + si) * vol, l32, Z[si]);
    stg(pr + NEQ * vol, l32, sqrt(c2));
    stg(pr + (NEQ + 1) * vol, l32, rho);

    if constexpr (D > 1) {
        // dilatation and vorticity on the 2-ghost box, from the neighbours'
        // velocities computed as above
        if (!in_sensor_box<D>(B, i, j, kk)) return;
        const uint32_t qs[3] = {1u, uint32_t(P->n[0] + 2 * P->gq[0]),
                                uint32_t((P->n[0] + 2 * P->gq[0]) * (P->n[1] + 2 * P->gq[1]))};
        sensor_cell<D>(*P, B, i, j, kk, [&](int a, int side, double (&v)[3]) {
            const uint32_t at = side > 0 ? iq + qs[a] : iq - qs[a];
            double rn = double(0);
#pragma unroll
            for (int si = 0; si < NS; si++) rn += q(si, at);
            const double rrn = rcp_rn(rn);
#pragma unroll
            for (int b = 0; b < D; b++) v[b] = quot(q(NS + b, at), rn, rrn);
        });
    }
}

// ----------------------------------------------------------------------------
// Fused midpoint flux of one face (multi-species models).
// ----------------------------------------------------------------------------
// Bounds of one side of a face normal to axis A.
template <int D, int NS, bool FOUR, int A>
__device__ __forceinline__ bool bounded_multi(const double* V, const KParams& k) {
    if constexpr (FOUR) {
        // FlowModelBasicUtilitiesFourEqnConservative.cpp:3737-4511 (bounds in the .hpp:24-25):
        // mass fractions in (-0.001, 1.001), rho > 0, p > 0.
        double rho = double(0);
#pragma unroll
        for (int si = 0; si < NS; si++) rho += V[si];
        bool ok = true;
        const double rrho = rcp_rn(rho);
#pragma unroll
        for (int si = 0; si < NS; si++) {
            const double Y = quot(V[si], rho, rrho);
            ok = ok && (Y > double(-0.001) && Y < double(1.001));
        }
        return ok && (rho > double(0)) && (V[NS + D] > double(0));
    }
    // FlowModelBasicUtilitiesFiveEqnAllaire.cpp:5446-7347 (bounds in the .hpp:24-27).
    // The sound-speed test differs by axis there: x-faces add Y_i Psi_i to
    // c^2 = Gamma p / rho (:5819, 6091, 6678), y- and z-faces assign
    // c^2 = Gamma p / rho again in the species loop (:6359, 6992, 7305).
    bool ok = true;
    double Z[NS];
    Z[NS - 1] = double(1);
#pragma unroll
    for (int si = 0; si < NS - 1; si++) {
        Z[si] = V[NS + D + 1 + si];
        Z[NS - 1] -= Z[si];
        ok = ok && (Z[si] > double(-1000.0) && Z[si] < double(1000.0));
    }
    ok = ok && (Z[NS - 1] > double(-1000.0) && Z[NS - 1] < double(1000.0));
    double rho = double(0);
#pragma unroll
    for (int si = 0; si < NS; si++) rho += V[si];
    double Y[NS];
    const double rrho = rcp_rn(rho);
#pragma unroll
    for (int si = 0; si < NS; si++) {
        Y[si] = quot(V[si], rho, rrho);
        ok = ok && (Y[si] > double(-0.001) && Y[si] < double(1.001));
    }
#pragma unroll
    for (int si = 0; si < NS; si++) ok = ok && (V[si] > double(0));
    const double p = V[NS + D];
    const double Gamma = mix_gamma_all<NS>(k, Z) - double(1);
    double c2 = quot(Gamma * p, rho, rrho);
    if constexpr (A == 0) {
#pragma unroll
        for (int si = 0; si < NS; si++) {
            const double Psi = quot(p, rho, rrho);
            c2 += Y[si] * Psi;
        }
    }
    return ok && (c2 > double(0));
}

// Midpoint flux Fo of the face normal to axis A between cells (c0, c1, c2)
// - e_A and (c0, c1, c2), and its HLLC normal velocity u* (five-equation).
// The prim and sensor arrays are read at 32-bit offsets from per-component
// bases (every patch's ghost box is below 2^29 cells).
template <int D, int NS, bool FOUR, int A, int SCHEME, int PP, int QQ>
__device__ __forceinline__ void midflux_multi_at(const DevPatch& P, const Boxes<D>& B, int c0, int c1, int c2,
                                                 const KParams& k, double (&Fo)[Multi<D, NS, FOUR>::NEQ],
                                                 double& ustar, double* wl) {
    using M = Multi<D, NS, FOUR>;
    constexpr int NEQ = M::NEQ;
    constexpr int NZ = M::NZ;
    constexpr int IU = NS + A;   // normal velocity slot
    constexpr int IP = NS + D;   // pressure slot
    constexpr int IZ = NS + D + 1;

    const int64_t v4 = B.vol(kGc);
    const int64_t st = (A == 0) ? 1 : (A == 1) ? int64_t(B.ext(0, kGc)) : int64_t(B.ext(0, kGc)) * B.ext(1, kGc);
    const int64_t iR = B.idx(kGc, c0, c1, c2);
    const double* pr = P.prim;
    auto ld = [&](int comp, int m) { return pr[comp * v4 + iR + (m - 3) * st]; };

    // Projection variables (FlowModelBasicUtilitiesFiveEqnAllaire.cpp:7474-8207).
    double Zra[NS];
#pragma unroll
    for (int si = 0; si < NS; si++) Zra[si] = double(1) / double(2) * (ld(si, 2) + ld(si, 3));
    const double ra = double(1) / double(2) * (ld(NEQ + 1, 2) + ld(NEQ + 1, 3));
    const double ca = double(1) / double(2) * (ld(NEQ, 2) + ld(NEQ, 3));

    // Characteristic fields (FlowModelBasicUtilitiesFiveEqnAllaire.cpp:8232-9178),
    // one at a time: a runtime loop keeps a single field's WENO interpolation
    // live, so the kernel fits the register file with room for more waves.
    // Each field's two reconstructed values go to this thread's LDS slots wl[f],
    // wl[NEQ + f] (stride kBlock): the field index is wave-uniform, so the store
    // is one address add, where storing into a register array by a runtime index
    // would select over every slot.
    const LDConst ldk{k.p, k.q, k.C, k.alpha_tau};
#pragma unroll 1
    for (int f = 0; f < NEQ; f++) {
        Six s;
        if (f == 0) {
#pragma unroll
            for (int m = 0; m < 6; m++) s.v[m] = ld(IU, m) - div_rn(double(1), ra * ca) * ld(IP, m);
        } else if (f <= NS) {
            const int si = f - 1;
            double z = Zra[0];
#pragma unroll
            for (int q = 1; q < NS; q++)
                if (q == si) z = Zra[q];
#pragma unroll
            for (int m = 0; m < 6; m++) s.v[m] = ld(si, m) - div_rn(z, ra * ca * ca) * ld(IP, m);
        } else if (f < NS + D) {
            const int t = f - NS - 1;             // transverse slot
            const int b = (t < A) ? t : t + 1;    // its axis
#pragma unroll
            for (int m = 0; m < 6; m++) s.v[m] = ld(NS + b, m);
        } else if (f < NEQ - 1) {
#pragma unroll
            for (int m = 0; m < 6; m++) s.v[m] = ld(IZ + (f - NS - D), m);
        } else {
#pragma unroll
            for (int m = 0; m < 6; m++) s.v[m] = ld(IU, m) + div_rn(double(1), ra * ca) * ld(IP, m);
        }
        double wm, wp;
        reconstruct_pair<SCHEME, PP, QQ>(s, ldk, wm, wp);
        wl[f * kBlock] = wm;
        wl[(NEQ + f) * kBlock] = wp;
    }
    double Wm[NEQ], Wp[NEQ];
#pragma unroll
    for (int e = 0; e < NEQ; e++) {
        Wm[e] = wl[e * kBlock];
        Wp[e] = wl[(NEQ + e) * kBlock];
    }

    // Back-projection (FlowModelBasicUtilitiesFiveEqnAllaire.cpp:9201-9974).
    double VL[NEQ], VR[NEQ];
    // (1/2 Z_i rho_i) / c: the forward's quotient, shared by both sides; the
    // minus form (-1/2 Z_i rho_i) / c is its exact negative.
    double zc[NS];
    {
        const double rca = rcp_rn(ca);
#pragma unroll
        for (int si = 0; si < NS; si++) zc[si] = quot(double(1) / double(2) * Zra[si], ca, rca);
    }
    auto back = [&](const double* Wc, double* Vo) {
#pragma unroll
        for (int si = 0; si < NS; si++) Vo[si] = -zc[si] * Wc[0] + Wc[si + 1] + zc[si] * Wc[NEQ - 1];
        Vo[IU] = double(1) / double(2) * Wc[0] + double(1) / double(2) * Wc[NEQ - 1];
        int t = NS + 1;
#pragma unroll
        for (int b = 0; b < D; b++) {
            if (b == A) continue;
            Vo[NS + b] = Wc[t++];
        }
        Vo[IP] = -double(1) / double(2) * ra * ca * Wc[0] + double(1) / double(2) * ra * ca * Wc[NEQ - 1];
#pragma unroll
        for (int si = 0; si < NZ; si++) Vo[IZ + si] = Wc[NS + D + si];
    };
    back(Wm, VL);
    back(Wp, VR);

    // Bounds on both sides, else first-order values (WCNS56-HLLC-HLL.cpp:947-1032).
    if (!(bounded_multi<D, NS, FOUR, A>(VL, k) && bounded_multi<D, NS, FOUR, A>(VR, k))) {
#pragma unroll
        for (int e = 0; e < NEQ; e++) {
            VL[e] = ld(e, 2);
            VR[e] = ld(e, 3);
        }
    }

    // Face EOS (HLLC drivers, FlowModelRiemannSolverFiveEqnAllaireHLLC.cpp:4824-7716).
    auto face_eos = [&](const double* V, double& rho, double& c, double& eps_int) {
        rho = double(0);
#pragma unroll
        for (int si = 0; si < NS; si++) rho += V[si];
        if constexpr (FOUR) {
            // FlowModelRiemannSolverFourEqnConservativeHLLC.cpp:4345-4700
            double Y[NS];
            const double rrho = rcp_rn(rho);
#pragma unroll
            for (int si = 0; si < NS; si++) Y[si] = quot(V[si], rho, rrho);
            c = sqrt(sound_speed_sq_four<NS>(k, rho, V[IP], Y, eps_int));
            return;
        }
        const double gmix = mix_gamma_last<NS>(k, V + IZ);
        const double p = V[IP];
        const double Gamma = gmix - double(1);
        const double rrho = rcp_rn(rho);
        double c2 = quot(Gamma * p, rho, rrho);
#pragma unroll
        for (int si = 0; si < NS; si++) {
            const double Psi = quot(p, rho, rrho);
            c2 += quot(V[si], rho, rrho) * Psi;
        }
        c = sqrt(c2);
        eps_int = div_rn(p, (gmix - double(1)) * rho);
    };
    double rhoL, rhoR, cL, cR, epsL, epsR;
    face_eos(VL, rhoL, cL, epsL);
    face_eos(VR, rhoR, cR, epsR);
    const double uL = VL[IU], uR = VR[IU], pL = VL[IP], pR = VR[IP];
    const double u_avg = double(1) / double(2) * (uL + uR);
    const double c_avg = double(1) / double(2) * (cL + cR);
    const double sL = fmin(u_avg - c_avg, uL - cL);
    const double sR = fmax(u_avg + c_avg, uR + cR);
    const double s_minus = fmin(double(0), sL);
    const double s_plus = fmax(double(0), sR);
    const double s_star =
        div_rn(pR - pL + rhoL * uL * (sL - uL) - rhoR * uR * (sR - uR), rhoL * (sL - uL) - rhoR * (sR - uR));

    // Sensor (WCNS56-HLLC-HLL.cpp:1066-1113).
    bool use_hll = false;
    if constexpr (D > 1) {
        const int64_t vs = B.vol(kGs);
        const uint32_t tR = uint32_t(B.idx(kGs, c0, c1, c2));
        const uint32_t sts = (A == 0) ? 1u : (A == 1) ? uint32_t(B.ext(0, kGs)) : uint32_t(B.ext(0, kGs) * B.ext(1, kGs));
        const uint32_t tL = tR - sts;
        const double th_avg = 0.5 * (ldg(P.sens, tL) + ldg(P.sens, tR));
        const double om_avg = 0.5 * (ldg(P.sens + vs, tL) + ldg(P.sens + vs, tR));
        const double sensor = div_rn(-th_avg, fabs(th_avg) + om_avg + kEps);
        use_hll = sensor > 0.65;
    }

    // Conservative state and physical flux of one side
    // (FlowModelRiemannSolverFiveEqnAllaireHLLC.cpp:883-921).
    auto cons_flux = [&](const double* V, double rho, double eps_int, double* Qs, double* Fs) {
#pragma unroll
        for (int si = 0; si < NS; si++) Qs[si] = V[si];
#pragma unroll
        for (int b = 0; b < D; b++) Qs[NS + b] = rho * V[NS + b];
        double ke;
        if constexpr (D == 1)
            ke = V[NS] * V[NS];
        else if constexpr (D == 2)
            ke = V[NS] * V[NS] + V[NS + 1] * V[NS + 1];
        else
            ke = V[NS] * V[NS] + V[NS + 1] * V[NS + 1] + V[NS + 2] * V[NS + 2];
        Qs[IP] = rho * (eps_int + double(1) / double(2) * ke);
#pragma unroll
        for (int si = 0; si < NZ; si++) Qs[IZ + si] = V[IZ + si];
        const double un = V[IU];
#pragma unroll
        for (int si = 0; si < NS; si++) Fs[si] = un * V[si];
#pragma unroll
        for (int b = 0; b < D; b++) Fs[NS + b] = (b == A) ? un * Qs[NS + b] + V[IP] : un * Qs[NS + b];
        Fs[IP] = un * (Qs[IP] + V[IP]);
#pragma unroll
        for (int si = 0; si < NZ; si++) Fs[IZ + si] = un * V[IZ + si];
    };

    const bool left = s_star > double(0);
    double QL[NEQ], QR[NEQ], FL[NEQ], FR[NEQ];
    if (use_hll) {
        cons_flux(VL, rhoL, epsL, QL, FL);
        cons_flux(VR, rhoR, epsR, QR, FR);
    } else if (left) {
        cons_flux(VL, rhoL, epsL, QL, FL);
    } else {
        cons_flux(VR, rhoR, epsR, QR, FR);
    }
    {
        const double* Vs = left ? VL : VR;
        const double* Qs = left ? QL : QR;
        const double* Fs = left ? FL : FR;
        const double rho = left ? rhoL : rhoR;
        const double sK = left ? sL : sR;
        const double spm = left ? s_minus : s_plus;
        const double chi = div_rn(sK - Vs[IU], sK - s_star);
        double Qst[NEQ];
#pragma unroll
        for (int si = 0; si < NS; si++) Qst[si] = chi * Vs[si];
#pragma unroll
        for (int b = 0; b < D; b++) Qst[NS + b] = (b == A) ? chi * rho * s_star : chi * Qs[NS + b];
        Qst[IP] = chi * (Qs[IP] + (s_star - Vs[IU]) * (rho * s_star + div_rn(Vs[IP], sK - Vs[IU])));
#pragma unroll
        for (int si = 0; si < NZ; si++) Qst[IZ + si] = chi * Vs[IZ + si];
#pragma unroll
        for (int e = 0; e < NEQ; e++) Fo[e] = Fs[e] + spm * (Qst[e] - Qs[e]);
        ustar = Vs[IU] + spm * (chi - double(1));  // FlowModelRiemannSolverFiveEqnAllaireHLLC.cpp:5300-5310
    }
    if constexpr (D > 1) {
        if (use_hll) {
            // HLL on partial densities, transverse momentum and volume fractions
            // (FlowModelRiemannSolverFiveEqnAllaireHLLC-HLL.cpp:1283-1504).
            const double rdsl = rcp_rn(sR - sL);
            auto hll = [&](int e) {
                double h = quot(sR * FL[e] - sL * FR[e] + sR * sL * (QR[e] - QL[e]), sR - sL, rdsl);
                if (sL > double(0)) h = FL[e];
                if (sR < double(0)) h = FR[e];
                return h;
            };
            double dsq;
            {
                const double du = VR[NS] - VL[NS];
                const double dv = VR[NS + 1] - VL[NS + 1];
                if constexpr (D == 2) {
                    dsq = du * du + dv * dv;
                } else {
                    const double dw = VR[NS + 2] - VL[NS + 2];
                    dsq = du * du + dv * dv + dw * dw;
                }
            }
            const double vel_mag = sqrt(dsq);
            double al1, al2;
            if (vel_mag < kEps) {
                al1 = double(1);
                al2 = double(0);
            } else {
                al1 = div_rn(fabs(VR[IU] - VL[IU]), vel_mag);
                al2 = sqrt(double(1) - al1 * al1);
            }
            const double be1 = double(1) / double(2) * (double(1) + div_rn(al1, al1 + al2));
            const double be2 = double(1) - be1;
#pragma unroll
            for (int si = 0; si < NS; si++) Fo[si] = be1 * Fo[si] + be2 * hll(si);
#pragma unroll
            for (int b = 0; b < D; b++)
                if (b != A) Fo[NS + b] = be1 * Fo[NS + b] + be2 * hll(NS + b);
#pragma unroll
            for (int si = 0; si < NZ; si++) Fo[IZ + si] = be1 * Fo[IZ + si] + be2 * hll(IZ + si);
        }
    }

}

// Midpoint face `loc` (numbered with x fastest over faces -1..n_A+1 along A)
// of the unfused pipeline: its flux to fmid[A], u* to ustar[A].
template <int D, int NS, bool FOUR, int A, int SCHEME, int PP, int QQ>
__device__ __forceinline__ void midflux_multi_face(const DevPatch& P, int64_t loc, const KParams& k, double* wl) {
    constexpr int NEQ = Multi<D, NS, FOUR>::NEQ;
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
    double Fo[NEQ];
    double ustar;
    midflux_multi_at<D, NS, FOUR, A, SCHEME, PP, QQ>(P, B, c0, c1, c2, k, Fo, ustar, wl);
    const int64_t mv = B.mid_vol(A);
    double* out = P.fmid[A];
    const uint32_t l32 = uint32_t(loc);
#pragma unroll
    for (int e = 0; e < NEQ; e++) stg(out + e * mv, l32, Fo[e]);
    if constexpr (!FOUR) stg(P.ustar[A], l32, ustar);
}

// The unfused midpoint-flux stage as a persistent kernel (Schedule::persistent):
// a grid of as many blocks as the device holds at once takes the stage's
// blocks of 256 midpoint faces in order from a queue (KParams::queue, one
// atomic fetch per block of work, cleared by k_queue_reset just before the
// launch), and every block leaves once its fetch is past the last block of
// work; the midpoint fluxes go to memory for k_faceflux_multi.
template <int D, int NS, bool FOUR, int SCHEME, int PP, int QQ, class W>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(HM_MIDFLUX_MULTI_WAVES))) void
k_midflux_multi_queue(W w, KParams k, uint32_t nwork) {
    __shared__ double wlds[2 * Multi<D, NS, FOUR>::NEQ * kBlock];
    __shared__ uint32_t ticket;
    double* wl = wlds + threadIdx.x;
    for (;;) {
        if (threadIdx.x == 0) ticket = atomicAdd(k.queue, 1u);
        __syncthreads();
        const uint32_t b = ticket;
        __syncthreads();  // every thread has read the ticket before the next fetch
        if (b >= nwork) return;  // the same for the whole block
        const int64_t gid = int64_t(b) * kBlock + threadIdx.x;
        const DevPatch* P;
        int ax;
        int64_t loc;
        if (w.find(gid, P, ax, loc)) {
            if (ax == 0) {
                midflux_multi_face<D, NS, FOUR, 0, SCHEME, PP, QQ>(*P, loc, k, wl);
            } else if constexpr (D > 1) {
                if (ax == 1) {
                    midflux_multi_face<D, NS, FOUR, 1, SCHEME, PP, QQ>(*P, loc, k, wl);
                } else if constexpr (D > 2) {
                    midflux_multi_face<D, NS, FOUR, 2, SCHEME, PP, QQ>(*P, loc, k, wl);
                }
            }
        }
    }
}

namespace {
// Clears the work-queue counter (a kernel, so that a captured step holds no memset).
__global__ void k_queue_reset(unsigned* queue) {
    if (threadIdx.x == 0) *queue = 0u;
}
}  // namespace

// ----------------------------------------------------------------------------
// Midpoint-to-face flux with the model's node flux
// (FlowModelFiveEqnAllaire.cpp:4877-5901, FlowModelFourEqnConservative.cpp:5499-6338).
// ----------------------------------------------------------------------------
// Node flux of equation e at a cell (FlowModelFiveEqnAllaire.cpp:4877-5101,
// FlowModelFourEqnConservative.cpp:5499-5700) along A, from its conservative
// variables at `q` in the Q box (below 2^29 cells) and its velocity ua along A
// and pressure p.
template <int D, int NS, int A>
__device__ __forceinline__ double node_flux_multi(const DevPatch& P, uint32_t q, double ua, double p, int e) {
    constexpr int IP = NS + D;
    if (e == IP) return ua * (ldg(P.Q[IP], q) + p);
    if (e == NS + A) return ua * ldg(P.Q[e], q) + p;
    // partial densities, transverse momentum and volume fractions: u_a * Q
    return ua * ldg(P.Q[e], q);
}

// Output flux of face (c0, c1, c2) along A from the midpoint fluxes of faces
// f - 1, f, f + 1 (m(side, e): side -1, 0, +1) and the node fluxes of cells
// f - 1 and f (nf(side, e): side -1, 0): dt (1/30 (m_{f+1} + m_{f-1}) -
// 3/10 (n_f + n_{f-1}) + 23/15 m_f) (WCNS56-HLLC-HLL.cpp:461-485, 1200-1282,
// 2327-2490).
template <int D, int NS, bool FOUR, int A, class MF, class NF>
__device__ __forceinline__ void combine_face_multi(const DevPatch& P, const Boxes<D>& B, int c0, int c1, int c2,
                                                   double dt, MF&& m, NF&& nf) {
    constexpr int NEQ = Multi<D, NS, FOUR>::NEQ;
    const int64_t fv = B.face_vol(A);
    const uint32_t fi = uint32_t(B.face_idx(A, c0, c1, c2));
    double* out = P.F[A];
#pragma unroll
    for (int e = 0; e < NEQ; e++) {
        const double nL = nf(-1, e), nR = nf(0, e);
        const double mR = m(1, e), mL = m(-1, e), mM = m(0, e);
        stg(out + e * fv, fi,
            dt * (double(1) / double(30) * (mR + mL) - double(3) / double(10) * (nR + nL) +
                  double(23) / double(15) * mM));
    }
}

// Node fluxes of the cell (c0, c1, c2) along A, nf[e] for every equation.
template <int D, int NS, bool FOUR, int A>
__device__ __forceinline__ void cell_node_flux_multi(const DevPatch& P, const Boxes<D>& B, int c0, int c1, int c2,
                                                     double* nf, int stride) {
    constexpr int NEQ = Multi<D, NS, FOUR>::NEQ;
    constexpr int IP = NS + D;
    const int64_t v4 = B.vol(kGc);
    const uint32_t pc = uint32_t(B.idx(kGc, c0, c1, c2));
    const uint32_t q = uint32_t(q_idx<D>(P, c0, c1, c2));
    const double ua = ldg(P.prim + (NS + A) * v4, pc), pr = ldg(P.prim + IP * v4, pc);
#pragma unroll
    for (int e = 0; e < NEQ; e++) nf[e * stride] = node_flux_multi<D, NS, A>(P, q, ua, pr, e);
}

// Output flux of face (c0, c1, c2) along A with the node fluxes formed here.
template <int D, int NS, bool FOUR, int A, class MF>
__device__ __forceinline__ void store_face_multi(const DevPatch& P, const Boxes<D>& B, int c0, int c1, int c2,
                                                 double dt, MF&& m) {
    constexpr int IP = NS + D;
    const int64_t v4 = B.vol(kGc);
    const uint32_t st4 = (A == 0) ? 1u : (A == 1) ? uint32_t(B.ext(0, kGc)) : uint32_t(B.ext(0, kGc) * B.ext(1, kGc));
    const uint32_t pR = uint32_t(B.idx(kGc, c0, c1, c2)), pL = pR - st4;
    const uint32_t qst = (A == 0) ? 1u
                         : (A == 1) ? uint32_t(P.n[0] + 2 * P.gq[0])
                                    : uint32_t((P.n[0] + 2 * P.gq[0]) * (P.n[1] + 2 * P.gq[1]));
    const uint32_t qR = uint32_t(q_idx<D>(P, c0, c1, c2)), qL = qR - qst;
    const double uaL = ldg(P.prim + (NS + A) * v4, pL), uaR = ldg(P.prim + (NS + A) * v4, pR);
    const double prL = ldg(P.prim + IP * v4, pL), prR = ldg(P.prim + IP * v4, pR);
    combine_face_multi<D, NS, FOUR, A>(P, B, c0, c1, c2, dt, m, [&](int side, int e) {
        return side < 0 ? node_flux_multi<D, NS, A>(P, qL, uaL, prL, e) : node_flux_multi<D, NS, A>(P, qR, uaR, prR, e);
    });
}

// Output face `loc` of the unfused pipeline, from the midpoint fluxes in fmid[A].
template <int D, int NS, bool FOUR, int A>
__device__ __forceinline__ void faceflux_multi_face(const DevPatch& P, int64_t loc, const KParams& k) {
    const Boxes<D> B(P);
    const int f0 = B.face_ext(A, 0), f1 = B.face_ext(A, 1);
    // 32-bit index arithmetic (a work segment has fewer than 2^31 items).
    const uint32_t r = uint32_t(loc) / uint32_t(f0);
    const int c0 = int(uint32_t(loc) - r * uint32_t(f0));
    const int c1 = (D > 1) ? int((D > 2) ? r % uint32_t(f1) : r) : 0;
    const int c2 = (D > 2) ? int(r / uint32_t(f1)) : 0;
    const int64_t mv = B.mid_vol(A);
    const uint32_t mC = uint32_t(B.mid_idx(A, c0, c1, c2));
    const uint32_t mst = (A == 0) ? 1u : (A == 1) ? uint32_t(B.mid_ext(A, 0)) : uint32_t(B.mid_ext(A, 0) * B.mid_ext(A, 1));
    const double* fm = P.fmid[A];
    store_face_multi<D, NS, FOUR, A>(P, B, c0, c1, c2, kdt(k),
                                     [&](int side, int e) { return ldg(fm + e * mv, mC + uint32_t(side) * mst); });
}

template <int D, int NS, bool FOUR, class W>
__global__ __launch_bounds__(kBlock) void k_faceflux_multi(W w, KParams k) {
    const int64_t gid = int64_t(rhs_block(k)) * kBlock + threadIdx.x;
    const DevPatch* P;
    int ax;
    int64_t loc;
    if (!w.find(gid, P, ax, loc)) return;
    if (ax == 0) {
        faceflux_multi_face<D, NS, FOUR, 0>(*P, loc, k);
    } else if constexpr (D > 1) {
        if (ax == 1) {
            faceflux_multi_face<D, NS, FOUR, 1>(*P, loc, k);
        } else if constexpr (D > 2) {
            faceflux_multi_face<D, NS, FOUR, 2>(*P, loc, k);
        }
    }
}

// ----------------------------------------------------------------------------
// Fused midpoint and face flux (tiles of mtile_shape): the block's tile `loc >> 8`
// of the patch's axis-A segment, thread `loc & 255` one of its midpoint faces.
// Midpoint face j (0 .. n_A + 2, the face between cells j - 2 and j - 1 along
// A) of column t: tile (tA, tw) = (tile % ntA, tile / ntA) holds faces
// j = tA (R - 2) + r, r < R, of columns t = tw W + w, w < W, thread
// r + w R on axis 0 (consecutive threads along x), w + r W on axes 1 and 2.
// Phase 1 computes the midpoint flux and the node fluxes of the cell right of
// the face (cell j - 1) into the thread's LDS slots (those the reconstruction
// staged its values in); after the barrier, a thread with 1 <= r <= R - 2 and
// j <= n_A + 1 writes the output face j - 1 from its own and its two
// neighbours' midpoint fluxes and its own and its lower neighbour's node fluxes.  Every output face has one such
// thread; the five-equation u* of a midpoint face is written by one tile (its
// first R - 2 faces, and the last tile of a column all of its faces).
static_assert(kBlock == 256, "a fused multi-species tile is one 256-thread block");

template <int D, int NS, bool FOUR, int A, int SCHEME, int PP, int QQ>
__device__ __forceinline__ void flux_multi_tile(const DevPatch& P, int64_t loc, const KParams& k, double* wlds) {
    constexpr int NEQ = Multi<D, NS, FOUR>::NEQ;
    const Boxes<D> B(P);
    const int nA = B.n[A];
    int R, W, ntA;
    mtile_shape(nA, R, W, ntA);
    const uint32_t tile = uint32_t(loc) >> 8, lane = uint32_t(loc) & 255u;
    const uint32_t tA = tile % uint32_t(ntA), tw = tile / uint32_t(ntA);
    uint32_t r, w;
    if constexpr (A == 0) {
        w = lane / uint32_t(R);
        r = lane - w * uint32_t(R);
    } else {
        r = lane / uint32_t(W);
        w = lane - r * uint32_t(W);
    }
    const int j = int(tA) * (R - 2) + int(r);
    // columns: positions across the axis, x fastest over the other axes
    constexpr int U = (A == 0) ? 1 : 0;  // first axis across
    constexpr int V = (A == 2) ? 1 : 2;  // second axis across
    const uint32_t t = tw * uint32_t(W) + w;
    const uint32_t nU = uint32_t(B.n[U]);
    const uint32_t tv = t / nU;
    const bool in = lane < uint32_t(R * W) && j < nA + 3 && tv < uint32_t(B.n[V]);
    int c[3];
    c[A] = j - 1;
    c[U] = int(t - tv * nU);
    c[V] = int(tv);
    double* wl = wlds + threadIdx.x;
    if (in) {
        double Fo[NEQ];
        double ustar;
        midflux_multi_at<D, NS, FOUR, A, SCHEME, PP, QQ>(P, B, c[0], c[1], c[2], k, Fo, ustar, wl);
#pragma unroll
        for (int e = 0; e < NEQ; e++) wl[e * kBlock] = Fo[e];
        // node fluxes of the cell right of the face (the right cell of output face j - 1, the left of face j)
        cell_node_flux_multi<D, NS, FOUR, A>(P, B, c[0], c[1], c[2], wl + NEQ * kBlock, kBlock);
        if constexpr (!FOUR) {
            if (int(r) < R - 2 || int(tA) == ntA - 1)
                stg(P.ustar[A], uint32_t(B.mid_idx(A, c[0], c[1], c[2])), ustar);
        }
    }
    __syncthreads();
    if (in && r >= 1u && int(r) <= R - 2 && j <= nA + 1) {
        const int step = (A == 0) ? 1 : W;  // thread distance of the neighbouring midpoint faces
        const int me = int(threadIdx.x);
        combine_face_multi<D, NS, FOUR, A>(
            P, B, c[0], c[1], c[2], kdt(k), [&](int side, int e) { return wlds[e * kBlock + me + side * step]; },
            [&](int side, int e) { return wlds[(NEQ + e) * kBlock + me + side * step]; });
    }
}

// The fused flux stage: every block is one tile (the stage's segments hold
// whole tiles), so all threads of a block reach the barrier.
template <int D, int NS, bool FOUR, int SCHEME, int PP, int QQ, class W>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(
    D == 3 ? HM_MIDFLUX_MULTI_WAVES_3D : HM_MIDFLUX_MULTI_WAVES))) void k_flux_multi(W w, KParams k) {
    __shared__ double wlds[2 * Multi<D, NS, FOUR>::NEQ * kBlock];
    const int64_t gid = int64_t(rhs_block(k)) * kBlock + threadIdx.x;
    const DevPatch* P;
    int ax;
    int64_t loc;
    if (!w.find(gid, P, ax, loc)) return;  // whole blocks only
    if (ax == 0) {
        flux_multi_tile<D, NS, FOUR, 0, SCHEME, PP, QQ>(*P, loc, k, wlds);
    } else if constexpr (D > 1) {
        if (ax == 1) {
            flux_multi_tile<D, NS, FOUR, 1, SCHEME, PP, QQ>(*P, loc, k, wlds);
        } else if constexpr (D > 2) {
            flux_multi_tile<D, NS, FOUR, 2, SCHEME, PP, QQ>(*P, loc, k, wlds);
        }
    }
}

// ----------------------------------------------------------------------------
// Advective source of the volume-fraction equations (WCNS56-HLLC-HLL.cpp:491-535,
// 1288-1369, 2492-2650).
// ----------------------------------------------------------------------------
template <int D, int NS, class W>
__global__ __launch_bounds__(kBlock) void k_source_five(W w, KParams k) {
    const int64_t gid = int64_t(rhs_block(k)) * kBlock + threadIdx.x;
    const DevPatch* P;
    int ax;
    int64_t loc;
    if (!w.find(gid, P, ax, loc)) return;
    const Boxes<D> B(*P);
    // 32-bit index arithmetic (a work segment has fewer than 2^31 items).
    const uint32_t r = uint32_t(loc) / uint32_t(B.n[0]);
    const int i = int(uint32_t(loc) - r * uint32_t(B.n[0]));
    const int j = (D > 1) ? int((D > 2) ? r % uint32_t(B.n[1]) : r) : 0;
    const int kk = (D > 2) ? int(r / uint32_t(B.n[1])) : 0;
    const int64_t v4 = B.vol(kGc);
    const int64_t pc = B.idx(kGc, i, j, kk);
    double acc = 0.0;
#pragma unroll
    for (int a = 0; a < D; a++) {
        const int64_t st4 = (a == 0) ? 1 : (a == 1) ? int64_t(B.ext(0, kGc)) : int64_t(B.ext(0, kGc)) * B.ext(1, kGc);
        const int64_t mst = (a == 0) ? 1 : (a == 1) ? int64_t(B.mid_ext(a, 0)) : int64_t(B.mid_ext(a, 0)) * B.mid_ext(a, 1);
        const int64_t m = B.mid_idx(a, i, j, kk);  // face between cells i-1 and i along a
        const double* us = P->ustar[a];
        const double uLL = us[m - mst], uL = us[m], uR = us[m + mst], uRR = us[m + 2 * mst];
        const double* u = P->prim + (NS + a) * v4;
        const double cL = u[pc - st4], cR = u[pc + st4];
        const double term = (double(3) / double(2) * (uR - uL) - double(3) / double(10) * (cR - cL) +
                             double(1) / double(30) * (uRR - uLL)) /
                            P->dx[a];
        acc = (a == 0) ? term : acc + term;
    }
    const int64_t svol = int64_t(B.n[0]) * B.n[1] * B.n[2];
    const int64_t iq = q_idx<D>(*P, i, j, kk);
#pragma unroll
    for (int si = 0; si < NS - 1; si++) {
        const int e = NS + D + 1 + si;
        P->S[e * svol + loc] += kdt(k) * P->Q[e][iq] * acc;
    }
}

// ----------------------------------------------------------------------------
// Host-side launch of a multi-species pipeline.
// ----------------------------------------------------------------------------
namespace {

template <class W>
unsigned nblocks(const Plan<W>& plan, int stage) {
    return unsigned((plan.items[stage] + kBlock - 1) / kBlock);
}

template <int D, int NS, bool FOUR, int SCHEME, int PP, int QQ, class W>
hipError_t launch_multi_dsq(const Plan<W>& plan, const KParams& k, hipStream_t s) {
    if (plan.items[0] > 0)
        hipLaunchKernelGGL((k_derived_multi<D, NS, FOUR, W>), dim3(nblocks(plan, 0)), dim3(kBlock), 0, s,
                           plan.work[0], k);
    // (stage 1 is empty: the derived-data kernel computes the dilatation and vorticity)
    if (plan.items[2] > 0) {
        if (plan.persistent) {
            // unfused: midpoint faces from the work queue here, output faces in stage 3
            if (!k.queue) return hipErrorInvalidValue;
            // resident blocks of this kernel on the device (queried once per instance)
            static const unsigned resident = [] {
                int dev = 0, cus = 0, per_cu = 0;
                (void)hipGetDevice(&dev);
                (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
                (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(
                    &per_cu, reinterpret_cast<const void*>(&k_midflux_multi_queue<D, NS, FOUR, SCHEME, PP, QQ, W>),
                    kBlock, 0);
                return unsigned(std::max(1, cus) * std::max(1, per_cu));
            }();
            const unsigned nwork = nblocks(plan, 2);
            hipLaunchKernelGGL(k_queue_reset, dim3(1), dim3(64), 0, s, k.queue);
            hipLaunchKernelGGL((k_midflux_multi_queue<D, NS, FOUR, SCHEME, PP, QQ, W>),
                               dim3(std::min(nwork, resident)), dim3(kBlock), 0, s, plan.work[2], k, uint32_t(nwork));
        } else {
            // fused midpoint and face flux over tiles (stage 3 is then empty)
            hipLaunchKernelGGL((k_flux_multi<D, NS, FOUR, SCHEME, PP, QQ, W>), dim3(nblocks(plan, 2)), dim3(kBlock),
                               0, s, plan.work[2], k);
        }
    }
    if (plan.items[3] > 0)
        hipLaunchKernelGGL((k_faceflux_multi<D, NS, FOUR, W>), dim3(nblocks(plan, 3)), dim3(kBlock), 0, s,
                           plan.work[3], k);
    if constexpr (!FOUR) {
        if (plan.items[4] > 0)
            hipLaunchKernelGGL((k_source_five<D, NS, W>), dim3(nblocks(plan, 4)), dim3(kBlock), 0, s, plan.work[4],
                               k);
    }
    return hipGetLastError();
}

template <int D, int NS, bool FOUR, class W>
hipError_t launch_multi_dn(const Plan<W>& plan, const KParams& k, int scheme, hipStream_t s) {
    const bool dflt = (k.p == 2 && k.q == 4);
    if (scheme == kWCNS6_LD)
        return dflt ? launch_multi_dsq<D, NS, FOUR, kWCNS6_LD, 2, 4>(plan, k, s)
                    : launch_multi_dsq<D, NS, FOUR, kWCNS6_LD, -1, -1>(plan, k, s);
    return (k.p == 2) ? launch_multi_dsq<D, NS, FOUR, kWCNS5_JS, 2, 4>(plan, k, s)
                      : launch_multi_dsq<D, NS, FOUR, kWCNS5_JS, -1, -1>(plan, k, s);
}

}  // namespace

}  // namespace hm
