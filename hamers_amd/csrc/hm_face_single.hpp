// hm_face_single.hpp -- midpoint flux of one face, single-species model
// (host+device): characteristic projection, WCNS reconstruction of both
// sides, back-projection, positivity fallback, HLLC or HLLC-HLL.
//
// V[m] are the primitive variables (rho, u.., p) of the stencil cells i-3+m of
// the face between cells i-1 and i along axis A, cL and cR the sound speeds of
// cells i-1 and i.  References: FlowModelBasicUtilitiesSingleSpecies.cpp:4638-5134
// (projection), 6194-6242 (characteristic), 7279-7323 (back), 3205-3222
// (bounds); WCNS56-HLLC-HLL.cpp:947-1032 (fallback);
// FlowModelRiemannSolverSingleSpeciesHLLC.cpp:519-1081, ...HLLC-HLL.cpp:804-1629.
#pragma once

#include "hm_numerics.hpp"

namespace hm {

// HLLC / HLLC-HLL midpoint flux from the left and right primitive states.
template <int D, int A>
HM_HD void riemann_flux_single(const double* VL, const double* VR, double gam, bool use_hll, double* Fo) {
    constexpr int NEQ = D + 2;
    constexpr int IP = D + 1;
    // Riemann problem at the midpoint.
    const double rhoL = VL[0], rhoR = VR[0];
    const double pL = VL[IP], pR = VR[IP];
    const double uL = VL[1 + A], uR = VR[1 + A];
    // Quotients through div_rn / quot: correctly rounded, as a / b.
    const double csL = sqrt(div_rn(gam * pL, rhoL));             // EquationOfStateIdealGas.cpp:5806
    const double csR = sqrt(div_rn(gam * pR, rhoR));
    const double epsL = div_rn(pL, (gam - double(1)) * rhoL);    // EquationOfStateIdealGas.cpp:6135
    const double epsR = div_rn(pR, (gam - double(1)) * rhoR);

    // Wave speeds and contact speed (FlowModelRiemannSolverSingleSpeciesHLLC.cpp:619-631).
    const double u_avg = double(1) / double(2) * (uL + uR);
    const double c_avg = double(1) / double(2) * (csL + csR);
    const double sL = fmin(u_avg - c_avg, uL - csL);
    const double sR = fmax(u_avg + c_avg, uR + csR);
    const double s_minus = fmin(double(0), sL);
    const double s_plus = fmax(double(0), sR);
    const double s_star =
        div_rn(pR - pL + rhoL * uL * (sL - uL) - rhoR * uR * (sR - uR), rhoL * (sL - uL) - rhoR * (sR - uR));

    // Conservative state and physical flux of one side (:637-648).
    auto cons_flux = [&](const double* Vs, double eps_int, double* Qs, double* Fs) {
        Qs[0] = Vs[0];
#pragma unroll
        for (int b = 0; b < D; b++) Qs[1 + b] = Vs[0] * Vs[1 + b];
        double ke;
        if constexpr (D == 1)
            ke = Vs[1] * Vs[1];
        else if constexpr (D == 2)
            ke = Vs[1] * Vs[1] + Vs[2] * Vs[2];
        else
            ke = Vs[1] * Vs[1] + Vs[2] * Vs[2] + Vs[3] * Vs[3];
        Qs[IP] = Vs[0] * (eps_int + double(1) / double(2) * ke);
        Fs[0] = Qs[1 + A];
#pragma unroll
        for (int b = 0; b < D; b++) Fs[1 + b] = (b == A) ? Qs[1 + A] * Vs[1 + b] + Vs[IP] : Qs[1 + A] * Vs[1 + b];
        Fs[IP] = Vs[1 + A] * (Qs[IP] + Vs[IP]);
    };

    double QL[NEQ], QR[NEQ], FL[NEQ], FR[NEQ];
    const bool left = s_star > double(0);
    if (use_hll) {
        cons_flux(VL, epsL, QL, FL);
        cons_flux(VR, epsR, QR, FR);
    } else if (left) {
        cons_flux(VL, epsL, QL, FL);
    } else {
        cons_flux(VR, epsR, QR, FR);
    }
    {
        // HLLC (:650-688): star state of the upwind side of the contact.
        const double* Vs = left ? VL : VR;
        const double* Qs = left ? QL : QR;
        const double* Fs = left ? FL : FR;
        const double sK = left ? sL : sR;
        const double spm = left ? s_minus : s_plus;
        const double chi = div_rn(sK - Vs[1 + A], sK - s_star);
        double Qst[NEQ];
        Qst[0] = chi * Vs[0];
#pragma unroll
        for (int b = 0; b < D; b++) Qst[1 + b] = (b == A) ? chi * Vs[0] * s_star : chi * Qs[1 + b];
        Qst[IP] = chi * (Qs[IP] + (s_star - Vs[1 + A]) * (Vs[0] * s_star + div_rn(Vs[IP], sK - Vs[1 + A])));
#pragma unroll
        for (int e = 0; e < NEQ; e++) Fo[e] = Fs[e] + spm * (Qst[e] - Qs[e]);
    }
    if constexpr (D > 1) {
        if (use_hll) {
            // HLL on mass and transverse momentum, blended by the velocity-jump
            // angle (FlowModelRiemannSolverSingleSpeciesHLLC-HLL.cpp:961-1022).
            const double rdsl = rcp_rn(sR - sL);
            auto hll = [&](int e) {
                double h = quot(sR * FL[e] - sL * FR[e] + sR * sL * (QR[e] - QL[e]), sR - sL, rdsl);
                if (sL > double(0)) h = FL[e];
                if (sR < double(0)) h = FR[e];
                return h;
            };
            double dsq;
            {
                const double du = VR[1] - VL[1];
                const double dv = VR[2] - VL[2];
                if constexpr (D == 2) {
                    dsq = du * du + dv * dv;
                } else {
                    const double dw = VR[3] - VL[3];
                    dsq = du * du + dv * dv + dw * dw;
                }
            }
            const double vel_mag = sqrt(dsq);
            double al1, al2;
            if (vel_mag < kEps) {
                al1 = double(1);
                al2 = double(0);
            } else {
                al1 = div_rn(fabs(VR[1 + A] - VL[1 + A]), vel_mag);
                al2 = sqrt(double(1) - al1 * al1);
            }
            const double be1 = double(1) / double(2) * (double(1) + div_rn(al1, al1 + al2));
            const double be2 = double(1) - be1;
            Fo[0] = be1 * Fo[0] + be2 * hll(0);
#pragma unroll
            for (int b = 0; b < D; b++)
                if (b != A) Fo[1 + b] = be1 * Fo[1 + b] + be2 * hll(1 + b);
        }
    }
}

// Face-state sink of face_flux_single that keeps nothing.
struct NoFaceStates {
    HM_HD void operator()(const double*, const double*) const {}
};

// sink(W-, W+) receives the reconstructed characteristic face values (a
// recording forward pass keeps them for the adjoint).
template <int D, int A, int SCHEME, int PP, int QQ, class Sink = NoFaceStates>
HM_HD void face_flux_single(const double (&V)[6][D + 2], double cL, double cR, double gam, const LDConst& ld,
                            bool use_hll, double* Fo, const Sink& sink = Sink{}) {
    constexpr int NEQ = D + 2;
    constexpr int IP = D + 1;  // pressure slot of V

    // Projection variables: simple averages (FlowModelBasicUtilitiesSingleSpecies.cpp:4638-5134).
    const double ra = double(1) / double(2) * (V[2][0] + V[3][0]);
    const double ca = double(1) / double(2) * (cL + cR);

    // Characteristic fields (FlowModelBasicUtilitiesSingleSpecies.cpp:6194-6242 and 3D analogues),
    // reconstructed one at a time.
    double Wm[NEQ], Wp[NEQ];
    {
        Six s;
#pragma unroll
        for (int m = 0; m < 6; m++)
            s.v[m] = -double(1) / double(2) * ra * ca * V[m][1 + A] + double(1) / double(2) * V[m][IP];
        reconstruct_pair<SCHEME, PP, QQ>(s, ld, Wm[0], Wp[0]);
    }
    {
        Six s;
#pragma unroll
        for (int m = 0; m < 6; m++) s.v[m] = V[m][0] - div_rn(double(1), ca * ca) * V[m][IP];
        reconstruct_pair<SCHEME, PP, QQ>(s, ld, Wm[1], Wp[1]);
    }
    {
        int t = 2;
#pragma unroll
        for (int b = 0; b < D; b++) {
            if (b == A) continue;
            Six s;
#pragma unroll
            for (int m = 0; m < 6; m++) s.v[m] = V[m][1 + b];
            reconstruct_pair<SCHEME, PP, QQ>(s, ld, Wm[t], Wp[t]);
            t++;
        }
    }
    {
        Six s;
#pragma unroll
        for (int m = 0; m < 6; m++)
            s.v[m] = double(1) / double(2) * ra * ca * V[m][1 + A] + double(1) / double(2) * V[m][IP];
        reconstruct_pair<SCHEME, PP, QQ>(s, ld, Wm[IP], Wp[IP]);
    }
    sink(Wm, Wp);

    // Back to primitive variables (FlowModelBasicUtilitiesSingleSpecies.cpp:7279-7284 and analogues).
    double VL[NEQ], VR[NEQ];
    auto back = [&](const double* W, double* Vo) {
        const double ic2 = div_rn(double(1), ca * ca), irc = div_rn(double(1), ra * ca);
        Vo[0] = ic2 * W[0] + W[1] + ic2 * W[IP];
        Vo[1 + A] = -irc * W[0] + irc * W[IP];
        int t = 2;
#pragma unroll
        for (int b = 0; b < D; b++) {
            if (b == A) continue;
            Vo[1 + b] = W[t++];
        }
        Vo[IP] = W[0] + W[IP];
    };
    back(Wm, VL);
    back(Wp, VR);

    // Positivity of rho and p on both sides, else first-order values
    // (FlowModelBasicUtilitiesSingleSpecies.cpp:3205-3222, WCNS56-HLLC-HLL.cpp:947-1032).
    const bool ok = (VL[0] > double(0)) && (VL[IP] > double(0)) && (VR[0] > double(0)) && (VR[IP] > double(0));
    if (!ok) {
#pragma unroll
        for (int e = 0; e < NEQ; e++) {
            VL[e] = V[2][e];
            VR[e] = V[3][e];
        }
    }

    riemann_flux_single<D, A>(VL, VR, gam, use_hll, Fo);
}

}  // namespace hm
