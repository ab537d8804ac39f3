// hm_adjoint_face.hpp -- reverse (adjoint) sweep of the pointwise numerics of
// one midpoint face, single-species model: characteristic projection, WCNS6-LD
// or WCNS5-JS reconstruction of both sides, back-projection, positivity
// fallback and the HLLC / HLLC-HLL flux.
//
// Each *_adj routine recomputes its forward intermediates with the forward
// code's expression trees (so every branch -- R_tau > alpha_tau, s* > 0, the
// sensor, the fallback, fmin/fmax selections -- is taken as in the forward
// pass) and then accumulates the transposed local Jacobian times the incoming
// adjoint.  Branch predicates are treated as constants, the usual discrete
// adjoint of a piecewise-smooth map.
//
// Forward references: WCNS6-LD-HLLC-HLL.cpp:9-319, WCNS5-JS-HLLC-HLL.cpp:9-164,
// FlowModelBasicUtilitiesSingleSpecies.cpp:3205-3222 (bounds), 4638-5134
// (projection), 6194-6242 (characteristic), 7279-7323 (back-projection),
// FlowModelRiemannSolverSingleSpeciesHLLC.cpp:519-1081, ...HLLC-HLL.cpp:804-1629.
#pragma once

#include "hm_numerics.hpp"

namespace hm {

HM_HD double sgn(double x) { return x > 0.0 ? 1.0 : (x < 0.0 ? -1.0 : 0.0); }

// The reverse sweep multiplies by reciprocals of shared denominators
// (rcp_rn): its outputs carry no bit-exactness requirement, and one multiply
// replaces a quotient.  Quotients that decide a branch (R_tau > alpha_tau,
// t1 >= t2, s* > 0, ...) are recomputed with the forward pass's correctly
// rounded quotients (div_rn, quot), so every branch goes the forward's way.

// d/dx of x^E (E >= 1) times g.
template <int E>
HM_HD double dpow_times(double x, int e_rt, double g) {
    const int e = (E > 0) ? E : e_rt;
    double r = 1.0;
    for (int i = 1; i < e; i++) r *= x;
    return double(e) * r * g;
}

// Gradient of the three sub-stencil indicators, weighted by bb0..bb2, added to sb.
HM_HD void beta3_adj(const Six& s, double bb0, double bb1, double bb2, double* sb) {
    const double t = double(1) / double(3);
    const double* v = s.v;
    sb[0] += t * bb0 * (8.0 * v[0] - 19.0 * v[1] + 11.0 * v[2]);
    sb[1] += t * (bb0 * (-19.0 * v[0] + 50.0 * v[1] - 31.0 * v[2]) + bb1 * (8.0 * v[1] - 13.0 * v[2] + 5.0 * v[3]));
    sb[2] += t * (bb0 * (11.0 * v[0] - 31.0 * v[1] + 20.0 * v[2]) + bb1 * (-13.0 * v[1] + 26.0 * v[2] - 13.0 * v[3]) +
                  bb2 * (20.0 * v[2] - 31.0 * v[3] + 11.0 * v[4]));
    sb[3] += t * (bb1 * (5.0 * v[1] - 13.0 * v[2] + 8.0 * v[3]) + bb2 * (-31.0 * v[2] + 50.0 * v[3] - 19.0 * v[4]));
    sb[4] += t * bb2 * (11.0 * v[2] - 19.0 * v[3] + 8.0 * v[4]);
}

// Gradient of the six-point indicator (a symmetric quadratic form), times bb3.
HM_HD void beta6_adj(const Six& s, double bb3, double* sb) {
    // Upper-triangle coefficients c_ij of sum_{i<=j} c_ij s_i s_j.
    const double c00 = 525910327.0, c01 = -4562164630.0, c02 = 7799501420.0, c03 = -6610694540.0,
                 c04 = 2794296070.0, c05 = -472758974.0;
    const double c11 = 5.0 * 2146987907.0, c12 = -5.0 * 7722406988.0, c13 = 5.0 * 6763559276.0,
                 c14 = -5.0 * 2926461814.0, c15 = 5.0 * 503766638.0;
    const double c22 = 20.0 * 1833221603.0, c23 = -20.0 * 3358664662.0, c24 = 20.0 * 1495974539.0,
                 c25 = -20.0 * 263126407.0;
    const double c33 = 20.0 * 1607794163.0, c34 = -20.0 * 1486026707.0, c35 = 20.0 * 268747951.0;
    const double c44 = 5.0 * 1432381427.0, c45 = -5.0 * 536951582.0;
    const double c55 = 263126407.0;
    const double g = bb3 * (1.0 / 232243200.0);
    const double* v = s.v;
    sb[0] += g * (2.0 * c00 * v[0] + c01 * v[1] + c02 * v[2] + c03 * v[3] + c04 * v[4] + c05 * v[5]);
    sb[1] += g * (c01 * v[0] + 2.0 * c11 * v[1] + c12 * v[2] + c13 * v[3] + c14 * v[4] + c15 * v[5]);
    sb[2] += g * (c02 * v[0] + c12 * v[1] + 2.0 * c22 * v[2] + c23 * v[3] + c24 * v[4] + c25 * v[5]);
    sb[3] += g * (c03 * v[0] + c13 * v[1] + c23 * v[2] + 2.0 * c33 * v[3] + c34 * v[4] + c35 * v[5]);
    sb[4] += g * (c04 * v[0] + c14 * v[1] + c24 * v[2] + c34 * v[3] + 2.0 * c44 * v[4] + c45 * v[5]);
    sb[5] += g * (c05 * v[0] + c15 * v[1] + c25 * v[2] + c35 * v[3] + c45 * v[4] + 2.0 * c55 * v[5]);
}

// Rounding-noise floors of the smoothness indicators, per m^2 with
// m = max |s_i| over the stencil.  beta3 and beta6 are evaluated in the
// reference's expanded quadratic forms; their rounding error is below about
// 10 u times the sum of the magnitudes of their terms (u = 2^-53): at most
// 33.3 m^2 for each of beta0..beta2 and 1462 m^2 for beta6 (the coefficient
// magnitudes summed and divided by 232243200).  Where the difference tau6 =
// |beta6 - beta_avg| (or tau5 = |beta0 - beta2|) is below its floor, its sign
// and size are rounding noise: on near-uniform O(1) data (variations below
// ~1e-6 relative) the nonlinear weights then scatter with the last bits of the
// stencil and carry no derivative in s.  The reverse sweep holds such weights
// constant -- it propagates only what varies with s above the noise -- while
// the forward values and every branch stay as the forward pass computed them.
constexpr double kNoiseBeta3 = 3.7e-14;  // 10 u * 33.3
constexpr double kNoiseBeta6 = 1.6e-12;  // 10 u * 1462

HM_HD double stencil_mag2(const Six& s) {
    double m = 0.0;
#pragma unroll
    for (int i = 0; i < 6; i++) m = fmax(m, fabs(s.v[i]));
    return m * m;
}

// Shock measure sigma of ld_sigma: adjoint sigb into sb.
HM_HD void ld_sigma_adj(const Six& s, double sigb, double* sb) {
    if (sigb == 0.0) return;
    const double a1 = s.v[2] - s.v[1];
    const double a2 = s.v[3] - s.v[2];
    const double a3 = s.v[4] - s.v[3];
    const double x1 = a1 - a2, x2 = a2 - a3;
    const double den1 = fabs(a1) + fabs(a2) + kEps;
    const double den2 = fabs(a2) + fabs(a3) + kEps;
    const double t1 = div_rn(fabs(x1), den1);  // the branch below compares them as the forward does
    const double t2 = div_rn(fabs(x2), den2);
    double a1b = 0.0, a2b = 0.0, a3b = 0.0;
    if (t1 >= t2) {
        const double id = rcp_rn(den1);
        const double xb = sigb * sgn(x1) * id;
        const double db = -sigb * t1 * id;
        a1b += xb + db * sgn(a1);
        a2b += -xb + db * sgn(a2);
    } else {
        const double id = rcp_rn(den2);
        const double xb = sigb * sgn(x2) * id;
        const double db = -sigb * t2 * id;
        a2b += xb + db * sgn(a2);
        a3b += -xb + db * sgn(a3);
    }
    sb[1] -= a1b;
    sb[2] += a1b - a2b;
    sb[3] += a2b - a3b;
    sb[4] += a3b;
}

// Adjoint of ld_side<P,Q>(s, sigma, k): adds d(out)/ds * ob to sb and
// d(out)/dsigma * ob to sigb.  The forward quantities are recomputed with
// reciprocals of the shared denominators (e_k, the weight sums); only the
// R_tau > alpha_tau predicate keeps the forward's division.  The central
// (upwind) weights are held constant where tau6 (tau5) is below its rounding
// noise floor.
template <int P, int Q>
HM_HD void ld_side_adj(const Six& s, double sigma, const LDConst& k, double ob, double* sb, double& sigb) {
    if (ob == 0.0) return;
    double b0, b1, b2;
    beta3(s, b0, b1, b2);
    const double b3 = beta6(s);
    const double ie0 = rcp_rn(b0 + kEps), ie1 = rcp_rn(b1 + kEps), ie2 = rcp_rn(b2 + kEps),
                 ie3 = rcp_rn(b3 + kEps);
    const double bavg = double(1) / double(8) * (b0 + b2 + double(6) * b1);
    const double d3 = b3 - bavg;
    const double tau6 = fabs(d3);
    const bool blend = div_rn(tau6, bavg + kEps) > k.alpha_tau;
    const double q0 = tau6 * ie0, q1 = tau6 * ie1, q2 = tau6 * ie2, q3 = tau6 * ie3;
    const double w0 = double(1) / double(32) * (k.C + rep_pow<Q>(q0, k.q));
    const double w1 = double(15) / double(32) * (k.C + rep_pow<Q>(q1, k.q));
    const double w2 = double(15) / double(32) * (k.C + rep_pow<Q>(q2, k.q));
    const double w3 = double(1) / double(32) * (k.C + rep_pow<Q>(q3, k.q));
    const double iws = rcp_rn(w0 + w1 + w2 + w3);
    const double wh0 = w0 * iws, wh1 = w1 * iws, wh2 = w2 * iws, wh3 = w3 * iws;
    const double* v = s.v;
    const double Wb0 = ob * (3.0 / 8.0 * v[0] - 10.0 / 8.0 * v[1] + 15.0 / 8.0 * v[2]);
    const double Wb1 = ob * (-1.0 / 8.0 * v[1] + 6.0 / 8.0 * v[2] + 3.0 / 8.0 * v[3]);
    const double Wb2 = ob * (3.0 / 8.0 * v[2] + 6.0 / 8.0 * v[3] - 1.0 / 8.0 * v[4]);
    const double Wb3 = ob * (15.0 / 8.0 * v[3] - 10.0 / 8.0 * v[4] + 3.0 / 8.0 * v[5]);
    double W0 = wh0, W1 = wh1, W2 = wh2, W3 = wh3;
    double whb0 = Wb0, whb1 = Wb1, whb2 = Wb2, whb3 = Wb3;
    double bb0 = 0.0, bb1 = 0.0, bb2 = 0.0, bb3 = 0.0;
    const double mag2 = stencil_mag2(s);
    if (blend) {
        // upwind weights u_k = d_k (1 + r_k^P), r_k = tau5 / e_k, normalized
        const double d02 = b0 - b2;
        const double tau5 = fabs(d02);
        const double r0 = tau5 * ie0, r1 = tau5 * ie1, r2 = tau5 * ie2;
        const double u0 = double(1) / double(16) * (double(1) + rep_pow<P>(r0, k.p));
        const double u1 = double(5) / double(8) * (double(1) + rep_pow<P>(r1, k.p));
        const double u2 = double(5) / double(16) * (double(1) + rep_pow<P>(r2, k.p));
        const double ius = rcp_rn(u0 + u1 + u2);
        const double uh0 = u0 * ius, uh1 = u1 * ius, uh2 = u2 * ius;
        const double om = double(1) - sigma;
        W0 = sigma * uh0 + om * wh0;
        W1 = sigma * uh1 + om * wh1;
        W2 = sigma * uh2 + om * wh2;
        W3 = om * wh3;
        whb0 = om * Wb0;
        whb1 = om * Wb1;
        whb2 = om * Wb2;
        whb3 = om * Wb3;
        sigb += (uh0 - wh0) * Wb0 + (uh1 - wh1) * Wb1 + (uh2 - wh2) * Wb2 - wh3 * Wb3;
        if (tau5 > 2.0 * kNoiseBeta3 * mag2) {
            const double uhb0 = sigma * Wb0, uhb1 = sigma * Wb1, uhb2 = sigma * Wb2;
            const double dotu = uhb0 * uh0 + uhb1 * uh1 + uhb2 * uh2;
            const double ub0 = (uhb0 - dotu) * ius, ub1 = (uhb1 - dotu) * ius, ub2 = (uhb2 - dotu) * ius;
            const double rb0 = dpow_times<P>(r0, k.p, ub0 * (1.0 / 16.0)) * ie0;
            const double rb1 = dpow_times<P>(r1, k.p, ub1 * (5.0 / 8.0)) * ie1;
            const double rb2 = dpow_times<P>(r2, k.p, ub2 * (5.0 / 16.0)) * ie2;
            const double tau5b = rb0 + rb1 + rb2;
            bb0 -= rb0 * r0;
            bb1 -= rb1 * r1;
            bb2 -= rb2 * r2;
            const double sg = sgn(d02);
            bb0 += tau5b * sg;
            bb2 -= tau5b * sg;
        }
    }
    // out = sum of linear combinations of the stencil values with weights W.
    sb[0] += ob * (3.0 / 8.0 * W0);
    sb[1] += ob * (-10.0 / 8.0 * W0 - 1.0 / 8.0 * W1);
    sb[2] += ob * (15.0 / 8.0 * W0 + 6.0 / 8.0 * W1 + 3.0 / 8.0 * W2);
    sb[3] += ob * (3.0 / 8.0 * W1 + 6.0 / 8.0 * W2 + 15.0 / 8.0 * W3);
    sb[4] += ob * (-1.0 / 8.0 * W2 - 10.0 / 8.0 * W3);
    sb[5] += ob * (3.0 / 8.0 * W3);
    // central weights w_k = d_k (C + q_k^Q), q_k = tau6 / e_k, normalized
    if (!(tau6 > kNoiseBeta6 * mag2)) {
        if (bb0 != 0.0 || bb1 != 0.0 || bb2 != 0.0) beta3_adj(s, bb0, bb1, bb2, sb);
        return;
    }
    const double dotw = whb0 * wh0 + whb1 * wh1 + whb2 * wh2 + whb3 * wh3;
    const double wb0 = (whb0 - dotw) * iws, wb1 = (whb1 - dotw) * iws, wb2 = (whb2 - dotw) * iws,
                 wb3 = (whb3 - dotw) * iws;
    const double qb0 = dpow_times<Q>(q0, k.q, wb0 * (1.0 / 32.0)) * ie0;
    const double qb1 = dpow_times<Q>(q1, k.q, wb1 * (15.0 / 32.0)) * ie1;
    const double qb2 = dpow_times<Q>(q2, k.q, wb2 * (15.0 / 32.0)) * ie2;
    const double qb3 = dpow_times<Q>(q3, k.q, wb3 * (1.0 / 32.0)) * ie3;
    const double tau6b = qb0 + qb1 + qb2 + qb3;
    bb0 -= qb0 * q0;
    bb1 -= qb1 * q1;
    bb2 -= qb2 * q2;
    bb3 -= qb3 * q3;
    const double sg3 = sgn(d3);
    bb3 += tau6b * sg3;
    const double bavgb = -tau6b * sg3;
    bb0 += bavgb * (1.0 / 8.0);
    bb2 += bavgb * (1.0 / 8.0);
    bb1 += bavgb * (6.0 / 8.0);
    beta3_adj(s, bb0, bb1, bb2, sb);
    beta6_adj(s, bb3, sb);
}

// Adjoint of js_side<P>(s, p): w_k = d_k / e_k^p recomputed as d_k (1/e_k)^p.
// The weights are held constant where every indicator is below its rounding
// noise floor.
template <int P>
HM_HD void js_side_adj(const Six& s, int p_rt, double ob, double* sb) {
    if (ob == 0.0) return;
    double b0, b1, b2;
    beta3(s, b0, b1, b2);
    const double ie0 = rcp_rn(b0 + kEps), ie1 = rcp_rn(b1 + kEps), ie2 = rcp_rn(b2 + kEps);
    const double w0 = double(1.0) / double(16) * rep_pow<P>(ie0, p_rt);
    const double w1 = double(5.0) / double(8) * rep_pow<P>(ie1, p_rt);
    const double w2 = double(5.0) / double(16) * rep_pow<P>(ie2, p_rt);
    const double iws = rcp_rn(w0 + w1 + w2);
    const double W0 = w0 * iws, W1 = w1 * iws, W2 = w2 * iws;
    const double* v = s.v;
    sb[0] += ob * (3.0 / 8.0 * W0);
    sb[1] += ob * (-10.0 / 8.0 * W0 - 1.0 / 8.0 * W1);
    sb[2] += ob * (15.0 / 8.0 * W0 + 6.0 / 8.0 * W1 + 3.0 / 8.0 * W2);
    sb[3] += ob * (3.0 / 8.0 * W1 + 6.0 / 8.0 * W2);
    sb[4] += ob * (-1.0 / 8.0 * W2);
    if (!(fmax(fmax(b0, b1), b2) > kNoiseBeta3 * stencil_mag2(s))) return;
    const double Wb0 = ob * (3.0 / 8.0 * v[0] - 10.0 / 8.0 * v[1] + 15.0 / 8.0 * v[2]);
    const double Wb1 = ob * (-1.0 / 8.0 * v[1] + 6.0 / 8.0 * v[2] + 3.0 / 8.0 * v[3]);
    const double Wb2 = ob * (3.0 / 8.0 * v[2] + 6.0 / 8.0 * v[3] - 1.0 / 8.0 * v[4]);
    const double dot = Wb0 * W0 + Wb1 * W1 + Wb2 * W2;
    const double wb0 = (Wb0 - dot) * iws, wb1 = (Wb1 - dot) * iws, wb2 = (Wb2 - dot) * iws;
    const int p = (P > 0) ? P : p_rt;
    // w_k = d_k e_k^-p  ->  dw_k/de_k = -p w_k / e_k
    const double bb0 = -wb0 * double(p) * w0 * ie0;
    const double bb1 = -wb1 * double(p) * w1 * ie1;
    const double bb2 = -wb2 * double(p) * w2 * ie2;
    beta3_adj(s, bb0, bb1, bb2, sb);
}

// Adjoint of reconstruct_pair: (umb, upb) -> sb (six stencil values).
template <int SCHEME, int P, int Q>
HM_HD void reconstruct_pair_adj(const Six& s, const LDConst& k, double umb, double upb, double* sb) {
    if (umb == 0.0 && upb == 0.0) return;
    if constexpr (SCHEME == kWCNS6_LD) {
        const double sig = ld_sigma(s);
        double sigb = 0.0;
        // The two sides one after the other (a runtime loop): their reverse
        // sweeps are independent, and interleaved they would not fit the
        // register file.  The plus side works on the mirrored stencil.
#pragma unroll 1
        for (int side = 0; side < 2; side++) {
            Six ss;
#pragma unroll
            for (int m = 0; m < 6; m++) ss.v[m] = side ? s.v[5 - m] : s.v[m];
            double tb[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
            ld_side_adj<P, Q>(ss, sig, k, side ? upb : umb, tb, sigb);
#pragma unroll
            for (int m = 0; m < 6; m++) sb[m] += side ? tb[5 - m] : tb[m];
        }
        ld_sigma_adj(s, sigb, sb);
    } else {
        const Six f = flip(s);
        double fb[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        js_side_adj<P>(s, k.p, umb, sb);
        js_side_adj<P>(f, k.p, upb, fb);
#pragma unroll
        for (int m = 0; m < 6; m++) sb[m] += fb[5 - m];
    }
}

// Adjoint of the single-species HLLC / HLLC-HLL midpoint flux of one face
// along axis A (forward: midpoint_flux_at).  Fb -> VLb, VRb (accumulated).
template <int D, int A>
HM_HD void riemann_single_adj(const double* VL, const double* VR, double gam, bool use_hll, const double* Fb,
                              double* VLb, double* VRb) {
    constexpr int NEQ = D + 2;
    constexpr int IP = D + 1;
    const double rhoL = VL[0], rhoR = VR[0];
    const double pL = VL[IP], pR = VR[IP];
    const double uL = VL[1 + A], uR = VR[1 + A];
    // Forward quotients recomputed exactly as the forward pass rounds them.
    const double csL = sqrt(div_rn(gam * pL, rhoL));
    const double csR = sqrt(div_rn(gam * pR, rhoR));
    const double epsL = div_rn(pL, (gam - double(1)) * rhoL);
    const double epsR = div_rn(pR, (gam - double(1)) * rhoR);
    const double u_avg = double(1) / double(2) * (uL + uR);
    const double c_avg = double(1) / double(2) * (csL + csR);
    const double A1 = u_avg - c_avg, A2 = uL - csL;
    const double B1 = u_avg + c_avg, B2 = uR + csR;
    const double sL = fmin(A1, A2);
    const double sR = fmax(B1, B2);
    const double s_minus = fmin(double(0), sL);
    const double s_plus = fmax(double(0), sR);
    const double num = pR - pL + rhoL * uL * (sL - uL) - rhoR * uR * (sR - uR);
    const double den = rhoL * (sL - uL) - rhoR * (sR - uR);
    const double s_star = div_rn(num, den);

    // Conservative states and fluxes of both sides.
    double Q[2][NEQ], F[2][NEQ];
    const double* Vs[2] = {VL, VR};
    const double epss[2] = {epsL, epsR};
#pragma unroll
    for (int x = 0; x < 2; x++) {
        const double* V = Vs[x];
        Q[x][0] = V[0];
#pragma unroll
        for (int b = 0; b < D; b++) Q[x][1 + b] = V[0] * V[1 + b];
        double ke;
        if constexpr (D == 1)
            ke = V[1] * V[1];
        else if constexpr (D == 2)
            ke = V[1] * V[1] + V[2] * V[2];
        else
            ke = V[1] * V[1] + V[2] * V[2] + V[3] * V[3];
        Q[x][IP] = V[0] * (epss[x] + double(1) / double(2) * ke);
        F[x][0] = Q[x][1 + A];
#pragma unroll
        for (int b = 0; b < D; b++) F[x][1 + b] = (b == A) ? Q[x][1 + A] * V[1 + b] + V[IP] : Q[x][1 + A] * V[1 + b];
        F[x][IP] = V[1 + A] * (Q[x][IP] + V[IP]);
    }
    // Upwind side K of the contact, selected element by element (no
    // run-time indexing of private arrays).
    const bool left = s_star > double(0);
    const double sK = left ? sL : sR;
    const double spm = left ? s_minus : s_plus;
    const double uK = left ? uL : uR, rhoK = left ? rhoL : rhoR, pK = left ? pL : pR;
    double QK[NEQ], FK[NEQ];
#pragma unroll
    for (int e = 0; e < NEQ; e++) {
        QK[e] = left ? Q[0][e] : Q[1][e];
        FK[e] = left ? F[0][e] : F[1][e];
    }
    const double chi = div_rn(sK - uK, sK - s_star);
    const double g = sK - uK;
    const double h = rhoK * s_star + div_rn(pK, g);
    const double inner = QK[IP] + (s_star - uK) * h;
    double Qst[NEQ], Fh[NEQ];
    Qst[0] = chi * rhoK;
#pragma unroll
    for (int b = 0; b < D; b++) Qst[1 + b] = (b == A) ? chi * rhoK * s_star : chi * QK[1 + b];
    Qst[IP] = chi * inner;
#pragma unroll
    for (int e = 0; e < NEQ; e++) Fh[e] = FK[e] + spm * (Qst[e] - QK[e]);

    double Qb[2][NEQ], Fxb[2][NEQ];
#pragma unroll
    for (int x = 0; x < 2; x++)
#pragma unroll
        for (int e = 0; e < NEQ; e++) {
            Qb[x][e] = 0.0;
            Fxb[x][e] = 0.0;
        }
    double sLb = 0.0, sRb = 0.0, sstb = 0.0;
    double Vb[2][NEQ];
#pragma unroll
    for (int x = 0; x < 2; x++)
#pragma unroll
        for (int e = 0; e < NEQ; e++) Vb[x][e] = 0.0;

    double Fhb[NEQ];
#pragma unroll
    for (int e = 0; e < NEQ; e++) Fhb[e] = Fb[e];
    if constexpr (D > 1) {
        if (use_hll) {
            // Forward hybrid: mass and transverse momentum blend HLLC with HLL.
            double du[D];
            double dsq = 0.0;
            {
                const double d0 = VR[1] - VL[1];
                const double d1 = VR[2] - VL[2];
                du[0] = d0;
                du[1] = d1;
                if constexpr (D == 2) {
                    dsq = d0 * d0 + d1 * d1;
                } else {
                    const double d2 = VR[3] - VL[3];
                    du[2] = d2;
                    dsq = d0 * d0 + d1 * d1 + d2 * d2;
                }
            }
            const double vm = sqrt(dsq);
            double al1, al2;
            const bool tiny = vm < kEps;
            if (tiny) {
                al1 = double(1);
                al2 = double(0);
            } else {
                al1 = div_rn(fabs(VR[1 + A] - VL[1 + A]), vm);
                al2 = sqrt(double(1) - al1 * al1);
            }
            const double Ssum = al1 + al2;
            const double be1 = double(1) / double(2) * (double(1) + div_rn(al1, Ssum));
            const double be2 = double(1) - be1;
            double be1b = 0.0, be2b = 0.0;
            const double dsl = sR - sL;
            const double idsl = rcp_rn(dsl);
#pragma unroll
            for (int e = 0; e < NEQ; e++) {
                const bool hyb = (e == 0) || (e >= 1 && e <= D && e != 1 + A);
                if (!hyb) continue;
                // recompute the HLL flux of component e
                double hv;
                const double numh = sR * F[0][e] - sL * F[1][e] + sR * sL * (Q[1][e] - Q[0][e]);
                hv = quot(numh, dsl, idsl);
                if (sL > double(0)) hv = F[0][e];
                if (sR < double(0)) hv = F[1][e];
                const double fb = Fb[e];
                Fhb[e] = be1 * fb;
                const double hb = be2 * fb;
                be1b += fb * Fh[e];
                be2b += fb * hv;
                if (sR < double(0)) {
                    Fxb[1][e] += hb;
                } else if (sL > double(0)) {
                    Fxb[0][e] += hb;
                } else {
                    const double nb = hb * idsl;
                    const double db = -hb * hv * idsl;
                    sRb += nb * (F[0][e] + sL * (Q[1][e] - Q[0][e])) + db;
                    sLb += nb * (-F[1][e] + sR * (Q[1][e] - Q[0][e])) - db;
                    Fxb[0][e] += nb * sR;
                    Fxb[1][e] -= nb * sL;
                    Qb[1][e] += nb * sR * sL;
                    Qb[0][e] -= nb * sR * sL;
                }
            }
            be1b -= be2b;
            if (!tiny) {
                const double iS = rcp_rn(Ssum);
                const double al1b = be1b * 0.5 * al2 * (iS * iS);
                double al2b = -be1b * 0.5 * al1 * (iS * iS);
                double a1b = al1b;
                if (al2 > 0.0) a1b += al2b * (-al1 * rcp_rn(al2));
                const double dua = du[A];
                const double ivm = rcp_rn(vm);  // vm >= kEps here
                const double duab = a1b * sgn(dua) * ivm;
                const double vmb = -a1b * al1 * ivm;
                const double dsqb = vmb * 0.5 * ivm;
#pragma unroll
                for (int b = 0; b < D; b++) {
                    double dub = dsqb * 2.0 * du[b];
                    if (b == A) dub += duab;
                    Vb[1][1 + b] += dub;
                    Vb[0][1 + b] -= dub;
                }
            }
        }
    }
    // HLLC on the side K.
    double spmb = 0.0, chib = 0.0;
    double Qstb[NEQ], QKb[NEQ], FKb[NEQ];
#pragma unroll
    for (int e = 0; e < NEQ; e++) {
        FKb[e] = Fhb[e];
        spmb += Fhb[e] * (Qst[e] - QK[e]);
        Qstb[e] = spm * Fhb[e];
        QKb[e] = -spm * Fhb[e];
    }
    double rhoKb = 0.0, uKb = 0.0, pKb = 0.0, sKb = 0.0;
    chib += Qstb[0] * rhoK;
    rhoKb += Qstb[0] * chi;
#pragma unroll
    for (int b = 0; b < D; b++) {
        if (b == A) {
            chib += Qstb[1 + b] * rhoK * s_star;
            rhoKb += Qstb[1 + b] * chi * s_star;
            sstb += Qstb[1 + b] * chi * rhoK;
        } else {
            chib += Qstb[1 + b] * QK[1 + b];
            QKb[1 + b] += Qstb[1 + b] * chi;
        }
    }
    {
        chib += Qstb[IP] * inner;
        const double innerb = Qstb[IP] * chi;
        QKb[IP] += innerb;
        sstb += innerb * h;
        uKb -= innerb * h;
        const double hb = innerb * (s_star - uK);
        rhoKb += hb * s_star;
        sstb += hb * rhoK;
        const double ig = rcp_rn(g);
        pKb += hb * ig;
        const double gb = -hb * pK * (ig * ig);
        sKb += gb;
        uKb -= gb;
    }
    {
        const double ib2 = rcp_rn(sK - s_star);
        const double ab = chib * ib2;
        const double b2b = -chib * chi * ib2;
        sKb += ab + b2b;
        uKb -= ab;
        sstb -= b2b;
    }
    if (left) {
        if (sL < double(0)) sLb += spmb;
        sLb += sKb;
    } else {
        if (sR > double(0)) sRb += spmb;
        sRb += sKb;
    }
#pragma unroll
    for (int e = 0; e < NEQ; e++) {
        Qb[0][e] += left ? QKb[e] : 0.0;
        Qb[1][e] += left ? 0.0 : QKb[e];
        Fxb[0][e] += left ? FKb[e] : 0.0;
        Fxb[1][e] += left ? 0.0 : FKb[e];
    }
    Vb[0][0] += left ? rhoKb : 0.0;
    Vb[1][0] += left ? 0.0 : rhoKb;
    Vb[0][1 + A] += left ? uKb : 0.0;
    Vb[1][1 + A] += left ? 0.0 : uKb;
    Vb[0][IP] += left ? pKb : 0.0;
    Vb[1][IP] += left ? 0.0 : pKb;

    // Conservative states and fluxes of both sides.
    double epsb[2] = {0.0, 0.0};
#pragma unroll
    for (int x = 0; x < 2; x++) {
        const double* V = Vs[x];
        double* vb = Vb[x];
        const double* fb = Fxb[x];
        double* qb = Qb[x];
        // F[0] = Q[1+A]
        qb[1 + A] += fb[0];
#pragma unroll
        for (int b = 0; b < D; b++) {
            qb[1 + A] += fb[1 + b] * V[1 + b];
            vb[1 + b] += fb[1 + b] * Q[x][1 + A];
            if (b == A) vb[IP] += fb[1 + b];
        }
        vb[1 + A] += fb[IP] * (Q[x][IP] + V[IP]);
        qb[IP] += fb[IP] * V[1 + A];
        vb[IP] += fb[IP] * V[1 + A];
        // Q[0] = rho, Q[1+b] = rho u_b, Q[IP] = rho (eps + ke/2)
        vb[0] += qb[0];
#pragma unroll
        for (int b = 0; b < D; b++) {
            vb[0] += qb[1 + b] * V[1 + b];
            vb[1 + b] += qb[1 + b] * V[0];
        }
        double ke = 0.0;
#pragma unroll
        for (int b = 0; b < D; b++) ke += V[1 + b] * V[1 + b];
        vb[0] += qb[IP] * (epss[x] + 0.5 * ke);
        epsb[x] += qb[IP] * V[0];
#pragma unroll
        for (int b = 0; b < D; b++) vb[1 + b] += qb[IP] * V[0] * V[1 + b];
    }
    // s* = num / den
    {
        const double iden = rcp_rn(den);
        const double numb = sstb * iden;
        const double denb = -sstb * s_star * iden;
        Vb[1][IP] += numb;
        Vb[0][IP] -= numb;
        Vb[0][0] += numb * uL * (sL - uL) + denb * (sL - uL);
        Vb[0][1 + A] += numb * (rhoL * (sL - uL) - rhoL * uL) - denb * rhoL;
        sLb += numb * rhoL * uL + denb * rhoL;
        Vb[1][0] -= numb * uR * (sR - uR) + denb * (sR - uR);
        Vb[1][1 + A] -= numb * (rhoR * (sR - uR) - rhoR * uR) - denb * rhoR;
        sRb -= numb * rhoR * uR + denb * rhoR;
    }
    // sL = fmin(A1, A2), sR = fmax(B1, B2)
    double uavgb = 0.0, cavgb = 0.0, csb[2] = {0.0, 0.0};
    if (A1 <= A2) {
        uavgb += sLb;
        cavgb -= sLb;
    } else {
        Vb[0][1 + A] += sLb;
        csb[0] -= sLb;
    }
    if (B1 >= B2) {
        uavgb += sRb;
        cavgb += sRb;
    } else {
        Vb[1][1 + A] += sRb;
        csb[1] += sRb;
    }
    Vb[0][1 + A] += 0.5 * uavgb;
    Vb[1][1 + A] += 0.5 * uavgb;
    csb[0] += 0.5 * cavgb;
    csb[1] += 0.5 * cavgb;
    // c = sqrt(gam p / rho), eps = p / ((gam - 1) rho)
    const double css[2] = {csL, csR};
    const double igm1 = rcp_rn(gam - 1.0);
#pragma unroll
    for (int x = 0; x < 2; x++) {
        const double* V = Vs[x];
        const double irho = rcp_rn(V[0]);
        const double ab = csb[x] * 0.5 * rcp_rn(css[x]);
        Vb[x][IP] += ab * gam * irho + epsb[x] * igm1 * irho;
        Vb[x][0] -= ab * gam * V[IP] * (irho * irho) + epsb[x] * epss[x] * irho;
    }
#pragma unroll
    for (int e = 0; e < NEQ; e++) {
        VLb[e] += Vb[0][e];
        VRb[e] += Vb[1][e];
    }
}

// Characteristic field f of stencil point m for axis A (the forward
// projection's expression trees, FlowModelBasicUtilitiesSingleSpecies.cpp:6194-6242);
// vat(m, e) reads component e of stencil point m.
template <int D, int A, class VAt>
HM_HD double char_field_at(const VAt& vat, double ra, double ca, int f, int m) {
    constexpr int IP = D + 1;
    if (f == 0) return -double(1) / double(2) * ra * ca * vat(m, 1 + A) + double(1) / double(2) * vat(m, IP);
    if (f == 1) return vat(m, 0) - div_rn(double(1), ca * ca) * vat(m, IP);
    if (f == IP) return double(1) / double(2) * ra * ca * vat(m, 1 + A) + double(1) / double(2) * vat(m, IP);
    int t = 2;
    for (int b = 0; b < D; b++) {
        if (b == A) continue;
        if (t == f) return vat(m, 1 + b);
        t++;
    }
    return 0.0;
}

template <int D, int A>
HM_HD double char_field(const double (&V)[6][D + 2], double ra, double ca, int f, int m) {
    return char_field_at<D, A>([&](int mm, int e) { return V[mm][e]; }, ra, ca, f, m);
}

// Size of the per-face hand-over between the two halves of a face adjoint.
template <int D>
constexpr int face_adj_h_size() {
    return 2 * (D + 2) + 3;
}

// Back-projection of characteristic face values W to primitive variables Vo
// (the forward code's expression trees, FlowModelBasicUtilitiesSingleSpecies.cpp:7279-7323).
template <int D, int A>
HM_HD void back_project_single(const double* W, double ra, double ca, double* Vo) {
    constexpr int IP = D + 1;
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
}

// The part of the face's first adjoint half after the reconstruction: the
// positivity check and fallback, the Riemann-flux adjoint and the
// back-projection adjoint.  Wm, Wp are the reconstructed characteristic
// values (recomputed, or kept by a recording forward pass), V2, V3 the
// primitive variables of cells i-1 and i, ra, ca the projection averages.
template <int D, int A>
HM_HD void face_adj_single_after_recon(const double* Wm, const double* Wp, const double* V2, const double* V3,
                                       double ra, double ca, double gam, bool use_hll, const double* mb, double* h) {
    constexpr int NEQ = D + 2;
    constexpr int IP = D + 1;
    double VL[NEQ], VR[NEQ];
    back_project_single<D, A>(Wm, ra, ca, VL);
    back_project_single<D, A>(Wp, ra, ca, VR);
    const bool ok = (VL[0] > double(0)) && (VL[IP] > double(0)) && (VR[0] > double(0)) && (VR[IP] > double(0));
    // Element-wise selects (a conditional copy from V2, V3 would make the
    // compiler keep the stencil in scratch memory to address it).
#pragma unroll
    for (int e = 0; e < NEQ; e++) {
        const double l = V2[e], r = V3[e];
        VL[e] = ok ? VL[e] : l;
        VR[e] = ok ? VR[e] : r;
    }
    double VLb[NEQ], VRb[NEQ];
#pragma unroll
    for (int e = 0; e < NEQ; e++) {
        VLb[e] = 0.0;
        VRb[e] = 0.0;
    }
    riemann_single_adj<D, A>(VL, VR, gam, use_hll, mb, VLb, VRb);
    if (!ok) {
#pragma unroll
        for (int e = 0; e < NEQ; e++) {
            h[e] = VLb[e];
            h[NEQ + e] = VRb[e];
        }
        h[2 * NEQ] = 0.0;
        h[2 * NEQ + 1] = 0.0;
        h[2 * NEQ + 2] = 0.0;
        return;
    }
    // Back-projection adjoint for both sides.
    double rab = 0.0, cab = 0.0;
    const double ica = rcp_rn(ca), ira = rcp_rn(ra);
    auto back_adj = [&](const double* W, const double* Vob, double* Wb) {
        const double ic2 = div_rn(double(1), ca * ca);
        const double irc = div_rn(double(1), ra * ca);
        Wb[0] = Vob[0] * ic2 - Vob[1 + A] * irc + Vob[IP];
        Wb[1] = Vob[0];
        Wb[IP] = Vob[0] * ic2 + Vob[1 + A] * irc + Vob[IP];
        int t = 2;
#pragma unroll
        for (int b = 0; b < D; b++) {
            if (b == A) continue;
            Wb[t++] = Vob[1 + b];
        }
        cab += Vob[0] * (W[0] + W[IP]) * (-2.0 * ic2 * ica);
        cab += Vob[1 + A] * (W[IP] - W[0]) * (-irc * ica);
        rab += Vob[1 + A] * (W[IP] - W[0]) * (-irc * ira);
    };
    double Wmb[NEQ], Wpb[NEQ];
    back_adj(Wm, VLb, Wmb);
    back_adj(Wp, VRb, Wpb);
#pragma unroll
    for (int e = 0; e < NEQ; e++) {
        h[e] = Wmb[e];
        h[NEQ + e] = Wpb[e];
    }
    h[2 * NEQ] = rab;
    h[2 * NEQ + 1] = cab;
    h[2 * NEQ + 2] = 1.0;
}

// First half of the adjoint of one face's midpoint flux along axis A: the
// forward reconstruction, back-projection and bounds check, the Riemann-flux
// adjoint, then the back-projection adjoint.  V[m][e] are the primitive
// variables of stencil cells i-3+m, cL/cR the sound speeds of cells i-1 and
// i, mb the adjoint of the midpoint flux.  Writes h[0..NEQ) = adjoint of the
// minus-side characteristic values (or of V_L where the first-order fallback
// was taken), h[NEQ..2NEQ) = the plus side's (or V_R's), h[2NEQ], h[2NEQ+1] =
// the adjoints of rho_avg and c_avg so far, h[2NEQ+2] = 1 (reconstructed) or
// 0 (fallback).
template <int D, int A, int SCHEME, int P, int Q>
HM_HD void face_adj_single_flux(const double (&V)[6][D + 2], double cL, double cR, double gam, const LDConst& ld,
                                bool use_hll, const double* mb, double* h) {
    constexpr int NEQ = D + 2;
    const double ra = double(1) / double(2) * (V[2][0] + V[3][0]);
    const double ca = double(1) / double(2) * (cL + cR);
    double Wm[NEQ], Wp[NEQ];
#pragma unroll
    for (int f = 0; f < NEQ; f++) {
        Six s;
#pragma unroll
        for (int m = 0; m < 6; m++) s.v[m] = char_field<D, A>(V, ra, ca, f, m);
        reconstruct_pair<SCHEME, P, Q>(s, ld, Wm[f], Wp[f]);
    }
    face_adj_single_after_recon<D, A>(Wm, Wp, V[2], V[3], ra, ca, gam, use_hll, mb, h);
}

// Size of the per-face record the second half writes.
template <int D>
constexpr int face_adj_g_size() {
    return 6 * (D + 2) + 2;
}

// Second half: given the hand-over h (read through hat(c)) and the stencil's
// primitive variables (read through vat(m, e), so that they need not stay in
// registers), the reverse reconstruction, one field at a time (a runtime
// loop, so only one field's WENO adjoint is live).  Writes the face record
// through gout(c, value):
//   reconstructed face (h flag 1): g[f*6 + m] = adjoint of characteristic value
//     f at stencil point m, g[6 NEQ] = adjoint of rho_avg, g[6 NEQ + 1] = of c_avg;
//   fallback face (h flag 0): g[e] = adjoint of V_L, g[NEQ + e] = of V_R.
// The characteristic projection's adjoint is applied per stencil point by
// face_record_to_point (in the gather), so no per-point accumulators are live here.
template <int D, int A, int SCHEME, int P, int Q, class VAt, class HAt, class GOut>
HM_HD void face_adj_single_recon(const VAt& vat, double cL, double cR, const LDConst& ld, const HAt& hat,
                                 const GOut& gout) {
    constexpr int NEQ = D + 2;
    constexpr int IP = D + 1;
    if (hat(2 * NEQ + 2) == 0.0) {
#pragma unroll
        for (int e = 0; e < 2 * NEQ; e++) gout(e, hat(e));
        return;
    }
    const double ra = double(1) / double(2) * (vat(2, 0) + vat(3, 0));
    const double ca = double(1) / double(2) * (cL + cR);
    double rab = hat(2 * NEQ), cab = hat(2 * NEQ + 1);
    const double ic3 = rcp_rn(ca * ca * ca);
#pragma unroll 1
    for (int f = 0; f < NEQ; f++) {
        Six s;
#pragma unroll
        for (int m = 0; m < 6; m++) s.v[m] = char_field_at<D, A>(vat, ra, ca, f, m);
        double sb[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        reconstruct_pair_adj<SCHEME, P, Q>(s, ld, hat(f), hat(NEQ + f), sb);
        if (f == 0 || f == IP) {
            const double sgn_ = (f == 0) ? -0.5 : 0.5;
#pragma unroll
            for (int m = 0; m < 6; m++) {
                rab += sb[m] * sgn_ * ca * vat(m, 1 + A);
                cab += sb[m] * sgn_ * ra * vat(m, 1 + A);
            }
        } else if (f == 1) {
#pragma unroll
            for (int m = 0; m < 6; m++) cab += sb[m] * vat(m, IP) * 2.0 * ic3;
        }
#pragma unroll
        for (int m = 0; m < 6; m++) gout(f * 6 + m, sb[m]);
    }
    gout(6 * NEQ, rab);
    gout(6 * NEQ + 1, cab);
}

// Share of stencil point m (0..5, cells i-3..i+2) of a face record (read
// through gat(c)): the characteristic projection's adjoint, added to Vbm (the
// point's primitive-variable adjoint) and cb (its sound-speed adjoint).  flag
// is the hand-over's last entry (1 reconstructed, 0 fallback); ra, ca are the
// face's projection averages.
template <int D, int A, class GAt>
HM_HD void face_record_to_point(double flag, const GAt& gat, double ra, double ca, int m, double* Vbm, double& cb) {
    constexpr int NEQ = D + 2;
    constexpr int IP = D + 1;
    if (flag == 0.0) {
        if (m == 2)
            for (int e = 0; e < NEQ; e++) Vbm[e] += gat(e);
        if (m == 3)
            for (int e = 0; e < NEQ; e++) Vbm[e] += gat(NEQ + e);
        return;
    }
    const double g0 = gat(m), gl = gat(IP * 6 + m), g1 = gat(6 + m);
    Vbm[1 + A] += g0 * -0.5 * ra * ca;
    Vbm[IP] += 0.5 * g0;
    Vbm[0] += g1;
    Vbm[IP] -= g1 * rcp_rn(ca * ca);
    int t = 2;
    for (int b = 0; b < D; b++) {
        if (b == A) continue;
        Vbm[1 + b] += gat(t * 6 + m);
        t++;
    }
    Vbm[1 + A] += gl * 0.5 * ra * ca;
    Vbm[IP] += 0.5 * gl;
    if (m == 2 || m == 3) {
        Vbm[0] += 0.5 * gat(6 * NEQ);
        cb += 0.5 * gat(6 * NEQ + 1);
    }
}

// Adjoint of the whole midpoint-flux computation of one face along axis A
// (both halves and the per-point projection adjoint): adds d(flux)/dV * mb to
// Vb and d(flux)/dc * mb to cLb, cRb.
template <int D, int A, int SCHEME, int P, int Q>
HM_HD void face_adj_single(const double (&V)[6][D + 2], double cL, double cR, double gam, const LDConst& ld,
                           bool use_hll, const double* mb, double (&Vb)[6][D + 2], double& cLb, double& cRb) {
    double h[face_adj_h_size<D>()];
    face_adj_single_flux<D, A, SCHEME, P, Q>(V, cL, cR, gam, ld, use_hll, mb, h);
    double g[face_adj_g_size<D>()];
    for (int c = 0; c < face_adj_g_size<D>(); c++) g[c] = 0.0;
    face_adj_single_recon<D, A, SCHEME, P, Q>([&](int m, int e) { return V[m][e]; }, cL, cR, ld,
                                              [&](int c) { return h[c]; }, [&](int c, double v) { g[c] = v; });
    const double ra = double(1) / double(2) * (V[2][0] + V[3][0]);
    const double ca = double(1) / double(2) * (cL + cR);
    for (int m = 0; m < 6; m++) {
        double cb = 0.0;
        face_record_to_point<D, A>(h[2 * (D + 2) + 2], [&](int c) { return g[c]; }, ra, ca, m, Vb[m], cb);
        if (m == 2) cLb += cb;
        if (m == 3) cRb += cb;
    }
}

}  // namespace hm
