// hm_numerics.hpp -- pointwise numerics of the convective-flux RHS on the GPU.
//
// Every expression keeps the association order and reciprocal forms of the
// reference source lines it cites, so that a build without multiply-add
// contraction (-ffp-contract=off) rounds exactly as the reference's
// uncontracted x86-64 build does.  The knife-edge predicates of the scheme
// (R_tau > alpha_tau, s* > 0, sensor > 0.65, positivity) then take the same
// branch as the reference on identical inputs.
#pragma once

// Host+device so that host-side checks can compile the same pointwise code.
#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define HM_HD __host__ __device__ __forceinline__
#else
#include <cmath>
#define HM_HD inline
#endif

namespace hm {

constexpr double kEps = 1.0e-15;  // HAMERS_EPSILON, include/HAMeRS_config.hpp.in:16

// Correctly rounded FP64 division in fewer instructions than the compiler's
// general sequence, for operands whose quotient and remainder are normal
// numbers (every quotient of the scheme: denominators are state values or
// sums with a positive floor).  rcp_rn refines the hardware reciprocal
// estimate by two Newton-Raphson steps r <- r + r (1 - b r); quot forms
// q0 = a r and returns q0 + r (a - b q0) with two FMAs (Markstein's
// correction, IBM J. Res. Develop. 34 (1990) 111-119), which rounds to the
// nearest FP64 number like a / b.  One reciprocal serves every quotient with
// the same denominator.  It skips the general sequence's range scaling and
// special-value fixup, so a zero dividend gives +0 for -0 and a zero divisor
// gives NaN; neither reaches a branch of the scheme.  Checked bit for bit
// against a / b on the GPU over 1.4e11 random operand pairs (exponents
// -80..79, all-ones significands included) and by the parity tests against
// the oracle's IEEE divisions.  Host builds divide.
HM_HD double rcp_rn(double b) {
#if defined(__HIP_DEVICE_COMPILE__)
    double r = __builtin_amdgcn_rcp(b);
    double e = fma(-b, r, 1.0);
    r = fma(r, e, r);
    e = fma(-b, r, 1.0);
    return fma(r, e, r);
#else
    return double(1) / b;
#endif
}

HM_HD double quot(double a, double b, double r) {
#if defined(__HIP_DEVICE_COMPILE__)
    const double q0 = a * r;
    return fma(fma(-b, q0, a), r, q0);
#else
    (void)r;
    return a / b;
#endif
}

HM_HD double div_rn(double a, double b) { return quot(a, b, rcp_rn(b)); }

// ipow of ConvectiveFluxReconstructorWCNS6-LD-HLLC-HLL.cpp:9-18: repeated
// multiplication, so base^4 is ((b*b)*b)*b.  E < 0 means "exponent known only
// at run time".
template <int E>
HM_HD double rep_pow(double b, int e_rt) {
    if constexpr (E > 0) {
        double r = b;
#pragma unroll
        for (int i = 1; i < E; i++) r *= b;
        return r;
    } else {
        double r = b;
        for (int i = 1; i < e_rt; i++) r *= b;
        return r;
    }
}

// The six stencil values of one characteristic field at one face, cells
// i-3 .. i+2 of the face between cells i-1 and i.
struct Six {
    double v[6];
};

HM_HD Six flip(const Six& s) { return Six{{s.v[5], s.v[4], s.v[3], s.v[2], s.v[1], s.v[0]}}; }

// Smoothness indicators of the three 3-point sub-stencils
// (WCNS6-LD-HLLC-HLL.cpp:55-68, identical in WCNS5-JS-HLLC-HLL.cpp:32-44).
HM_HD void beta3(const Six& s, double& b0, double& b1, double& b2) {
    b0 = double(1) / double(3) *
         (s.v[0] * (double(4) * s.v[0] - double(19) * s.v[1] + double(11) * s.v[2]) +
          s.v[1] * (double(25) * s.v[1] - double(31) * s.v[2]) + double(10) * s.v[2] * s.v[2]);
    b1 = double(1) / double(3) *
         (s.v[1] * (double(4) * s.v[1] - double(13) * s.v[2] + double(5) * s.v[3]) +
          double(13) * s.v[2] * (s.v[2] - s.v[3]) + double(4) * s.v[3] * s.v[3]);
    b2 = double(1) / double(3) *
         (s.v[2] * (double(10) * s.v[2] - double(31) * s.v[3] + double(11) * s.v[4]) +
          s.v[3] * (double(25) * s.v[3] - double(19) * s.v[4]) + double(4) * s.v[4] * s.v[4]);
}

// Six-point smoothness indicator (WCNS6-LD-HLLC-HLL.cpp:70-83).
HM_HD double beta6(const Six& s) {
    return double(1) / double(232243200) *
           (s.v[0] * (double(525910327) * s.v[0] - double(4562164630) * s.v[1] + double(7799501420) * s.v[2] -
                    double(6610694540) * s.v[3] + double(2794296070) * s.v[4] - double(472758974) * s.v[5]) +
            double(5) * s.v[1] *
                (double(2146987907) * s.v[1] - double(7722406988) * s.v[2] + double(6763559276) * s.v[3] -
                 double(2926461814) * s.v[4] + double(503766638) * s.v[5]) +
            double(20) * s.v[2] *
                (double(1833221603) * s.v[2] - double(3358664662) * s.v[3] + double(1495974539) * s.v[4] -
                 double(263126407) * s.v[5]) +
            double(20) * s.v[3] * (double(1607794163) * s.v[3] - double(1486026707) * s.v[4] + double(268747951) * s.v[5]) +
            double(5) * s.v[4] * (double(1432381427) * s.v[4] - double(536951582) * s.v[5]) +
            double(263126407) * s.v[5] * s.v[5]);
}

// Shock measure sigma (WCNS6-LD-HLLC-HLL.cpp:24-41); both sides use the
// unflipped stencil.
HM_HD double ld_sigma(const Six& s) {
    const double a1 = s.v[2] - s.v[1];
    const double a2 = s.v[3] - s.v[2];
    const double a3 = s.v[4] - s.v[3];
    const double t1 = div_rn(fabs(a1 - a2), fabs(a1) + fabs(a2) + kEps);
    const double t2 = div_rn(fabs(a2 - a3), fabs(a2) + fabs(a3) + kEps);
    return fmax(t1, t2);
}

struct LDConst {
    int p, q;
    double C, alpha_tau;
};

// One side of the WCNS6-LD nonlinear midpoint interpolation
// (WCNS6-LD-HLLC-HLL.cpp:133-223; the plus side 229-319 is this on flip(s)).
// The upwind weights u_k and the shock measure sigma enter only where the
// ratio R_tau = tau6 / (beta_avg + eps) exceeds alpha_tau, so they are formed
// only there (the same expressions, evaluated on demand): sigma of the
// unflipped stencil s0 is computed once per field, into sig (< 0: not yet).
template <int P, int Q>
HM_HD double ld_side(const Six& s, const Six& s0, double& sig, const LDConst& k) {
    double b0, b1, b2;
    beta3(s, b0, b1, b2);
    const double b3 = beta6(s);

    // e_k = b_k + eps divides both tau5 and tau6: one reciprocal each.
    const double e0 = b0 + kEps, e1 = b1 + kEps, e2 = b2 + kEps, e3 = b3 + kEps;
    const double r0 = rcp_rn(e0), r1 = rcp_rn(e1), r2 = rcp_rn(e2);

    const double bavg = double(1) / double(8) * (b0 + b2 + double(6) * b1);
    const double tau6 = fabs(b3 - bavg);
    double w0 = double(1) / double(32) * (k.C + rep_pow<Q>(quot(tau6, e0, r0), k.q));
    double w1 = double(15) / double(32) * (k.C + rep_pow<Q>(quot(tau6, e1, r1), k.q));
    double w2 = double(15) / double(32) * (k.C + rep_pow<Q>(quot(tau6, e2, r2), k.q));
    double w3 = double(1) / double(32) * (k.C + rep_pow<Q>(div_rn(tau6, e3), k.q));
    const double ws = w0 + w1 + w2 + w3;
    const double rws = rcp_rn(ws);
    w0 = quot(w0, ws, rws);
    w1 = quot(w1, ws, rws);
    w2 = quot(w2, ws, rws);
    w3 = quot(w3, ws, rws);

    if (div_rn(tau6, bavg + kEps) > k.alpha_tau) {
        const double tau5 = fabs(b0 - b2);
        double u0 = double(1) / double(16) * (double(1) + rep_pow<P>(quot(tau5, e0, r0), k.p));
        double u1 = double(5) / double(8) * (double(1) + rep_pow<P>(quot(tau5, e1, r1), k.p));
        double u2 = double(5) / double(16) * (double(1) + rep_pow<P>(quot(tau5, e2, r2), k.p));
        const double us = u0 + u1 + u2;
        const double rus = rcp_rn(us);
        u0 = quot(u0, us, rus);
        u1 = quot(u1, us, rus);
        u2 = quot(u2, us, rus);
        if (sig < double(0)) sig = ld_sigma(s0);
        const double sigma = sig;
        w0 = sigma * u0 + (double(1) - sigma) * w0;
        w1 = sigma * u1 + (double(1) - sigma) * w1;
        w2 = sigma * u2 + (double(1) - sigma) * w2;
        w3 = (double(1) - sigma) * w3;
    }
    return double(3) / double(8) * w0 * s.v[0] + (-double(10) / double(8) * w0 - double(1) / double(8) * w1) * s.v[1] +
           (double(15) / double(8) * w0 + double(6) / double(8) * w1 + double(3) / double(8) * w2) * s.v[2] +
           (double(3) / double(8) * w1 + double(6) / double(8) * w2 + double(15) / double(8) * w3) * s.v[3] +
           (-double(1) / double(8) * w2 - double(10) / double(8) * w3) * s.v[4] + double(3) / double(8) * w3 * s.v[5];
}

// One side of WCNS5-JS (WCNS5-JS-HLLC-HLL.cpp:78-118; plus side 124-164 on flip(s)).
template <int P>
HM_HD double js_side(const Six& s, int p_rt) {
    double b0, b1, b2;
    beta3(s, b0, b1, b2);
    double w0 = div_rn(double(1.0) / double(16), rep_pow<P>((b0 + kEps), p_rt));
    double w1 = div_rn(double(5.0) / double(8), rep_pow<P>((b1 + kEps), p_rt));
    double w2 = div_rn(double(5.0) / double(16), rep_pow<P>((b2 + kEps), p_rt));
    const double ws = w0 + w1 + w2;
    const double rws = rcp_rn(ws);
    w0 = quot(w0, ws, rws);
    w1 = quot(w1, ws, rws);
    w2 = quot(w2, ws, rws);
    return double(3) / double(8) * w0 * s.v[0] + (-double(10) / double(8) * w0 - double(1) / double(8) * w1) * s.v[1] +
           (double(15) / double(8) * w0 + double(6) / double(8) * w1 + double(3) / double(8) * w2) * s.v[2] +
           (double(3) / double(8) * w1 + double(6) / double(8) * w2) * s.v[3] - double(1) / double(8) * w2 * s.v[4];
}

// Scheme selector: both sides of one characteristic field.
enum { kWCNS5_JS = 0, kWCNS6_LD = 1 };

template <int SCHEME, int P, int Q>
HM_HD void reconstruct_pair(const Six& s, const LDConst& k, double& um, double& up) {
    if constexpr (SCHEME == kWCNS6_LD) {
        double sig = double(-1);  // sigma >= 0, formed on the first side that blends
        // The two sides one after the other (a runtime loop), so that only one
        // side's indicators and weights are live at a time.
#pragma unroll 1
        for (int side = 0; side < 2; side++) {
            Six ss;
#pragma unroll
            for (int m = 0; m < 6; m++) ss.v[m] = side ? s.v[5 - m] : s.v[m];
            const double r = ld_side<P, Q>(ss, s, sig, k);
            if (side)
                up = r;
            else
                um = r;
        }
    } else {
        um = js_side<P>(s, k.p);
        up = js_side<P>(flip(s), k.p);
    }
}

}  // namespace hm
