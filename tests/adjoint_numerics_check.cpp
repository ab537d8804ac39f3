// Host-side check of the hand-written adjoint of one midpoint face
// (hamers_amd/csrc/hm_adjoint_face.hpp) against central differences of the
// forward code it reverses (hm_numerics.hpp, hm_face_single.hpp).
//
// For random inputs x, directions dx and output weights g, the adjoint
// gives <J^T g, dx>; central differences give <g, (f(x + h dx) - f(x - h dx)) / 2h>,
// Richardson-extrapolated from steps h and h/2 (and h/2, h/4).  A trial whose
// two extrapolants disagree straddles a kink (a branch of the scheme) and is
// skipped.  Prints one summary line per
// group and exits nonzero if any compared trial disagrees.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>

#include "hm_adjoint_face.hpp"
#include "hm_face_single.hpp"

using namespace hm;

static std::mt19937_64 rng(12345);
static double U(double a, double b) { return std::uniform_real_distribution<double>(a, b)(rng); }

struct Tally {
    const char* name;
    int compared = 0, skipped = 0, failed = 0, loose = 0;
    double worst = 0.0;
    template <class F>
    void add(F&& f, double h, double ad, double scale) {
        auto cd = [&](double s) { return (f(s) - f(-s)) / (2 * s); };
        const double d1 = cd(h), d2 = cd(h / 2), d4 = cd(h / 4);
        const double fd2 = (4 * d2 - d1) / 3, fd1 = (4 * d4 - d2) / 3;
        const double kink = std::fabs(fd1 - fd2) > 1e-7 * (std::fabs(fd1) + scale);
        if (kink) {
            skipped++;
            return;
        }
        compared++;
        const double err = std::fabs(fd1 - ad) / (std::fabs(fd1) + std::fabs(ad) + 1e-3 * scale);
        if (err > worst) worst = err;
        // Difference quotients of the stiffest weights (smoothness indicators
        // near epsilon) carry errors near 1e-6 of the scale; a wrong adjoint
        // term shows as an O(1) error.
        if (err > 1e-6) loose++;
        if (err > 1e-4) {
            failed++;
            if (failed <= 3) std::printf("  %s mismatch: fd %.12e ad %.12e\n", name, fd1, ad);
        }
    }
    bool report() const {
        std::printf("%-28s compared %5d skipped %4d loose %3d failed %d worst %.2e\n", name, compared, skipped, loose,
                    failed, worst);
        return failed == 0 && compared >= 50 && loose * 100 < compared;
    }
};

static Six random_stencil(int kind) {
    Six s;
    switch (kind % 4) {
        case 0:  // smooth
        {
            const double a = U(-1, 1), b = U(-1, 1), c = U(-0.3, 0.3), d = U(-0.1, 0.1);
            for (int m = 0; m < 6; m++) {
                const double x = m - 2.5;
                s.v[m] = a + b * x + c * x * x + d * x * x * x;
            }
            break;
        }
        case 1:  // random
            for (int m = 0; m < 6; m++) s.v[m] = U(-1, 1);
            break;
        case 2:  // step
        {
            const int j = 1 + int(U(0, 4.999));
            const double lo = U(-1, 1), hi = U(-1, 1);
            for (int m = 0; m < 6; m++) s.v[m] = (m < j ? lo : hi) + 1e-3 * U(-1, 1);
            break;
        }
        default:  // near-constant
        {
            const double a = U(-1, 1);
            for (int m = 0; m < 6; m++) s.v[m] = a + 1e-4 * U(-1, 1);
        }
    }
    return s;
}

template <int SCHEME, int P, int Q>
static bool check_weno(const char* name, const LDConst& k) {
    Tally t{name};
    for (int trial = 0; trial < 4000; trial++) {
        const Six s = random_stencil(trial);
        double dir[6];
        double scale = 0.0;
        for (int m = 0; m < 6; m++) {
            dir[m] = U(-1, 1);
            scale = std::fmax(scale, std::fabs(s.v[m]));
        }
        scale = std::fmax(scale, 1e-3);
        const double gm = U(-1, 1), gp = U(-1, 1);
        auto f = [&](double h) {
            Six x = s;
            for (int m = 0; m < 6; m++) x.v[m] += h * dir[m];
            double um, up;
            reconstruct_pair<SCHEME, P, Q>(x, k, um, up);
            return gm * um + gp * up;
        };
        const double h = 1e-4 * scale;
        double sb[6] = {0, 0, 0, 0, 0, 0};
        reconstruct_pair_adj<SCHEME, P, Q>(s, k, gm, gp, sb);
        double ad = 0.0;
        for (int m = 0; m < 6; m++) ad += sb[m] * dir[m];
        t.add(f, h, ad, std::fabs(gm) + std::fabs(gp));
    }
    return t.report();
}

template <int D, int A, int SCHEME, int P, int Q>
static bool check_face(const char* name, bool use_hll, int kinds) {
    constexpr int NEQ = D + 2;
    const double gam = 1.4;
    const LDConst ld{2, 4, 1e9, 35.0};
    Tally t{name};
    for (int trial = 0; trial < 1500; trial++) {
        double V[6][NEQ];
        const int kind = trial % kinds;
        // rho, u.., p per stencil cell
        for (int m = 0; m < 6; m++) {
            if (kind == 0) {  // smooth
                const double x = m - 2.5;
                V[m][0] = 1.0 + 0.2 * std::sin(0.3 * x + trial);
                for (int b = 0; b < D; b++) V[m][1 + b] = 0.3 * std::cos(0.2 * x + b + trial);
                V[m][NEQ - 1] = 1.0 + 0.1 * std::cos(0.25 * x + 2 * trial);
            } else if (kind == 1) {  // random states
                V[m][0] = U(0.5, 1.5);
                for (int b = 0; b < D; b++) V[m][1 + b] = U(-1, 1);
                V[m][NEQ - 1] = U(0.5, 1.5);
            } else {  // strong jump that may trigger the positivity fallback
                const bool hi = m >= 3;
                V[m][0] = hi ? 1e-3 * U(0.5, 1.5) : U(1, 2);
                for (int b = 0; b < D; b++) V[m][1 + b] = U(-2, 2);
                V[m][NEQ - 1] = hi ? 1e-3 * U(0.5, 1.5) : U(1, 2);
            }
        }
        double cs[6];
        for (int m = 0; m < 6; m++) cs[m] = std::sqrt(gam * V[m][NEQ - 1] / V[m][0]);
        double dir[6][NEQ], dc[2], g[NEQ];
        for (int m = 0; m < 6; m++)
            for (int e = 0; e < NEQ; e++) dir[m][e] = U(-1, 1) * (e == 0 || e == NEQ - 1 ? V[m][e] : 1.0);
        dc[0] = U(-1, 1) * cs[2];
        dc[1] = U(-1, 1) * cs[3];
        for (int e = 0; e < NEQ; e++) g[e] = U(-1, 1);
        auto f = [&](double h) {
            double X[6][NEQ];
            for (int m = 0; m < 6; m++)
                for (int e = 0; e < NEQ; e++) X[m][e] = V[m][e] + h * dir[m][e];
            double Fo[NEQ];
            face_flux_single<D, A, SCHEME, P, Q>(X, cs[2] + h * dc[0], cs[3] + h * dc[1], gam, ld, use_hll, Fo);
            double r = 0.0;
            for (int e = 0; e < NEQ; e++) r += g[e] * Fo[e];
            return r;
        };
        const double h = 2e-6;
        double Vb[6][NEQ];
        for (int m = 0; m < 6; m++)
            for (int e = 0; e < NEQ; e++) Vb[m][e] = 0.0;
        double cLb = 0.0, cRb = 0.0;
        face_adj_single<D, A, SCHEME, P, Q>(V, cs[2], cs[3], gam, ld, use_hll, g, Vb, cLb, cRb);
        double ad = cLb * dc[0] + cRb * dc[1];
        for (int m = 0; m < 6; m++)
            for (int e = 0; e < NEQ; e++) ad += Vb[m][e] * dir[m][e];
        double scale = 0.0;
        for (int e = 0; e < NEQ; e++) scale += std::fabs(g[e]);
        t.add(f, h, ad, scale);
    }
    return t.report();
}

template <int D, int A>
static bool check_riemann(const char* name, bool use_hll) {
    constexpr int NEQ = D + 2;
    const double gam = 1.4;
    Tally t{name};
    for (int trial = 0; trial < 3000; trial++) {
        double VL[NEQ], VR[NEQ], dL[NEQ], dR[NEQ], g[NEQ];
        const int kind = trial % 3;
        for (int x = 0; x < 2; x++) {
            double* V = x ? VR : VL;
            V[0] = U(0.2, 2.0);
            for (int b = 0; b < D; b++) V[1 + b] = (kind == 2) ? U(-3, 3) : U(-1, 1);
            V[NEQ - 1] = U(0.2, 2.0);
        }
        if (kind == 1)  // a weak jump
            for (int e = 0; e < NEQ; e++) VR[e] = VL[e] * (1.0 + 0.05 * U(-1, 1)) + (e >= 1 && e <= D ? 0.05 * U(-1, 1) : 0.0);
        for (int e = 0; e < NEQ; e++) {
            dL[e] = U(-1, 1);
            dR[e] = U(-1, 1);
            g[e] = U(-1, 1);
        }
        auto f = [&](double h) {
            double XL[NEQ], XR[NEQ], Fo[NEQ];
            for (int e = 0; e < NEQ; e++) {
                XL[e] = VL[e] + h * dL[e];
                XR[e] = VR[e] + h * dR[e];
            }
            riemann_flux_single<D, A>(XL, XR, gam, use_hll, Fo);
            double r = 0.0;
            for (int e = 0; e < NEQ; e++) r += g[e] * Fo[e];
            return r;
        };
        double bL[NEQ] = {}, bR[NEQ] = {};
        riemann_single_adj<D, A>(VL, VR, gam, use_hll, g, bL, bR);
        double ad = 0.0;
        for (int e = 0; e < NEQ; e++) ad += bL[e] * dL[e] + bR[e] * dR[e];
        double scale = 0.0;
        for (int e = 0; e < NEQ; e++) scale += std::fabs(g[e]);
        t.add(f, 1e-5, ad, scale);
    }
    return t.report();
}

int main() {
    bool ok = true;
    const LDConst k{2, 4, 1e9, 35.0};
    const LDConst krt{3, 6, 1e9, 35.0};
    ok &= check_weno<kWCNS6_LD, 2, 4>("WCNS6-LD pair (2,4)", k);
    ok &= check_weno<kWCNS6_LD, -1, -1>("WCNS6-LD pair (3,6) runtime", krt);
    ok &= check_weno<kWCNS5_JS, 2, 4>("WCNS5-JS pair (2)", k);
    ok &= check_weno<kWCNS5_JS, -1, -1>("WCNS5-JS pair (3) runtime", krt);
    ok &= check_riemann<1, 0>("riemann 1D HLLC", false);
    ok &= check_riemann<2, 0>("riemann 2D x HLLC", false);
    ok &= check_riemann<2, 1>("riemann 2D y HLLC", false);
    ok &= check_riemann<2, 0>("riemann 2D x HLLC-HLL", true);
    ok &= check_riemann<2, 1>("riemann 2D y HLLC-HLL", true);
    ok &= check_riemann<3, 0>("riemann 3D x HLLC-HLL", true);
    ok &= check_riemann<3, 2>("riemann 3D z HLLC-HLL", true);
    ok &= check_face<1, 0, kWCNS6_LD, 2, 4>("face 1D x WCNS6-LD HLLC", false, 3);
    ok &= check_face<2, 0, kWCNS6_LD, 2, 4>("face 2D x WCNS6-LD HLLC", false, 3);
    ok &= check_face<2, 1, kWCNS6_LD, 2, 4>("face 2D y WCNS6-LD HLLC", false, 3);
    ok &= check_face<2, 0, kWCNS6_LD, 2, 4>("face 2D x WCNS6-LD HLLC-HLL", true, 3);
    ok &= check_face<2, 1, kWCNS5_JS, 2, 4>("face 2D y WCNS5-JS HLLC-HLL", true, 3);
    ok &= check_face<3, 0, kWCNS5_JS, 2, 4>("face 3D x WCNS5-JS HLLC", false, 3);
    ok &= check_face<3, 1, kWCNS6_LD, 2, 4>("face 3D y WCNS6-LD HLLC-HLL", true, 3);
    ok &= check_face<3, 2, kWCNS6_LD, 2, 4>("face 3D z WCNS6-LD HLLC-HLL", true, 3);
    std::printf("%s\n", ok ? "ALL OK" : "FAILURES");
    return ok ? 0 : 1;
}
