// hamers_oracle.cpp -- CPU restatement of HAMeRS' convective-flux RHS hot path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the parity oracle: only tests/,
// __graft_entry__.smoke() and bench.py's CPU-baseline leg may load it, and
// only as the checker (or as the timed CPU baseline).  The product path
// (hamers_amd/) never calls it.
//
// It restates, in operation order, the reference loops of
//   src/flow/convective_flux_reconstructors/WCNS56/ConvectiveFluxReconstructorWCNS56-HLLC-HLL.cpp:39-2656
//   src/flow/convective_flux_reconstructors/WCNS56/ConvectiveFluxReconstructorWCNS6-LD-HLLC-HLL.cpp:9-319
//   src/flow/convective_flux_reconstructors/WCNS56/ConvectiveFluxReconstructorWCNS5-JS-HLLC-HLL.cpp:9-164
//   src/flow/flow_models/single-species/* (derived data, projection, bounds, Riemann solvers)
//   src/flow/flow_models/four-eqn_conservative/*, five-eqn_Allaire/* (the same, multi-species)
//   src/util/mixing_rules/equations_of_state/ideal_gas/* (mixture EOS)
//   src/util/derivatives/DerivativeFirstOrder.cpp:216-228 (2nd-order central first derivative)
//   src/apps/Euler/Euler.cpp:1003-1680 (Runge-Kutta stage update)
//   src/algs/integrator/RungeKuttaLevelIntegrator.cpp:1672-1745, 3907-3924 (SSP-RK3 stage loop, weights)
// pass by pass over whole-patch arrays, as the reference does, so that it is
// structurally independent of the fused GPU kernels it checks.  Each
// arithmetic expression keeps the reference's expression tree (association
// order, reciprocal-then-multiply forms) so that an uncontracted build
// (-ffp-contract=off, no fast-math) reproduces the reference's rounding.
//
// Layouts (SAMRAI CellData / SideData, x fastest, component-major):
//   Q  : neq components over the ghost box (n+2g)^d
//   F_a: neq components over the axis-a face box (n_a+1) x prod_{b!=a} n_b, 0 ghosts
//   S  : neq components over the interior box, accumulated (+=)
//
// Pinned by tests/test_oracle.py (closed forms, invariants, convergence orders
// of tests/2D_convergence_test_single_species/convergence_test.py:13-16).

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>
#include <algorithm>
#ifdef _OPENMP
#include <omp.h>
#endif

#define HMO_EPS 1.0e-15  // include/HAMeRS_config.hpp.in:16 (HAMERS_EPSILON)

extern "C" {

typedef struct hmo_params {
    int dim;           // 1, 2 or 3
    int model;         // 0 = SINGLE_SPECIES, 1 = FOUR_EQN_CONSERVATIVE, 2 = FIVE_EQN_ALLAIRE
    int num_species;   // 1 for SINGLE_SPECIES
    int scheme;        // 0 = WCNS5_JS_HLLC_HLL, 1 = WCNS6_LD_HLLC_HLL
    int p, q;          // WCNS exponents (defaults 2, 4)
    double C;          // WCNS6-LD constant C (default 1e9)
    double alpha_tau;  // WCNS6-LD constant alpha_tau (default 35)
    double gamma[8];   // species ratios of specific heats
    double R[8];       // species gas constants (four-eqn mixing)
    int mutation;      // test-only: plants a known mistake (0 = none)
} hmo_params;

// Constant transport properties per species (the "constant" mixing rules of
// equations_of_{shear_viscosity,bulk_viscosity,thermal_conductivity,
// mass_diffusivity}/constant/*MixingRulesConstant.cpp): mu_i, mu_v,i, kappa_i,
// D_i and the molecular weights M_i of the 1/sqrt(M_i) mixing weights.
typedef struct hmo_transport {
    double mu[8];
    double mu_v[8];
    double kappa[8];
    double D[8];
    double M[8];
} hmo_transport;

}  // extern "C"

namespace {

constexpr double EPS = HMO_EPS;

// ---------------------------------------------------------------------------
// Branch tape (single-species path, test instrumentation).  The scheme is only
// piecewise smooth: the Ducros sensor's HLLC/HLLC-HLL choice, WCNS6-LD's
// R_tau > alpha_tau switch and the bounds fallback jump at thresholds (the
// HLLC-HLL |du| < eps switch too, at rounding-level jumps only), and the
// HLLC side s* > 0 and the fmin/fmax wave speeds have kinks (exact ties are
// common where a field is uniform along a sweep).  A discrete adjoint
// differentiates the smooth piece selected at the evaluation point; central
// differences that cross a switch do not converge to it.  Mode 1 records every
// decision in evaluation order, mode 2 replays the record, so that differences
// of a replayed run are differences of that one smooth piece.  Mode 0 is the
// plain scheme (bit-identical to the code without the tape).  The tape is a
// single global sequence: use it from one thread only.
// ---------------------------------------------------------------------------
// Every decision belongs to a category (kTape*); a replay can leave chosen
// categories free (decided afresh, their tape entries still consumed), which
// measures how much of the gap between free and replayed differences one kind
// of switch accounts for.
enum {
    kTapeOther = 0,
    kTapeRtau = 1,      // WCNS6-LD R_tau > alpha_tau
    kTapeSigma = 2,     // WCNS6-LD sigma = fmax(t1, t2)
    kTapeWaveSpeed = 3, // sL = fmin, sR = fmax
    kTapeSpm = 4,       // s_minus = fmin(0, sL), s_plus = fmax(0, sR)
    kTapeSstar = 5,     // HLLC side s* > 0
    kTapeHll = 6,       // HLL upwind sL > 0, sR < 0
    kTapeFallback = 7,  // bounds fallback
    kTapeSensor = 8,    // sensor > 0.65
};
struct BranchTape {
    int mode = 0;
    std::vector<uint64_t> bits;  // one bit per decision
    size_t size = 0;             // decisions recorded
    size_t pos = 0;
    bool overrun = false;
    int free_mask = 0;  // replay: bit c set = category c decided afresh
};
BranchTape g_tape;

inline int tape_sel(int natural, int cat) {
    if (g_tape.mode == 1) {
        if ((g_tape.size & 63) == 0) g_tape.bits.push_back(0);
        if (natural) g_tape.bits.back() |= uint64_t(1) << (g_tape.size & 63);
        g_tape.size++;
    } else if (g_tape.mode == 2) {
        if (g_tape.pos >= g_tape.size) {
            g_tape.overrun = true;
            return natural;
        }
        const int r = int((g_tape.bits[g_tape.pos >> 6] >> (g_tape.pos & 63)) & 1);
        g_tape.pos++;
        return ((g_tape.free_mask >> cat) & 1) ? natural : r;
    }
    return natural;
}

// |x| is continuous and stays untaped: central differences of |x| at x = 0
// give 0, the derivative the adjoint takes there, while a frozen sign would let
// WCNS6-LD's sigma denominator |a1| + |a2| + eps pass through zero.
inline double t_abs(double x) { return std::fabs(x); }
// fmin(a, b) / fmax(a, b); a tie selects a, as the adjoint does.
inline double t_min(double a, double b, int cat) {
    if (g_tape.mode == 0) return std::fmin(a, b);
    return tape_sel(a <= b ? 1 : 0, cat) ? a : b;
}
inline double t_max(double a, double b, int cat) {
    if (g_tape.mode == 0) return std::fmax(a, b);
    return tape_sel(a >= b ? 1 : 0, cat) ? a : b;
}
// A plain threshold decision.
inline bool t_if(bool c, int cat) { return g_tape.mode == 0 ? c : tape_sel(c ? 1 : 0, cat) != 0; }

// ---------------------------------------------------------------------------
// Pointwise reconstruction (WCNS6-LD: ...WCNS6-LD-HLLC-HLL.cpp:9-319,
// WCNS5-JS: ...WCNS5-JS-HLLC-HLL.cpp:9-164).  U[0..5] are the six stencil
// values of one characteristic field for one face.
// ---------------------------------------------------------------------------

inline double ipow_ref(double base, int e) {  // repeated multiplication, WCNS6-LD:9-18
    double r = base;
    for (int i = 1; i < e; i++) r *= base;
    return r;
}

// Jiang-Shu style smoothness indicators on sub-stencils (0,1,2), (1,2,3),
// (2,3,4) and the six-point indicator (WCNS6-LD:47-84).  `a` maps the
// stencil: a(k) = U[k] (minus side) or U[5-k] (plus side, WCNS6-LD:90-127).
struct Stencil6 {
    double u0, u1, u2, u3, u4, u5;
};

inline Stencil6 mirror(const Stencil6& s) { return {s.u5, s.u4, s.u3, s.u2, s.u1, s.u0}; }

inline void smoothness3(const Stencil6& s, double& b0, double& b1, double& b2) {
    b0 = double(1) / double(3) *
         (s.u0 * (double(4) * s.u0 - double(19) * s.u1 + double(11) * s.u2) +
          s.u1 * (double(25) * s.u1 - double(31) * s.u2) + double(10) * s.u2 * s.u2);
    b1 = double(1) / double(3) *
         (s.u1 * (double(4) * s.u1 - double(13) * s.u2 + double(5) * s.u3) +
          double(13) * s.u2 * (s.u2 - s.u3) + double(4) * s.u3 * s.u3);
    b2 = double(1) / double(3) *
         (s.u2 * (double(10) * s.u2 - double(31) * s.u3 + double(11) * s.u4) +
          s.u3 * (double(25) * s.u3 - double(19) * s.u4) + double(4) * s.u4 * s.u4);
}

inline double smoothness6(const Stencil6& s) {
    return double(1) / double(232243200) *
           (s.u0 * (double(525910327) * s.u0 - double(4562164630) * s.u1 + double(7799501420) * s.u2 -
                    double(6610694540) * s.u3 + double(2794296070) * s.u4 - double(472758974) * s.u5) +
            double(5) * s.u1 *
                (double(2146987907) * s.u1 - double(7722406988) * s.u2 + double(6763559276) * s.u3 -
                 double(2926461814) * s.u4 + double(503766638) * s.u5) +
            double(20) * s.u2 *
                (double(1833221603) * s.u2 - double(3358664662) * s.u3 + double(1495974539) * s.u4 -
                 double(263126407) * s.u5) +
            double(20) * s.u3 * (double(1607794163) * s.u3 - double(1486026707) * s.u4 + double(268747951) * s.u5) +
            double(5) * s.u4 * (double(1432381427) * s.u4 - double(536951582) * s.u5) +
            double(263126407) * s.u5 * s.u5);
}

// WCNS6-LD one side.  `s` is already oriented (mirrored for the plus side);
// `sig` is the shock measure computed on the unmirrored stencil
// (WCNS6-LD:24-41 is called with U_array unchanged on both sides).
inline double wcns6ld_side(const Stencil6& s, double sig, int p, int q, double C, double alpha_tau, int mut) {
    double b0, b1, b2;
    smoothness3(s, b0, b1, b2);
    const double b3 = smoothness6(s);

    const double tau5 = t_abs(b0 - b2);
    double wu0 = double(1) / double(16) * (double(1) + ipow_ref(tau5 / (b0 + EPS), p));
    double wu1 = double(5) / double(8) * (double(1) + ipow_ref(tau5 / (b1 + EPS), p));
    double wu2 = double(5) / double(16) * (double(1) + ipow_ref(tau5 / (b2 + EPS), p));
    const double wus = wu0 + wu1 + wu2;
    wu0 = wu0 / wus;
    wu1 = wu1 / wus;
    wu2 = wu2 / wus;

    const double bavg = double(1) / double(8) * (b0 + b2 + double(6) * b1);
    const double tau6 = t_abs(b3 - bavg);
    double w0 = double(1) / double(32) * (C + ipow_ref(tau6 / (b0 + EPS), q));
    double w1 = double(15) / double(32) * (C + ipow_ref(tau6 / (b1 + EPS), q));
    double w2 = double(15) / double(32) * (C + ipow_ref(tau6 / (b2 + EPS), q));
    double w3 = double(1) / double(32) * (C + ipow_ref(tau6 / (b3 + EPS), q));
    const double ws = w0 + w1 + w2 + w3;
    w0 = w0 / ws;
    w1 = w1 / ws;
    w2 = w2 / ws;
    w3 = w3 / ws;

    const double Rtau = tau6 / (bavg + EPS);
    if (t_if(Rtau > alpha_tau, kTapeRtau)) {
        w0 = sig * wu0 + (double(1) - sig) * w0;
        w1 = sig * wu1 + (double(1) - sig) * w1;
        w2 = sig * wu2 + (double(1) - sig) * w2;
        w3 = (double(1) - sig) * w3;
    }
    if (mut == 1) {  // planted mistake: drop the sixth-point contribution
        return double(3) / double(8) * w0 * s.u0 + (-double(10) / double(8) * w0 - double(1) / double(8) * w1) * s.u1 +
               (double(15) / double(8) * w0 + double(6) / double(8) * w1 + double(3) / double(8) * w2) * s.u2 +
               (double(3) / double(8) * w1 + double(6) / double(8) * w2 + double(15) / double(8) * w3) * s.u3 +
               (-double(1) / double(8) * w2 - double(10) / double(8) * w3) * s.u4;
    }
    return double(3) / double(8) * w0 * s.u0 + (-double(10) / double(8) * w0 - double(1) / double(8) * w1) * s.u1 +
           (double(15) / double(8) * w0 + double(6) / double(8) * w1 + double(3) / double(8) * w2) * s.u2 +
           (double(3) / double(8) * w1 + double(6) / double(8) * w2 + double(15) / double(8) * w3) * s.u3 +
           (-double(1) / double(8) * w2 - double(10) / double(8) * w3) * s.u4 + double(3) / double(8) * w3 * s.u5;
}

inline double wcns6ld_sigma(const Stencil6& s) {  // WCNS6-LD:24-41
    const double a1 = s.u2 - s.u1;
    const double a2 = s.u3 - s.u2;
    const double a3 = s.u4 - s.u3;
    const double t1 = t_abs(a1 - a2) / (t_abs(a1) + t_abs(a2) + EPS);
    const double t2 = t_abs(a2 - a3) / (t_abs(a2) + t_abs(a3) + EPS);
    return t_max(t1, t2, kTapeSigma);
}

// WCNS5-JS one side (WCNS5-JS:78-118); the plus side is the mirrored stencil.
inline double wcns5js_side(const Stencil6& s, int p, int mut) {
    double b0, b1, b2;
    smoothness3(s, b0, b1, b2);
    double w0 = double(1.0) / double(16) / ipow_ref((b0 + EPS), p);
    double w1 = double(5.0) / double(8) / ipow_ref((b1 + EPS), p);
    double w2 = double(5.0) / double(16) / ipow_ref((b2 + EPS), p);
    if (mut == 2) w2 = double(5.0) / double(16) / ipow_ref((b1 + EPS), p);  // planted mistake: wrong beta index
    const double ws = w0 + w1 + w2;
    w0 = w0 / ws;
    w1 = w1 / ws;
    w2 = w2 / ws;
    return double(3) / double(8) * w0 * s.u0 + (-double(10) / double(8) * w0 - double(1) / double(8) * w1) * s.u1 +
           (double(15) / double(8) * w0 + double(6) / double(8) * w1 + double(3) / double(8) * w2) * s.u2 +
           (double(3) / double(8) * w1 + double(6) / double(8) * w2) * s.u3 - double(1) / double(8) * w2 * s.u4;
}

// ---------------------------------------------------------------------------
// Patch geometry helpers.
// ---------------------------------------------------------------------------

struct Box {
    int lo[3], hi[3];  // inclusive-exclusive per axis
    int64_t size() const {
        return int64_t(hi[0] - lo[0]) * (hi[1] - lo[1]) * (hi[2] - lo[2]);
    }
};

struct CellArray {  // a component-major array over a box with given ghost width
    int n[3];
    int g[3];
    int ext[3];
    int nc;
    std::vector<double> v;
    void init(const int* nn, int gw, int dim, int ncomp) {
        for (int a = 0; a < 3; a++) {
            n[a] = nn[a];
            g[a] = (a < dim) ? gw : 0;
            ext[a] = n[a] + 2 * g[a];
        }
        nc = ncomp;
        v.assign(size_t(nc) * ext[0] * ext[1] * ext[2], 0.0);
    }
    int64_t vol() const { return int64_t(ext[0]) * ext[1] * ext[2]; }
    int64_t idx(int i, int j, int k) const {
        return (i + g[0]) + int64_t(j + g[1]) * ext[0] + int64_t(k + g[2]) * ext[0] * ext[1];
    }
    double* c(int comp) { return v.data() + comp * vol(); }
    const double* c(int comp) const { return v.data() + comp * vol(); }
};

// Midpoint face array for axis a: faces i_a in [-1, n_a+1], transverse
// indices in [0, n_b).  (The reference allocates these with one ghost on
// every axis, but only this sub-box is ever read, WCNS6-LD:486-496.)
struct FaceArray {
    int a;
    int lo[3], ext[3];
    int nc;
    std::vector<double> v;
    void init(const int* n, int axis, int ncomp) {
        a = axis;
        for (int b = 0; b < 3; b++) {
            lo[b] = (b == axis) ? -1 : 0;
            ext[b] = (b == axis) ? n[b] + 3 : n[b];
        }
        nc = ncomp;
        v.assign(size_t(nc) * vol(), 0.0);
    }
    int64_t vol() const { return int64_t(ext[0]) * ext[1] * ext[2]; }
    int64_t idx(int i, int j, int k) const {
        return (i - lo[0]) + int64_t(j - lo[1]) * ext[0] + int64_t(k - lo[2]) * ext[0] * ext[1];
    }
    double* c(int comp) { return v.data() + comp * vol(); }
    const double* c(int comp) const { return v.data() + comp * vol(); }
};

struct Model {
    int dim, model, ns, neq;
    int nproj;            // number of projection variables
    double gamma[8], R[8];
    bool advective[16];
    bool has_advective;
    // Equation/primitive-variable layout (single-species: rho, u.., p).
    int iv_vel0;  // index of first velocity component in V
    int iv_p;     // index of pressure in V
};

Model make_model(const hmo_params& prm) {
    Model m{};
    m.dim = prm.dim;
    m.model = prm.model;
    m.ns = prm.num_species;
    for (int i = 0; i < 8; i++) {
        m.gamma[i] = prm.gamma[i];
        m.R[i] = prm.R[i];
    }
    if (m.model == 2) {
        m.neq = m.dim + 2 * m.ns;  // FlowModelFiveEqnAllaire.cpp:30
        m.nproj = m.ns + 2;
    } else if (m.model == 1) {
        m.neq = m.ns + m.dim + 1;  // FlowModelBasicUtilitiesFourEqnConservative.hpp:20
        m.nproj = m.ns + 2;
    } else if (m.model == 0) {
        m.ns = 1;
        m.neq = m.dim + 2;  // FlowModelSingleSpecies: rho, rho*u, E
        m.nproj = 2;        // FlowModelBasicUtilitiesSingleSpecies.cpp:3526-3529
        m.iv_vel0 = 1;
        m.iv_p = m.dim + 1;
    } else {
        m.neq = -1;  // other models are restated separately
    }
    m.has_advective = false;
    for (int e = 0; e < 16; e++) m.advective[e] = false;
    return m;
}

// Derived cell data on the 4-ghost box (FlowModelSingleSpecies.cpp:2649-3183,
// EquationOfStateIdealGas.cpp:5477, 5806): V = (rho, u.., p), c, and the node
// convective flux along every axis (FlowModelSingleSpecies.cpp:3233-3883).
struct Derived {
    CellArray V;     // neq components
    CellArray c;     // sound speed
    CellArray Fn[3]; // node convective flux per axis, neq comps
};

void compute_derived_single(const Model& m, const int* n, int g, const double* Q, Derived& D) {
    const int d = m.dim;
    const double gam = m.gamma[0];
    CellArray Qv;  // view description only
    Qv.init(n, g, d, 0);
    D.V.init(n, 4, d, m.neq);
    D.c.init(n, 4, d, 1);
    for (int a = 0; a < d; a++) D.Fn[a].init(n, 4, d, m.neq);
    const int64_t qvol = Qv.vol();
    const int lo2 = (d > 2) ? -4 : 0, hi2 = (d > 2) ? n[2] + 4 : 1;
    const int lo1 = (d > 1) ? -4 : 0, hi1 = (d > 1) ? n[1] + 4 : 1;
    // velocity (FlowModelSingleSpecies.cpp:2649-2848)
    for (int k = lo2; k < hi2; k++)
        for (int j = lo1; j < hi1; j++)
            for (int i = -4; i < n[0] + 4; i++) {
                const int64_t iq = Qv.idx(i, j, k), iv = D.V.idx(i, j, k);
                const double rho = Q[iq];
                D.V.c(0)[iv] = rho;
                for (int b = 0; b < d; b++) D.V.c(1 + b)[iv] = Q[(1 + b) * qvol + iq] / rho;
            }
    // internal energy, pressure, sound speed (FlowModelSingleSpecies.cpp:2849-3179,
    // EquationOfStateIdealGas.cpp:5477, 5806)
    for (int k = lo2; k < hi2; k++)
        for (int j = lo1; j < hi1; j++)
            for (int i = -4; i < n[0] + 4; i++) {
                const int64_t iq = Qv.idx(i, j, k), iv = D.V.idx(i, j, k);
                const double rho = Q[iq];
                const double E = Q[(d + 1) * qvol + iq];
                double ke;
                const double u = D.V.c(1)[iv];
                if (d == 1) {
                    ke = u * u;
                } else if (d == 2) {
                    const double v = D.V.c(2)[iv];
                    ke = u * u + v * v;
                } else {
                    const double v = D.V.c(2)[iv], w = D.V.c(3)[iv];
                    ke = u * u + v * v + w * w;
                }
                const double eps_int = E / rho - double(1) / double(2) * ke;
                const double p = (gam - double(1)) * rho * eps_int;
                D.V.c(d + 1)[iv] = p;
                D.c.c(0)[iv] = std::sqrt(gam * p / rho);
            }
    // node convective flux per axis (FlowModelSingleSpecies.cpp:3329-3331, 3388-3391, 3470-3474, ...)
    for (int a = 0; a < d; a++)
        for (int k = lo2; k < hi2; k++)
            for (int j = lo1; j < hi1; j++)
                for (int i = -4; i < n[0] + 4; i++) {
                    const int64_t iq = Qv.idx(i, j, k), iv = D.V.idx(i, j, k);
                    const double ua = D.V.c(1 + a)[iv];
                    const double p = D.V.c(d + 1)[iv];
                    const double E = Q[(d + 1) * qvol + iq];
                    D.Fn[a].c(0)[iv] = Q[(1 + a) * qvol + iq];
                    for (int b = 0; b < d; b++) {
                        if (b == a)
                            D.Fn[a].c(1 + b)[iv] = ua * Q[(1 + b) * qvol + iq] + p;
                        else
                            D.Fn[a].c(1 + b)[iv] = ua * Q[(1 + b) * qvol + iq];
                    }
                    D.Fn[a].c(d + 1)[iv] = ua * (E + p);
                }
}

// Dilatation and vorticity magnitude on the 2-ghost box from second-order
// central velocity derivatives (WCNS56-HLLC-HLL.cpp:626-731 (2D),
// 1470-1656 (3D); DerivativeFirstOrder.cpp:216-228).
void compute_theta_omega(const Model& m, const int* n, const Derived& D, const double* dx, CellArray& theta,
                         CellArray& Omega) {
    const int d = m.dim;
    theta.init(n, 2, d, 1);
    Omega.init(n, 2, d, 1);
    if (d == 1) return;
    // derivative array du_b/dx_a, layout [b*d + a] as in the reference (dudx, dudy, ..).
    CellArray dv;
    dv.init(n, 2, d, d * d);
    const int lo2 = (d > 2) ? -2 : 0, hi2 = (d > 2) ? n[2] + 2 : 1;
    for (int b = 0; b < d; b++)
        for (int a = 0; a < d; a++) {
            const double* u = D.V.c(1 + b);
            double* out = dv.c(b * d + a);
            const int off[3] = {a == 0, a == 1, a == 2};
            for (int k = lo2; k < hi2; k++)
                for (int j = -2; j < n[1] + 2; j++)
                    for (int i = -2; i < n[0] + 2; i++) {
                        const double uR = u[D.V.idx(i + off[0], j + off[1], k + off[2])];
                        const double uL = u[D.V.idx(i - off[0], j - off[1], k - off[2])];
                        out[dv.idx(i, j, k)] = (double(1) / double(2) * (uR - uL)) / dx[a];
                    }
        }
    for (int k = lo2; k < hi2; k++)
        for (int j = -2; j < n[1] + 2; j++)
            for (int i = -2; i < n[0] + 2; i++) {
                const int64_t id = dv.idx(i, j, k);
                if (d == 2) {
                    const double dudx = dv.c(0)[id], dudy = dv.c(1)[id], dvdx = dv.c(2)[id], dvdy = dv.c(3)[id];
                    theta.c(0)[id] = dudx + dvdy;
                    Omega.c(0)[id] = std::fabs(dvdx - dudy);
                } else {
                    const double dudx = dv.c(0)[id], dudy = dv.c(1)[id], dudz = dv.c(2)[id];
                    const double dvdx = dv.c(3)[id], dvdy = dv.c(4)[id], dvdz = dv.c(5)[id];
                    const double dwdx = dv.c(6)[id], dwdy = dv.c(7)[id], dwdz = dv.c(8)[id];
                    theta.c(0)[id] = dudx + dvdy + dwdz;
                    const double ox = dwdy - dvdz;
                    const double oy = dudz - dwdx;
                    const double oz = dvdx - dudy;
                    Omega.c(0)[id] = std::sqrt(ox * ox + oy * oy + oz * oz);
                }
            }
}

// Midpoint flux of one face for the single-species model: HLLC
// (FlowModelRiemannSolverSingleSpeciesHLLC.cpp:519-1081), replaced in 2D/3D
// by the HLLC-HLL flux (FlowModelRiemannSolverSingleSpeciesHLLC-HLL.cpp:889-1629)
// where the Ducros-type sensor selects it (WCNS56-HLLC-HLL.cpp:1090-1113).
// Sound speed and internal energy on faces come from the ideal-gas EOS
// (EquationOfStateIdealGas.cpp:5806, 6135).  When `ustar` is non-null it
// receives the HLLC normal velocity (FlowModelRiemannSolverSingleSpeciesHLLC.cpp:3213-3220).
void riemann_face_single(int d, int a, double gam, bool use_hll, const double* VL, const double* VR, double* Fout,
                         double* ustar, int mutation = 0) {
    const int neq = d + 2;
    const double rhoL = VL[0], rhoR = VR[0];
    const double pL = VL[neq - 1], pR = VR[neq - 1];
    const double cL = std::sqrt(gam * pL / rhoL);
    const double cR = std::sqrt(gam * pR / rhoR);
    const double epsL = pL / ((gam - double(1)) * rhoL);
    const double epsR = pR / ((gam - double(1)) * rhoR);
    const double uL = VL[1 + a], uR = VR[1 + a];

    const double u_avg = double(1) / double(2) * (uL + uR);
    const double c_av = double(1) / double(2) * (cL + cR);
    const double sL = t_min(u_avg - c_av, uL - cL, kTapeWaveSpeed);
    const double sR = t_max(u_avg + c_av, uR + cR, kTapeWaveSpeed);
    const double sm = (g_tape.mode == 0) ? std::fmin(double(0), sL) : (t_if(sL < double(0), kTapeSpm) ? sL : double(0));
    const double sp = (g_tape.mode == 0) ? std::fmax(double(0), sR) : (t_if(sR > double(0), kTapeSpm) ? sR : double(0));
    const double sstar =
        (pR - pL + rhoL * uL * (sL - uL) - rhoR * uR * (sR - uR)) / (rhoL * (sL - uL) - rhoR * (sR - uR));

    // Conservative states and physical fluxes on both sides.
    auto cons = [&](const double* V, double eps_int, double* Qs) {
        Qs[0] = V[0];
        for (int b = 0; b < d; b++) Qs[1 + b] = V[0] * V[1 + b];
        double ke;
        if (d == 1)
            ke = V[1] * V[1];
        else if (d == 2)
            ke = V[1] * V[1] + V[2] * V[2];
        else
            ke = V[1] * V[1] + V[2] * V[2] + V[3] * V[3];
        Qs[d + 1] = V[0] * (eps_int + double(1) / double(2) * ke);
    };
    auto flux = [&](const double* V, const double* Qs, double* Fs) {
        Fs[0] = Qs[1 + a];
        for (int b = 0; b < d; b++) {
            if (b == a)
                Fs[1 + b] = Qs[1 + a] * V[1 + b] + V[d + 1];
            else
                Fs[1 + b] = Qs[1 + a] * V[1 + b];
        }
        Fs[d + 1] = V[1 + a] * (Qs[d + 1] + V[d + 1]);
    };
    double QL[5], QR[5], FL[5], FR[5];
    cons(VL, epsL, QL);
    cons(VR, epsR, QR);
    flux(VL, QL, FL);
    flux(VR, QR, FR);

    double Fhllc[5];
    const bool left = t_if(sstar > double(0), kTapeSstar);
    {
        const double* V = left ? VL : VR;
        const double* Qs = left ? QL : QR;
        const double* Fs = left ? FL : FR;
        const double sK = left ? sL : sR;
        const double spm = left ? sm : sp;
        const double chi = (sK - V[1 + a]) / (sK - sstar);
        double Qst[5];
        Qst[0] = chi * V[0];
        for (int b = 0; b < d; b++) {
            if (b == a)
                Qst[1 + b] = chi * V[0] * sstar;
            else
                Qst[1 + b] = chi * Qs[1 + b];
        }
        Qst[d + 1] = chi * (Qs[d + 1] + (sstar - V[1 + a]) * (V[0] * sstar + V[d + 1] / (sK - V[1 + a])));
        for (int e = 0; e < neq; e++) Fhllc[e] = Fs[e] + spm * (Qst[e] - Qs[e]);
        if (ustar) *ustar = V[1 + a] + spm * (chi - double(1));
    }
    for (int e = 0; e < neq; e++) Fout[e] = Fhllc[e];
    if (d > 1 && use_hll) {
        // HLL flux on mass and transverse momentum.
        double Fhll[5];
        for (int e = 0; e < neq; e++) Fhll[e] = 0.0;
        auto hll = [&](int e) {
            double r = (sR * FL[e] - sL * FR[e] + sR * sL * (QR[e] - QL[e])) / (sR - sL);
            if (t_if(sL > double(0), kTapeHll)) r = FL[e];
            if (t_if(sR < double(0), kTapeHll)) r = FR[e];
            return r;
        };
        Fhll[0] = hll(0);
        for (int b = 0; b < d; b++)
            if (b != a) Fhll[1 + b] = hll(1 + b);
        double dsq;
        {
            const double du = VR[1] - VL[1];
            if (d == 2) {
                const double dv = VR[2] - VL[2];
                dsq = du * du + dv * dv;
            } else {
                const double dv = VR[2] - VL[2], dw = VR[3] - VL[3];
                dsq = du * du + dv * dv + dw * dw;
            }
        }
        const double vel_mag = std::sqrt(dsq);
        double al1, al2;
        if (vel_mag < EPS) {  // untaped: a replayed "not tiny" with vel_mag == 0 would divide 0 by 0
            al1 = double(1);
            al2 = double(0);
        } else {
            al1 = t_abs(VR[1 + a] - VL[1 + a]) / vel_mag;
            al2 = std::sqrt(double(1) - al1 * al1);
        }
        double be1 = double(1) / double(2) * (double(1) + al1 / (al1 + al2));
        if (mutation == 8) be1 = double(1) / double(2) * (double(1) + al2 / (al1 + al2));  // planted: alphas swapped
        const double be2 = double(1) - be1;
        Fout[0] = be1 * Fhllc[0] + be2 * Fhll[0];
        for (int b = 0; b < d; b++) {
            if (b == a)
                Fout[1 + b] = Fhllc[1 + b];
            else
                Fout[1 + b] = be1 * Fhllc[1 + b] + be2 * Fhll[1 + b];
        }
        Fout[d + 1] = Fhllc[d + 1];
    }
}

// Characteristic variables of one stencil cell for axis a
// (FlowModelBasicUtilitiesSingleSpecies.cpp:5882-6451).
inline void char_from_prim_single(int d, int a, double ra, double ca, const double* V, double* W) {
    const double un = V[1 + a];
    const double p = V[d + 1];
    W[0] = -double(1) / double(2) * ra * ca * un + double(1) / double(2) * p;
    W[1] = V[0] - double(1) / (ca * ca) * p;
    int t = 2;
    for (int b = 0; b < d; b++)
        if (b != a) W[t++] = V[1 + b];
    W[d + 1] = double(1) / double(2) * ra * ca * un + double(1) / double(2) * p;
}

// Primitive variables back from characteristic ones
// (FlowModelBasicUtilitiesSingleSpecies.cpp:7023-7473).
inline void prim_from_char_single(int d, int a, double ra, double ca, const double* W, double* V) {
    const double w0 = W[0], wl = W[d + 1];
    V[0] = double(1) / (ca * ca) * w0 + W[1] + double(1) / (ca * ca) * wl;
    V[1 + a] = -double(1) / (ra * ca) * w0 + double(1) / (ra * ca) * wl;
    int t = 2;
    for (int b = 0; b < d; b++)
        if (b != a) V[1 + b] = W[t++];
    V[d + 1] = w0 + wl;
}

// One axis of the WCNS56 driver for the single-species model
// (WCNS56-HLLC-HLL.cpp:879-1282 in 2D, 1746-2490 in 3D), pass by pass.
// Test probe of one direction pass: the sensor value and the midpoint flux of
// every midpoint face (FaceArray index space: i_a in [-1, n_a+1], others
// interior).  Filled only when passed.
struct FaceProbe {
    double* sensor;  // [faces]
    double* fmid;    // [neq][faces]
};

void direction_pass_single(const hmo_params& prm, const Model& m, const int* n, int a, const Derived& D,
                           const CellArray& theta, const CellArray& Omega, double dt, double* F_out,
                           FaceProbe* probe = nullptr) {
    const int d = m.dim, neq = m.neq;
    const double gam = m.gamma[0];
    const int sh[3] = {a == 0, a == 1, a == 2};

    FaceArray rho_avg, c_avg;
    rho_avg.init(n, a, 1);
    c_avg.init(n, a, 1);
    auto for_faces = [&](auto&& fn) {
        for (int k = rho_avg.lo[2]; k < rho_avg.lo[2] + rho_avg.ext[2]; k++)
            for (int j = rho_avg.lo[1]; j < rho_avg.lo[1] + rho_avg.ext[1]; j++)
                for (int i = rho_avg.lo[0]; i < rho_avg.lo[0] + rho_avg.ext[0]; i++) fn(i, j, k);
    };

    // Projection variables: simple face averages (FlowModelBasicUtilitiesSingleSpecies.cpp:4638-5134).
    for_faces([&](int i, int j, int k) {
        const int64_t iL = D.V.idx(i - sh[0], j - sh[1], k - sh[2]);
        const int64_t iR = D.V.idx(i, j, k);
        const int64_t f = rho_avg.idx(i, j, k);
        rho_avg.v[f] = double(1) / double(2) * (D.V.c(0)[iL] + D.V.c(0)[iR]);
        c_avg.v[f] = double(1) / double(2) * (D.c.c(0)[iL] + D.c.c(0)[iR]);
    });

    // Characteristic variables for the six stencil offsets m-3, m = 0..5 (driver WCNS56:886-893).
    std::vector<FaceArray> W(6 * neq);
    for (int mm = 0; mm < 6; mm++) {
        for (int e = 0; e < neq; e++) W[mm * neq + e].init(n, a, 1);
        const int o = mm - 3;
        for_faces([&](int i, int j, int k) {
            const int64_t ic = D.V.idx(i + o * sh[0], j + o * sh[1], k + o * sh[2]);
            const int64_t f = rho_avg.idx(i, j, k);
            double V[5], Wc[5];
            for (int e = 0; e < neq; e++) V[e] = D.V.c(e)[ic];
            char_from_prim_single(d, a, rho_avg.v[f], c_avg.v[f], V, Wc);
            for (int e = 0; e < neq; e++) W[mm * neq + e].v[f] = Wc[e];
        });
    }

    // WENO interpolation of each characteristic field (WCNS6-LD:459-954 / WCNS5-JS:234-737).
    std::vector<FaceArray> Wm(neq), Wp(neq);
    for (int e = 0; e < neq; e++) {
        Wm[e].init(n, a, 1);
        Wp[e].init(n, a, 1);
        for_faces([&](int i, int j, int k) {
            const int64_t f = rho_avg.idx(i, j, k);
            Stencil6 s{W[0 * neq + e].v[f], W[1 * neq + e].v[f], W[2 * neq + e].v[f],
                       W[3 * neq + e].v[f], W[4 * neq + e].v[f], W[5 * neq + e].v[f]};
            if (prm.scheme == 1) {
                const double sig = wcns6ld_sigma(s);
                Wm[e].v[f] = wcns6ld_side(s, sig, prm.p, prm.q, prm.C, prm.alpha_tau, prm.mutation);
                Wp[e].v[f] = wcns6ld_side(mirror(s), sig, prm.p, prm.q, prm.C, prm.alpha_tau, prm.mutation);
            } else {
                Wm[e].v[f] = wcns5js_side(s, prm.p, prm.mutation);
                Wp[e].v[f] = wcns5js_side(mirror(s), prm.p, prm.mutation);
            }
        });
    }

    // Back to primitive variables.
    std::vector<FaceArray> Vm(neq), Vp(neq);
    for (int e = 0; e < neq; e++) {
        Vm[e].init(n, a, 1);
        Vp[e].init(n, a, 1);
    }
    for_faces([&](int i, int j, int k) {
        const int64_t f = rho_avg.idx(i, j, k);
        double Wc[5], V[5];
        for (int e = 0; e < neq; e++) Wc[e] = Wm[e].v[f];
        prim_from_char_single(d, a, rho_avg.v[f], c_avg.v[f], Wc, V);
        for (int e = 0; e < neq; e++) Vm[e].v[f] = V[e];
        for (int e = 0; e < neq; e++) Wc[e] = Wp[e].v[f];
        prim_from_char_single(d, a, rho_avg.v[f], c_avg.v[f], Wc, V);
        for (int e = 0; e < neq; e++) Vp[e].v[f] = V[e];
    });

    // Bounds check and first-order fallback (FlowModelBasicUtilitiesSingleSpecies.cpp:3013-3439,
    // driver WCNS56:947-1032): rho > 0 and p > 0 on both sides.
    for_faces([&](int i, int j, int k) {
        const int64_t f = rho_avg.idx(i, j, k);
        int fm = 1, fp = 1;
        fm &= (Vm[0].v[f] > double(0)) ? 1 : 0;
        fm &= (Vm[neq - 1].v[f] > double(0)) ? 1 : 0;
        fp &= (Vp[0].v[f] > double(0)) ? 1 : 0;
        fp &= (Vp[neq - 1].v[f] > double(0)) ? 1 : 0;
        if (prm.mutation == 3) fm = fp = 1;  // planted mistake: bounds check dropped
        if (prm.mutation == 5) fm = 0;       // test hook: first-order face states on every face
        if (t_if(fm == 0 || fp == 0, kTapeFallback)) {
            const int64_t iL = D.V.idx(i - sh[0], j - sh[1], k - sh[2]);
            const int64_t iR = D.V.idx(i, j, k);
            for (int e = 0; e < neq; e++) {
                Vm[e].v[f] = D.V.c(e)[iL];
                Vp[e].v[f] = D.V.c(e)[iR];
            }
        }
    });

    // Riemann solvers and sensor select.
    FaceArray Fmid;
    Fmid.init(n, a, neq);
    for_faces([&](int i, int j, int k) {
        const int64_t f = rho_avg.idx(i, j, k);
        double VL[5], VR[5];
        for (int e = 0; e < neq; e++) {
            VL[e] = Vm[e].v[f];
            VR[e] = Vp[e].v[f];
        }
        bool use_hll = false;
        if (d > 1) {
            // Ducros-type sensor from the cells either side (WCNS56:1066-1088).
            const int64_t tL = theta.idx(i - sh[0], j - sh[1], k - sh[2]);
            const int64_t tR = theta.idx(i, j, k);
            double th_avg = 0.5 * (theta.c(0)[tL] + theta.c(0)[tR]);
            double om_avg = 0.5 * (Omega.c(0)[tL] + Omega.c(0)[tR]);
            if (prm.mutation == 7) {  // planted mistake: the right cell's values, no face average
                th_avg = theta.c(0)[tR];
                om_avg = Omega.c(0)[tR];
            }
            double sensor = -th_avg / (std::fabs(th_avg) + om_avg + EPS);
            if (prm.mutation == 6) sensor = -sensor;  // planted mistake: sign of the dilatation
            if (probe) probe->sensor[f] = sensor;
            use_hll = t_if(sensor > 0.65, kTapeSensor);
        }
        double Fo[5];
        riemann_face_single(d, a, gam, use_hll, VL, VR, Fo, nullptr, prm.mutation);
        for (int e = 0; e < neq; e++) Fmid.c(e)[f] = Fo[e];
        if (probe)
            for (int e = 0; e < neq; e++) probe->fmid[e * Fmid.vol() + f] = Fo[e];
    });

    // Midpoint-to-face sixth-order flux, scaled by dt (WCNS56:1197-1282, 2327-2490).
    int fext[3];
    for (int b = 0; b < 3; b++) fext[b] = (b == a) ? n[b] + 1 : n[b];
    const int64_t fvol = int64_t(fext[0]) * fext[1] * fext[2];
    for (int e = 0; e < neq; e++) {
        const double* Fm = Fmid.c(e);
        const double* Fn = D.Fn[a].c(e);
        double* out = F_out + e * fvol;
        for (int k = 0; k < fext[2]; k++)
            for (int j = 0; j < fext[1]; j++)
                for (int i = 0; i < fext[0]; i++) {
                    const int64_t of = i + int64_t(j) * fext[0] + int64_t(k) * fext[0] * fext[1];
                    const double mR = Fm[Fmid.idx(i + sh[0], j + sh[1], k + sh[2])];
                    const double mL = Fm[Fmid.idx(i - sh[0], j - sh[1], k - sh[2])];
                    const double mC = Fm[Fmid.idx(i, j, k)];
                    const double nR = Fn[D.V.idx(i, j, k)];
                    const double nL = Fn[D.V.idx(i - sh[0], j - sh[1], k - sh[2])];
                    out[of] = dt * (double(1) / double(30) * (mR + mL) - double(3) / double(10) * (nR + nL) +
                                    double(23) / double(15) * mC);
                }
    }
}

// ===========================================================================
// Five-equation Allaire model (FlowModelFiveEqnAllaire*, ISOBARIC ideal-gas
// mixing).  num_eqn = d + 2 ns (FlowModelFiveEqnAllaire.cpp:30); the last
// ns-1 equations (volume fractions) are ADVECTIVE (:94-113).
//   Q (input, neq+1 comps): Z_i rho_i (ns), rho u (d), E, Z_i (all ns, the
//       last one stored but not evolved, WCNS56:223-239)
//   V (primitive, neq comps): Z_i rho_i (ns), u (d), p, Z_i (ns-1)
// ===========================================================================

// The same structure serves the four-equation conservative model
// (FlowModelFourEqnConservative.cpp:16-113): partial densities rho Y_i (ns),
// momentum (d), total energy; no volume fractions (nz = 0), the mixture EOS
// mixes c_p and c_v with mass fractions, and no equation is ADVECTIVE.
struct Five {
    int d, ns, neq;
    int nz;     // evolved volume fractions: ns-1 (five-equation), 0 (four-equation)
    bool four;  // four-equation conservative model
    int mutation;
    double gam[8];
    double cp[8], cv[8];  // species heat capacities (four-equation mixing)
    double R[8];          // species gas constants
    int iu() const { return ns; }          // first velocity slot of V
    int ip() const { return ns + d; }      // pressure slot of V
    int iz() const { return ns + d + 1; }  // first volume-fraction slot of V
};

// Mixture ratio of specific heats and c_v from mass fractions, ISOTHERMAL_AND_ISOBARIC
// closure (EquationOfStateMixingRulesIdealGas.cpp:4677-4737, 6946-7150):
// c_p = sum Y_i c_p,i, c_v = sum Y_i c_v,i, gamma = c_p / c_v.
inline void mix_mass(const Five& f, const double* Y, double& gamma, double& c_v) {
    double c_p = double(0);
    c_v = double(0);
    for (int si = 0; si < f.ns; si++) {
        c_p += Y[si] * f.cp[si];
        c_v += Y[si] * f.cv[si];
    }
    gamma = c_p / c_v;
    if (f.mutation == 4) {  // planted mistake: mass-fraction average of the species gammas
        gamma = double(0);
        for (int si = 0; si < f.ns; si++) gamma += Y[si] * f.gam[si];
    }
}

// Squared sound speed of the four-equation model from density, pressure and
// mass fractions: Gamma p / rho + sum Y_i Psi_i with Gamma = gamma - 1 and
// Psi_i = ((c_p,i - gamma c_v,i) / c_v + gamma - 1) eps, eps = p / ((gamma - 1) rho)
// (FlowModelFourEqnConservative.cpp:5087-5440; EquationOfStateIdealGas.cpp:8059,
// 6291; EquationOfStateMixingRulesIdealGas.cpp:6367-6564).  Also returns eps.
inline double sound_speed_sq_four(const Five& f, double rho, double p, const double* Y, double& eps) {
    double gamma, c_v;
    mix_mass(f, Y, gamma, c_v);
    const double Gamma = gamma - double(1);
    eps = p / ((gamma - double(1)) * rho);
    double c2 = Gamma * p / rho;
    for (int si = 0; si < f.ns; si++) {
        const double Psi = ((f.cp[si] - gamma * f.cv[si]) / c_v + gamma - double(1)) * eps;
        c2 += Y[si] * Psi;
    }
    return c2;
}

// Mixture gamma from all ns volume fractions (EquationOfStateMixingRulesIdealGas.cpp:7404-7598):
// xi = sum Z_i * (1/(gamma_i - 1)), gamma = 1/xi + 1.
inline double mix_gamma_all(const Five& f, const double* Z) {
    double xi = double(0);
    for (int si = 0; si < f.ns; si++) {
        const double one_over_denominator = double(1) / (f.gam[si] - double(1));
        xi += Z[si] * one_over_denominator;
    }
    return double(1) / xi + double(1);
}

// Mixture gamma from ns-1 volume fractions, the last one being 1 - sum
// (EquationOfStateMixingRulesIdealGas.cpp:7599-7800).
inline double mix_gamma_last(const Five& f, const double* Z) {
    double xi = double(0);
    double Z_last = double(1);
    for (int si = 0; si < f.ns - 1; si++) {
        const double one_over_denominator = double(1) / (f.gam[si] - double(1));
        xi += Z[si] * one_over_denominator;
        Z_last -= Z[si];
    }
    xi += Z_last / (f.gam[f.ns - 1] - double(1));
    return double(1) / xi + double(1);
}

struct DerivedFive {
    CellArray V;      // neq comps
    CellArray c;      // sound speed
    CellArray rho;    // mixture density
    CellArray Fn[3];  // node convective flux per axis
};

// Derived cell data (FlowModelFiveEqnAllaire.cpp:3734-4876): density
// (EquationOfStateMixingRules.cpp:735-860), mass fractions, velocity, internal
// energy, pressure (ISOBARIC, EquationOfStateIdealGas.cpp:5633), sound speed
// c^2 = Gamma p / rho + sum Y_i Psi_i with Psi_i = p / rho (:4512-4876,
// EquationOfStateMixingRulesIdealGas.cpp:6564-6610), node flux (:4877-5901).
void compute_derived_five(const Five& f, const int* n, int g, const double* Q, DerivedFive& D) {
    const int d = f.d, ns = f.ns, neq = f.neq;
    CellArray Qv;
    Qv.init(n, g, d, 0);
    const int64_t qvol = Qv.vol();
    D.V.init(n, 4, d, neq);
    D.c.init(n, 4, d, 1);
    D.rho.init(n, 4, d, 1);
    for (int a = 0; a < d; a++) D.Fn[a].init(n, 4, d, neq);
    const int lo2 = (d > 2) ? -4 : 0, hi2 = (d > 2) ? n[2] + 4 : 1;
    const int lo1 = (d > 1) ? -4 : 0, hi1 = (d > 1) ? n[1] + 4 : 1;
    for (int k = lo2; k < hi2; k++)
        for (int j = lo1; j < hi1; j++)
            for (int i = -4; i < n[0] + 4; i++) {
                const int64_t iq = Qv.idx(i, j, k), iv = D.V.idx(i, j, k);
                double Zr[8], Z[8];
                for (int si = 0; si < ns; si++) Zr[si] = Q[si * qvol + iq];
                if (!f.four)
                    for (int si = 0; si < ns; si++) Z[si] = Q[(ns + d + 1 + si) * qvol + iq];
                double rho = double(0);
                for (int si = 0; si < ns; si++) rho += Zr[si];
                double Y[8];
                for (int si = 0; si < ns; si++) Y[si] = Zr[si] / rho;
                double u[3] = {0.0, 0.0, 0.0};
                for (int b = 0; b < d; b++) u[b] = Q[(ns + b) * qvol + iq] / rho;
                double ke;
                if (d == 1)
                    ke = u[0] * u[0];
                else if (d == 2)
                    ke = u[0] * u[0] + u[1] * u[1];
                else
                    ke = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
                const double E = Q[(ns + d) * qvol + iq];
                const double eps_int = E / rho - double(1) / double(2) * ke;
                double p, c2;
                if (f.four) {
                    // FlowModelFourEqnConservative.cpp:5034-5086 (pressure), 5087-5440 (sound speed)
                    double gmix, c_v, eps_psi;
                    mix_mass(f, Y, gmix, c_v);
                    p = (gmix - double(1)) * rho * eps_int;
                    c2 = sound_speed_sq_four(f, rho, p, Y, eps_psi);
                } else {
                    const double gmix = mix_gamma_all(f, Z);
                    p = (gmix - double(1)) * rho * eps_int;
                    const double Gamma = gmix - double(1);
                    c2 = Gamma * p / rho;
                    for (int si = 0; si < ns; si++) {
                        const double Psi = p / rho;
                        c2 += Y[si] * Psi;
                    }
                }
                for (int si = 0; si < ns; si++) D.V.c(si)[iv] = Zr[si];
                for (int b = 0; b < d; b++) D.V.c(ns + b)[iv] = u[b];
                D.V.c(ns + d)[iv] = p;
                for (int si = 0; si < f.nz; si++) D.V.c(ns + d + 1 + si)[iv] = Z[si];
                D.c.c(0)[iv] = std::sqrt(c2);
                D.rho.c(0)[iv] = rho;
                for (int a = 0; a < d; a++) {
                    const double ua = u[a];
                    for (int si = 0; si < ns; si++) D.Fn[a].c(si)[iv] = ua * Zr[si];
                    for (int b = 0; b < d; b++) {
                        if (b == a)
                            D.Fn[a].c(ns + b)[iv] = ua * Q[(ns + b) * qvol + iq] + p;
                        else
                            D.Fn[a].c(ns + b)[iv] = ua * Q[(ns + b) * qvol + iq];
                    }
                    D.Fn[a].c(ns + d)[iv] = ua * (E + p);
                    for (int si = 0; si < f.nz; si++) D.Fn[a].c(ns + d + 1 + si)[iv] = ua * Z[si];
                }
            }
}

// Characteristic variables (FlowModelBasicUtilitiesFiveEqnAllaire.cpp:8232-9178).
inline void char_from_prim_five(const Five& f, int a, const double* Zra, double ra, double ca, const double* V,
                                double* W) {
    const int ns = f.ns, d = f.d;
    const double un = V[ns + a], p = V[ns + d];
    W[0] = un - double(1) / (ra * ca) * p;
    for (int si = 0; si < ns; si++) W[1 + si] = V[si] - Zra[si] / (ra * ca * ca) * p;
    int t = ns + 1;
    for (int b = 0; b < d; b++)
        if (b != a) W[t++] = V[ns + b];
    for (int si = 0; si < f.nz; si++) W[ns + d + si] = V[ns + d + 1 + si];
    W[f.neq - 1] = un + double(1) / (ra * ca) * p;
}

// Back-projection (FlowModelBasicUtilitiesFiveEqnAllaire.cpp:9201-9974).
inline void prim_from_char_five(const Five& f, int a, const double* Zra, double ra, double ca, const double* W,
                                double* V) {
    const int ns = f.ns, d = f.d, neq = f.neq;
    for (int si = 0; si < ns; si++)
        V[si] = -double(1) / double(2) * Zra[si] / ca * W[0] + W[si + 1] +
                double(1) / double(2) * Zra[si] / ca * W[neq - 1];
    V[ns + a] = double(1) / double(2) * W[0] + double(1) / double(2) * W[neq - 1];
    int t = ns + 1;
    for (int b = 0; b < d; b++)
        if (b != a) V[ns + b] = W[t++];
    V[ns + d] = -double(1) / double(2) * ra * ca * W[0] + double(1) / double(2) * ra * ca * W[neq - 1];
    for (int si = 0; si < f.nz; si++) V[ns + d + 1 + si] = W[ns + d + si];
}

// Bounds of one side of a face normal to axis a
// (FlowModelBasicUtilitiesFiveEqnAllaire.cpp:5446-7347, bounds
// FlowModelBasicUtilitiesFiveEqnAllaire.hpp:24-27).  The sound-speed check
// differs by axis in the reference: on x-faces the species loop adds
// Y_i Psi_i to c^2 = Gamma p / rho (:5819, 6091, 6678), on y- and z-faces it
// assigns c^2 = Gamma p / rho again (:6359, 6992, 7305), so there the test is
// Gamma p / rho > 0.
inline int bounded_five(const Five& f, const double* V, int a) {
    const int ns = f.ns, d = f.d;
    int ok = 1;
    if (f.four) {
        // FlowModelBasicUtilitiesFourEqnConservative.cpp:3737-4511 (bounds in the .hpp:24-25):
        // mass fractions in (-0.001, 1.001), rho > 0, p > 0.
        double rho = double(0);
        for (int si = 0; si < ns; si++) rho += V[si];
        for (int si = 0; si < ns; si++) {
            const double Y = V[si] / rho;
            ok &= (Y > double(-0.001) && Y < double(1.001)) ? 1 : 0;
        }
        ok &= (rho > double(0)) ? 1 : 0;
        ok &= (V[ns + d] > double(0)) ? 1 : 0;
        return ok;
    }
    double Z[8];
    Z[ns - 1] = double(1);
    for (int si = 0; si < ns - 1; si++) {
        Z[si] = V[ns + d + 1 + si];
        Z[ns - 1] -= Z[si];
        ok &= (Z[si] > double(-1000.0) && Z[si] < double(1000.0)) ? 1 : 0;
    }
    ok &= (Z[ns - 1] > double(-1000.0) && Z[ns - 1] < double(1000.0)) ? 1 : 0;
    double rho = double(0);
    for (int si = 0; si < ns; si++) rho += V[si];
    double Y[8];
    for (int si = 0; si < ns; si++) {
        Y[si] = V[si] / rho;
        ok &= (Y[si] > double(-0.001) && Y[si] < double(1.001)) ? 1 : 0;
    }
    for (int si = 0; si < ns; si++) ok &= (V[si] > double(0)) ? 1 : 0;
    const double p = V[ns + d];
    const double Gamma = mix_gamma_all(f, Z) - double(1);
    double c2 = Gamma * p / rho;
    for (int si = 0; si < ns; si++) {
        if (a == 0) {
            const double Psi = p / rho;
            c2 += Y[si] * Psi;
        } else {
            c2 = Gamma * p / rho;
        }
    }
    ok &= (c2 > double(0)) ? 1 : 0;
    return ok;
}

// HLLC (FlowModelRiemannSolverFiveEqnAllaireHLLC.cpp:711-1591; face EOS in the
// drivers :4824-7716) and HLLC-HLL (...HLLC-HLL.cpp:1146-2439) of one face.
void riemann_face_five(const Five& f, int a, bool use_hll, const double* VL, const double* VR, double* Fout,
                       double* ustar) {
    const int ns = f.ns, d = f.d, neq = f.neq;
    const int iu = ns + a, ip = ns + d;
    auto face_eos = [&](const double* V, double& rho, double& c, double& eps_int) {
        rho = double(0);
        for (int si = 0; si < ns; si++) rho += V[si];
        double Y[8];
        for (int si = 0; si < ns; si++) Y[si] = V[si] / rho;
        if (f.four) {
            // FlowModelRiemannSolverFourEqnConservativeHLLC.cpp:4345-4700 (c, eps on faces)
            c = std::sqrt(sound_speed_sq_four(f, rho, V[ip], Y, eps_int));
            return;
        }
        const double gmix = mix_gamma_last(f, V + ns + d + 1);
        const double p = V[ip];
        const double Gamma = gmix - double(1);
        double c2 = Gamma * p / rho;
        for (int si = 0; si < ns; si++) {
            const double Psi = p / rho;
            c2 += Y[si] * Psi;
        }
        c = std::sqrt(c2);
        eps_int = p / ((gmix - double(1)) * rho);
    };
    double rhoL, rhoR, cL, cR, epsL, epsR;
    face_eos(VL, rhoL, cL, epsL);
    face_eos(VR, rhoR, cR, epsR);
    const double uL = VL[iu], uR = VR[iu], pL = VL[ip], pR = VR[ip];
    const double u_avg = double(1) / double(2) * (uL + uR);
    const double c_av = double(1) / double(2) * (cL + cR);
    const double sL = std::fmin(u_avg - c_av, uL - cL);
    const double sR = std::fmax(u_avg + c_av, uR + cR);
    const double sm = std::fmin(double(0), sL);
    const double sp = std::fmax(double(0), sR);
    const double sstar =
        (pR - pL + rhoL * uL * (sL - uL) - rhoR * uR * (sR - uR)) / (rhoL * (sL - uL) - rhoR * (sR - uR));

    auto cons_flux = [&](const double* V, double rho, double eps_int, double* Qs, double* Fs) {
        for (int si = 0; si < ns; si++) Qs[si] = V[si];
        for (int b = 0; b < d; b++) Qs[ns + b] = rho * V[ns + b];
        double ke;
        if (d == 1)
            ke = V[ns] * V[ns];
        else if (d == 2)
            ke = V[ns] * V[ns] + V[ns + 1] * V[ns + 1];
        else
            ke = V[ns] * V[ns] + V[ns + 1] * V[ns + 1] + V[ns + 2] * V[ns + 2];
        Qs[ns + d] = rho * (eps_int + double(1) / double(2) * ke);
        for (int si = 0; si < f.nz; si++) Qs[ns + d + 1 + si] = V[ns + d + 1 + si];
        const double un = V[iu];
        for (int si = 0; si < ns; si++) Fs[si] = un * V[si];
        for (int b = 0; b < d; b++) {
            if (b == a)
                Fs[ns + b] = un * Qs[ns + b] + V[ip];
            else
                Fs[ns + b] = un * Qs[ns + b];
        }
        Fs[ns + d] = un * (Qs[ns + d] + V[ip]);
        for (int si = 0; si < f.nz; si++) Fs[ns + d + 1 + si] = un * V[ns + d + 1 + si];
    };
    double QL[12], QR[12], FL[12], FR[12];
    cons_flux(VL, rhoL, epsL, QL, FL);
    cons_flux(VR, rhoR, epsR, QR, FR);

    double Fhllc[12];
    {
        const bool left = (sstar > double(0));
        const double* V = left ? VL : VR;
        const double* Qs = left ? QL : QR;
        const double* Fs = left ? FL : FR;
        const double rho = left ? rhoL : rhoR;
        const double sK = left ? sL : sR;
        const double spm = left ? sm : sp;
        const double chi = (sK - V[iu]) / (sK - sstar);
        double Qst[12];
        for (int si = 0; si < ns; si++) Qst[si] = chi * V[si];
        for (int b = 0; b < d; b++) {
            if (b == a)
                Qst[ns + b] = chi * rho * sstar;
            else
                Qst[ns + b] = chi * Qs[ns + b];
        }
        Qst[ns + d] = chi * (Qs[ns + d] + (sstar - V[iu]) * (rho * sstar + V[ip] / (sK - V[iu])));
        for (int si = 0; si < f.nz; si++) Qst[ns + d + 1 + si] = chi * V[ns + d + 1 + si];
        for (int e = 0; e < neq; e++) Fhllc[e] = Fs[e] + spm * (Qst[e] - Qs[e]);
        if (ustar) *ustar = V[iu] + spm * (chi - double(1));
    }
    for (int e = 0; e < neq; e++) Fout[e] = Fhllc[e];
    if (d > 1 && use_hll) {
        auto hll = [&](int e) {
            double r = (sR * FL[e] - sL * FR[e] + sR * sL * (QR[e] - QL[e])) / (sR - sL);
            if (sL > double(0)) r = FL[e];
            if (sR < double(0)) r = FR[e];
            return r;
        };
        double dsq;
        {
            const double du = VR[ns] - VL[ns];
            const double dv = VR[ns + 1] - VL[ns + 1];
            if (d == 2) {
                dsq = du * du + dv * dv;
            } else {
                const double dw = VR[ns + 2] - VL[ns + 2];
                dsq = du * du + dv * dv + dw * dw;
            }
        }
        const double vel_mag = std::sqrt(dsq);
        double al1, al2;
        if (vel_mag < EPS) {
            al1 = double(1);
            al2 = double(0);
        } else {
            al1 = std::fabs(VR[iu] - VL[iu]) / vel_mag;
            al2 = std::sqrt(double(1) - al1 * al1);
        }
        const double be1 = double(1) / double(2) * (double(1) + al1 / (al1 + al2));
        const double be2 = double(1) - be1;
        for (int si = 0; si < ns; si++) Fout[si] = be1 * Fhllc[si] + be2 * hll(si);
        for (int b = 0; b < d; b++)
            if (b != a) Fout[ns + b] = be1 * Fhllc[ns + b] + be2 * hll(ns + b);
        for (int si = 0; si < f.nz; si++) {
            const int e = ns + d + 1 + si;
            Fout[e] = be1 * Fhllc[e] + be2 * hll(e);
        }
    }
}

// One axis of the WCNS56 driver for the five-equation model, pass by pass.
// Also returns the HLLC normal velocity on the midpoint faces (for the source).
void direction_pass_five(const hmo_params& prm, const Five& f, const int* n, int a, const DerivedFive& D,
                         const CellArray& theta, const CellArray& Omega, double dt, double* F_out, FaceArray& ustar) {
    const int d = f.d, ns = f.ns, neq = f.neq;
    const int sh[3] = {a == 0, a == 1, a == 2};
    FaceArray avg;  // ns + 2 projection variables: Z_i rho_i, rho, c averages
    avg.init(n, a, ns + 2);
    auto for_faces = [&](auto&& fn) {
        for (int k = avg.lo[2]; k < avg.lo[2] + avg.ext[2]; k++)
            for (int j = avg.lo[1]; j < avg.lo[1] + avg.ext[1]; j++)
                for (int i = avg.lo[0]; i < avg.lo[0] + avg.ext[0]; i++) fn(i, j, k);
    };
    // Projection variables (FlowModelBasicUtilitiesFiveEqnAllaire.cpp:7474-8207).
    for_faces([&](int i, int j, int k) {
        const int64_t iL = D.V.idx(i - sh[0], j - sh[1], k - sh[2]);
        const int64_t iR = D.V.idx(i, j, k);
        const int64_t fc = avg.idx(i, j, k);
        for (int si = 0; si < ns; si++)
            avg.c(si)[fc] = double(1) / double(2) * (D.V.c(si)[iL] + D.V.c(si)[iR]);
        avg.c(ns)[fc] = double(1) / double(2) * (D.rho.c(0)[iL] + D.rho.c(0)[iR]);
        avg.c(ns + 1)[fc] = double(1) / double(2) * (D.c.c(0)[iL] + D.c.c(0)[iR]);
    });
    auto proj = [&](int64_t fc, double* Zra, double& ra, double& ca) {
        for (int si = 0; si < ns; si++) Zra[si] = avg.c(si)[fc];
        ra = avg.c(ns)[fc];
        ca = avg.c(ns + 1)[fc];
    };
    std::vector<FaceArray> W(6 * neq);
    for (int mm = 0; mm < 6; mm++) {
        for (int e = 0; e < neq; e++) W[mm * neq + e].init(n, a, 1);
        const int o = mm - 3;
        for_faces([&](int i, int j, int k) {
            const int64_t ic = D.V.idx(i + o * sh[0], j + o * sh[1], k + o * sh[2]);
            const int64_t fc = avg.idx(i, j, k);
            double V[12], Wc[12], Zra[8], ra, ca;
            for (int e = 0; e < neq; e++) V[e] = D.V.c(e)[ic];
            proj(fc, Zra, ra, ca);
            char_from_prim_five(f, a, Zra, ra, ca, V, Wc);
            for (int e = 0; e < neq; e++) W[mm * neq + e].v[fc] = Wc[e];
        });
    }
    std::vector<FaceArray> Wm(neq), Wp(neq);
    for (int e = 0; e < neq; e++) {
        Wm[e].init(n, a, 1);
        Wp[e].init(n, a, 1);
        for_faces([&](int i, int j, int k) {
            const int64_t fc = avg.idx(i, j, k);
            Stencil6 s{W[0 * neq + e].v[fc], W[1 * neq + e].v[fc], W[2 * neq + e].v[fc],
                       W[3 * neq + e].v[fc], W[4 * neq + e].v[fc], W[5 * neq + e].v[fc]};
            if (prm.scheme == 1) {
                const double sig = wcns6ld_sigma(s);
                Wm[e].v[fc] = wcns6ld_side(s, sig, prm.p, prm.q, prm.C, prm.alpha_tau, prm.mutation);
                Wp[e].v[fc] = wcns6ld_side(mirror(s), sig, prm.p, prm.q, prm.C, prm.alpha_tau, prm.mutation);
            } else {
                Wm[e].v[fc] = wcns5js_side(s, prm.p, prm.mutation);
                Wp[e].v[fc] = wcns5js_side(mirror(s), prm.p, prm.mutation);
            }
        });
    }
    std::vector<FaceArray> Vm(neq), Vp(neq);
    for (int e = 0; e < neq; e++) {
        Vm[e].init(n, a, 1);
        Vp[e].init(n, a, 1);
    }
    for_faces([&](int i, int j, int k) {
        const int64_t fc = avg.idx(i, j, k);
        double Wc[12], V[12], Zra[8], ra, ca;
        proj(fc, Zra, ra, ca);
        for (int e = 0; e < neq; e++) Wc[e] = Wm[e].v[fc];
        prim_from_char_five(f, a, Zra, ra, ca, Wc, V);
        for (int e = 0; e < neq; e++) Vm[e].v[fc] = V[e];
        for (int e = 0; e < neq; e++) Wc[e] = Wp[e].v[fc];
        prim_from_char_five(f, a, Zra, ra, ca, Wc, V);
        for (int e = 0; e < neq; e++) Vp[e].v[fc] = V[e];
    });
    for_faces([&](int i, int j, int k) {
        const int64_t fc = avg.idx(i, j, k);
        double VL[12], VR[12];
        for (int e = 0; e < neq; e++) {
            VL[e] = Vm[e].v[fc];
            VR[e] = Vp[e].v[fc];
        }
        int fm = bounded_five(f, VL, a), fp = bounded_five(f, VR, a);
        if (prm.mutation == 3) fm = fp = 1;
        if (fm == 0 || fp == 0) {
            const int64_t iL = D.V.idx(i - sh[0], j - sh[1], k - sh[2]);
            const int64_t iR = D.V.idx(i, j, k);
            for (int e = 0; e < neq; e++) {
                Vm[e].v[fc] = D.V.c(e)[iL];
                Vp[e].v[fc] = D.V.c(e)[iR];
            }
        }
    });
    FaceArray Fmid;
    Fmid.init(n, a, neq);
    ustar.init(n, a, 1);
    for_faces([&](int i, int j, int k) {
        const int64_t fc = avg.idx(i, j, k);
        double VL[12], VR[12];
        for (int e = 0; e < neq; e++) {
            VL[e] = Vm[e].v[fc];
            VR[e] = Vp[e].v[fc];
        }
        bool use_hll = false;
        if (d > 1) {
            const int64_t tL = theta.idx(i - sh[0], j - sh[1], k - sh[2]);
            const int64_t tR = theta.idx(i, j, k);
            const double th_avg = 0.5 * (theta.c(0)[tL] + theta.c(0)[tR]);
            const double om_avg = 0.5 * (Omega.c(0)[tL] + Omega.c(0)[tR]);
            const double sensor = -th_avg / (std::fabs(th_avg) + om_avg + EPS);
            use_hll = sensor > 0.65;
        }
        double Fo[12], us;
        riemann_face_five(f, a, use_hll, VL, VR, Fo, &us);
        for (int e = 0; e < neq; e++) Fmid.c(e)[fc] = Fo[e];
        ustar.v[fc] = us;
    });
    int fext[3];
    for (int b = 0; b < 3; b++) fext[b] = (b == a) ? n[b] + 1 : n[b];
    const int64_t fvol = int64_t(fext[0]) * fext[1] * fext[2];
    for (int e = 0; e < neq; e++) {
        const double* Fm = Fmid.c(e);
        const double* Fn = D.Fn[a].c(e);
        double* out = F_out + e * fvol;
        for (int k = 0; k < fext[2]; k++)
            for (int j = 0; j < fext[1]; j++)
                for (int i = 0; i < fext[0]; i++) {
                    const int64_t of = i + int64_t(j) * fext[0] + int64_t(k) * fext[0] * fext[1];
                    const double mR = Fm[Fmid.idx(i + sh[0], j + sh[1], k + sh[2])];
                    const double mL = Fm[Fmid.idx(i - sh[0], j - sh[1], k - sh[2])];
                    const double mC = Fm[Fmid.idx(i, j, k)];
                    const double nR = Fn[D.V.idx(i, j, k)];
                    const double nL = Fn[D.V.idx(i - sh[0], j - sh[1], k - sh[2])];
                    out[of] = dt * (double(1) / double(30) * (mR + mL) - double(3) / double(10) * (nR + nL) +
                                    double(23) / double(15) * mC);
                }
    }
}

// Advective source of the volume-fraction equations (WCNS56:491-535, 1288-1369, 2492-2650):
// S_i += dt Q_i sum_a [3/2 (u*_{i+1/2} - u*_{i-1/2}) - 3/10 (u_{i+1} - u_{i-1})
//                      + 1/30 (u*_{i+3/2} - u*_{i-3/2})] / dx_a
void source_five(const Five& f, const int* n, int g, const double* dx, double dt, const double* Q,
                 const DerivedFive& D, const FaceArray* ustar, double* S) {
    const int d = f.d, ns = f.ns;
    CellArray Qv;
    Qv.init(n, g, d, 0);
    const int64_t qvol = Qv.vol();
    const int64_t svol = int64_t(n[0]) * n[1] * n[2];
    for (int si = 0; si < ns - 1; si++) {
        const int ei = ns + d + 1 + si;
        for (int k = 0; k < n[2]; k++)
            for (int j = 0; j < n[1]; j++)
                for (int i = 0; i < n[0]; i++) {
                    double acc = 0.0;
                    for (int a = 0; a < d; a++) {
                        const int s0 = a == 0, s1 = a == 1, s2 = a == 2;
                        const FaceArray& us = ustar[a];
                        const double* u = D.V.c(ns + a);
                        const double uLL = us.v[us.idx(i - s0, j - s1, k - s2)];
                        const double uL = us.v[us.idx(i, j, k)];
                        const double uR = us.v[us.idx(i + s0, j + s1, k + s2)];
                        const double uRR = us.v[us.idx(i + 2 * s0, j + 2 * s1, k + 2 * s2)];
                        const double cL = u[D.V.idx(i - s0, j - s1, k - s2)];
                        const double cR = u[D.V.idx(i + s0, j + s1, k + s2)];
                        const double term = (double(3) / double(2) * (uR - uL) - double(3) / double(10) * (cR - cL) +
                                             double(1) / double(30) * (uRR - uLL)) /
                                            dx[a];
                        acc = (a == 0) ? term : acc + term;
                    }
                    const int64_t is = i + int64_t(j) * n[0] + int64_t(k) * n[0] * n[1];
                    S[ei * svol + is] += dt * Q[ei * qvol + Qv.idx(i, j, k)] * acc;
                }
    }
}

// theta / Omega for the five-equation model (velocity in V slots ns..ns+d-1).
void compute_theta_omega_five(const Five& f, const int* n, const DerivedFive& D, const double* dx, CellArray& theta,
                              CellArray& Omega) {
    // Build a single-species-like view of the velocity so the same derivative code runs.
    Model m{};
    m.dim = f.d;
    Derived tmp;
    tmp.V.init(n, 4, f.d, f.d + 2);
    for (int b = 0; b < f.d; b++) std::memcpy(tmp.V.c(1 + b), D.V.c(f.ns + b), sizeof(double) * D.V.vol());
    compute_theta_omega(m, n, tmp, dx, theta, Omega);
}

// Model constants of the multi-species restatement; returns false for bad species counts.
bool make_multi(const hmo_params* prm, Five& f) {
    f = Five{};
    f.d = prm->dim;
    f.ns = prm->num_species;
    f.four = prm->model == 1;
    f.neq = f.four ? f.ns + f.d + 1 : f.d + 2 * f.ns;
    f.nz = f.four ? 0 : f.ns - 1;
    f.mutation = prm->mutation;
    if (f.ns < (f.four ? 1 : 2) || f.ns > 8) return false;
    for (int i = 0; i < 8; i++) {
        f.gam[i] = prm->gamma[i];
        f.R[i] = prm->R[i];
        // EquationOfStateMixingRulesIdealGas.cpp:108-120
        f.cp[i] = (i < f.ns) ? prm->gamma[i] / (prm->gamma[i] - double(1)) * prm->R[i] : 0.0;
        f.cv[i] = (i < f.ns) ? double(1) / (prm->gamma[i] - double(1)) * prm->R[i] : 0.0;
    }
    return true;
}

int rhs_patch_five(const hmo_params* prm, const int* n, int g, const double* dx, double dt, const double* Q,
                   double* const* F, double* S) {
    Five f;
    if (!make_multi(prm, f)) return 5;
    DerivedFive D;
    compute_derived_five(f, n, g, Q, D);
    CellArray theta, Omega;
    compute_theta_omega_five(f, n, D, dx, theta, Omega);
    FaceArray ustar[3];
    for (int a = 0; a < f.d; a++) direction_pass_five(*prm, f, n, a, D, theta, Omega, dt, F[a], ustar[a]);
    if (S && !f.four) source_five(f, n, g, dx, dt, Q, D, ustar, S);
    return 0;
}

// ---------------------------------------------------------------------------
// Diffusive flux of the four-equation model with the node-based SIXTH_ORDER
// reconstructor (DiffusiveFluxReconstructorNode.cpp:32-1738,
// DiffusiveFluxReconstructorNodeSixthOrder.cpp:65-939), the variables and
// diffusivities of FlowModelDiffusiveFluxUtilitiesFourEqnConservative.cpp
// (764-1821, 1824-2953, 5673-6425) and constant transport mixing rules.
// ---------------------------------------------------------------------------

constexpr int kDiffGhosts = 6;  // DiffusiveFluxReconstructorNodeSixthOrder.cpp:22

struct DiffCells {  // derived cell data over the 6-ghost box
    CellArray rho, Y, u, T, h, mu, mu_v, kappa;
};

// Derived data of the four-equation model (density, mass fractions,
// velocity, pressure as in compute_derived_five), temperature
// T = p / ((gamma - 1) c_v rho) (EquationOfStateIdealGas.cpp:6951),
// species densities rho_i = p / (R_i T) (:8357) and enthalpies
// h_i = gamma_i p / ((gamma_i - 1) rho_i) (:6464), and the mixture shear and
// bulk viscosities and conductivity sum(x_i w_i) / sum(w_i), w_i = Y_i / sqrt(M_i)
// (EquationOfShearViscosityMixingRulesConstant.cpp:497-528 and the bulk /
// conductivity analogues).
void compute_diffusive_cells(const Five& f, const hmo_transport& tr, const int* n, int g, const double* Q,
                             DiffCells& C) {
    const int d = f.d, ns = f.ns;
    CellArray Qv;
    Qv.init(n, g, d, 0);
    const int64_t qvol = Qv.vol();
    const int G = kDiffGhosts;
    C.rho.init(n, G, d, 1);
    C.Y.init(n, G, d, ns);
    C.u.init(n, G, d, d);
    C.T.init(n, G, d, 1);
    C.h.init(n, G, d, ns);
    C.mu.init(n, G, d, 1);
    C.mu_v.init(n, G, d, 1);
    C.kappa.init(n, G, d, 1);
    double wfac[8];
    for (int si = 0; si < ns; si++) wfac[si] = double(1) / (std::sqrt(tr.M[si]));
    const int lo2 = (d > 2) ? -G : 0, hi2 = (d > 2) ? n[2] + G : 1;
    const int lo1 = (d > 1) ? -G : 0, hi1 = (d > 1) ? n[1] + G : 1;
    for (int k = lo2; k < hi2; k++)
        for (int j = lo1; j < hi1; j++)
            for (int i = -G; i < n[0] + G; i++) {
                const int64_t iq = Qv.idx(i, j, k), ic = C.rho.idx(i, j, k);
                double rho = double(0);
                for (int si = 0; si < ns; si++) rho += Q[si * qvol + iq];
                double Y[8];
                for (int si = 0; si < ns; si++) Y[si] = Q[si * qvol + iq] / rho;
                double u[3] = {0.0, 0.0, 0.0};
                for (int b = 0; b < d; b++) u[b] = Q[(ns + b) * qvol + iq] / rho;
                double ke;
                if (d == 1)
                    ke = u[0] * u[0];
                else if (d == 2)
                    ke = u[0] * u[0] + u[1] * u[1];
                else
                    ke = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
                const double E = Q[(ns + d) * qvol + iq];
                const double eps_int = E / rho - double(1) / double(2) * ke;
                double gmix, c_v;
                mix_mass(f, Y, gmix, c_v);
                const double p = (gmix - double(1)) * rho * eps_int;
                const double T = p / ((gmix - double(1)) * c_v * rho);
                double mu_n = double(0), mu_d = double(0), mv_n = double(0), mv_d = double(0);
                double ka_n = double(0), ka_d = double(0);
                for (int si = 0; si < ns; si++) {
                    const double rho_i = p / (f.R[si] * T);
                    C.h.c(si)[ic] = f.gam[si] * p / ((f.gam[si] - double(1)) * rho_i);
                    const double w = Y[si] * wfac[si];
                    mu_n += tr.mu[si] * w;
                    mu_d += w;
                    mv_n += tr.mu_v[si] * w;
                    mv_d += w;
                    ka_n += tr.kappa[si] * w;
                    ka_d += w;
                }
                C.rho.c(0)[ic] = rho;
                for (int si = 0; si < ns; si++) C.Y.c(si)[ic] = Y[si];
                for (int b = 0; b < d; b++) C.u.c(b)[ic] = u[b];
                C.T.c(0)[ic] = T;
                C.mu.c(0)[ic] = mu_n / mu_d;
                C.mu_v.c(0)[ic] = mv_n / mv_d;
                C.kappa.c(0)[ic] = ka_n / ka_d;
            }
}

// Diffusivity kinds of FlowModelDiffusiveFluxUtilitiesFourEqnConservative.cpp:5673-6425.
enum DiffKind {
    DK_MASS_SELF,   // -rho D_si
    DK_MASS_CROSS,  // rho Y_si D_sj
    DK_A0,          // -(4/3 mu + mu_v)
    DK_A1,          // 2/3 mu - mu_v
    DK_A2,          // -mu
    DK_U43,         // -u_c (4/3 mu + mu_v)
    DK_U23,         // u_c (2/3 mu - mu_v)
    DK_UMU,         // -u_c mu
    DK_KAP,         // -kappa
    DK_ENT_SELF,    // -rho D_si h_si
    DK_ENT_CROSS    // rho Y_si D_sj h_si
};
enum DiffVar { DV_Y, DV_U, DV_T };

struct DiffTerm {
    int var, comp;       // variable whose derivative is taken (mass fraction, velocity component, T)
    int kind, c, si, sj; // diffusivity
};

// (variable, diffusivity) terms of equation e of the axis-a flux whose
// derivative is along axis b, in the reference's order
// (getCellDataOfDiffusiveFluxVariablesForDerivative:764-1821 and
// getCellDataOfDiffusiveFluxDiffusivities:1824-2953).
int diffusive_terms(const Five& f, int a, int b, int e, DiffTerm* out) {
    const int ns = f.ns, d = f.d;
    int t = 0;
    if (e < ns) {
        if (b == a) {
            out[t++] = {DV_Y, e, DK_MASS_SELF, 0, e, 0};
            for (int sj = 0; sj < ns; sj++) out[t++] = {DV_Y, sj, DK_MASS_CROSS, 0, e, sj};
        }
    } else if (e < ns + d) {
        const int c = e - ns;
        if (b == a) {
            if (c == a)
                out[t++] = {DV_U, a, DK_A0, 0, 0, 0};
            else
                out[t++] = {DV_U, c, DK_A2, 0, 0, 0};
        } else if (c == a) {
            out[t++] = {DV_U, b, DK_A1, 0, 0, 0};
        } else if (c == b) {
            out[t++] = {DV_U, a, DK_A2, 0, 0, 0};
        }
    } else {
        if (b == a) {
            for (int c = 0; c < d; c++) out[t++] = {DV_U, c, c == a ? DK_U43 : DK_UMU, c, 0, 0};
            out[t++] = {DV_T, 0, DK_KAP, 0, 0, 0};
            for (int si = 0; si < ns; si++) {
                out[t++] = {DV_Y, si, DK_ENT_SELF, 0, si, 0};
                for (int sj = 0; sj < ns; sj++) out[t++] = {DV_Y, sj, DK_ENT_CROSS, 0, si, sj};
            }
        } else if (a < b) {
            out[t++] = {DV_U, a, DK_UMU, b, 0, 0};
            out[t++] = {DV_U, b, DK_U23, a, 0, 0};
        } else {
            out[t++] = {DV_U, b, DK_U23, a, 0, 0};
            out[t++] = {DV_U, a, DK_UMU, b, 0, 0};
        }
    }
    return t;
}

inline double diffusivity(const DiffTerm& t, const DiffCells& C, const hmo_transport& tr, int64_t ic, int mut) {
    const double mu = C.mu.c(0)[ic], mu_v = C.mu_v.c(0)[ic];
    switch (t.kind) {
        case DK_MASS_SELF: return -C.rho.c(0)[ic] * tr.D[t.si];
        case DK_MASS_CROSS:
            if (mut == 5) return double(0);  // planted mistake: the mass-flux correction dropped
            return C.rho.c(0)[ic] * C.Y.c(t.si)[ic] * tr.D[t.sj];
        case DK_A0: return -(double(4) / double(3) * mu + mu_v);
        case DK_A1:
            if (mut == 6) return -(double(2) / double(3) * mu - mu_v);  // planted mistake: sign of the dilatation term
            return double(2) / double(3) * mu - mu_v;
        case DK_A2: return -mu;
        case DK_U43: return -C.u.c(t.c)[ic] * (double(4) / double(3) * mu + mu_v);
        case DK_U23: return C.u.c(t.c)[ic] * (double(2) / double(3) * mu - mu_v);
        case DK_UMU: return -C.u.c(t.c)[ic] * mu;
        case DK_KAP: return -C.kappa.c(0)[ic];
        case DK_ENT_SELF: return -C.rho.c(0)[ic] * tr.D[t.si] * C.h.c(t.si)[ic];
        default: return C.rho.c(0)[ic] * C.Y.c(t.si)[ic] * tr.D[t.sj] * C.h.c(t.si)[ic];
    }
}

// Sixth-order central first derivative along axis b at cell (i, j, k)
// (DiffusiveFluxReconstructorNodeSixthOrder.cpp:77-112).
inline double first_derivative6(const CellArray& A, int comp, int b, int i, int j, int k, double dx_inv) {
    const double a_n = double(3) / double(4);
    const double b_n = -double(3) / double(20);
    const double c_n = double(1) / double(60);
    const int s[3] = {b == 0, b == 1, b == 2};
    const double* u = A.c(comp);
    auto at = [&](int o) { return u[A.idx(i + o * s[0], j + o * s[1], k + o * s[2])]; };
    return (a_n * (at(1) - at(-1)) + b_n * (at(2) - at(-2)) + c_n * (at(3) - at(-3))) * dx_inv;
}

// computeDiffusiveFluxOnPatch for the four-equation model: node fluxes
// F_node[e] = sum_b sum_terms D * d(var)/dx_b on nodes [-3, n_a + 3) along a,
// then F_face = dt (a_r (F_L + F_R) + b_r (F_LL + F_RR) + c_r (F_LLL + F_RRR))
// (DiffusiveFluxReconstructorNodeSixthOrder.cpp:507-690).
int diffusive_patch_four(const hmo_params* prm, const hmo_transport* tr, const int* n, int g, const double* dx,
                         double dt, const double* Q, double* const* F) {
    Five f;
    if (!make_multi(prm, f) || !f.four) return 5;
    if (g < kDiffGhosts) return 3;
    DiffCells C;
    compute_diffusive_cells(f, *tr, n, g, Q, C);
    const int d = f.d, neq = f.neq;
    double dx_inv[3];
    for (int b = 0; b < d; b++) dx_inv[b] = double(1) / dx[b];
    const double a_n = double(3) / double(4);
    const double b_n = -double(3) / double(20);
    const double c_n = double(1) / double(60);
    const double a_r = a_n + b_n + c_n;
    const double b_r = b_n + c_n;
    const double c_r = c_n;
    DiffTerm terms[96];
    for (int a = 0; a < d; a++) {
        const int sh[3] = {a == 0, a == 1, a == 2};
        // node box: [-3, n_a + 3) along a, interior across
        int lo[3] = {0, 0, 0}, hi[3] = {n[0], n[1], n[2]};
        lo[a] = -3;
        hi[a] = n[a] + 3;
        int next[3];
        for (int b = 0; b < 3; b++) next[b] = hi[b] - lo[b];
        const int64_t nvol = int64_t(next[0]) * next[1] * next[2];
        std::vector<double> Fn(size_t(neq) * nvol, 0.0);
        auto nidx = [&](int i, int j, int k) {
            return int64_t(i - lo[0]) + int64_t(j - lo[1]) * next[0] + int64_t(k - lo[2]) * next[0] * next[1];
        };
        for (int e = 0; e < neq; e++)
            for (int b = 0; b < d; b++) {
                const int nt = diffusive_terms(f, a, b, e, terms);
                for (int ti = 0; ti < nt; ti++) {
                    const DiffTerm& t = terms[ti];
                    const CellArray& A = (t.var == DV_Y) ? C.Y : (t.var == DV_U) ? C.u : C.T;
                    for (int k = lo[2]; k < hi[2]; k++)
                        for (int j = lo[1]; j < hi[1]; j++)
                            for (int i = lo[0]; i < hi[0]; i++) {
                                const int64_t ic = C.rho.idx(i, j, k);
                                const double D = diffusivity(t, C, *tr, ic, prm->mutation);
                                const double der = first_derivative6(A, t.comp, b, i, j, k, dx_inv[b]);
                                Fn[e * nvol + nidx(i, j, k)] += D * der;
                            }
                }
            }
        int fext[3];
        for (int b = 0; b < 3; b++) fext[b] = (b == a) ? n[b] + 1 : n[b];
        const int64_t fvol = int64_t(fext[0]) * fext[1] * fext[2];
        for (int e = 0; e < neq; e++) {
            const double* Fe = Fn.data() + e * nvol;
            double* out = F[a] + e * fvol;
            for (int k = 0; k < fext[2]; k++)
                for (int j = 0; j < fext[1]; j++)
                    for (int i = 0; i < fext[0]; i++) {
                        auto node = [&](int o) { return Fe[nidx(i + o * sh[0], j + o * sh[1], k + o * sh[2])]; };
                        const int64_t of = i + int64_t(j) * fext[0] + int64_t(k) * fext[0] * fext[1];
                        out[of] = double(0);
                        out[of] += dt * (a_r * (node(-1) + node(0)) + b_r * (node(-2) + node(1)) +
                                         c_r * (node(-3) + node(2)));
                    }
        }
    }
    return 0;
}

int rhs_patch_impl(const hmo_params* prm, const int* n_in, int g, const double* dx, double dt, const double* Q,
                   double* const* F, double* S) {
    (void)S;
    if (!prm || !n_in || !dx || !Q || !F) return 1;
    Model m = make_model(*prm);
    if (g < 4) return 3;
    int n[3] = {1, 1, 1};
    for (int a = 0; a < m.dim; a++) {
        n[a] = n_in[a];
        if (n[a] < 1) return 4;
    }
    if (prm->model == 1 || prm->model == 2) return rhs_patch_five(prm, n, g, dx, dt, Q, F, S);
    Derived D;
    compute_derived_single(m, n, g, Q, D);
    CellArray theta, Omega;
    compute_theta_omega(m, n, D, dx, theta, Omega);
    for (int a = 0; a < m.dim; a++) direction_pass_single(*prm, m, n, a, D, theta, Omega, dt, F[a]);
    return 0;
}

}  // namespace

extern "C" {

int hmo_num_eqn(const hmo_params* prm) {
    Model m = make_model(*prm);
    return m.neq;
}

// ConvectiveFluxReconstructorWCNS56::computeConvectiveFluxAndSourceOnPatch
// (WCNS56-HLLC-HLL.cpp:39-2656) for one patch with host arrays.
// F0/F1/F2 are overwritten; S is accumulated for ADVECTIVE equations only.
int hmo_rhs_patch(const hmo_params* prm, const int* n, int g, const double* dx, double dt, const double* Q,
                  double* F0, double* F1, double* F2, double* S) {
    double* F[3] = {F0, F1, F2};
    return rhs_patch_impl(prm, n, g, dx, dt, Q, F, S);
}

typedef struct hmo_patch {
    int n[3];
    int g;
    double dx[3];
    const double* Q;
    double* F[3];
    double* S;
} hmo_patch;

// Many independent patches, parallel over patches with OpenMP (the analogue of
// one MPI rank per core, halo exchange excluded).  Returns the first nonzero
// status, or 0.
int hmo_rhs_batch(const hmo_params* prm, int npatch, const hmo_patch* patches, double dt, int nthreads) {
    int status = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1) reduction(max : status) if (g_tape.mode == 0)
#endif
    for (int ip = 0; ip < npatch; ip++) {  // serial while the branch tape is on
        const hmo_patch& P = patches[ip];
        double* F[3] = {P.F[0], P.F[1], P.F[2]};
        const int st = rhs_patch_impl(prm, P.n, P.g, P.dx, dt, P.Q, F, P.S);
        if (st > status) status = st;
    }
    return status;
}

// ---- pointwise entry points (used only to pin the oracle in tests) ----

// Both sides of one characteristic field: WCNS6-LD (scheme 1) or WCNS5-JS (0).
void hmo_weno_pair(const hmo_params* prm, const double* U, double* um, double* up) {
    Stencil6 s{U[0], U[1], U[2], U[3], U[4], U[5]};
    if (prm->scheme == 1) {
        const double sig = wcns6ld_sigma(s);
        *um = wcns6ld_side(s, sig, prm->p, prm->q, prm->C, prm->alpha_tau, prm->mutation);
        *up = wcns6ld_side(mirror(s), sig, prm->p, prm->q, prm->C, prm->alpha_tau, prm->mutation);
    } else {
        *um = wcns5js_side(s, prm->p, prm->mutation);
        *up = wcns5js_side(mirror(s), prm->p, prm->mutation);
    }
}

// Smoothness indicators beta_0..beta_3 of the minus side.
void hmo_smoothness(const double* U, double* beta) {
    Stencil6 s{U[0], U[1], U[2], U[3], U[4], U[5]};
    smoothness3(s, beta[0], beta[1], beta[2]);
    beta[3] = smoothness6(s);
}

// Characteristic round trip of one primitive state (single species).
void hmo_char_roundtrip_single(int dim, int axis, double ra, double ca, const double* V, double* W, double* Vback) {
    char_from_prim_single(dim, axis, ra, ca, V, W);
    prim_from_char_single(dim, axis, ra, ca, W, Vback);
}

// Midpoint flux of one face from left/right primitive states (single species).
void hmo_riemann_single(int dim, int axis, double gamma, int use_hll, const double* VL, const double* VR, double* F,
                        double* ustar) {
    riemann_face_single(dim, axis, gamma, use_hll != 0, VL, VR, F, ustar);
}

// The same with a planted mistake (Params.mutation; 8: HLLC-HLL blend weights
// alpha_1 and alpha_2 swapped), for the tests that show the pins catch it.
void hmo_riemann_single_mutated(int dim, int axis, double gamma, int use_hll, const double* VL, const double* VR,
                                double* F, double* ustar, int mutation) {
    riemann_face_single(dim, axis, gamma, use_hll != 0, VL, VR, F, ustar, mutation);
}

// Sensor value and midpoint flux of every midpoint face of axis `axis` of one
// single-species patch (faces i_a in [-1, n_a+1] x interior, x fastest; fmid
// component-major), from the same pass hmo_rhs_patch runs.
int hmo_midpoint_probe_single(const hmo_params* prm, const int* n_in, int g, const double* dx, const double* Q,
                              int axis, double* sensor, double* fmid) {
    if (!prm || !n_in || !dx || !Q || !sensor || !fmid || prm->model != 0) return 1;
    Model m = make_model(*prm);
    if (g < 4 || axis < 0 || axis >= m.dim) return 3;
    int n[3] = {1, 1, 1};
    for (int a = 0; a < m.dim; a++) {
        n[a] = n_in[a];
        if (n[a] < 1) return 4;
    }
    Derived D;
    compute_derived_single(m, n, g, Q, D);
    CellArray theta, Omega;
    compute_theta_omega(m, n, D, dx, theta, Omega);
    int64_t fvol = 1;
    for (int b = 0; b < 3; b++) fvol *= (b == axis) ? n[b] + 1 : n[b];
    std::vector<double> F(size_t(m.neq) * fvol);
    FaceProbe probe{sensor, fmid};
    direction_pass_single(*prm, m, n, axis, D, theta, Omega, 1.0, F.data(), &probe);
    return 0;
}

// Euler::computeSpectralRadiusesAndStableDtOnPatch (Euler.cpp:489-888) of one
// patch, over its interior cells: RungeKuttaLevelIntegrator's default
// use_ghosts_to_compute_dt = false (RungeKuttaLevelIntegrator.cpp:3787) reads
// the current data, which has no ghost cells.  Maximum wave speed
// lambda_a = |u_a| + c (FlowModelSingleSpecies.cpp:3884-4070; the
// multi-species models use their mixture sound speed); out[a] = max over cells
// of lambda_a / dx_a (a < dim) and out[dim] = 1 / max over cells of the sum of
// those (Euler.cpp:592-605, 713-729, 861-890), the maxima by fmax from 0.
int hmo_spectral_radiuses_and_dt(const hmo_params* prm, const int* n_in, int g, const double* dx, const double* Q,
                                 double* out) {
    if (!prm || !n_in || !dx || !Q || !out) return 1;
    if (g < 4) return 3;  // the derived data covers the 4-cell halo
    const int d = prm->dim;
    int n[3] = {1, 1, 1};
    for (int a = 0; a < d; a++) {
        n[a] = n_in[a];
        if (n[a] < 1) return 4;
    }
    CellArray V, C;
    int iu = 1;  // first velocity slot of V
    if (prm->model == 0) {
        Model m = make_model(*prm);
        Derived D;
        compute_derived_single(m, n, g, Q, D);
        V = D.V;
        C = D.c;
    } else {
        Five f;
        if (!make_multi(prm, f)) return 5;
        DerivedFive D;
        compute_derived_five(f, n, g, Q, D);
        V = D.V;
        C = D.c;
        iu = f.ns;
    }
    double s[3] = {double(0), double(0), double(0)}, ssum = double(0);
    for (int k = 0; k < n[2]; k++)
        for (int j = 0; j < n[1]; j++)
            for (int i = 0; i < n[0]; i++) {
                const int64_t ic = V.idx(i, j, k);
                const double c = C.c(0)[C.idx(i, j, k)];
                double sr[3] = {double(0), double(0), double(0)};
                for (int a = 0; a < d; a++) {
                    const double lambda = std::fabs(V.c(iu + a)[ic]) + c;
                    sr[a] = lambda / dx[a];
                    s[a] = std::fmax(s[a], sr[a]);
                }
                double sum = sr[0];
                if (d == 2) sum = sr[0] + sr[1];
                if (d == 3) sum = sr[0] + sr[1] + sr[2];
                if (prm->mutation == 9) sum = std::fmax(sr[0], std::fmax(sr[1], sr[2]));  // planted: max, not sum
                ssum = std::fmax(ssum, sum);
            }
    for (int a = 0; a < d; a++) out[a] = s[a];
    out[d] = double(1) / ssum;
    return 0;
}

// Bounds check of one side of a face normal to `axis`, four- or five-equation
// model (FlowModelBasicUtilities{FourEqnConservative,FiveEqnAllaire}.cpp):
// 1 bounded, 0 not, -1 bad arguments.
int hmo_bounded_multi(const hmo_params* prm, int axis, const double* V) {
    Five f;
    if (!prm || !V || (prm->model != 1 && prm->model != 2) || !make_multi(prm, f)) return -1;
    return bounded_five(f, V, axis);
}

// Midpoint flux of one face from left/right primitive states, four- or
// five-equation model (V = partial densities, velocity, pressure[, volume fractions]).
int hmo_riemann_multi(const hmo_params* prm, int axis, int use_hll, const double* VL, const double* VR, double* F,
                      double* ustar) {
    Five f;
    if (!prm || (prm->model != 1 && prm->model != 2) || !make_multi(prm, f)) return 1;
    riemann_face_five(f, axis, use_hll != 0, VL, VR, F, ustar);
    return 0;
}

// DiffusiveFluxReconstructorNode::computeDiffusiveFluxOnPatch
// (DiffusiveFluxReconstructorNode.cpp:32-1738, SIXTH_ORDER) for one patch of
// the four-equation model: F0/F1/F2 (diffusive flux SideData, 0 ghosts) are
// overwritten with dt-scaled face fluxes.  Q needs a ghost width >= 6.
int hmo_diffusive_patch(const hmo_params* prm, const hmo_transport* tr, const int* n_in, int g, const double* dx,
                        double dt, const double* Q, double* F0, double* F1, double* F2) {
    if (!prm || !tr || !n_in || !dx || !Q || !F0) return 1;
    int n[3] = {1, 1, 1};
    for (int a = 0; a < prm->dim; a++) {
        n[a] = n_in[a];
        if (n[a] < 1) return 4;
    }
    double* F[3] = {F0, F1, F2};
    return diffusive_patch_four(prm, tr, n, g, dx, dt, Q, F);
}

// Branch tape control (single-species path): mode 1 records (clearing the
// tape), mode 2 replays from its start, mode 0 switches it off.  Returns the
// tape length.
int64_t hmo_branch_tape(int mode) {
    if (mode == 1) {
        g_tape.bits.clear();
        g_tape.size = 0;
    }
    g_tape.mode = mode;
    g_tape.pos = 0;
    g_tape.overrun = false;
    return int64_t(g_tape.size);
}

// Categories (bit c = kTape c) that a replay decides afresh; 0 replays all.
void hmo_branch_tape_free_mask(int mask) { g_tape.free_mask = mask; }

// Replay position and whether a replay ran past the end of the tape.
void hmo_branch_tape_status(int64_t* length, int64_t* position, int* overrun) {
    *length = int64_t(g_tape.size);
    *position = int64_t(g_tape.pos);
    *overrun = g_tape.overrun ? 1 : 0;
}

int hmo_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

}  // extern "C"
