// hm_diffusive.hip -- kernels and C-ABI of the diffusive flux of the
// four-equation conservative model (node-based SIXTH_ORDER reconstructor,
// DiffusiveFluxReconstructorNode.cpp:32-1738 with
// DiffusiveFluxReconstructorNodeSixthOrder.cpp:65-939; variables and
// diffusivities of FlowModelDiffusiveFluxUtilitiesFourEqnConservative.cpp:
// 764-1821, 1824-2953, 5673-6425).  Every expression keeps the reference's
// tree, so that with uncontracted arithmetic the result is bit-identical.
#include "hm_ctx.hpp"

namespace hm {

namespace {

template <int D>
__device__ __forceinline__ int dext(const DiffPatch& P, int a) {
    return (a < D) ? P.n[a] + 2 * kGd : 1;
}

template <int D>
__device__ __forceinline__ int64_t dcell(const DiffPatch& P, int i, int j, int k) {
    const int e0 = dext<D>(P, 0), e1 = dext<D>(P, 1);
    return (i + kGd) + int64_t(j + ((D > 1) ? kGd : 0)) * e0 + int64_t(k + ((D > 2) ? kGd : 0)) * e0 * e1;
}

template <int D>
__device__ __forceinline__ int64_t dvol(const DiffPatch& P) {
    return int64_t(dext<D>(P, 0)) * dext<D>(P, 1) * dext<D>(P, 2);
}

template <int D>
__device__ __forceinline__ int64_t dq_idx(const DiffPatch& P, int i, int j, int k) {
    const int e0 = P.n[0] + 2 * P.gq[0];
    const int e1 = (D > 1) ? P.n[1] + 2 * P.gq[1] : 1;
    const int gy = (D > 1) ? P.gq[1] : 0, gz = (D > 2) ? P.gq[2] : 0;
    return (i + P.gq[0]) + int64_t(j + gy) * e0 + int64_t(k + gz) * e0 * e1;
}

// Component slots of the derived cell data.
template <int D, int NS>
struct DiffSlots {
    static constexpr int RHO = 0, Y = 1, U = 1 + NS, T = 1 + NS + D, H = 2 + NS + D, MU = 2 + 2 * NS + D,
                         MUV = 3 + 2 * NS + D, KAP = 4 + 2 * NS + D;
};

// Cells of the 6-ghost box: density, mass fractions, velocity, pressure of the
// four-equation model (FlowModelFourEqnConservative.cpp:5034-5086), temperature
// p / ((gamma - 1) c_v rho) (EquationOfStateIdealGas.cpp:6951), species
// densities p / (R_i T) (:8357) and enthalpies gamma_i p / ((gamma_i - 1) rho_i)
// (:6464), mixture viscosities and conductivity sum(x_i w_i) / sum(w_i),
// w_i = Y_i / sqrt(M_i) (EquationOfShearViscosityMixingRulesConstant.cpp:497-528).
template <int D, int NS, class W>
__global__ __launch_bounds__(kBlock) void k_diff_cells(W w, DParams k) {
    const int64_t gid = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    const DiffPatch* P;
    int ax;
    int64_t loc;
    if (!w.find(gid, P, ax, loc)) return;
    using S = DiffSlots<D, NS>;
    const int e0 = dext<D>(*P, 0), e1 = dext<D>(*P, 1);
    // 32-bit index arithmetic (a work segment has fewer than 2^31 items).
    const uint32_t r = uint32_t(loc) / uint32_t(e0);
    const int i = int(uint32_t(loc) - r * uint32_t(e0)) - kGd;
    const int j = (D > 1) ? int((D > 2) ? r % uint32_t(e1) : r) - kGd : 0;
    const int kk = (D > 2) ? int(r / uint32_t(e1)) - kGd : 0;
    const int64_t iq = dq_idx<D>(*P, i, j, kk);
    double rho = double(0);
    double Y[NS];
#pragma unroll
    for (int si = 0; si < NS; si++) rho += P->Q[si][iq];
#pragma unroll
    for (int si = 0; si < NS; si++) Y[si] = P->Q[si][iq] / rho;
    double u[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int b = 0; b < D; b++) u[b] = P->Q[NS + b][iq] / rho;
    double ke;
    if constexpr (D == 1)
        ke = u[0] * u[0];
    else if constexpr (D == 2)
        ke = u[0] * u[0] + u[1] * u[1];
    else
        ke = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    const double E = P->Q[NS + D][iq];
    const double eps_int = E / rho - double(1) / double(2) * ke;
    double c_p = double(0), c_v = double(0);
#pragma unroll
    for (int si = 0; si < NS; si++) {
        c_p += Y[si] * k.cp[si];
        c_v += Y[si] * k.cv[si];
    }
    const double gmix = c_p / c_v;
    const double p = (gmix - double(1)) * rho * eps_int;
    const double T = p / ((gmix - double(1)) * c_v * rho);
    const int64_t vol = dvol<D>(*P);
    const int64_t c = dcell<D>(*P, i, j, kk);
    double* out = P->cell;
    double mu_n = double(0), mu_d = double(0), mv_n = double(0), mv_d = double(0), ka_n = double(0),
           ka_d = double(0);
#pragma unroll
    for (int si = 0; si < NS; si++) {
        const double rho_i = p / (k.R[si] * T);
        out[(S::H + si) * vol + c] = k.gamma[si] * p / ((k.gamma[si] - double(1)) * rho_i);
        const double wgt = Y[si] * k.wfac[si];
        mu_n += k.mu[si] * wgt;
        mu_d += wgt;
        mv_n += k.mu_v[si] * wgt;
        mv_d += wgt;
        ka_n += k.kappa[si] * wgt;
        ka_d += wgt;
    }
    out[S::RHO * vol + c] = rho;
#pragma unroll
    for (int si = 0; si < NS; si++) out[(S::Y + si) * vol + c] = Y[si];
#pragma unroll
    for (int b = 0; b < D; b++) out[(S::U + b) * vol + c] = u[b];
    out[S::T * vol + c] = T;
    out[S::MU * vol + c] = mu_n / mu_d;
    out[S::MUV * vol + c] = mv_n / mv_d;
    out[S::KAP * vol + c] = ka_n / ka_d;
}

// Sixth-order central first derivative along axis B of cell-data component
// `comp` at cell c (DiffusiveFluxReconstructorNodeSixthOrder.cpp:77-112).
__device__ __forceinline__ double deriv6(const double* f, int64_t c, int64_t st, double dxinv) {
    const double a_n = double(3) / double(4);
    const double b_n = -double(3) / double(20);
    const double c_n = double(1) / double(60);
    return (a_n * (f[c + st] - f[c - st]) + b_n * (f[c + 2 * st] - f[c - 2 * st]) +
            c_n * (f[c + 3 * st] - f[c - 3 * st])) *
           dxinv;
}

// Node fluxes of axis A on nodes [-3, n_A + 3) x interior: per equation, the
// terms of derivative axes b = 0..D-1 in the reference's order, each
// diffusivity * derivative added to a sum that starts at zero
// (DiffusiveFluxReconstructorNode.cpp:920-1070).
template <int D, int NS, class W>
__global__ __launch_bounds__(kBlock) void k_diff_nodes(W w, DParams k) {
    const int64_t gid = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    const DiffPatch* P;
    int A;
    int64_t loc;
    if (!w.find(gid, P, A, loc)) return;
    using S = DiffSlots<D, NS>;
    constexpr int NEQ = NS + D + 1;
    int ext[3];
#pragma unroll
    for (int b = 0; b < 3; b++) ext[b] = (b < D) ? P->n[b] + ((b == A) ? 6 : 0) : 1;
    int idx[3];
    // 32-bit index arithmetic (a work segment has fewer than 2^31 items).
    const uint32_t r = uint32_t(loc) / uint32_t(ext[0]);
    idx[0] = int(uint32_t(loc) - r * uint32_t(ext[0]));
    idx[1] = int(r % uint32_t(ext[1]));
    idx[2] = int(r / uint32_t(ext[1]));
    idx[A] -= 3;
    const int64_t c = dcell<D>(*P, idx[0], idx[1], idx[2]);
    const int64_t vol = dvol<D>(*P);
    int64_t st[3];
    st[0] = 1;
    st[1] = dext<D>(*P, 0);
    st[2] = int64_t(dext<D>(*P, 0)) * dext<D>(*P, 1);
    const double* cd = P->cell;
    const double rho = cd[S::RHO * vol + c];
    const double mu = cd[S::MU * vol + c];
    const double mu_v = cd[S::MUV * vol + c];
    double Y[NS], h[NS], u[D];
#pragma unroll
    for (int si = 0; si < NS; si++) {
        Y[si] = cd[(S::Y + si) * vol + c];
        h[si] = cd[(S::H + si) * vol + c];
    }
#pragma unroll
    for (int b = 0; b < D; b++) u[b] = cd[(S::U + b) * vol + c];

    // Derivatives along A of Y_i, u_c, T; along b != A of u_A and u_b.
    const double dA = P->dxinv[A];
    double dY[NS], dU[D], dT;
#pragma unroll
    for (int si = 0; si < NS; si++) dY[si] = deriv6(cd + (S::Y + si) * vol, c, st[A], dA);
#pragma unroll
    for (int b = 0; b < D; b++) dU[b] = deriv6(cd + (S::U + b) * vol, c, st[A], dA);
    dT = deriv6(cd + S::T * vol, c, st[A], dA);
    double dUa[3] = {0.0, 0.0, 0.0}, dUb[3] = {0.0, 0.0, 0.0};  // d u_A / dx_b, d u_b / dx_b
#pragma unroll
    for (int b = 0; b < D; b++) {
        if (b == A) continue;
        dUa[b] = deriv6(cd + (S::U + A) * vol, c, st[b], P->dxinv[b]);
        dUb[b] = deriv6(cd + (S::U + b) * vol, c, st[b], P->dxinv[b]);
    }

    double Fn[NEQ];
    // Mass equations: -rho D_si dY_si + sum_j rho Y_si D_sj dY_sj (derivative along A only).
#pragma unroll
    for (int si = 0; si < NS; si++) {
        double acc = double(0);
        acc += -rho * k.D[si] * dY[si];
#pragma unroll
        for (int sj = 0; sj < NS; sj++) acc += rho * Y[si] * k.D[sj] * dY[sj];
        Fn[si] = acc;
    }
    // Momentum equations.
#pragma unroll
    for (int cc = 0; cc < D; cc++) {
        double acc = double(0);
#pragma unroll
        for (int b = 0; b < D; b++) {
            if (b == A) {
                if (cc == A)
                    acc += -(double(4) / double(3) * mu + mu_v) * dU[cc];
                else
                    acc += -mu * dU[cc];
            } else if (cc == A) {
                acc += (double(2) / double(3) * mu - mu_v) * dUb[b];
            } else if (cc == b) {
                acc += -mu * dUa[b];
            }
        }
        Fn[NS + cc] = acc;
    }
    // Energy equation.
    {
        double acc = double(0);
#pragma unroll
        for (int b = 0; b < D; b++) {
            if (b == A) {
#pragma unroll
                for (int cc = 0; cc < D; cc++) {
                    if (cc == A)
                        acc += -u[cc] * (double(4) / double(3) * mu + mu_v) * dU[cc];
                    else
                        acc += -u[cc] * mu * dU[cc];
                }
                acc += -cd[S::KAP * vol + c] * dT;
#pragma unroll
                for (int si = 0; si < NS; si++) {
                    acc += -rho * k.D[si] * h[si] * dY[si];
#pragma unroll
                    for (int sj = 0; sj < NS; sj++) acc += rho * Y[si] * k.D[sj] * h[si] * dY[sj];
                }
            } else if (A < b) {
                acc += -u[b] * mu * dUa[b];
                acc += u[A] * (double(2) / double(3) * mu - mu_v) * dUb[b];
            } else {
                acc += u[A] * (double(2) / double(3) * mu - mu_v) * dUb[b];
                acc += -u[b] * mu * dUa[b];
            }
        }
        Fn[NS + D] = acc;
    }
    int64_t nvol = 1;
#pragma unroll
    for (int b = 0; b < 3; b++) nvol *= ext[b];
    double* out = P->fnode[A];
#pragma unroll
    for (int e = 0; e < NEQ; e++) out[e * nvol + loc] = Fn[e];
}

// Face fluxes of axis A (DiffusiveFluxReconstructorNodeSixthOrder.cpp:507-690):
// F_face = 0 + dt (a_r (F_L + F_R) + b_r (F_LL + F_RR) + c_r (F_LLL + F_RRR)).
template <int D, int NS, class W>
__global__ __launch_bounds__(kBlock) void k_diff_faces(W w, DParams k) {
    const int64_t gid = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    const DiffPatch* P;
    int A;
    int64_t loc;
    if (!w.find(gid, P, A, loc)) return;
    constexpr int NEQ = NS + D + 1;
    int fext[3], next[3];
#pragma unroll
    for (int b = 0; b < 3; b++) {
        fext[b] = (b < D) ? P->n[b] + ((b == A) ? 1 : 0) : 1;
        next[b] = (b < D) ? P->n[b] + ((b == A) ? 6 : 0) : 1;
    }
    int idx[3];
    // 32-bit index arithmetic (a work segment has fewer than 2^31 items).
    const uint32_t r = uint32_t(loc) / uint32_t(fext[0]);
    idx[0] = int(uint32_t(loc) - r * uint32_t(fext[0]));
    idx[1] = int(r % uint32_t(fext[1]));
    idx[2] = int(r / uint32_t(fext[1]));
    // node of cell m along A sits at m + 3; the face f reads cells f-3 .. f+2
    idx[A] += 3;
    const int64_t nst = (A == 0) ? 1 : (A == 1) ? int64_t(next[0]) : int64_t(next[0]) * next[1];
    const int64_t n0 = idx[0] + int64_t(idx[1]) * next[0] + int64_t(idx[2]) * next[0] * next[1];
    const int64_t nvol = int64_t(next[0]) * next[1] * next[2];
    const int64_t fvol = int64_t(fext[0]) * fext[1] * fext[2];
    const double a_n = double(3) / double(4);
    const double b_n = -double(3) / double(20);
    const double c_n = double(1) / double(60);
    const double a_r = a_n + b_n + c_n;
    const double b_r = b_n + c_n;
    const double c_r = c_n;
    const double* Fn = P->fnode[A];
    double* out = P->F[A];
    const double dt = k.dtp ? *k.dtp : k.dt;
#pragma unroll
    for (int e = 0; e < NEQ; e++) {
        const double* f = Fn + e * nvol + n0;
        double acc = double(0);
        acc += dt * (a_r * (f[-nst] + f[0]) + b_r * (f[-2 * nst] + f[nst]) + c_r * (f[-3 * nst] + f[2 * nst]));
        out[e * fvol + loc] = acc;
    }
}

template <class W>
unsigned dblocks(const DPlan<W>& plan, int stage) {
    return unsigned((plan.items[stage] + kBlock - 1) / kBlock);
}

template <int D, int NS, class W>
hipError_t launch_dn(const DPlan<W>& plan, const DParams& k, hipStream_t s, int first, int last) {
    if (first <= 0 && last > 0 && plan.items[0] > 0)
        hipLaunchKernelGGL((k_diff_cells<D, NS, W>), dim3(dblocks(plan, 0)), dim3(kBlock), 0, s, plan.work[0], k);
    if (first <= 1 && last > 1 && plan.items[1] > 0)
        hipLaunchKernelGGL((k_diff_nodes<D, NS, W>), dim3(dblocks(plan, 1)), dim3(kBlock), 0, s, plan.work[1], k);
    if (first <= 2 && last > 2 && plan.items[2] > 0)
        hipLaunchKernelGGL((k_diff_faces<D, NS, W>), dim3(dblocks(plan, 2)), dim3(kBlock), 0, s, plan.work[2], k);
    return hipGetLastError();
}

template <int D, class W>
hipError_t launch_d(const DPlan<W>& plan, const DParams& k, hipStream_t s, int first, int last) {
    switch (k.num_species) {
        case 1: return launch_dn<D, 1, W>(plan, k, s, first, last);
        case 2: return launch_dn<D, 2, W>(plan, k, s, first, last);
        case 3: return launch_dn<D, 3, W>(plan, k, s, first, last);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace

template <class W>
hipError_t launch_diffusive_four(int dim, const DPlan<W>& plan, const DParams& k, hipStream_t s, int first,
                                 int last) {
    switch (dim) {
        case 1: return launch_d<1, W>(plan, k, s, first, last);
        case 2: return launch_d<2, W>(plan, k, s, first, last);
        case 3: return launch_d<3, W>(plan, k, s, first, last);
        default: return hipErrorInvalidValue;
    }
}

template hipError_t launch_diffusive_four<SegSingleT<DiffPatch>>(int, const DPlan<SegSingleT<DiffPatch>>&,
                                                                 const DParams&, hipStream_t, int, int);
template hipError_t launch_diffusive_four<SegTableT<DiffPatch>>(int, const DPlan<SegTableT<DiffPatch>>&,
                                                                const DParams&, hipStream_t, int, int);

}  // namespace hm

extern "C" {

int hm_set_diffusive_transport(hm_ctx* ctx, const hm_transport* tr) {
    if (!ctx || !ctx->rec) return HM_ERR_INVALID_ARGUMENT;
    if (!tr) return ctx->rec->fail(HM_ERR_INVALID_ARGUMENT, "NULL transport properties");
    const hm_config& c = ctx->rec->config();
    if (c.model != HM_MODEL_FOUR_EQN_CONSERVATIVE)
        return ctx->rec->fail(HM_ERR_UNSUPPORTED, "the diffusive flux is built for the four-equation model");
    for (int i = 0; i < c.num_species; i++) {
        if (!(tr->species_M[i] > 0.0)) return ctx->rec->fail(HM_ERR_INVALID_ARGUMENT, "species_M must be positive");
        if (!(tr->species_mu[i] >= 0.0) || !(tr->species_mu_v[i] >= 0.0) || !(tr->species_kappa[i] >= 0.0) ||
            !(tr->species_D[i] >= 0.0))
            return ctx->rec->fail(HM_ERR_INVALID_ARGUMENT, "transport coefficients must be non-negative");
    }
    if (ctx->levels > 0)
        return ctx->rec->fail(HM_ERR_INVALID_ARGUMENT,
                              "set the transport properties before creating levels on this context");
    if (int st = ctx->rec->selectDevice()) return st;
    ctx->diff.reset();
    ctx->diff.reset(new (std::nothrow) hm::DiffusiveFluxReconstructorGPU(*ctx->rec, *tr));
    if (!ctx->diff) return ctx->rec->fail(HM_ERR_OUT_OF_MEMORY, "out of host memory");
    return HM_OK;
}

int hm_get_diffusive_num_ghosts(const hm_ctx* ctx, int* num_ghosts) {
    if (!ctx || !ctx->diff || !num_ghosts) return HM_ERR_INVALID_ARGUMENT;
    *num_ghosts = ctx->diff->getDiffusiveFluxNumberOfGhostCells();
    return HM_OK;
}

int hm_compute_diffusive_flux_on_patch(hm_ctx* ctx, const hm_patch_view* patch, double time, double dt,
                                       int rk_step, void* stream) {
    if (!ctx || !ctx->rec) return HM_ERR_INVALID_ARGUMENT;
    if (!ctx->diff) return ctx->rec->fail(HM_ERR_INVALID_ARGUMENT, "set the transport properties first");
    if (!patch) return ctx->rec->fail(HM_ERR_INVALID_ARGUMENT, "NULL patch view");
    return ctx->diff->computeDiffusiveFluxOnPatch(*patch, time, dt, rk_step, static_cast<hipStream_t>(stream));
}

int hm_compute_diffusive_flux_on_patches(hm_ctx* ctx, const hm_patch_view* patches, int num_patches, double time,
                                         double dt, int rk_step, void* stream) {
    if (!ctx || !ctx->rec) return HM_ERR_INVALID_ARGUMENT;
    if (!ctx->diff) return ctx->rec->fail(HM_ERR_INVALID_ARGUMENT, "set the transport properties first");
    return ctx->diff->computeDiffusiveFluxOnPatches(patches, num_patches, time, dt, rk_step,
                                                    static_cast<hipStream_t>(stream));
}

}  // extern "C"
