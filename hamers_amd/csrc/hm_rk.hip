// hm_rk.hip -- Runge-Kutta stage update of Euler::advanceSingleStepOnPatch
// (src/apps/Euler/Euler.cpp:1003-1680) for one or many patches.
//
// Per interior cell and equation, with the scratch value zero-filled first
// (FlowModel*::fillCellDataOfConservativeVariablesWithZero):
//   for n < num_coeffs:
//     if alpha[n] != 0:  Q += alpha[n] * Q_int[n]
//     if beta[n]  != 0:  Q += beta[n] * (-(F_x,R - F_x,L)/dx_0 - (F_y,T - F_y,B)/dx_1
//                                        - (F_z,F - F_z,B)/dx_2 + S_int[n])
//     if beta[n] != 0 (five-equation model): last volume fraction = 1 - sum of the others
//                        (FlowModelFiveEqnAllaire.cpp:1739-1890)
// and, where accumulation arrays are given, F_acc += gamma[n] F_int[n],
// S_acc += gamma[n] S_int[n] (Euler.cpp:1230-1270, 1360-1420, 1560-1660).
#include "hm_rk.hpp"

namespace hm {

template <int D, bool NS, class W>
__global__ __launch_bounds__(kBlock) void k_rk_update(W w, RKParams r) {
    // Blocks numbered per XCD: a cell's y and z face neighbours, read again by
    // the blocks n0 and n0 n1 cells on, then sit in the same L2.
    const int64_t gid = int64_t(xcd_block(blockIdx.x, gridDim.x)) * kBlock + threadIdx.x;
    const RKPatch* P;
    int ax;
    int64_t loc;
    if (!w.find(gid, P, ax, loc)) return;
    const int n0 = P->n[0], n1 = (D > 1) ? P->n[1] : 1;
    // 32-bit index arithmetic (a work segment has fewer than 2^31 items).
    const uint32_t rr = uint32_t(loc) / uint32_t(n0);
    const int i = int(uint32_t(loc) - rr * uint32_t(n0));
    const int j = (D > 1) ? int((D > 2) ? rr % uint32_t(n1) : rr) : 0;
    const int k = (D > 2) ? int(rr / uint32_t(n1)) : 0;
    // Scratch (output) and intermediate (input) ghost boxes.
    auto box_idx = [&](const int* g) -> int64_t {
        const int e0 = n0 + 2 * g[0];
        const int e1 = (D > 1) ? n1 + 2 * g[1] : 1;
        return (i + g[0]) + int64_t(j + ((D > 1) ? g[1] : 0)) * e0 + int64_t(k + ((D > 2) ? g[2] : 0)) * e0 * e1;
    };
    const int64_t iq = box_idx(P->gq);
    const int64_t ii = box_idx(P->gi);
    // Face indices of the two faces of this cell on each axis.
    int64_t fL[3], fR[3];
    {
        // x faces: i + j (n0+1) + k (n0+1) n1
        fL[0] = i + int64_t(j) * (n0 + 1) + int64_t(k) * (n0 + 1) * n1;
        fR[0] = fL[0] + 1;
        if (D > 1) {
            fL[1] = i + int64_t(j) * n0 + int64_t(k) * n0 * (n1 + 1);
            fR[1] = fL[1] + n0;
        }
        if (D > 2) {
            fL[2] = i + int64_t(j) * n0 + int64_t(k) * n0 * n1;
            fR[2] = fL[2] + int64_t(n0) * n1;
        }
    }
    int64_t fvol[3];
    fvol[0] = int64_t(n0 + 1) * n1 * ((D > 2) ? P->n[2] : 1);
    if (D > 1) fvol[1] = int64_t(n0) * (n1 + 1) * ((D > 2) ? P->n[2] : 1);
    if (D > 2) fvol[2] = int64_t(n0) * n1 * (P->n[2] + 1);
    const int64_t svol = int64_t(n0) * n1 * ((D > 2) ? P->n[2] : 1);

    const int nz = r.nz;
    double zl = double(1);  // last volume fraction: 1 - sum of the others, in species order
    for (int e = 0; e < r.neq; e++) {
        double q = double(0);
        double snap = double(0);
        for (int n = 0; n < r.nc; n++) {
            if (r.alpha[n] != 0.0) q += r.alpha[n] * P->Qi[n][e][ii];
            if (r.beta[n] != 0.0) {
                // Difference of the face fluxes on axis a: (F_R - F_L), or for the
                // Navier-Stokes form (F_R - F_L + Fd_R - Fd_L) in the reference's order.
                auto diff = [&](int a) {
                    const double* Fa = P->Fi[n][a] + e * fvol[a];
                    double dF = Fa[fR[a]] - Fa[fL[a]];
                    if (NS && P->Fdi[n][0]) {
                        const double* Fd = P->Fdi[n][a] + e * fvol[a];
                        dF = dF + Fd[fR[a]] - Fd[fL[a]];
                    }
                    return dF;
                };
                double t = -diff(0) / P->dx[0];
                if constexpr (D > 1) t = t - diff(1) / P->dx[1];
                if constexpr (D > 2) t = t - diff(2) / P->dx[2];
                q += r.beta[n] * (t + P->Si[n][e * svol + loc]);
                snap = q;
            }
        }
        P->Q[e][iq] = q;
        if (nz > 0 && e >= r.zfirst && e < r.zfirst + nz) zl -= snap;
    }
    // The zero-filled value stays if no coefficient beta is nonzero.
    if (nz > 0) P->Q[r.neq][iq] = (r.last_beta >= 0) ? zl : double(0);
}

// gamma-weighted accumulation of stage fluxes and sources: the terms t in order.
template <int D, class W>
__global__ __launch_bounds__(kBlock) void k_rk_accumulate(W w, AccTerms t) {
    const int64_t gid = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    const RKPatch* P;
    int ax;  // 0..D-1: faces of that axis, D: cells (source)
    int64_t loc;
    if (!w.find(gid, P, ax, loc)) return;
    const int n2 = (D > 2) ? P->n[2] : 1, n1 = (D > 1) ? P->n[1] : 1;
    int64_t vol;
    if (ax == D)
        vol = int64_t(P->n[0]) * n1 * n2;
    else
        vol = int64_t(P->n[0] + (ax == 0)) * (n1 + (ax == 1)) * (n2 + (ax == 2));
    for (int e = 0; e < t.neq; e++) {
        double* acc = (ax == D) ? P->Sacc : P->Facc[ax];
        if (acc) {
            double v = acc[e * vol + loc];
            for (int i = 0; i < t.n; i++) {
                const int n = t.src[i];
                const double* src = (ax == D) ? P->Si[n] : P->Fi[n][ax];
                v += t.gamma[i] * src[e * vol + loc];
            }
            acc[e * vol + loc] = v;
        }
        double* dacc = (ax == D) ? nullptr : P->Fdacc[ax];
        if (dacc) {
            double v = dacc[e * vol + loc];
            for (int i = 0; i < t.n; i++) {
                const int n = t.src[i];
                if (!P->Fdi[n][ax]) continue;
                v += t.gamma[i] * P->Fdi[n][ax][e * vol + loc];
            }
            dacc[e * vol + loc] = v;
        }
    }
}

template <class W>
hipError_t launch_rk_accumulate(int dim, const W& acc, int64_t acc_items, const AccTerms& terms, hipStream_t s) {
    if (acc_items <= 0 || terms.n == 0) return hipSuccess;
    const unsigned na = unsigned((acc_items + kBlock - 1) / kBlock);
    switch (dim) {
        case 1: hipLaunchKernelGGL((k_rk_accumulate<1, W>), dim3(na), dim3(kBlock), 0, s, acc, terms); break;
        case 2: hipLaunchKernelGGL((k_rk_accumulate<2, W>), dim3(na), dim3(kBlock), 0, s, acc, terms); break;
        case 3: hipLaunchKernelGGL((k_rk_accumulate<3, W>), dim3(na), dim3(kBlock), 0, s, acc, terms); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <class W>
hipError_t launch_rk_update(int dim, const W& cells, int64_t cell_items, const W* acc, int64_t acc_items,
                            const RKParams& r, hipStream_t s) {
    const unsigned nb = unsigned((cell_items + kBlock - 1) / kBlock);
    if (cell_items > 0) {
        // The Euler form is compiled without the diffusive-flux terms.
        switch (dim * 2 + (r.ns ? 1 : 0)) {
            case 2: hipLaunchKernelGGL((k_rk_update<1, false, W>), dim3(nb), dim3(kBlock), 0, s, cells, r); break;
            case 3: hipLaunchKernelGGL((k_rk_update<1, true, W>), dim3(nb), dim3(kBlock), 0, s, cells, r); break;
            case 4: hipLaunchKernelGGL((k_rk_update<2, false, W>), dim3(nb), dim3(kBlock), 0, s, cells, r); break;
            case 5: hipLaunchKernelGGL((k_rk_update<2, true, W>), dim3(nb), dim3(kBlock), 0, s, cells, r); break;
            case 6: hipLaunchKernelGGL((k_rk_update<3, false, W>), dim3(nb), dim3(kBlock), 0, s, cells, r); break;
            case 7: hipLaunchKernelGGL((k_rk_update<3, true, W>), dim3(nb), dim3(kBlock), 0, s, cells, r); break;
            default: return hipErrorInvalidValue;
        }
    }
    if (acc && acc_items > 0) {
        // this stage's terms: its intermediates with nonzero gamma, in order
        AccTerms t{};
        t.neq = r.neq;
        for (int n = 0; n < r.nc; n++)
            if (r.gamma[n] != 0.0) {
                t.src[t.n] = n;
                t.gamma[t.n] = r.gamma[n];
                t.n++;
            }
        if (hipError_t e = hipGetLastError(); e != hipSuccess) return e;
        return launch_rk_accumulate<W>(dim, *acc, acc_items, t, s);
    }
    return hipGetLastError();
}

template hipError_t launch_rk_update<SegSingleT<RKPatch>>(int, const SegSingleT<RKPatch>&, int64_t,
                                                          const SegSingleT<RKPatch>*, int64_t, const RKParams&,
                                                          hipStream_t);
template hipError_t launch_rk_update<SegTableT<RKPatch>>(int, const SegTableT<RKPatch>&, int64_t,
                                                         const SegTableT<RKPatch>*, int64_t, const RKParams&,
                                                         hipStream_t);
template hipError_t launch_rk_accumulate<SegTableT<RKPatch>>(int, const SegTableT<RKPatch>&, int64_t,
                                                             const AccTerms&, hipStream_t);

}  // namespace hm
