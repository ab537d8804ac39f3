// hm_adjoint.hip -- VJP kernels of the single-species convective-flux RHS and
// their host driver (see hm_adjoint.hpp for the structure).
#include <algorithm>
#include <cstring>

#include "hm_adjoint.hpp"
#include "hm_adjoint_face.hpp"
#include "hm_ctx.hpp"
#include "hm_tables.hpp"

// Waves per SIMD the two face-adjoint kernels are compiled for (a few bytes
// of spill at these settings cost less than the lower occupancy; the 3D flux
// half would spill over 100 bytes at 3 waves and stays at 2).
#ifndef HM_RECON_WAVES
#define HM_RECON_WAVES 2
#endif
#ifndef HM_ADJ_FLUX_WAVES
#define HM_ADJ_FLUX_WAVES 3
#endif

namespace hm {

// ----------------------------------------------------------------------------
// Kernels.
// ----------------------------------------------------------------------------

// Addressing shared by the two face-adjoint kernels: the face's cell indices
// (c0, c1, c2: the cell right of the face) from its midpoint-face number.
template <int D, int A>
__device__ __forceinline__ void mid_cell(const Boxes<D>& B, int64_t loc, int& c0, int& c1, int& c2) {
    const int m0 = B.mid_ext(A, 0), m1 = B.mid_ext(A, 1);
    // 32-bit index arithmetic (a work segment has fewer than 2^31 items).
    const uint32_t r = uint32_t(loc) / uint32_t(m0);
    c0 = int(uint32_t(loc) - r * uint32_t(m0));
    c1 = (D > 1) ? int((D > 2) ? r % uint32_t(m1) : r) : 0;
    c2 = (D > 2) ? int(r / uint32_t(m1)) : 0;
    if constexpr (A == 0) c0 -= 1;
    if constexpr (A == 1) c1 -= 1;
    if constexpr (A == 2) c2 -= 1;
}

template <int D, int A>
__device__ __forceinline__ void load_stencil(const DevPatch& F, const Boxes<D>& B, int c0, int c1, int c2,
                                             double (&V)[6][D + 2], double& cL, double& cR) {
    constexpr int NEQ = D + 2;
    const int64_t v4 = B.vol(kGc);
    const int64_t st = (A == 0) ? 1 : (A == 1) ? int64_t(B.ext(0, kGc)) : int64_t(B.ext(0, kGc)) * B.ext(1, kGc);
    const int64_t iR = B.idx(kGc, c0, c1, c2);
    const double* pr = F.prim;
#pragma unroll
    for (int m = 0; m < 6; m++)
#pragma unroll
        for (int e = 0; e < NEQ; e++) V[m][e] = pr[e * v4 + iR + (m - 3) * st];
    cL = pr[NEQ * v4 + iR - st];
    cR = pr[NEQ * v4 + iR];
}

// Level runs (KParams::lam_src >= 0): the stage adjoint along the line of
// axis a through cell (c0, c1, c2), and the face-flux adjoint it implies,
// F_bar_f = lam_scale (lam_f - lam_{f-1}) / dx_a for output faces f = 0..n_a
// (lam is zero outside the interior).
template <int D>
struct LamLine {
    const double* L;  // component 0 at the line's cell 0
    int64_t vol, st;
    int n;
    double sc;
    __device__ __forceinline__ LamLine(const AdjPatch& P, const Boxes<D>& B, const KParams& k, int a, int c0, int c1,
                                       int c2) {
        const int lg = P.lam.g;
        const int64_t le0 = B.n[0] + 2 * lg, le1 = (D > 1) ? B.n[1] + 2 * lg : 1;
        vol = le0 * le1 * ((D > 2) ? B.n[2] + 2 * lg : 1);
        st = (a == 0) ? 1 : (a == 1) ? le0 : le0 * le1;
        const int x0 = (a == 0) ? 0 : c0, x1 = (a == 1) ? 0 : c1, x2 = (a == 2) ? 0 : c2;
        const int64_t i0 = (x0 + lg) + int64_t(x1 + ((D > 1) ? lg : 0)) * le0 + int64_t(x2 + ((D > 2) ? lg : 0)) * le0 * le1;
        L = P.lam.L0 + int64_t(k.lam_src) * P.lam.stride + i0;
        n = B.n[a];
        sc = div_rn(k.lam_scale, P.f.dx[a]);
    }
    __device__ __forceinline__ double cell(int e, int x) const {
        return (x >= 0 && x < n) ? L[e * vol + int64_t(x) * st] : 0.0;
    }
    __device__ __forceinline__ double face(int e, int f) const {
        return (f >= 0 && f <= n) ? sc * (cell(e, f) - cell(e, f - 1)) : 0.0;
    }
};

// Per midpoint face of axis A: adjoint of the midpoint-to-face combination
// F_i = dt (1/30 (m_{i+1} + m_{i-1}) - 3/10 (n_i + n_{i-1}) + 23/15 m_i)
// (WCNS56-HLLC-HLL.cpp:1200-1282) with respect to the midpoint flux m_j, then
// the first half of the face's reverse sweep (Riemann flux, fallback,
// back-projection) into the hand-over buffer H; H's last entry is 2 when the
// face receives no adjoint at all.
template <int D, int A, int SCHEME, int PP, int QQ, bool STORED>
__device__ __forceinline__ void adj_face_flux(const AdjPatch& P, int64_t loc, const KParams& k) {
    constexpr int NEQ = D + 2;
    constexpr int NH = face_adj_h_size<D>();
    const DevPatch& F = P.f;
    const Boxes<D> B(F);
    int c0, c1, c2;
    mid_cell<D, A>(B, loc, c0, c1, c2);
    const int j = (A == 0) ? c0 : (A == 1) ? c1 : c2;
    const int nA = B.n[A];
    const int64_t fv = B.face_vol(A);
    auto fidx = [&](int jj) -> int64_t {
        return B.face_idx(A, (A == 0) ? jj : c0, (A == 1) ? jj : c1, (A == 2) ? jj : c2);
    };
    const double dt = k.dt;
    double mb[NEQ];
    bool any = false;
    if (k.lam_src >= 0) {
        const LamLine<D> Ll(P, B, k, A, c0, c1, c2);
#pragma unroll
        for (int e = 0; e < NEQ; e++) {
            const double fl = Ll.face(e, j - 1), fc = Ll.face(e, j), fr = Ll.face(e, j + 1);
            mb[e] = dt * (1.0 / 30.0 * (fl + fr) + 23.0 / 15.0 * fc);
            any = any || (mb[e] != 0.0);
        }
    } else {
        const double* Fb = P.Fb[A];
#pragma unroll
        for (int e = 0; e < NEQ; e++) {
            const double fl = (j - 1 >= 0 && j - 1 <= nA) ? Fb[e * fv + fidx(j - 1)] : 0.0;
            const double fr = (j + 1 >= 0 && j + 1 <= nA) ? Fb[e * fv + fidx(j + 1)] : 0.0;
            const double fc = (j >= 0 && j <= nA) ? Fb[e * fv + fidx(j)] : 0.0;
            mb[e] = dt * (1.0 / 30.0 * (fl + fr) + 23.0 / 15.0 * fc);
            any = any || (mb[e] != 0.0);
        }
    }
    const int64_t mv = B.mid_vol(A);
    double* H = P.H[A];
    if (!any) {
        H[(NH - 1) * mv + loc] = 2.0;
        return;
    }
    auto sensor_hll = [&]() -> bool {
        if constexpr (D > 1) {
            const int64_t vs = B.vol(kGs);
            const int64_t tR = B.idx(kGs, c0, c1, c2);
            const int64_t sts =
                (A == 0) ? 1 : (A == 1) ? int64_t(B.ext(0, kGs)) : int64_t(B.ext(0, kGs)) * B.ext(1, kGs);
            const int64_t tL = tR - sts;
            const double th_avg = 0.5 * (F.sens[tL] + F.sens[tR]);
            const double om_avg = 0.5 * (F.sens[vs + tL] + F.sens[vs + tR]);
            const double sensor = div_rn(-th_avg, fabs(th_avg) + om_avg + kEps);
            return sensor > 0.65;
        }
        return false;
    };
    double h[NH];
    if constexpr (STORED) {
        // W-, W+ as the recording forward pass reconstructed them.
        const int64_t v4 = B.vol(kGc);
        const int64_t st = (A == 0) ? 1 : (A == 1) ? int64_t(B.ext(0, kGc)) : int64_t(B.ext(0, kGc)) * B.ext(1, kGc);
        const double* pc = F.prim + B.idx(kGc, c0, c1, c2);  // cell right of the face
        const double* ws = k.wrec + F.wrec[A] + loc;
        double Wm[NEQ], Wp[NEQ], V2[NEQ], V3[NEQ];
#pragma unroll
        for (int e = 0; e < NEQ; e++) {
            Wm[e] = ws[e * mv];
            Wp[e] = ws[(NEQ + e) * mv];
            V2[e] = pc[e * v4 - st];
            V3[e] = pc[e * v4];
        }
        const bool use_hll = sensor_hll();
        const double ra = double(1) / double(2) * (V2[0] + V3[0]);
        const double ca = double(1) / double(2) * (pc[NEQ * v4 - st] + pc[NEQ * v4]);
        face_adj_single_after_recon<D, A>(Wm, Wp, V2, V3, ra, ca, k.gamma[0], use_hll, mb, h);
    } else {
        double V[6][NEQ];
        double cL, cR;
        load_stencil<D, A>(F, B, c0, c1, c2, V, cL, cR);
        const bool use_hll = sensor_hll();
        face_adj_single_flux<D, A, SCHEME, PP, QQ>(V, cL, cR, k.gamma[0], LDConst{k.p, k.q, k.C, k.alpha_tau},
                                                   use_hll, mb, h);
    }
#pragma unroll
    for (int c = 0; c < NH; c++) H[c * mv + loc] = h[c];
}

// Second half: the reverse reconstruction; writes the face record G (the
// characteristic-value adjoints of the six stencil points per field and the
// projection averages' adjoints; see face_adj_single_recon).  Faces without
// adjoint (H flag 2) write nothing: the gather skips them.
template <int D, int A, int SCHEME, int PP, int QQ>
__device__ __forceinline__ void adj_face_recon(const AdjPatch& P, int64_t loc, const KParams& k) {
    constexpr int NEQ = D + 2;
    constexpr int NH = face_adj_h_size<D>();
    const DevPatch& F = P.f;
    const Boxes<D> B(F);
    const uint32_t mv = uint32_t(B.mid_vol(A)), l32 = uint32_t(loc);
    const double* H = P.H[A];
    // component c of a face record starts at H + c * mv (64-bit base), so the
    // 32-bit offset stays below one component's size
    if (ldg(H + size_t(NH - 1) * mv, l32) == 2.0) return;
    int c0, c1, c2;
    mid_cell<D, A>(B, loc, c0, c1, c2);
    const int64_t v4 = B.vol(kGc);
    const uint32_t st = (A == 0) ? 1u : (A == 1) ? uint32_t(B.ext(0, kGc)) : uint32_t(B.ext(0, kGc) * B.ext(1, kGc));
    const uint32_t iR = uint32_t(B.idx(kGc, c0, c1, c2));  // cell right of the face
    const double* pr = F.prim;
    const double cL = ldg(pr + NEQ * v4, iR - st);
    const double cR = ldg(pr + NEQ * v4, iR);
    double* G = P.G[A];
    face_adj_single_recon<D, A, SCHEME, PP, QQ>(
        [&](int m, int e) { return ldg(pr + e * v4, iR + uint32_t(m - 3) * st); }, cL, cR,
        LDConst{k.p, k.q, k.C, k.alpha_tau}, [&](int c) { return ldg(H + size_t(c) * mv, l32); },
        [&](int c, double v) { stg(G + size_t(c) * mv, l32, v); });
}

template <int D, int SCHEME, int PP, int QQ, class W, bool STORED = false>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(D < 3 ? HM_ADJ_FLUX_WAVES : 2))) void k_adj_face_flux(
    W w, KParams k) {
    const int64_t gid = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    const AdjPatch* P;
    int ax;
    int64_t loc;
    if (!w.find(gid, P, ax, loc)) return;
    if (ax == 0) {
        adj_face_flux<D, 0, SCHEME, PP, QQ, STORED>(*P, loc, k);
    } else if constexpr (D > 1) {
        if (ax == 1) {
            adj_face_flux<D, 1, SCHEME, PP, QQ, STORED>(*P, loc, k);
        } else if constexpr (D > 2) {
            adj_face_flux<D, 2, SCHEME, PP, QQ, STORED>(*P, loc, k);
        }
    }
}

template <int D, int SCHEME, int PP, int QQ, class W>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(HM_RECON_WAVES))) void k_adj_face_recon(
    W w, KParams k) {
    const int64_t gid = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    const AdjPatch* P;
    int ax;
    int64_t loc;
    if (!w.find(gid, P, ax, loc)) return;
    if (ax == 0) {
        adj_face_recon<D, 0, SCHEME, PP, QQ>(*P, loc, k);
    } else if constexpr (D > 1) {
        if (ax == 1) {
            adj_face_recon<D, 1, SCHEME, PP, QQ>(*P, loc, k);
        } else if constexpr (D > 2) {
            adj_face_recon<D, 2, SCHEME, PP, QQ>(*P, loc, k);
        }
    }
}

// Per cell of the 4-ghost box: gather the stencil adjoints of every midpoint
// face whose stencil holds the cell, add the node-flux adjoint, and map V_bar,
// c_bar through the derived-data Jacobian (FlowModelSingleSpecies.cpp:2649-3179:
// u = m/rho, p = (gamma-1) rho (E/rho - |u|^2/2), c = sqrt(gamma p / rho)).
template <int D, class W>
__global__ __launch_bounds__(kBlock) void k_adj_cells(W w, KParams k) {
    const int64_t gid = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    const AdjPatch* P;
    int ax;
    int64_t loc;
    if (!w.find(gid, P, ax, loc)) return;
    constexpr int NEQ = D + 2;
    constexpr int IP = D + 1;
    const DevPatch& F = P->f;
    const Boxes<D> B(F);
    const int e0 = B.ext(0, kGc), e1 = B.ext(1, kGc);
    int c[3];
    // 32-bit index arithmetic (a work segment has fewer than 2^31 items).
    const uint32_t r = uint32_t(loc) / uint32_t(e0);
    c[0] = int(uint32_t(loc) - r * uint32_t(e0)) - kGc;
    c[1] = (D > 1) ? int((D > 2) ? r % uint32_t(e1) : r) - kGc : 0;
    c[2] = (D > 2) ? int(r / uint32_t(e1)) - kGc : 0;
    const double gam = k.gamma[0];
    const double dt = k.dt;
    const int64_t v4 = B.vol(kGc);
    const double rho = F.prim[loc];
    double u[D];
#pragma unroll
    for (int b = 0; b < D; b++) u[b] = F.prim[(1 + b) * v4 + loc];
    const double p = F.prim[IP * v4 + loc];
    const double cs = F.prim[NEQ * v4 + loc];
    const int64_t iq = q_idx<D>(F, c[0], c[1], c[2]);  // Q_bar index
    const int64_t iqr = iq + k.qoff;                    // Q index (reads)

    double Vb[NEQ], Qbd[NEQ];
#pragma unroll
    for (int e = 0; e < NEQ; e++) {
        Vb[e] = 0.0;
        Qbd[e] = 0.0;
    }
    double cb = 0.0;
#pragma unroll
    for (int a = 0; a < D; a++) {
        bool trans_ok = true;
#pragma unroll
        for (int b = 0; b < D; b++)
            if (b != a && (c[b] < 0 || c[b] >= B.n[b])) trans_ok = false;
        if (!trans_ok) continue;
        const int na = B.n[a];
        const int64_t mv = B.mid_vol(a);
        const int64_t sta = (a == 0) ? 1 : (a == 1) ? int64_t(e0) : int64_t(e0) * e1;
        constexpr int NH = face_adj_h_size<D>();
#pragma unroll
        for (int m = 0; m < 6; m++) {
            const int jf = c[a] + 3 - m;
            if (jf < -1 || jf > na + 1) continue;
            const int64_t fi = B.mid_idx(a, (a == 0) ? jf : c[0], (a == 1) ? jf : c[1], (a == 2) ? jf : c[2]);
            const uint32_t fi32 = uint32_t(fi), mv32 = uint32_t(mv);
            const double flag = ldg(P->H[a] + size_t(NH - 1) * mv32, fi32);
            if (flag == 2.0) continue;
            // projection averages of the face: its cells sit at stencil points 2 and 3
            const int64_t l2 = loc + (2 - m) * sta, l3 = loc + (3 - m) * sta;
            const double ra = double(1) / double(2) * (F.prim[l2] + F.prim[l3]);
            const double ca = double(1) / double(2) * (F.prim[NEQ * v4 + l2] + F.prim[NEQ * v4 + l3]);
            const double* G = P->G[a];
            auto gat = [&](int cc) { return ldg(G + size_t(cc) * mv32, fi32); };
            if (a == 0) {
                face_record_to_point<D, 0>(flag, gat, ra, ca, m, Vb, cb);
            } else if constexpr (D > 1) {
                if (a == 1)
                    face_record_to_point<D, 1>(flag, gat, ra, ca, m, Vb, cb);
                else if constexpr (D > 2)
                    face_record_to_point<D, 2>(flag, gat, ra, ca, m, Vb, cb);
            }
        }
        // Node flux of this cell enters the faces c_a and c_a + 1 with weight -3/10 dt.
        if (c[a] >= -1 && c[a] <= na) {
            const int64_t fv = B.face_vol(a);
            const double* Fb = P->Fb[a];
            auto fidx = [&](int jj) -> int64_t {
                return B.face_idx(a, (a == 0) ? jj : c[0], (a == 1) ? jj : c[1], (a == 2) ? jj : c[2]);
            };
            const bool hasL = c[a] >= 0, hasR = c[a] + 1 <= na;
            double nb[NEQ];
            if (k.lam_src >= 0) {
                const LamLine<D> Ll(*P, B, k, a, c[0], c[1], c[2]);
#pragma unroll
                for (int e = 0; e < NEQ; e++) nb[e] = -dt * (3.0 / 10.0) * (Ll.face(e, c[a]) + Ll.face(e, c[a] + 1));
            } else {
#pragma unroll
                for (int e = 0; e < NEQ; e++) {
                    const double fL = hasL ? Fb[e * fv + fidx(c[a])] : 0.0;
                    const double fR = hasR ? Fb[e * fv + fidx(c[a] + 1)] : 0.0;
                    nb[e] = -dt * (3.0 / 10.0) * (fL + fR);
                }
            }
            // n[0] = Q[1+a]; n[1+b] = u_a Q[1+b] (+ p if b == a); n[IP] = u_a (E + p)
            Qbd[1 + a] += nb[0];
#pragma unroll
            for (int b = 0; b < D; b++) {
                const double Qv = F.Q[1 + b][iqr];
                Qbd[1 + b] += u[a] * nb[1 + b];
                Vb[1 + a] += Qv * nb[1 + b];
                if (b == a) Vb[IP] += nb[1 + b];
            }
            const double E = F.Q[IP][iqr];
            Vb[1 + a] += (E + p) * nb[IP];
            Qbd[IP] += u[a] * nb[IP];
            Vb[IP] += u[a] * nb[IP];
        }
    }
    // Derived-data Jacobian.
    // Reciprocals instead of quotients: the adjoint carries no bit-exactness requirement.
    const double irho = rcp_rn(rho), i2cr = rcp_rn(2.0 * cs * rho);
    const double pb = Vb[IP] + cb * gam * i2cr;
    double rhob = Vb[0] - cb * gam * p * i2cr * irho;
    double ke = 0.0;
#pragma unroll
    for (int b = 0; b < D; b++) ke += u[b] * u[b];
    rhob += pb * (gam - 1.0) * 0.5 * ke;
    double* const* Qb = P->Qb;
#pragma unroll
    for (int b = 0; b < D; b++) {
        Qb[1 + b][iq] = -pb * (gam - 1.0) * u[b] + Vb[1 + b] * irho + Qbd[1 + b];
        rhob -= Vb[1 + b] * u[b] * irho;
    }
    Qb[0][iq] = rhob + Qbd[0];
    Qb[IP][iq] = pb * (gam - 1.0) + Qbd[IP];
}

// Zero the whole Q_bar ghost box (cells beyond the 4-ghost box stay zero).
template <int D, class W>
__global__ __launch_bounds__(kBlock) void k_adj_zero(W w) {
    const int64_t gid = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    const AdjPatch* P;
    int ax;
    int64_t loc;
    if (!w.find(gid, P, ax, loc)) return;
#pragma unroll
    for (int e = 0; e < D + 2; e++) P->Qb[e][loc] = 0.0;
}

namespace {

unsigned nblk(int64_t items) { return unsigned((items + kBlock - 1) / kBlock); }

template <int D, int SCHEME, int PP, int QQ>
hipError_t launch_adj_dsq(const AdjPlanDev& pl, const KParams& k, hipStream_t s) {
    using W = SegTableT<AdjPatch>;
    if (pl.n_qbox > 0) hipLaunchKernelGGL((k_adj_zero<D, W>), dim3(nblk(pl.n_qbox)), dim3(kBlock), 0, s, pl.qbox);
    if (pl.n_face > 0) {
        if (k.wrec)
            hipLaunchKernelGGL((k_adj_face_flux<D, SCHEME, PP, QQ, W, true>), dim3(nblk(pl.n_face)), dim3(kBlock), 0, s,
                               pl.face, k);
        else
            hipLaunchKernelGGL((k_adj_face_flux<D, SCHEME, PP, QQ, W>), dim3(nblk(pl.n_face)), dim3(kBlock), 0, s,
                               pl.face, k);
        hipLaunchKernelGGL((k_adj_face_recon<D, SCHEME, PP, QQ, W>), dim3(nblk(pl.n_face)), dim3(kBlock), 0, s,
                           pl.face, k);
    }
    if (pl.n_cell > 0)
        hipLaunchKernelGGL((k_adj_cells<D, W>), dim3(nblk(pl.n_cell)), dim3(kBlock), 0, s, pl.cell, k);
    return hipGetLastError();
}

template <int D>
hipError_t launch_adj_d(int scheme, const AdjPlanDev& pl, const KParams& k, hipStream_t s) {
    const bool dflt = (k.p == 2 && k.q == 4);
    if (scheme == kWCNS6_LD)
        return dflt ? launch_adj_dsq<D, kWCNS6_LD, 2, 4>(pl, k, s) : launch_adj_dsq<D, kWCNS6_LD, -1, -1>(pl, k, s);
    return (k.p == 2) ? launch_adj_dsq<D, kWCNS5_JS, 2, 4>(pl, k, s) : launch_adj_dsq<D, kWCNS5_JS, -1, -1>(pl, k, s);
}

}  // namespace

hipError_t launch_adjoint_single(int dim, int scheme, const AdjPlanDev& plan, const KParams& k, hipStream_t s) {
    switch (dim) {
        case 1: return launch_adj_d<1>(scheme, plan, k, s);
        case 2: return launch_adj_d<2>(scheme, plan, k, s);
        case 3: return launch_adj_d<3>(scheme, plan, k, s);
        default: return hipErrorInvalidValue;
    }
}

// ----------------------------------------------------------------------------
// Host driver.
// ----------------------------------------------------------------------------

AdjointSingle::~AdjointSingle() { release(cached_); }

void AdjointSingle::release(Owned& o) {
    if (o.scratch || o.tables) (void)hipDeviceSynchronize();
    rec_->releaseOwnedPlan(o.fwd);
    if (o.scratch) (void)hipFree(o.scratch);
    if (o.tables) (void)hipFree(o.tables);
    o.scratch = nullptr;
    o.tables = nullptr;
    o.plan = AdjPlanDev{};
}

int AdjointSingle::build(const hm_vjp_patch_view* views, int np, Owned& out, const int64_t* wrec_off,
                         const AdjLamSource* lam) {
    const hm_config& c = rec_->config();
    if (c.model != HM_MODEL_SINGLE_SPECIES)
        return rec_->fail(HM_ERR_UNSUPPORTED, "the adjoint is available for the single-species model");
    if (np < 1 || !views) return rec_->fail(HM_ERR_INVALID_ARGUMENT, "empty patch list");
    const int d = c.dim, neq = rec_->numEqn();
    std::vector<hm_patch_view> fv(np);
    for (int i = 0; i < np; i++) {
        const hm_vjp_patch_view& v = views[i];
        hm_patch_view w{};
        for (int a = 0; a < 3; a++) {
            w.n[a] = v.n[a];
            w.ghost[a] = v.ghost[a];
            w.dx[a] = v.dx[a];
        }
        for (int e = 0; e < HM_MAX_EQN; e++) w.Q[e] = v.Q[e];
        for (int a = 0; a < d; a++) {
            if (!v.F_bar[a]) return rec_->fail(HM_ERR_INVALID_ARGUMENT, "NULL flux adjoint");
            // the forward stages used here (derived data, sensor) never write F
            w.F[a] = const_cast<double*>(v.F_bar[a]);
        }
        for (int e = 0; e < neq; e++)
            if (!v.Q_bar[e]) return rec_->fail(HM_ERR_INVALID_ARGUMENT, "NULL Q adjoint array");
        fv[i] = w;
    }
    release(out);
    // Scratch: forward scratch of every patch, then the G buffers.
    const size_t fwd_doubles = rec_->scratchDoubles(fv.data(), np);
    std::vector<size_t> goff(size_t(np) * 3, 0);
    size_t gtot = 0;
    for (int i = 0; i < np; i++)
        for (int a = 0; a < d; a++) {
            int64_t mv = 1;
            for (int b = 0; b < d; b++) mv *= fv[i].n[b] + ((b == a) ? 3 : 0);
            goff[size_t(i) * 3 + a] = gtot;
            gtot += (size_t(6 * neq + 2) * size_t(mv) + 31) / 32 * 32;
            gtot += (size_t(2 * neq + 3) * size_t(mv) + 31) / 32 * 32;  // H
        }
    if (hipMalloc(&out.scratch, (fwd_doubles + gtot) * sizeof(double)) != hipSuccess) {
        out.scratch = nullptr;
        return rec_->fail(HM_ERR_OUT_OF_MEMORY, "adjoint scratch");
    }
    double* base = static_cast<double*>(out.scratch);
    if (int st = rec_->buildOwnedPlan(fv.data(), np, base, out.fwd)) return st;
    TableImage<AdjPatch> face, cell, qbox;
    face.items.resize(np);
    double* pb = base;
    for (int i = 0; i < np; i++) {
        AdjPatch& A = face.items[i];
        std::memset(&A, 0, sizeof A);
        A.f = rec_->devPatch(fv[i], pb, wrec_off ? wrec_off + size_t(i) * 3 : nullptr);
        if (lam) A.lam = lam[i];
        pb += rec_->patchScratchDoubles(fv[i]);
        for (int a = 0; a < 3; a++) {
            A.Fb[a] = (a < d) ? views[i].F_bar[a] : nullptr;
            A.G[a] = (a < d) ? base + fwd_doubles + goff[size_t(i) * 3 + a] : nullptr;
            if (a < d) {
                int64_t mv = 1;
                for (int b = 0; b < d; b++) mv *= fv[i].n[b] + ((b == a) ? 3 : 0);
                A.H[a] = A.G[a] + (size_t(6 * neq + 2) * size_t(mv) + 31) / 32 * 32;
            } else {
                A.H[a] = nullptr;
            }
        }
        for (int e = 0; e < HM_MAX_EQN; e++) A.Qb[e] = (e < neq) ? views[i].Q_bar[e] : nullptr;
        int64_t v4 = 1, vq = 1;
        for (int a = 0; a < d; a++) {
            v4 *= fv[i].n[a] + 2 * kGc;
            vq *= fv[i].n[a] + 2 * fv[i].ghost[a];
            int64_t mv = 1;
            for (int b = 0; b < d; b++) mv *= fv[i].n[b] + ((b == a) ? 3 : 0);
            face.add(i, a, mv);
        }
        cell.add(i, 0, v4);
        qbox.add(i, 0, vq);
    }
    cell.items = face.items;
    qbox.items = face.items;
    const size_t bytes = face.bytes() + cell.bytes() + qbox.bytes();
    if (hipMalloc(&out.tables, bytes) != hipSuccess) {
        out.tables = nullptr;
        return rec_->fail(HM_ERR_OUT_OF_MEMORY, "adjoint tables");
    }
    std::vector<char> h(bytes, 0);
    char* dv = static_cast<char*>(out.tables);
    size_t o = 0;
    out.plan.face = face.write(h.data() + o, dv + o);
    out.plan.n_face = face.total;
    o += face.bytes();
    out.plan.cell = cell.write(h.data() + o, dv + o);
    out.plan.n_cell = cell.total;
    o += cell.bytes();
    out.plan.qbox = qbox.write(h.data() + o, dv + o);
    // k_adj_cells writes every cell of the 4-ghost box; only wider ghost boxes need the zero fill.
    bool wider = false;
    for (int i = 0; i < np; i++)
        for (int a = 0; a < d; a++) wider = wider || fv[i].ghost[a] > kGc;
    out.plan.n_qbox = wider ? qbox.total : 0;
    if (hipMemcpy(out.tables, h.data(), bytes, hipMemcpyHostToDevice) != hipSuccess)
        return rec_->fail(HM_ERR_DEVICE, "adjoint table upload");
    return HM_OK;
}

int AdjointSingle::run(const Owned& o, double dt, hipStream_t s, const double* wst, int lam_src, double lam_scale,
                       int64_t qoff) {
    if (int st = rec_->runOwnedPlanStages(o.fwd, dt, 0, 1, s, nullptr, qoff)) return st;  // derived data and sensor
    KParams k = rec_->kernelParams();
    k.dt = dt;
    k.wrec = const_cast<double*>(wst);  // read only here
    k.lam_src = lam_src;
    k.lam_scale = lam_scale;
    k.qoff = qoff;
    const hm_config& c = rec_->config();
    const hipError_t e = launch_adjoint_single(c.dim, c.scheme, o.plan, k, s);
    if (e != hipSuccess) return rec_->fail(HM_ERR_DEVICE, std::string("adjoint launch failed: ") + hipGetErrorString(e));
    return HM_OK;
}

int AdjointSingle::vjp(const hm_vjp_patch_view* views, int np, double dt, hipStream_t s) {
    if (np < 0 || (np > 0 && !views)) return rec_->fail(HM_ERR_INVALID_ARGUMENT, "invalid patch list");
    if (np == 0) return HM_OK;
    if (int st = rec_->selectDevice()) return st;
    // calls on other streams finish with the cached scratch first
    if (order_.enter(s) != hipSuccess) return rec_->fail(HM_ERR_DEVICE, "hipStreamWaitEvent failed");
    const bool same = int(cached_views_.size()) == np && cached_.tables &&
                      std::memcmp(cached_views_.data(), views, sizeof(hm_vjp_patch_view) * size_t(np)) == 0;
    if (!same) {
        if (hipStreamSynchronize(s) != hipSuccess) return rec_->fail(HM_ERR_DEVICE, "stream synchronize failed");
        cached_views_.clear();
        if (int st = build(views, np, cached_)) return st;
        cached_views_.assign(views, views + np);
    }
    if (int st = run(cached_, dt, s)) return st;
    if (order_.leave(s) != hipSuccess) return rec_->fail(HM_ERR_DEVICE, "hipEventRecord failed");
    return HM_OK;
}

}  // namespace hm

extern "C" {

int hm_convective_flux_vjp_on_patches(hm_ctx* ctx, const hm_vjp_patch_view* patches, int num_patches, double time,
                                      double dt, int rk_step, void* stream) {
    (void)time;
    (void)rk_step;
    if (!ctx || !ctx->rec) return HM_ERR_INVALID_ARGUMENT;
    if (!ctx->adj) {
        ctx->adj.reset(new (std::nothrow) hm::AdjointSingle(ctx->rec.get()));
        if (!ctx->adj) return HM_ERR_OUT_OF_MEMORY;
    }
    return ctx->adj->vjp(patches, num_patches, dt, static_cast<hipStream_t>(stream));
}

int hm_convective_flux_vjp_on_patch(hm_ctx* ctx, const hm_vjp_patch_view* patch, double time, double dt, int rk_step,
                                    void* stream) {
    return hm_convective_flux_vjp_on_patches(ctx, patch, 1, time, dt, rk_step, stream);
}

}  // extern "C"
