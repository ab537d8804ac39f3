// hm_level.hip -- one AMR level advanced by the Runge-Kutta stage loop of
// RungeKuttaLevelIntegrator::advanceLevel (RungeKuttaLevelIntegrator.cpp:1457-1939,
// stage loop 1672-1745, default SSP-RK3 weights 3907-3924), entirely on the GPU:
//
//   fill ghosts of the scratch state (same level + periodic images,
//   then the physical boundary conditions)                              :1568
//   for each stage sn:
//     intermediate[sn] <- scratch (whole ghost box)                       :1677
//     sn > 0: fill ghosts of intermediate[sn] (the same fill)             :1686-1711
//     S[sn] <- 0, F[sn], S[sn] <- batched convective RHS of intermediate  Euler.cpp:904-1001
//     scratch interior <- stage update from intermediate[0..sn]           Euler.cpp:1003-1680
//
// One step is a fixed sequence of launches on one stream, so it is captured
// once into a hipGraph and replayed.  Ghost cells inside the domain that no
// same-level patch covers (coarse-fine boundaries) keep the values written by
// the caller, as the reference's intermediate fills do not use the coarser
// level; ghost cells outside the domain get the physical boundary conditions
// set for their sides (hm_level_set_boundary_conditions), else also keep the
// caller's values.
//
// Also here: the stage-update entry points for callers that keep their own
// patch data (per patch and batched).
#include <algorithm>
#include <cstring>
#include <map>

#include "hm_ctx.hpp"
#include "hm_dt.hpp"
#include "hm_rk.hpp"
#include "hm_tables.hpp"

namespace hm {

constexpr int kMaxBuf = 1 + kMaxRK;  // scratch + intermediates

struct LevelPatch {
    int n[3];
    int64_t gvol;                // ghost-box cells (component stride)
    double* Q[kMaxBuf];          // buffer b: nq components over the ghost box
    double* S[kMaxRK];           // source of stage n: neq components over the interior
};

struct HaloItem {
    int dst, src;
    int lo[3];   // first destination cell, in the destination's ghost-box coordinates (0-based)
    int ext[3];  // extent of the copied box
    int off[3];  // source ghost-box coordinates = destination + off
};

// Ghost-cell copy from same-level neighbours (and periodic images).
template <int D>
__global__ __launch_bounds__(kBlock) void k_halo(SegTableT<HaloItem> w, const LevelPatch* patches, int buf, int nq,
                                                 int g) {
    const int64_t gid = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    const HaloItem* it;
    int ax;
    int64_t loc;
    if (!w.find(gid, it, ax, loc)) return;
    // 32-bit index arithmetic (a work segment has fewer than 2^31 items).
    const uint32_t r = uint32_t(loc) / uint32_t(it->ext[0]);
    const int i = int(uint32_t(loc) - r * uint32_t(it->ext[0]));
    const int j = (D > 1) ? int((D > 2) ? r % uint32_t(it->ext[1]) : r) : 0;
    const int k = (D > 2) ? int(r / uint32_t(it->ext[1])) : 0;
    const LevelPatch& Pd = patches[it->dst];
    const LevelPatch& Ps = patches[it->src];
    auto gidx = [&](const LevelPatch& P, int a0, int a1, int a2) -> int64_t {
        const int e0 = P.n[0] + 2 * g;
        const int e1 = (D > 1) ? P.n[1] + 2 * g : 1;
        return a0 + int64_t(a1) * e0 + int64_t(a2) * e0 * e1;
    };
    const int d0 = it->lo[0] + i, d1 = it->lo[1] + j, d2 = it->lo[2] + k;
    const int64_t id = gidx(Pd, d0, d1, d2);
    const int64_t is = gidx(Ps, d0 + it->off[0], d1 + it->off[1], d2 + it->off[2]);
    for (int c = 0; c < nq; c++) Pd.Q[buf][c * Pd.gvol + id] = Ps.Q[buf][c * Ps.gvol + is];
}

// Physical boundary conditions of the ghost cells outside the domain
// (BasicCartesianBoundaryUtilities1.cpp, BasicCartesianBoundaryUtilities2.cpp:215-440
// edge and 590-900 node fills, BasicCartesianBoundaryUtilities3.cpp face, edge
// and node fills), applied to every conservative variable as
// EulerBoundaryConditions::setPhysicalBoundaryConditions does
// (EulerBoundaryConditions.cpp:749-1100): FLOW copies the boundary cell,
// REFLECT and SYMMETRY copy the mirror image about the boundary face, REFLECT
// negating the momentum component normal to it (scalar variables take
// SYMMETRY where the deck says REFLECT, EulerBoundaryConditions.cpp:196-345),
// DIRICHLET writes fixed values.  A cell outside the domain along several axes
// (an edge or node box) composes the rules of those axes -- the reference's
// node and edge conditions named after the x axis, whose pivot lies in the
// box of the next axis -- so every item reads only cells inside the domain
// (interior or same-level ghosts, filled before) and the items run in one
// launch.  With a DIRICHLET axis among them, the first such axis (x, y, z
// order) gives the values.
struct BcItem {
    int dst;
    int lo[3], ext[3];  // region in the patch's ghost-box coordinates
    int mode[3];        // per axis: 0 same coordinate, 1 mirror (source = par - x), 2 clamp (source = par)
    int par[3];
    int flip;           // bit b: momentum component b negated
    int side;           // >= 0: the DIRICHLET values of side `side` (2 axis + upper)
};

template <int D>
__global__ __launch_bounds__(kBlock) void k_bc(SegTableT<BcItem> w, const LevelPatch* patches, int buf, int nq,
                                               int g, int mom0, const double* dir) {
    const int64_t gid = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    const BcItem* it;
    int ax;
    int64_t loc;
    if (!w.find(gid, it, ax, loc)) return;
    const uint32_t r = uint32_t(loc) / uint32_t(it->ext[0]);
    const int x[3] = {it->lo[0] + int(uint32_t(loc) - r * uint32_t(it->ext[0])),
                      it->lo[1] + ((D > 1) ? int((D > 2) ? r % uint32_t(it->ext[1]) : r) : 0),
                      it->lo[2] + ((D > 2) ? int(r / uint32_t(it->ext[1])) : 0)};
    int y[3];
    for (int a = 0; a < 3; a++) y[a] = it->mode[a] == 0 ? x[a] : (it->mode[a] == 1 ? it->par[a] - x[a] : it->par[a]);
    const LevelPatch& P = patches[it->dst];
    const int e0 = P.n[0] + 2 * g;
    const int e1 = (D > 1) ? P.n[1] + 2 * g : 1;
    const int64_t id = x[0] + int64_t(x[1]) * e0 + int64_t(x[2]) * e0 * e1;
    const int64_t is = y[0] + int64_t(y[1]) * e0 + int64_t(y[2]) * e0 * e1;
    double* Q = P.Q[buf];
    for (int c = 0; c < nq; c++) {
        double v;
        if (it->side >= 0) {
            v = dir[it->side * nq + c];
        } else {
            v = Q[c * P.gvol + is];
            const int b = c - mom0;
            if (b >= 0 && b < D && ((it->flip >> b) & 1)) v = -v;
        }
        Q[c * P.gvol + id] = v;
    }
}

// Whole-ghost-box copy between two state buffers of every patch.
__global__ __launch_bounds__(kBlock) void k_copy_state(SegTableT<LevelPatch> w, int from, int to, int nq) {
    const int64_t gid = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    const LevelPatch* P;
    int ax;
    int64_t loc;
    if (!w.find(gid, P, ax, loc)) return;
    for (int c = 0; c < nq; c++) P->Q[to][c * P->gvol + loc] = P->Q[from][c * P->gvol + loc];
}

// Zero the source of one stage (Euler.cpp:913-926); a kernel, not a memset, so
// that the step can be captured in a graph.  Only the advective equations'
// components (from first on) are ever written by the RHS; the others stay at
// the zero the storage was allocated with.
__global__ __launch_bounds__(kBlock) void k_zero_source(SegTableT<LevelPatch> w, int stage, int first) {
    const int64_t gid = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    const LevelPatch* P;
    int ax;
    int64_t loc;
    if (!w.find(gid, P, ax, loc)) return;
    P->S[stage][int64_t(first) * P->n[0] * P->n[1] * P->n[2] + loc] = 0.0;
}


// ----------------------------------------------------------------------------
// Adjoint of the stage loop (discrete adjoint of one level's RK steps).
// ----------------------------------------------------------------------------

struct AdjLevelPatch {
    int n[3];
    int64_t gvol;               // ghost-box cells (component stride)
    double dx[3];
    double* lam[kMaxRK + 1];    // adjoints of the stage states (interior cells used, ghost-box layout)
    double* qbar;               // VJP output over the ghost box (neq comps)
    int rev_first, rev_count;   // halo items whose source is this patch, in the reverse list
};

template <int D>
__device__ __forceinline__ int64_t gbox_idx(const int* n, int g, int i, int j, int k) {
    const int e0 = n[0] + 2 * g;
    const int e1 = (D > 1) ? n[1] + 2 * g : 1;
    return (i + g) + int64_t(j + ((D > 1) ? g : 0)) * e0 + int64_t(k + ((D > 2) ? g : 0)) * e0 * e1;
}

// The VJP output of this patch at interior cell (i, j, k) plus the ghost-cell
// adjoints that the same-level halo copied from that cell (the reverse of
// k_halo, as a gather: deterministic, no atomics).
template <int D, int NEQ>
__device__ __forceinline__ void gather_qbar(const SegTableT<AdjLevelPatch>& w, const HaloItem* rev,
                                            const AdjLevelPatch* P, int i, int j, int k, int g, int64_t gi,
                                            double (&acc)[NEQ]) {
#pragma unroll
    for (int e = 0; e < NEQ; e++) acc[e] = P->qbar[e * P->gvol + gi];
    // A cell at least g cells from every face of its patch is no ghost
    // image: a ghost cell lies within g cells of its own patch, which is
    // disjoint from this one.
    bool deep = i >= g && i < P->n[0] - g;
    if constexpr (D > 1) deep = deep && j >= g && j < P->n[1] - g;
    if constexpr (D > 2) deep = deep && k >= g && k < P->n[2] - g;
    if (deep) return;
    const int x[3] = {i + g, (D > 1) ? j + g : 0, (D > 2) ? k + g : 0};  // ghost-box coordinates
    for (int q = P->rev_first; q < P->rev_first + P->rev_count; q++) {
        const HaloItem& it = rev[q];
        int y[3];
        bool in = true;
        for (int a = 0; a < 3; a++) {
            y[a] = x[a] - it.off[a];
            if (y[a] < it.lo[a] || y[a] >= it.lo[a] + it.ext[a]) in = false;
        }
        if (!in) continue;
        const AdjLevelPatch& Pd = w.patches[it.dst];
        const int e0 = Pd.n[0] + 2 * g;
        const int e1 = (D > 1) ? Pd.n[1] + 2 * g : 1;
        const int64_t di = y[0] + int64_t(y[1]) * e0 + int64_t(y[2]) * e0 * e1;
#pragma unroll
        for (int e = 0; e < NEQ; e++) acc[e] += Pd.qbar[e * Pd.gvol + di];
    }
}

// lam[m] (interior) += the gathered VJP output (a stage whose update has
// several nonzero beta: every VJP but the last of the stage).
template <int D>
__global__ __launch_bounds__(kBlock) void k_halo_reverse(SegTableT<AdjLevelPatch> w, const HaloItem* rev, int m,
                                                         int g) {
    const int64_t gid = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    const AdjLevelPatch* P;
    int ax;
    int64_t loc;
    if (!w.find(gid, P, ax, loc)) return;
    // 32-bit index arithmetic (a work segment has fewer than 2^31 items).
    const uint32_t r = uint32_t(loc) / uint32_t(P->n[0]);
    const int i = int(uint32_t(loc) - r * uint32_t(P->n[0]));
    const int j = (D > 1) ? int((D > 2) ? r % uint32_t(P->n[1]) : r) : 0;
    const int k = (D > 2) ? int(r / uint32_t(P->n[1])) : 0;
    const int64_t gi = gbox_idx<D>(P->n, g, i, j, k);
    constexpr int NEQ = D + 2;  // the level adjoint is built for the single-species model
    double acc[NEQ];
    gather_qbar<D, NEQ>(w, rev, P, i, j, k, g, gi, acc);
#pragma unroll
    for (int e = 0; e < NEQ; e++) P->lam[m][e * P->gvol + gi] += acc[e];
}

// The stage adjoints' update after the VJPs of stage sn, one launch per stage.
// Stage sn computes Q_{sn+1} = sum_m alpha_m Q_m + sum_m beta_m dt L(Q_m), so
// lam[m] receives alpha_m lam[sn+1] and, for the stage input whose VJP ran
// last, the gathered VJP output.  Entry t writes
//   lam[dst_t] = (init_t ? 0 : lam[m_t]) + (t == q ? gathered VJP : 0) + a_t lam[sn+1]
// (the addition order of a zero fill, the halo gather and the alpha update in
// sequence, so the rounding is the same).  init_t marks the first write to
// lam[m_t] in the step's reverse sweep, which replaces a zero fill; dst_t
// differs from m_t only for the step input in the last stage, which lands
// directly in the step's output slot lam[nst].
struct StageAdj {
    int nt;                 // entries
    int src;                // sn + 1
    int q;                  // entry receiving the VJP output, -1 if none
    int m[kMaxRK], dst[kMaxRK], init[kMaxRK];
    double a[kMaxRK];
};

template <int D>
__global__ __launch_bounds__(kBlock) void k_stage_adjoint(SegTableT<AdjLevelPatch> w, const HaloItem* rev,
                                                          StageAdj sa, int g) {
    const int64_t gid = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    const AdjLevelPatch* P;
    int ax;
    int64_t loc;
    if (!w.find(gid, P, ax, loc)) return;
    // 32-bit index arithmetic (a work segment has fewer than 2^31 items).
    const uint32_t r = uint32_t(loc) / uint32_t(P->n[0]);
    const int i = int(uint32_t(loc) - r * uint32_t(P->n[0]));
    const int j = (D > 1) ? int((D > 2) ? r % uint32_t(P->n[1]) : r) : 0;
    const int k = (D > 2) ? int(r / uint32_t(P->n[1])) : 0;
    const int64_t gi = gbox_idx<D>(P->n, g, i, j, k);
    constexpr int NEQ = D + 2;
    double acc[NEQ];
    if (sa.q >= 0) gather_qbar<D, NEQ>(w, rev, P, i, j, k, g, gi, acc);
    const double* S = P->lam[sa.src];
    double sv[NEQ];
#pragma unroll
    for (int e = 0; e < NEQ; e++) sv[e] = S[e * P->gvol + gi];
    for (int t = 0; t < sa.nt; t++) {
        const double* L = P->lam[sa.m[t]];
        double* O = P->lam[sa.dst[t]];
        const double a = sa.a[t];
#pragma unroll
        for (int e = 0; e < NEQ; e++) {
            const int64_t idx = e * P->gvol + gi;
            double v = sa.init[t] ? 0.0 : L[idx];
            if (t == sa.q) v += acc[e];
            if (a != 0.0) v += a * sv[e];
            O[idx] = v;
        }
    }
}

// Interior updates of the stage adjoints: kLamCopy: lam[dst] = lam[src];
// kLamZeroBelow: lam[0 .. dst) = 0.
enum { kLamCopy = 2, kLamZeroBelow = 3 };
template <int D>
__global__ __launch_bounds__(kBlock) void k_lam_op(SegTableT<AdjLevelPatch> w, int op, int dst, int src, int neq,
                                                   int g) {
    const int64_t gid = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    const AdjLevelPatch* P;
    int ax;
    int64_t loc;
    if (!w.find(gid, P, ax, loc)) return;
    // 32-bit index arithmetic (a work segment has fewer than 2^31 items).
    const uint32_t r = uint32_t(loc) / uint32_t(P->n[0]);
    const int i = int(uint32_t(loc) - r * uint32_t(P->n[0]));
    const int j = (D > 1) ? int((D > 2) ? r % uint32_t(P->n[1]) : r) : 0;
    const int k = (D > 2) ? int(r / uint32_t(P->n[1])) : 0;
    const int64_t gi = gbox_idx<D>(P->n, g, i, j, k);
    double* L = P->lam[dst];
    const double* S = P->lam[src];
    if (op == kLamZeroBelow) {
        for (int mm = 0; mm < dst; mm++)
            for (int e = 0; e < neq; e++) P->lam[mm][e * P->gvol + gi] = 0.0;
        return;
    }
    for (int e = 0; e < neq; e++) L[e * P->gvol + gi] = S[e * P->gvol + gi];  // kLamCopy
}

// Whole-ghost-box copy between a state buffer and a storage slot
// (patch p's part of a slot starts at slot + off[p]).  dir 0: buffer -> slot.
__global__ __launch_bounds__(kBlock) void k_store(SegTableT<LevelPatch> w, int buf, double* slot, const int64_t* off,
                                                  int nq, int dir) {
    const int64_t gid = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    const LevelPatch* P;
    int ax;
    int64_t loc;
    if (!w.find(gid, P, ax, loc)) return;
    double* sp = slot + off[P - w.patches];
    for (int c = 0; c < nq; c++) {
        if (dir == 0)
            sp[c * P->gvol + loc] = P->Q[buf][c * P->gvol + loc];
        else
            P->Q[buf][c * P->gvol + loc] = sp[c * P->gvol + loc];
    }
}

namespace {

unsigned blocks_for(int64_t items) { return unsigned((items + kBlock - 1) / kBlock); }

}  // namespace

class Level {
   public:
    // diff: the context's diffusive flux reconstructor (four-equation model
    // with transport properties), or NULL for the Euler form.  With it the
    // level keeps 6 ghost cells and its steps add the diffusive flux to every
    // stage in the Navier-Stokes form (NavierStokes.cpp:1108-1247, 1704-1750).
    Level(ConvectiveFluxReconstructorGPU* rec, DiffusiveFluxReconstructorGPU* diff = nullptr) : rec_(rec), diff_(diff) {}

    ~Level() {
        (void)rec_->selectDevice();
        (void)hipDeviceSynchronize();
        for (auto& kv : graphs_) (void)hipGraphExecDestroy(kv.second);
        for (auto& p : rhs_plan_) rec_->releaseOwnedPlan(p);
        if (diff_)
            for (auto& p : diff_plan_) diff_->releaseOwnedPlan(p);
        if (diff_scratch_) (void)hipFree(diff_scratch_);
        if (adj_)
            adj_->release(vjp_plan_);
        if (adj_mem_) (void)hipFree(adj_mem_);
        if (store_) (void)hipFree(store_);
        if (wst_) (void)hipFree(wst_);
        if (rhs_scratch_) (void)hipFree(rhs_scratch_);
        if (dt_mem_) (void)hipFree(dt_mem_);
        if (cap_stream_) (void)hipStreamDestroy(cap_stream_);
        if (ev_a_) (void)hipEventDestroy(ev_a_);
        if (ev_b_) (void)hipEventDestroy(ev_b_);
        if (tables_) (void)hipFree(tables_);
        if (storage_) (void)hipFree(storage_);
        if (acc_mem_) (void)hipFree(acc_mem_);
    }

    int init(int np, const int* lo, const int* n, const int* dlo, const int* dn, const int* periodic, const double* dx) {
        const hm_config& c = rec_->config();
        d_ = c.dim;
        neq_ = rec_->numEqn();
        nq_ = rec_->numConservativeArrays();
        g_ = diff_ ? kGd : kGc;
        np_ = np;
        if (np < 1) return rec_->fail(HM_ERR_INVALID_ARGUMENT, "a level needs at least one patch");
        for (int a = 0; a < 3; a++) {
            dlo_[a] = a < d_ ? dlo[a] : 0;
            dn_[a] = a < d_ ? dn[a] : 1;
            periodic_[a] = a < d_ ? periodic[a] != 0 : false;
            dx_[a] = a < d_ ? dx[a] : 1.0;
        }
        lo_.assign(size_t(np) * 3, 0);
        n_.assign(size_t(np) * 3, 1);
        for (int p = 0; p < np; p++)
            for (int a = 0; a < d_; a++) {
                lo_[3 * p + a] = lo[3 * p + a];
                n_[3 * p + a] = n[3 * p + a];
                if (n_[3 * p + a] < 1) return rec_->fail(HM_ERR_INVALID_ARGUMENT, "patch extent below 1");
                if (lo_[3 * p + a] < dlo_[a] || lo_[3 * p + a] + n_[3 * p + a] > dlo_[a] + dn_[a])
                    return rec_->fail(HM_ERR_INVALID_ARGUMENT, "patch outside the domain");
            }
        for (int a = 0; a < d_; a++)
            if (periodic_[a] && dn_[a] < g_)
                return rec_->fail(HM_ERR_INVALID_ARGUMENT, "periodic domain narrower than the ghost width");
        // default SSP-RK3 (RungeKuttaLevelIntegrator.cpp:3907-3924)
        const double a3[3][3] = {{1.0, 0, 0}, {3.0 / 4.0, 1.0 / 4.0, 0}, {1.0 / 3.0, 0.0, 2.0 / 3.0}};
        const double b3[3][3] = {{1.0, 0, 0}, {0.0, 1.0 / 4.0, 0}, {0.0, 0.0, 2.0 / 3.0}};
        const double c3[3][3] = {{1.0 / 6.0, 0, 0}, {0.0, 1.0 / 6.0, 0}, {0.0, 0.0, 2.0 / 3.0}};
        nst_ = 3;
        for (int s = 0; s < kMaxRK; s++)
            for (int m = 0; m < kMaxRK; m++) {
                alpha_[s][m] = (s < 3 && m < 3) ? a3[s][m] : 0.0;
                beta_[s][m] = (s < 3 && m < 3) ? b3[s][m] : 0.0;
                gamma_[s][m] = (s < 3 && m < 3) ? c3[s][m] : 0.0;
            }
        plan_stage_adjoint();
        return allocate();
    }

    int setCoefficients(int nst, const double* a, const double* b, const double* c) {
        if (nst < 1 || nst > kMaxRK) return rec_->fail(HM_ERR_INVALID_ARGUMENT, "1..5 Runge-Kutta stages");
        nst_ = nst;
        for (int s = 0; s < kMaxRK; s++)
            for (int m = 0; m < kMaxRK; m++) {
                const bool in = s < nst && m <= s;
                alpha_[s][m] = in ? a[s * nst + m] : 0.0;
                beta_[s][m] = in ? b[s * nst + m] : 0.0;
                gamma_[s][m] = in && c ? c[s * nst + m] : 0.0;
            }
        plan_stage_adjoint();
        // Launches already queued read the tables rewritten below.
        if (hipDeviceSynchronize() != hipSuccess) return rec_->fail(HM_ERR_DEVICE, "device synchronize failed");
        drop_graphs();
        return build_rk_tables();
    }

    // Interior cells only: ghost cells keep their values.
    int writePatch(int p, const double* Q) {
        if (p < 0 || p >= np_ || !Q) return rec_->fail(HM_ERR_INVALID_ARGUMENT, "bad patch index or array");
        std::vector<double> h(size_t(nq_) * gvol(p));
        if (hipDeviceSynchronize() != hipSuccess ||
            hipMemcpy(h.data(), patch_[p].Q[0], h.size() * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
            return rec_->fail(HM_ERR_DEVICE, "hipMemcpy from the level failed");
        const double* src = Q;
        const int64_t vol = ivol(p);
        for (int c = 0; c < nq_; c++)
            for (int64_t l = 0; l < vol; l++) h[size_t(c) * gvol(p) + gidx_of_interior(p, l)] = src[c * vol + l];
        if (hipMemcpy(patch_[p].Q[0], h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
            return rec_->fail(HM_ERR_DEVICE, "hipMemcpy to the level failed");
        return HM_OK;
    }

    int readPatch(int p, double* Q) {
        if (p < 0 || p >= np_ || !Q) return rec_->fail(HM_ERR_INVALID_ARGUMENT, "bad patch index or array");
        std::vector<double> h(size_t(nq_) * gvol(p));
        if (hipDeviceSynchronize() != hipSuccess) return rec_->fail(HM_ERR_DEVICE, "device synchronize failed");
        if (hipMemcpy(h.data(), patch_[p].Q[0], h.size() * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
            return rec_->fail(HM_ERR_DEVICE, "hipMemcpy from the level failed");
        const int64_t vol = ivol(p);
        for (int c = 0; c < nq_; c++)
            for (int64_t l = 0; l < vol; l++) Q[c * vol + l] = h[size_t(c) * gvol(p) + gidx_of_interior(p, l)];
        return HM_OK;
    }

    int patchArray(int p, double** Q, int* ghost) {
        if (p < 0 || p >= np_ || !Q) return rec_->fail(HM_ERR_INVALID_ARGUMENT, "bad patch index");
        *Q = patch_[p].Q[0];
        if (ghost) *ghost = g_;
        return HM_OK;
    }

    int fillGhosts(hipStream_t s) { return launch_halo(0, s); }

    // Physical boundary conditions of the sides of non-periodic axes
    // (HM_BC_*, xlo, xhi, ylo, yhi, zlo, zhi); dirichlet: [2 dim][nq] values.
    int setBoundaryConditions(const int* cond, const double* dirichlet) {
        int c[6] = {HM_BC_NONE, HM_BC_NONE, HM_BC_NONE, HM_BC_NONE, HM_BC_NONE, HM_BC_NONE};
        for (int sd = 0; sd < 2 * d_; sd++) {
            c[sd] = cond[sd];
            if (c[sd] < HM_BC_NONE || c[sd] > HM_BC_DIRICHLET)
                return rec_->fail(HM_ERR_INVALID_ARGUMENT, "unknown boundary condition");
            if (c[sd] != HM_BC_NONE && periodic_[sd / 2])
                return rec_->fail(HM_ERR_INVALID_ARGUMENT, "a periodic axis takes no boundary condition");
            if (c[sd] != HM_BC_NONE && dn_[sd / 2] < g_)
                return rec_->fail(HM_ERR_INVALID_ARGUMENT, "domain narrower than the ghost width on a boundary axis");
            if (c[sd] == HM_BC_DIRICHLET && !dirichlet)
                return rec_->fail(HM_ERR_INVALID_ARGUMENT, "DIRICHLET sides need their values");
        }
        // Launches already queued read the tables rewritten below.
        if (hipDeviceSynchronize() != hipSuccess) return rec_->fail(HM_ERR_DEVICE, "device synchronize failed");
        drop_graphs();
        for (int sd = 0; sd < 6; sd++) {
            bc_cond_[sd] = c[sd];
            for (int e = 0; e < nq_; e++)
                bc_values_[sd * nq_ + e] = (c[sd] == HM_BC_DIRICHLET) ? dirichlet[sd * nq_ + e] : 0.0;
        }
        has_bc_ = false;
        for (int sd = 0; sd < 6; sd++) has_bc_ = has_bc_ || bc_cond_[sd] != HM_BC_NONE;
        return build_tables();
    }
    bool hasBoundaryConditions() const { return has_bc_; }

    // Autotuner of the level's batched RHS: every candidate schedule is timed
    // on the first stage's RHS of the current state (ghosts filled), the
    // fastest becomes the schedule of every stage's RHS.  Captured steps are
    // dropped (their launches used the old tables).
    int autotune(double dt, int reps, Schedule& best, double& best_ms, double& default_ms) {
        if (!(dt > 0.0)) return rec_->fail(HM_ERR_INVALID_ARGUMENT, "bad dt");
        if (int st = rec_->selectDevice()) return st;
        hipStream_t s = cap_stream_;
        if (hipDeviceSynchronize() != hipSuccess) return rec_->fail(HM_ERR_DEVICE, "device synchronize failed");
        drop_graphs();
        hipLaunchKernelGGL(k_copy_state, dim3(blocks_for(copy_items_)), dim3(kBlock), 0, s, copy_, 0, 1, nq_);
        if (int st = launch_halo(1, s)) return st;
        if (zero_items_ > 0)
            hipLaunchKernelGGL(k_zero_source, dim3(blocks_for(zero_items_)), dim3(kBlock), 0, s, zero_, 0, src_first_);
        const bool single = rec_->config().model == HM_MODEL_SINGLE_SPECIES;
        ConvectiveFluxReconstructorGPU::OwnedPlan trial;
        const Schedule dflt = rec_->defaultSchedule(views_[0].data(), np_);
        int st = rec_->tuneSchedules(
            s, reps,
            [&](const Schedule& sc) {
                if (hipStreamSynchronize(s) != hipSuccess) return rec_->fail(HM_ERR_DEVICE, "stream synchronize failed");
                return rec_->buildOwnedPlan(views_[0].data(), np_, static_cast<double*>(rhs_scratch_), trial,
                                            single ? wst_off_.data() : nullptr, &sc);
            },
            [&]() {
                if (int e = rec_->runOwnedPlan(trial, dt, s)) return e;
                if (diff_) return diff_->runOwnedPlan(diff_plan_[0], dt, s);
                return int(HM_OK);
            },
            best, best_ms, default_ms, dflt);
        if (hipStreamSynchronize(s) != hipSuccess && st == HM_OK) st = rec_->fail(HM_ERR_DEVICE, "stream synchronize failed");
        rec_->releaseOwnedPlan(trial);
        if (st) return st;
        return setSchedule(best);
    }

    int setSchedule(const Schedule& sc) {
        if (*rec_->scheduleError(sc)) return rec_->fail(HM_ERR_INVALID_ARGUMENT, rec_->scheduleError(sc));
        if (hipDeviceSynchronize() != hipSuccess) return rec_->fail(HM_ERR_DEVICE, "device synchronize failed");
        drop_graphs();
        sched_ = sc;
        return build_rhs_plans();
    }
    const Schedule& schedule() const { return sched_; }

    // Flux sums for the synchronization of a two-level hierarchy: F_acc of
    // every axis and S_acc accumulate gamma_s F_s and gamma_s S_s over the
    // stages of every step (Euler.cpp:1230-1270, 1360-1420, 1560-1660) until
    // clearFluxSums; in the Navier-Stokes form also Fd_acc the diffusive
    // fluxes (a flux variable of its own, NavierStokes.cpp:393-402).  Drops
    // the captured steps (their tables change).
    int enableFluxSums() {
        if (acc_mem_) return HM_OK;
        size_t total = 0;
        std::vector<size_t> off(size_t(np_) * 4), doff(size_t(np_) * 3, 0);
        for (int p = 0; p < np_; p++) {
            for (int a = 0; a < d_; a++) {
                off[size_t(p) * 4 + a] = total;
                total += up256(sizeof(double) * size_t(neq_) * fvol(p, a));
            }
            off[size_t(p) * 4 + 3] = total;
            total += up256(sizeof(double) * size_t(neq_) * ivol(p));
            if (diff_)
                for (int a = 0; a < d_; a++) {
                    doff[size_t(p) * 3 + a] = total;
                    total += up256(sizeof(double) * size_t(neq_) * fvol(p, a));
                }
        }
        if (hipDeviceSynchronize() != hipSuccess) return rec_->fail(HM_ERR_DEVICE, "device synchronize failed");
        if (hipMalloc(&acc_mem_, total) != hipSuccess) {
            acc_mem_ = nullptr;
            return rec_->fail(HM_ERR_OUT_OF_MEMORY, "level flux sums");
        }
        if (hipMemset(acc_mem_, 0, total) != hipSuccess) return rec_->fail(HM_ERR_DEVICE, "flux-sum clear");
        acc_bytes_ = total;
        facc_.assign(size_t(np_) * 3, nullptr);
        fdacc_.assign(size_t(np_) * 3, nullptr);
        sacc_.assign(size_t(np_), nullptr);
        char* b = static_cast<char*>(acc_mem_);
        for (int p = 0; p < np_; p++) {
            for (int a = 0; a < d_; a++) {
                facc_[size_t(p) * 3 + a] = reinterpret_cast<double*>(b + off[size_t(p) * 4 + a]);
                if (diff_) fdacc_[size_t(p) * 3 + a] = reinterpret_cast<double*>(b + doff[size_t(p) * 3 + a]);
            }
            sacc_[p] = reinterpret_cast<double*>(b + off[size_t(p) * 4 + 3]);
        }
        drop_graphs();
        return build_tables();
    }

    int clearFluxSums(hipStream_t s) {
        if (!acc_mem_) return HM_OK;
        return hipMemsetAsync(acc_mem_, 0, acc_bytes_, s) == hipSuccess ? HM_OK
                                                                        : rec_->fail(HM_ERR_DEVICE, "flux-sum clear");
    }

    // Pieces of the level a two-level hierarchy drives.
    int stepOnce(double dt, hipStream_t s) { return step(dt, s); }
    int fillGhostsOf(int buf, hipStream_t s) { return launch_halo(buf, s); }
    const LevelPatch* devPatches() const { return patch_dev(); }
    const LevelPatch& hostPatch(int p) const { return patch_[p]; }
    double* fluxSum(int p, int a) const { return facc_[size_t(p) * 3 + a]; }
    double* diffusiveFluxSum(int p, int a) const { return fdacc_.empty() ? nullptr : fdacc_[size_t(p) * 3 + a]; }
    double* sourceSum(int p) const { return sacc_[p]; }
    int dim() const { return d_; }
    bool diffusive() const { return diff_ != nullptr; }
    int numEqn() const { return neq_; }
    const int* patchLo(int p) const { return &lo_[3 * p]; }
    const int* patchN(int p) const { return &n_[3 * p]; }
    const int* domainLo() const { return dlo_; }
    const int* domainN() const { return dn_; }
    bool periodic(int a) const { return periodic_[a]; }
    const double* dx() const { return dx_; }
    int64_t ghostVolume(int p) const { return gvol(p); }
    ConvectiveFluxReconstructorGPU* reconstructor() const { return rec_; }
    // The patch whose interior holds cell x (global index, wrapped on
    // periodic axes), or -1.  Patches are looked up through buckets of
    // kBucket^d cells over the domain, built on first use.
    int patchOf(const int* x) {
        if (bucket_start_.empty()) build_buckets();
        int v[3] = {0, 0, 0};
        int64_t b = 0, st = 1;
        for (int a = 0; a < d_; a++) {
            v[a] = x[a];
            if (periodic_[a]) v[a] = dlo_[a] + ((v[a] - dlo_[a]) % dn_[a] + dn_[a]) % dn_[a];
            if (v[a] < dlo_[a] || v[a] >= dlo_[a] + dn_[a]) return -1;
            b += int64_t((v[a] - dlo_[a]) / kBucket) * st;
            st *= nbucket_[a];
        }
        for (int i = bucket_start_[b]; i < bucket_start_[b + 1]; i++) {
            const int p = bucket_items_[i];
            bool in = true;
            for (int a = 0; a < d_ && in; a++) in = v[a] >= lo_[3 * p + a] && v[a] < lo_[3 * p + a] + n_[3 * p + a];
            if (in) return p;
        }
        return -1;
    }

    static constexpr int kBucket = 16;
    void build_buckets() {
        int64_t nb = 1;
        for (int a = 0; a < 3; a++) {
            nbucket_[a] = a < d_ ? (dn_[a] + kBucket - 1) / kBucket : 1;
            nb *= nbucket_[a];
        }
        std::vector<int> count(size_t(nb) + 1, 0);
        auto each = [&](int p, auto&& f) {  // the buckets patch p's box (clipped to the domain) overlaps
            int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
            for (int a = 0; a < 3; a++) {
                if (a < d_) {
                    const int l = std::max(lo_[3 * p + a], dlo_[a]) - dlo_[a];
                    const int h = std::min(lo_[3 * p + a] + n_[3 * p + a], dlo_[a] + dn_[a]) - dlo_[a];
                    if (h <= l) return;
                    lo[a] = l / kBucket;
                    hi[a] = (h - 1) / kBucket;
                }
            }
            for (int z = lo[2]; z <= hi[2]; z++)
                for (int y = lo[1]; y <= hi[1]; y++)
                    for (int x = lo[0]; x <= hi[0]; x++) f((int64_t(z) * nbucket_[1] + y) * nbucket_[0] + x);
        };
        for (int p = 0; p < np_; p++) each(p, [&](int64_t b) { count[size_t(b) + 1]++; });
        for (int64_t b = 0; b < nb; b++) count[size_t(b) + 1] += count[size_t(b)];
        bucket_start_ = count;
        bucket_items_.assign(size_t(count[size_t(nb)]), 0);
        for (int p = 0; p < np_; p++) each(p, [&](int64_t b) { bucket_items_[size_t(count[size_t(b)]++)] = p; });
    }
    int nbucket_[3] = {1, 1, 1};
    std::vector<int> bucket_start_, bucket_items_;

    // num_steps steps of the configured Runge-Kutta scheme with fixed dt.
    int advance(double dt, int nsteps, bool use_graph, hipStream_t s) {
        if (nsteps < 0 || !(dt > 0.0)) return rec_->fail(HM_ERR_INVALID_ARGUMENT, "bad dt or step count");
        if (int st = rec_->selectDevice()) return st;
        if (!use_graph) {
            for (int it = 0; it < nsteps; it++)
                if (int st = step(dt, s)) return st;
            return HM_OK;
        }
        if (nsteps == 0) return HM_OK;
        auto itg = graphs_.find(dt);
        if (itg == graphs_.end()) {
            // Every table and scratch array exists since init, so the captured
            // step allocates and uploads nothing.
            if (int st = capture(dt)) return st;
            itg = graphs_.find(dt);
        }
        if (hipEventRecord(ev_a_, s) != hipSuccess || hipStreamWaitEvent(cap_stream_, ev_a_, 0) != hipSuccess)
            return rec_->fail(HM_ERR_DEVICE, "stream ordering failed");
        for (int it = 0; it < nsteps; it++)
            if (hipGraphLaunch(itg->second, cap_stream_) != hipSuccess)
                return rec_->fail(HM_ERR_DEVICE, "hipGraphLaunch failed");
        if (hipEventRecord(ev_b_, cap_stream_) != hipSuccess || hipStreamWaitEvent(s, ev_b_, 0) != hipSuccess)
            return rec_->fail(HM_ERR_DEVICE, "stream ordering failed");
        return HM_OK;
    }

    // num_steps steps whose dt the device computes at the start of each step
    // from the level state (the stable dt of RungeKuttaLevelIntegrator::getLevelDt
    // times cfl, hm_dt.hpp): the RHS kernels read it from device memory, so the
    // host never waits for it and one captured graph serves every step.
    int advanceCfl(double cfl, int nsteps, bool use_graph, hipStream_t s) {
        if (nsteps < 0 || !(cfl > 0.0)) return rec_->fail(HM_ERR_INVALID_ARGUMENT, "bad cfl or step count");
        if (diff_)
            return rec_->fail(HM_ERR_UNSUPPORTED,
                              "the stable dt of the Navier-Stokes form (maximum diffusivity, NavierStokes.cpp:"
                              "846-893) is not built: advance the level with a given dt");
        if (int st = rec_->selectDevice()) return st;
        if (int st = ensure_dt()) return st;
        if (launch_dt_reset_time(dt_buf_, s) != hipSuccess) return rec_->fail(HM_ERR_DEVICE, "dt launch failed");
        if (!use_graph) {
            for (int it = 0; it < nsteps; it++)
                if (int st = step_cfl(cfl, s)) return st;
            return HM_OK;
        }
        if (nsteps == 0) return HM_OK;
        const double key = -cfl;  // fixed-dt graphs are keyed by dt > 0
        auto itg = graphs_.find(key);
        if (itg == graphs_.end()) {
            if (int st = capture_with(key, [&](hipStream_t x) { return step_cfl(cfl, x); })) return st;
            itg = graphs_.find(key);
        }
        if (hipEventRecord(ev_a_, s) != hipSuccess || hipStreamWaitEvent(cap_stream_, ev_a_, 0) != hipSuccess)
            return rec_->fail(HM_ERR_DEVICE, "stream ordering failed");
        for (int it = 0; it < nsteps; it++)
            if (hipGraphLaunch(itg->second, cap_stream_) != hipSuccess)
                return rec_->fail(HM_ERR_DEVICE, "hipGraphLaunch failed");
        if (hipEventRecord(ev_b_, cap_stream_) != hipSuccess || hipStreamWaitEvent(s, ev_b_, 0) != hipSuccess)
            return rec_->fail(HM_ERR_DEVICE, "stream ordering failed");
        return HM_OK;
    }

    // The last advanceCfl step's reduction (kDtSlots doubles: spectral
    // radiuses, dt of the patches, of the level, of the step, the sum of the
    // step dts of that call); waits for the device.
    int dtInfo(double* out) {
        if (!out) return rec_->fail(HM_ERR_INVALID_ARGUMENT, "NULL output");
        if (!dt_mem_) return rec_->fail(HM_ERR_INVALID_ARGUMENT, "no stable-dt step has run on this level");
        if (hipDeviceSynchronize() != hipSuccess ||
            hipMemcpy(out, dt_buf_.res, sizeof(double) * kDtSlots, hipMemcpyDeviceToHost) != hipSuccess)
            return rec_->fail(HM_ERR_DEVICE, "dt read failed");
        return HM_OK;
    }

    // ---------------------------------------------------------------------
    // Discrete adjoint of the stage loop with checkpointing.
    // ---------------------------------------------------------------------

    // Forward steps that keep a checkpoint of the level state every `every`
    // steps, for a later adjoint().
    int advanceRecorded(double dt, int nsteps, int every, bool use_graph, hipStream_t s) {
        if (nsteps < 1 || every < 1 || !(dt > 0.0)) return rec_->fail(HM_ERR_INVALID_ARGUMENT, "bad dt, step count or interval");
        if (has_bc_)
            return rec_->fail(HM_ERR_UNSUPPORTED,
                              "the adjoint of the physical boundary conditions is not built (periodic and "
                              "same-level ghost fills only)");
        if (int st = rec_->selectDevice()) return st;
        if (int st = ensure_adjoint(nsteps, every)) return st;
        hipStream_t x = s;
        hipGraphExec_t exec = nullptr;
        if (use_graph) {
            auto itg = graphs_.find(dt);
            if (itg == graphs_.end()) {
                if (int st = capture(dt)) return st;
                itg = graphs_.find(dt);
            }
            exec = itg->second;
            x = cap_stream_;
            if (hipEventRecord(ev_a_, s) != hipSuccess || hipStreamWaitEvent(cap_stream_, ev_a_, 0) != hipSuccess)
                return rec_->fail(HM_ERR_DEVICE, "stream ordering failed");
        }
        for (int n = 0; n < nsteps; n++) {
            if (n % every == 0)
                if (int st = store_op(0, 1 + n / every, 0, x)) return st;
            if (exec) {
                if (hipGraphLaunch(exec, x) != hipSuccess) return rec_->fail(HM_ERR_DEVICE, "hipGraphLaunch failed");
            } else if (int st = step(dt, x)) {
                return st;
            }
        }
        if (use_graph &&
            (hipEventRecord(ev_b_, cap_stream_) != hipSuccess || hipStreamWaitEvent(s, ev_b_, 0) != hipSuccess))
            return rec_->fail(HM_ERR_DEVICE, "stream ordering failed");
        rec_dt_ = dt;
        rec_nsteps_ = nsteps;
        every_ = every;
        recorded_ = true;
        return HM_OK;
    }

    // Adjoint of the last advanceRecorded: the adjoint of the final state
    // (written with writeAdjoint) becomes the adjoint of the initial state,
    // segment by segment from the checkpoints.  The level state is restored.
    int adjoint(hipStream_t s) {
        if (!recorded_) return rec_->fail(HM_ERR_INVALID_ARGUMENT, "no recorded forward run");
        if (int st = rec_->selectDevice()) return st;
        if (int st = store_op(0, 0, 0, s)) return st;  // keep the final state
        const int nseg = (rec_nsteps_ + every_ - 1) / every_;
        const int traj0 = 1 + nckpt_;
        for (int seg = nseg - 1; seg >= 0; seg--) {
            const int n0 = seg * every_, n1 = std::min(n0 + every_, rec_nsteps_);
            if (int st = store_op(0, 1 + seg, 1, s)) return st;
            for (int t = n0; t < n1; t++)
                if (int st = step(rec_dt_, s, traj0 + (t - n0) * nst_, (t - n0) * nst_)) return st;
            for (int t = n1 - 1; t >= n0; t--)
                if (int st = adjoint_step(traj0 + (t - n0) * nst_, (t - n0) * nst_, s)) return st;
        }
        return store_op(0, 0, 1, s);
    }

    int writeAdjoint(int p, const double* lam) { return adjoint_io(p, const_cast<double*>(lam), true); }
    int readAdjoint(int p, double* lam) { return adjoint_io(p, lam, false); }

    int numPatches() const { return np_; }
    int ghost() const { return g_; }
    int numArrays() const { return nq_; }

   private:
    int64_t gvol(int p) const {
        int64_t v = 1;
        for (int a = 0; a < d_; a++) v *= n_[3 * p + a] + 2 * g_;
        return v;
    }
    int64_t ivol(int p) const {
        int64_t v = 1;
        for (int a = 0; a < d_; a++) v *= n_[3 * p + a];
        return v;
    }
    int64_t fvol(int p, int ax) const {
        int64_t v = 1;
        for (int a = 0; a < d_; a++) v *= n_[3 * p + a] + (a == ax);
        return v;
    }
    int64_t gidx_of_interior(int p, int64_t l) const {
        const int* n = &n_[3 * p];
        int c[3] = {0, 0, 0};
        c[0] = int(l % n[0]);
        if (d_ > 1) c[1] = int((l / n[0]) % n[1]);
        if (d_ > 2) c[2] = int(l / (int64_t(n[0]) * n[1]));
        int64_t idx = 0, st = 1;
        for (int a = 0; a < d_; a++) {
            idx += (c[a] + g_) * st;
            st *= n[a] + 2 * g_;
        }
        return idx;
    }

    int allocate() {
        // Storage: per patch, kMaxBuf state buffers, kMaxRK flux and source sets.
        size_t total = 0;
        std::vector<size_t> off(np_);
        for (int p = 0; p < np_; p++) {
            off[p] = total;
            total += up256(sizeof(double) * size_t(nq_) * gvol(p)) * kMaxBuf;
            for (int s = 0; s < kMaxRK; s++) {
                for (int a = 0; a < d_; a++) total += up256(sizeof(double) * size_t(neq_) * fvol(p, a));
                total += up256(sizeof(double) * size_t(neq_) * ivol(p));
                if (diff_)
                    for (int a = 0; a < d_; a++) total += up256(sizeof(double) * size_t(neq_) * fvol(p, a));
            }
        }
        if (hipMalloc(&storage_, total) != hipSuccess) return rec_->fail(HM_ERR_OUT_OF_MEMORY, "level storage");
        if (hipMemset(storage_, 0, total) != hipSuccess) return rec_->fail(HM_ERR_DEVICE, "level storage clear");
        patch_.resize(np_);
        fl_.assign(size_t(np_) * kMaxRK * 3, nullptr);
        fdl_.assign(size_t(np_) * kMaxRK * 3, nullptr);
        for (int p = 0; p < np_; p++) {
            char* b = static_cast<char*>(storage_) + off[p];
            LevelPatch& L = patch_[p];
            for (int a = 0; a < 3; a++) L.n[a] = a < d_ ? n_[3 * p + a] : 1;
            L.gvol = gvol(p);
            for (int q = 0; q < kMaxBuf; q++) {
                L.Q[q] = reinterpret_cast<double*>(b);
                b += up256(sizeof(double) * size_t(nq_) * gvol(p));
            }
            for (int s = 0; s < kMaxRK; s++) {
                for (int a = 0; a < d_; a++) {
                    fl_[(size_t(p) * kMaxRK + s) * 3 + a] = reinterpret_cast<double*>(b);
                    b += up256(sizeof(double) * size_t(neq_) * fvol(p, a));
                }
                L.S[s] = reinterpret_cast<double*>(b);
                b += up256(sizeof(double) * size_t(neq_) * ivol(p));
                if (diff_)
                    for (int a = 0; a < d_; a++) {
                        fdl_[(size_t(p) * kMaxRK + s) * 3 + a] = reinterpret_cast<double*>(b);
                        b += up256(sizeof(double) * size_t(neq_) * fvol(p, a));
                    }
            }
        }
        // RHS views of every intermediate buffer.
        views_.assign(kMaxRK, std::vector<hm_patch_view>(np_));
        for (int s = 0; s < kMaxRK; s++)
            for (int p = 0; p < np_; p++) {
                hm_patch_view v{};
                for (int a = 0; a < 3; a++) {
                    v.n[a] = a < d_ ? n_[3 * p + a] : 1;
                    v.ghost[a] = a < d_ ? g_ : 0;
                    v.dx[a] = dx_[a];
                }
                for (int e = 0; e < nq_; e++) v.Q[e] = patch_[p].Q[1 + s] + size_t(e) * gvol(p);
                for (int a = 0; a < d_; a++) v.F[a] = fl_[(size_t(p) * kMaxRK + s) * 3 + a];
                v.S = patch_[p].S[s];
                views_[s][p] = v;
            }
        // RHS scratch: the stages run one after another, so they share it.
        const size_t sd = rec_->scratchDoubles(views_[0].data(), np_);
        if (hipMalloc(&rhs_scratch_, sd * sizeof(double)) != hipSuccess)
            return rec_->fail(HM_ERR_OUT_OF_MEMORY, "level RHS scratch");
        // Face-state slot layout (single-species model): per patch and axis,
        // 2 NEQ values per midpoint face.
        const bool single = rec_->config().model == HM_MODEL_SINGLE_SPECIES;
        wst_off_.assign(size_t(np_) * 3, -1);
        wst_slot_doubles_ = 0;
        if (single)
            for (int p = 0; p < np_; p++)
                for (int a = 0; a < d_; a++) {
                    int64_t mv = 1;
                    for (int b = 0; b < d_; b++) mv *= n_[3 * p + b] + (b == a ? 3 : 0);
                    wst_off_[size_t(p) * 3 + a] = int64_t(wst_slot_doubles_);
                    wst_slot_doubles_ += up256(sizeof(double) * size_t(2 * neq_) * size_t(mv)) / sizeof(double);
                }
        sched_ = rec_->defaultSchedule(views_[0].data(), np_);
        if (int st = build_rhs_plans()) return st;
        if (diff_) {
            // Diffusive RHS of every intermediate buffer into the stage's
            // diffusive fluxes; the stages share one scratch region.
            std::vector<hm_patch_view> dv(np_);
            for (int s = 0; s < kMaxRK; s++) {
                for (int p = 0; p < np_; p++) {
                    dv[p] = views_[s][p];
                    for (int a = 0; a < d_; a++) dv[p].F[a] = fdl_[(size_t(p) * kMaxRK + s) * 3 + a];
                    dv[p].S = nullptr;
                }
                if (s == 0) {
                    const size_t dd = diff_->scratchDoubles(dv.data(), np_);
                    if (hipMalloc(&diff_scratch_, dd * sizeof(double)) != hipSuccess) {
                        diff_scratch_ = nullptr;
                        return rec_->fail(HM_ERR_OUT_OF_MEMORY, "level diffusive scratch");
                    }
                }
                if (int st = diff_->buildOwnedPlan(dv.data(), np_, static_cast<double*>(diff_scratch_), diff_plan_[s]))
                    return st;
            }
        }
        if (hipStreamCreateWithFlags(&cap_stream_, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&ev_a_, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&ev_b_, hipEventDisableTiming) != hipSuccess)
            return rec_->fail(HM_ERR_DEVICE, "stream or event creation failed");
        return build_tables();
    }

    // Work tables of every stage's batched RHS with the level's schedule.
    int build_rhs_plans() {
        const bool single = rec_->config().model == HM_MODEL_SINGLE_SPECIES;
        for (int s = 0; s < kMaxRK; s++)
            if (int st = rec_->buildOwnedPlan(views_[s].data(), np_, static_cast<double*>(rhs_scratch_), rhs_plan_[s],
                                              single ? wst_off_.data() : nullptr, &sched_))
                return st;
        return HM_OK;
    }

    // Same-level ghost copies, periodic images included.
    void build_halo(TableImage<HaloItem>& img) {
        int shifts[3][3];
        int nsh[3];
        for (int a = 0; a < 3; a++) {
            nsh[a] = 1;
            shifts[a][0] = 0;
            if (a < d_ && periodic_[a]) {
                shifts[a][1] = -dn_[a];
                shifts[a][2] = dn_[a];
                nsh[a] = 3;
            }
        }
        for (int p = 0; p < np_; p++) {
            int glo[3], ghi[3];  // ghost box of p (inclusive, global index space)
            for (int a = 0; a < 3; a++) {
                const bool used = a < d_;
                glo[a] = used ? lo_[3 * p + a] - g_ : 0;
                ghi[a] = used ? lo_[3 * p + a] + n_[3 * p + a] - 1 + g_ : 0;
            }
            for (int q = 0; q < np_; q++)
                for (int s0 = 0; s0 < nsh[0]; s0++)
                    for (int s1 = 0; s1 < nsh[1]; s1++)
                        for (int s2 = 0; s2 < nsh[2]; s2++) {
                            const int sh[3] = {shifts[0][s0], shifts[1][s1], shifts[2][s2]};
                            if (q == p && sh[0] == 0 && sh[1] == 0 && sh[2] == 0) continue;
                            int blo[3], bhi[3];
                            bool empty = false;
                            for (int a = 0; a < 3; a++) {
                                const bool used = a < d_;
                                const int qlo = used ? lo_[3 * q + a] + sh[a] : 0;
                                const int qhi = used ? lo_[3 * q + a] + n_[3 * q + a] - 1 + sh[a] : 0;
                                blo[a] = std::max(glo[a], qlo);
                                bhi[a] = std::min(ghi[a], qhi);
                                if (blo[a] > bhi[a]) empty = true;
                            }
                            if (empty) continue;
                            HaloItem it{};
                            it.dst = p;
                            it.src = q;
                            int64_t cells = 1;
                            for (int a = 0; a < 3; a++) {
                                const bool used = a < d_;
                                it.lo[a] = used ? blo[a] - (lo_[3 * p + a] - g_) : 0;
                                it.ext[a] = used ? bhi[a] - blo[a] + 1 : 1;
                                // source ghost coordinate of global cell x: x - sh - (lo_q - g)
                                it.off[a] = used ? (lo_[3 * p + a] - g_) - sh[a] - (lo_[3 * q + a] - g_) : 0;
                                cells *= it.ext[a];
                            }
                            img.items.push_back(it);
                            img.add(int(img.items.size()) - 1, 0, cells);
                        }
        }
    }

    // Ghost regions outside the domain along the axes with a physical
    // boundary condition: for each patch, the ghost box split per axis into
    // below / inside / above the domain, one item per region outside along at
    // least one such axis.
    void build_bc(TableImage<BcItem>& img) {
        for (int p = 0; p < np_; p++) {
            int rlo[3][3], rhi[3][3], nr[3];  // ranges per axis (global, inclusive); range kind 0 inside, 1 below, 2 above
            int kind[3][3];
            for (int a = 0; a < 3; a++) {
                nr[a] = 0;
                if (a >= d_) {
                    rlo[a][0] = rhi[a][0] = 0;
                    kind[a][0] = 0;
                    nr[a] = 1;
                    continue;
                }
                const int glo = lo_[3 * p + a] - g_, ghi = lo_[3 * p + a] + n_[3 * p + a] - 1 + g_;
                const int dlo = dlo_[a], dhi = dlo_[a] + dn_[a] - 1;
                const bool bcl = !periodic_[a] && bc_cond_[2 * a] != HM_BC_NONE;
                const bool bch = !periodic_[a] && bc_cond_[2 * a + 1] != HM_BC_NONE;
                const int ilo = bcl ? std::max(glo, dlo) : glo, ihi = bch ? std::min(ghi, dhi) : ghi;
                if (bcl && glo < dlo) {
                    rlo[a][nr[a]] = glo, rhi[a][nr[a]] = dlo - 1, kind[a][nr[a]] = 1;
                    nr[a]++;
                }
                if (ilo <= ihi) {
                    rlo[a][nr[a]] = ilo, rhi[a][nr[a]] = ihi, kind[a][nr[a]] = 0;
                    nr[a]++;
                }
                if (bch && ghi > dhi) {
                    rlo[a][nr[a]] = dhi + 1, rhi[a][nr[a]] = ghi, kind[a][nr[a]] = 2;
                    nr[a]++;
                }
            }
            for (int r0 = 0; r0 < nr[0]; r0++)
                for (int r1 = 0; r1 < nr[1]; r1++)
                    for (int r2 = 0; r2 < nr[2]; r2++) {
                        const int rr[3] = {r0, r1, r2};
                        if (kind[0][r0] == 0 && kind[1][r1] == 0 && kind[2][r2] == 0) continue;
                        BcItem it{};
                        it.dst = p;
                        it.side = -1;
                        int64_t cells = 1;
                        for (int a = 0; a < 3; a++) {
                            const int k = kind[a][rr[a]];
                            const int base = (a < d_) ? lo_[3 * p + a] - g_ : 0;
                            it.lo[a] = rlo[a][rr[a]] - base;
                            it.ext[a] = rhi[a][rr[a]] - rlo[a][rr[a]] + 1;
                            cells *= it.ext[a];
                            if (k == 0) continue;
                            const int sd = 2 * a + (k == 2 ? 1 : 0);
                            const int face = (k == 1) ? 2 * dlo_[a] - 1 : 2 * (dlo_[a] + dn_[a] - 1) + 1;
                            switch (bc_cond_[sd]) {
                                case HM_BC_FLOW:
                                    it.mode[a] = 2;
                                    it.par[a] = ((k == 1) ? dlo_[a] : dlo_[a] + dn_[a] - 1) - base;
                                    break;
                                case HM_BC_REFLECT:
                                    it.flip |= 1 << a;
                                    it.mode[a] = 1;
                                    it.par[a] = face - 2 * base;
                                    break;
                                case HM_BC_SYMMETRY:
                                    it.mode[a] = 1;
                                    it.par[a] = face - 2 * base;
                                    break;
                                default:  // HM_BC_DIRICHLET
                                    if (it.side < 0) it.side = sd;
                                    break;
                            }
                        }
                        // every source cell lies in the patch's ghost box (the kernel does not check)
                        for (int a = 0; a < 3; a++) {
                            const int ge = (a < d_) ? n_[3 * p + a] + 2 * g_ : 1;
                            for (int x : {it.lo[a], it.lo[a] + it.ext[a] - 1}) {
                                const int y = it.mode[a] == 0 ? x : (it.mode[a] == 1 ? it.par[a] - x : it.par[a]);
                                if (x < 0 || x >= ge || y < 0 || y >= ge) bc_bad_ = true;
                            }
                        }
                        img.items.push_back(it);
                        img.add(int(img.items.size()) - 1, 0, cells);
                    }
        }
    }

    int build_tables() {
        TableImage<HaloItem> halo;
        build_halo(halo);
        TableImage<BcItem> bc;
        bc_bad_ = false;
        build_bc(bc);
        if (bc_bad_) return rec_->fail(HM_ERR_INVALID_ARGUMENT, "boundary-condition source outside a ghost box");
        TableImage<LevelPatch> copy, zero;
        copy.items = patch_;
        zero.items = patch_;
        // advective equations: the five-equation model's volume fractions
        const hm_config& cfg = rec_->config();
        const int nadv = cfg.model == HM_MODEL_FIVE_EQN_ALLAIRE ? cfg.num_species - 1 : 0;
        src_first_ = cfg.model == HM_MODEL_FIVE_EQN_ALLAIRE ? cfg.num_species + d_ + 1 : 0;
        for (int p = 0; p < np_; p++) {
            copy.add(p, 0, gvol(p));
            if (nadv > 0) zero.add(p, 0, int64_t(nadv) * ivol(p));
        }
        halo_host_ = halo.items;
        halo_items_ = halo.total;
        bc_items_ = bc.total;
        copy_items_ = copy.total;
        zero_items_ = zero.total;
        const size_t dir_bytes = up256(sizeof(double) * 6 * size_t(nq_));
        const size_t bytes = halo.bytes() + bc.bytes() + dir_bytes + copy.bytes() + zero.bytes() + rk_bytes();
        if (tables_) (void)hipFree(tables_);
        tables_ = nullptr;
        if (hipMalloc(&tables_, bytes) != hipSuccess) return rec_->fail(HM_ERR_OUT_OF_MEMORY, "level tables");
        std::vector<char> h(bytes, 0);
        char* d = static_cast<char*>(tables_);
        size_t o = 0;
        halo_ = halo.write(h.data() + o, d + o);
        o += halo.bytes();
        bc_table_ = bc.write(h.data() + o, d + o);
        o += bc.bytes();
        std::memcpy(h.data() + o, bc_values_, sizeof(double) * 6 * size_t(nq_));
        bc_dir_ = reinterpret_cast<const double*>(d + o);
        o += dir_bytes;
        copy_ = copy.write(h.data() + o, d + o);
        o += copy.bytes();
        zero_ = zero.write(h.data() + o, d + o);
        o += zero.bytes();
        rk_off_ = o;
        if (hipMemcpy(tables_, h.data(), o, hipMemcpyHostToDevice) != hipSuccess)
            return rec_->fail(HM_ERR_DEVICE, "level table upload");
        return build_rk_tables();
    }

    size_t rk_bytes() const {
        TableImage<RKPatch> t;
        t.items.resize(np_);
        for (int p = 0; p < np_; p++) t.add(p, 0, ivol(p));
        size_t b = t.bytes() * kMaxRK;
        if (acc_mem_) b += acc_table(0).bytes() * kMaxRK;
        return b;
    }

    // Work table of the gamma accumulation (k_rk_accumulate): per patch the
    // faces of each axis, then the cells (axis code dim).
    TableImage<RKPatch> acc_table(int) const {
        TableImage<RKPatch> t;
        t.items.resize(np_);
        for (int p = 0; p < np_; p++) {
            for (int a = 0; a < d_; a++) t.add(p, a, fvol(p, a));
            t.add(p, d_, ivol(p));
        }
        return t;
    }

    // One stage-update table per stage (the intermediate pointers differ).
    int build_rk_tables() {
        const hm_config& c = rec_->config();
        const int ns = c.num_species;
        std::vector<char> h(rk_bytes(), 0);
        char* d = static_cast<char*>(tables_) + rk_off_;
        size_t o = 0;
        for (int s = 0; s < kMaxRK; s++) {
            TableImage<RKPatch> t;
            t.items.resize(np_);
            for (int p = 0; p < np_; p++) {
                RKPatch& R = t.items[p];
                std::memset(&R, 0, sizeof R);
                for (int a = 0; a < 3; a++) {
                    R.n[a] = a < d_ ? n_[3 * p + a] : 1;
                    R.gq[a] = R.gi[a] = a < d_ ? g_ : 0;
                    R.dx[a] = dx_[a];
                }
                for (int e = 0; e < nq_; e++) R.Q[e] = patch_[p].Q[0] + size_t(e) * gvol(p);
                for (int m = 0; m <= s && m < kMaxRK; m++) {
                    for (int e = 0; e < nq_; e++) R.Qi[m][e] = patch_[p].Q[1 + m] + size_t(e) * gvol(p);
                    for (int a = 0; a < d_; a++) R.Fi[m][a] = fl_[(size_t(p) * kMaxRK + m) * 3 + a];
                    if (diff_)
                        for (int a = 0; a < d_; a++) R.Fdi[m][a] = fdl_[(size_t(p) * kMaxRK + m) * 3 + a];
                    R.Si[m] = patch_[p].S[m];
                }
                if (acc_mem_) {
                    for (int a = 0; a < d_; a++) {
                        R.Facc[a] = facc_[size_t(p) * 3 + a];
                        if (diff_) R.Fdacc[a] = fdacc_[size_t(p) * 3 + a];
                    }
                    R.Sacc = sacc_[p];
                }
                t.add(p, 0, ivol(p));
            }
            rk_[s] = t.write(h.data() + o, d + o);
            rk_items_ = t.total;
            o += t.bytes();
            if (acc_mem_) {
                TableImage<RKPatch> at = acc_table(s);
                at.items = t.items;
                acc_[s] = at.write(h.data() + o, d + o);
                acc_items_ = at.total;
                o += at.bytes();
            }
            RKParams& r = rkp_[s];
            r.nc = s + 1;
            r.neq = neq_;
            const bool five = c.model == HM_MODEL_FIVE_EQN_ALLAIRE;
            r.nz = five ? ns - 1 : 0;
            r.zfirst = five ? ns + d_ + 1 : 0;
            r.last_beta = -1;
            r.ns = diff_ ? 1 : 0;  // Navier-Stokes form with the diffusive fluxes, else Euler
            for (int m = 0; m < kMaxRK; m++) {
                r.alpha[m] = alpha_[s][m];
                r.beta[m] = beta_[s][m];
                r.gamma[m] = gamma_[s][m];
                if (m <= s && beta_[s][m] != 0.0) r.last_beta = m;
            }
        }
        // the flux sums' terms of a whole step: stage by stage, intermediates in order
        step_terms_ = AccTerms{};
        step_terms_.neq = neq_;
        for (int s = 0; s < nst_; s++)
            for (int m = 0; m <= s; m++)
                if (gamma_[s][m] != 0.0) {
                    step_terms_.src[step_terms_.n] = m;
                    step_terms_.gamma[step_terms_.n] = gamma_[s][m];
                    step_terms_.n++;
                }
        if (hipMemcpy(d, h.data(), h.size(), hipMemcpyHostToDevice) != hipSuccess)
            return rec_->fail(HM_ERR_DEVICE, "level table upload");
        return HM_OK;
    }

    // Ghost fill of state buffer buf: same-level copies (periodic images
    // included), then the physical boundary conditions, which read them.
    int launch_halo(int buf, hipStream_t s) {
        if (halo_items_ > 0) {
            switch (d_) {
                case 1: hipLaunchKernelGGL(k_halo<1>, dim3(blocks_for(halo_items_)), dim3(kBlock), 0, s, halo_, patch_dev(), buf, nq_, g_); break;
                case 2: hipLaunchKernelGGL(k_halo<2>, dim3(blocks_for(halo_items_)), dim3(kBlock), 0, s, halo_, patch_dev(), buf, nq_, g_); break;
                default: hipLaunchKernelGGL(k_halo<3>, dim3(blocks_for(halo_items_)), dim3(kBlock), 0, s, halo_, patch_dev(), buf, nq_, g_); break;
            }
        }
        if (bc_items_ > 0) {
            const int mom0 = rec_->config().model == HM_MODEL_SINGLE_SPECIES ? 1 : rec_->config().num_species;
            switch (d_) {
                case 1: hipLaunchKernelGGL(k_bc<1>, dim3(blocks_for(bc_items_)), dim3(kBlock), 0, s, bc_table_, patch_dev(), buf, nq_, g_, mom0, bc_dir_); break;
                case 2: hipLaunchKernelGGL(k_bc<2>, dim3(blocks_for(bc_items_)), dim3(kBlock), 0, s, bc_table_, patch_dev(), buf, nq_, g_, mom0, bc_dir_); break;
                default: hipLaunchKernelGGL(k_bc<3>, dim3(blocks_for(bc_items_)), dim3(kBlock), 0, s, bc_table_, patch_dev(), buf, nq_, g_, mom0, bc_dir_); break;
            }
        }
        return hipGetLastError() == hipSuccess ? HM_OK : rec_->fail(HM_ERR_DEVICE, "ghost fill launch failed");
    }


    const LevelPatch* patch_dev() const { return copy_.patches; }

    // One Runge-Kutta step (the sequence the graph captures).  With
    // record_slot >= 0 the ghost-filled input of stage sn is also stored in
    // slot record_slot + sn (recompute pass of the adjoint), and with
    // face_slot >= 0 and a face-state store its RHS keeps the face states in
    // face-state slot face_slot + sn.
    int step(double dt, hipStream_t s, int record_slot = -1, int face_slot = -1, const double* dtp = nullptr) {
        if (int st = launch_halo(0, s)) return st;
        for (int sn = 0; sn < nst_; sn++) {
            hipLaunchKernelGGL(k_copy_state, dim3(blocks_for(copy_items_)), dim3(kBlock), 0, s, copy_, 0, 1 + sn, nq_);
            if (sn > 0)
                if (int st = launch_halo(1 + sn, s)) return st;
            if (record_slot >= 0)
                if (int st = store_op(1 + sn, record_slot + sn, 0, s)) return st;
            if (zero_items_ > 0)
                hipLaunchKernelGGL(k_zero_source, dim3(blocks_for(zero_items_)), dim3(kBlock), 0, s, zero_, sn,
                                   src_first_);
            if (hipGetLastError() != hipSuccess) return rec_->fail(HM_ERR_DEVICE, "level launch failed");
            if (int st = rec_->runOwnedPlan(rhs_plan_[sn], dt, s, face_slot >= 0 ? face_states(face_slot + sn) : nullptr,
                                            dtp))
                return st;
            if (diff_)
                if (int st = diff_->runOwnedPlan(diff_plan_[sn], dt, s, dtp)) return st;
            if (launch_rk_update<SegTableT<RKPatch>>(d_, rk_[sn], rk_items_, nullptr, 0, rkp_[sn], s) != hipSuccess)
                return rec_->fail(HM_ERR_DEVICE, "stage update launch failed");
        }
        // Flux sums: the whole step's gamma terms in one pass over the stages'
        // fluxes (the same additions, in the same order, as one pass per stage).
        if (acc_mem_ &&
            launch_rk_accumulate<SegTableT<RKPatch>>(d_, acc_[nst_ - 1], acc_items_, step_terms_, s) != hipSuccess)
            return rec_->fail(HM_ERR_DEVICE, "flux-sum launch failed");
        return HM_OK;
    }

    int capture(double dt) {
        return capture_with(dt, [&](hipStream_t x) { return step(dt, x); });
    }

    // Stable dt of the level state, then one step with that dt (both on s).
    int step_cfl(double cfl, hipStream_t s) {
        if (launch_stable_dt(rec_->config(), dt_work_, dt_items_, rec_->kernelParams(), cfl, dt_buf_, s) != hipSuccess)
            return rec_->fail(HM_ERR_DEVICE, "stable dt launch failed");
        return step(0.0, s, -1, -1, dt_buf_.res + kDtStep);
    }

    // Work table of the stable-dt kernel over the interiors of the scratch
    // state, and its accumulators (built on first use).
    int ensure_dt() {
        if (dt_mem_) return HM_OK;
        TableImage<DevPatch> tab;
        tab.items.resize(np_);
        for (int p = 0; p < np_; p++) {
            DevPatch& d = tab.items[p];
            std::memset(&d, 0, sizeof d);
            int64_t cells = 1;
            for (int a = 0; a < 3; a++) {
                d.n[a] = a < d_ ? n_[3 * p + a] : 1;
                d.gq[a] = a < d_ ? g_ : 0;
                d.dx[a] = dx_[a];
                d.wrec[a] = -1;
                cells *= d.n[a];
            }
            for (int e = 0; e < nq_; e++) d.Q[e] = patch_[p].Q[0] + size_t(e) * gvol(p);
            tab.add(p, 0, cells);
        }
        const size_t tb = tab.bytes();
        const size_t bytes = tb + up256(sizeof(unsigned long long) * kDtAcc) + up256(sizeof(double) * kDtSlots);
        if (hipMalloc(&dt_mem_, bytes) != hipSuccess) {
            dt_mem_ = nullptr;
            return rec_->fail(HM_ERR_OUT_OF_MEMORY, "level dt tables");
        }
        std::vector<char> img(bytes, 0);
        dt_work_ = tab.write(img.data(), static_cast<char*>(dt_mem_));
        dt_items_ = tab.total;
        dt_buf_.acc = reinterpret_cast<unsigned long long*>(static_cast<char*>(dt_mem_) + tb);
        dt_buf_.res = reinterpret_cast<double*>(static_cast<char*>(dt_mem_) + tb +
                                                up256(sizeof(unsigned long long) * kDtAcc));
        if (hipMemcpy(dt_mem_, img.data(), bytes, hipMemcpyHostToDevice) != hipSuccess)
            return rec_->fail(HM_ERR_DEVICE, "level dt table upload failed");
        return HM_OK;
    }

    template <class F>
    int capture_with(double key, F&& body) {
        hipGraph_t graph = nullptr;
        if (hipStreamBeginCapture(cap_stream_, hipStreamCaptureModeThreadLocal) != hipSuccess)
            return rec_->fail(HM_ERR_DEVICE, "hipStreamBeginCapture failed");
        const int st = body(cap_stream_);
        const hipError_t e = hipStreamEndCapture(cap_stream_, &graph);
        if (st != HM_OK) {
            if (graph) (void)hipGraphDestroy(graph);
            return st;
        }
        if (e != hipSuccess || !graph) return rec_->fail(HM_ERR_DEVICE, "hipStreamEndCapture failed");
        hipGraphExec_t exec = nullptr;
        const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ei != hipSuccess) return rec_->fail(HM_ERR_DEVICE, "hipGraphInstantiate failed");
        if (graphs_.size() >= 4) drop_graphs();
        graphs_[key] = exec;
        return HM_OK;
    }

    // Level state buffer `buf` <-> storage slot; dir 0 stores, 1 restores.
    int store_op(int buf, int slot, int dir, hipStream_t s) {
        hipLaunchKernelGGL(k_store, dim3(blocks_for(copy_items_)), dim3(kBlock), 0, s, copy_, buf,
                           store_ + size_t(slot) * slot_doubles_, slot_off_dev_, nq_, dir);
        return hipGetLastError() == hipSuccess ? HM_OK : rec_->fail(HM_ERR_DEVICE, "store launch failed");
    }

    int lam_op(int op, int dst, int src, hipStream_t s) {
        switch (d_) {
            case 1: hipLaunchKernelGGL(k_lam_op<1>, dim3(blocks_for(adj_cell_items_)), dim3(kBlock), 0, s, adj_cells_, op, dst, src, neq_, g_); break;
            case 2: hipLaunchKernelGGL(k_lam_op<2>, dim3(blocks_for(adj_cell_items_)), dim3(kBlock), 0, s, adj_cells_, op, dst, src, neq_, g_); break;
            default: hipLaunchKernelGGL(k_lam_op<3>, dim3(blocks_for(adj_cell_items_)), dim3(kBlock), 0, s, adj_cells_, op, dst, src, neq_, g_); break;
        }
        return hipGetLastError() == hipSuccess ? HM_OK : rec_->fail(HM_ERR_DEVICE, "adjoint launch failed");
    }

    // Reverse of one step whose stage inputs are stored from slot `slot0`.
    // lam[nst] holds the adjoint of the step's output on entry and of its
    // input on return.  Per stage: the VJP of each stage input with nonzero
    // beta, then one k_stage_adjoint launch for the alpha terms and the last
    // VJP's output (earlier VJPs of the same stage, which only tableaux with
    // several nonzero beta per stage have, add theirs with k_halo_reverse).
    int adjoint_step(int slot0, int face_slot0, hipStream_t s) {
        if (zero_first_) {
            if (int st = lam_op(kLamZeroBelow, nst_, 0, s)) return st;
        }
        for (int sn = nst_ - 1; sn >= 0; sn--) {
            int last_vjp = -1;
            for (int m = 0; m <= sn; m++)
                if (beta_[sn][m] != 0.0) last_vjp = m;
            for (int m = 0; m <= sn; m++) {
                if (beta_[sn][m] == 0.0) continue;
                // The VJP reads stage input m from its trajectory slot and forms
                // the face-flux adjoints beta (lam_f - lam_{f-1}) / dx of stage
                // adjoint sn + 1 itself (the transposed stage update).
                if (int st = adj_->run(vjp_plan_, rec_dt_, s, face_states(face_slot0 + m), sn + 1, beta_[sn][m],
                                       int64_t(slot0 + m) * int64_t(slot_doubles_)))
                    return st;
                if (m == last_vjp) break;
                switch (d_) {
                    case 1: hipLaunchKernelGGL(k_halo_reverse<1>, dim3(blocks_for(adj_cell_items_)), dim3(kBlock), 0, s, adj_cells_, rev_dev_, m, g_); break;
                    case 2: hipLaunchKernelGGL(k_halo_reverse<2>, dim3(blocks_for(adj_cell_items_)), dim3(kBlock), 0, s, adj_cells_, rev_dev_, m, g_); break;
                    default: hipLaunchKernelGGL(k_halo_reverse<3>, dim3(blocks_for(adj_cell_items_)), dim3(kBlock), 0, s, adj_cells_, rev_dev_, m, g_); break;
                }
                if (hipGetLastError() != hipSuccess) return rec_->fail(HM_ERR_DEVICE, "adjoint launch failed");
            }
            const StageAdj& sa = stage_adj_[sn];
            if (sa.nt == 0) continue;
            switch (d_) {
                case 1: hipLaunchKernelGGL(k_stage_adjoint<1>, dim3(blocks_for(adj_cell_items_)), dim3(kBlock), 0, s, adj_cells_, rev_dev_, sa, g_); break;
                case 2: hipLaunchKernelGGL(k_stage_adjoint<2>, dim3(blocks_for(adj_cell_items_)), dim3(kBlock), 0, s, adj_cells_, rev_dev_, sa, g_); break;
                default: hipLaunchKernelGGL(k_stage_adjoint<3>, dim3(blocks_for(adj_cell_items_)), dim3(kBlock), 0, s, adj_cells_, rev_dev_, sa, g_); break;
            }
            if (hipGetLastError() != hipSuccess) return rec_->fail(HM_ERR_DEVICE, "adjoint launch failed");
        }
        if (!out_in_place_) return lam_op(kLamCopy, nst_, 0, s);
        return HM_OK;
    }

    // Entries of the per-stage adjoint updates (k_stage_adjoint) for the
    // tableau alpha_, beta_: which stage adjoints each stage writes, with
    // which alpha, which is first written there (no zero fill needed), and
    // whether the step input's adjoint can land directly in lam[nst].
    void plan_stage_adjoint() {
        bool touched[kMaxRK] = {};
        zero_first_ = false;
        for (int sn = nst_ - 1; sn >= 0; sn--) {
            StageAdj& sa = stage_adj_[sn];
            sa = StageAdj{};
            sa.src = sn + 1;
            sa.q = -1;
            int last_vjp = -1;
            for (int m = 0; m <= sn; m++)
                if (beta_[sn][m] != 0.0) last_vjp = m;
            // VJPs of this stage before the last add into lam[m] with
            // k_halo_reverse, which needs lam[m] to hold a value already.
            for (int m = 0; m < last_vjp; m++)
                if (beta_[sn][m] != 0.0) {
                    if (!touched[m]) zero_first_ = true;
                    touched[m] = true;
                }
            for (int m = 0; m <= sn; m++) {
                if (alpha_[sn][m] == 0.0 && m != last_vjp) continue;
                const int t = sa.nt++;
                sa.m[t] = m;
                sa.dst[t] = m;
                sa.a[t] = alpha_[sn][m];
                sa.init[t] = touched[m] ? 0 : 1;
                if (m == last_vjp) sa.q = t;
                touched[m] = true;
            }
        }
        // A stage adjoint read before any write (an unused stage) needs the zero fill.
        for (int m = 0; m < nst_; m++)
            if (!touched[m]) zero_first_ = true;
        if (zero_first_)
            for (int sn = 0; sn < nst_; sn++)
                for (int t = 0; t < stage_adj_[sn].nt; t++) stage_adj_[sn].init[t] = 0;
        // The step input's last update is in stage 0 (its only entry there is
        // m = 0): write it straight to lam[nst].
        out_in_place_ = false;
        if (stage_adj_[0].nt == 1 && stage_adj_[0].m[0] == 0) {
            stage_adj_[0].dst[0] = nst_;
            out_in_place_ = true;
        }
    }

    int adjoint_io(int p, double* host, bool write) {
        if (p < 0 || p >= np_ || !host) return rec_->fail(HM_ERR_INVALID_ARGUMENT, "bad patch index or array");
        if (int st = ensure_adjoint(0, 0)) return st;
        std::vector<double> h(size_t(neq_) * gvol(p), 0.0);
        if (hipDeviceSynchronize() != hipSuccess) return rec_->fail(HM_ERR_DEVICE, "device synchronize failed");
        double* dev = lam_base_[size_t(p) * (kMaxRK + 1) + nst_];
        const int64_t vol = ivol(p);
        if (write) {
            for (int c = 0; c < neq_; c++)
                for (int64_t l = 0; l < vol; l++) h[size_t(c) * gvol(p) + gidx_of_interior(p, l)] = host[c * vol + l];
            if (hipMemcpy(dev, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
                return rec_->fail(HM_ERR_DEVICE, "adjoint upload failed");
        } else {
            if (hipMemcpy(h.data(), dev, h.size() * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
                return rec_->fail(HM_ERR_DEVICE, "adjoint download failed");
            for (int c = 0; c < neq_; c++)
                for (int64_t l = 0; l < vol; l++) host[c * vol + l] = h[size_t(c) * gvol(p) + gidx_of_interior(p, l)];
        }
        return HM_OK;
    }

    // Buffers and tables of the adjoint (built once), and storage slots for
    // nsteps steps with a checkpoint every `every` steps (grown on demand):
    // slot 0 final state, slots 1.. checkpoints, then every*nst stage inputs.
    int ensure_adjoint(int nsteps, int every) {
        if (rec_->config().model != HM_MODEL_SINGLE_SPECIES)
            return rec_->fail(HM_ERR_UNSUPPORTED, "the adjoint is available for the single-species model");
        if (!adj_mem_) {
            // lam (kMaxRK+1 per patch) and qbar over ghost boxes
            size_t tot = 0;
            std::vector<size_t> off_lam(np_), off_q(np_);
            for (int p = 0; p < np_; p++) {
                off_lam[p] = tot;
                tot += up256(sizeof(double) * size_t(neq_) * gvol(p)) / sizeof(double) * (kMaxRK + 1);
                off_q[p] = tot;
                tot += up256(sizeof(double) * size_t(neq_) * gvol(p)) / sizeof(double);
            }
            // per-patch offsets within a storage slot
            std::vector<int64_t> soff(np_);
            size_t sd = 0;
            for (int p = 0; p < np_; p++) {
                soff[p] = int64_t(sd);
                sd += up256(sizeof(double) * size_t(nq_) * gvol(p)) / sizeof(double);
            }
            slot_doubles_ = sd;
            // tables: AdjLevelPatch cells, reverse halo list, slot offsets
            std::vector<std::vector<int>> by_src(np_);
            for (size_t q = 0; q < halo_host_.size(); q++) by_src[halo_host_[q].src].push_back(int(q));
            std::vector<HaloItem> rev;
            TableImage<AdjLevelPatch> cells;
            std::vector<AdjLevelPatch> items(np_);
            lam_base_.assign(size_t(np_) * (kMaxRK + 1), nullptr);
            for (int p = 0; p < np_; p++) {
                AdjLevelPatch& A = items[p];
                std::memset(&A, 0, sizeof A);
                for (int a = 0; a < 3; a++) {
                    A.n[a] = a < d_ ? n_[3 * p + a] : 1;
                    A.dx[a] = dx_[a];
                }
                A.gvol = gvol(p);
                A.rev_first = int(rev.size());
                for (int q : by_src[p]) rev.push_back(halo_host_[q]);
                A.rev_count = int(rev.size()) - A.rev_first;
                cells.add(p, 0, ivol(p));
            }
            cells.items.resize(np_);
            const size_t tb = cells.bytes() + up256(sizeof(HaloItem) * (rev.size() + 1)) +
                              up256(sizeof(int64_t) * np_);
            if (hipMalloc(&adj_mem_, tot * sizeof(double) + tb) != hipSuccess) {
                adj_mem_ = nullptr;
                return rec_->fail(HM_ERR_OUT_OF_MEMORY, "adjoint buffers");
            }
            double* base = static_cast<double*>(adj_mem_);
            if (hipMemset(adj_mem_, 0, tot * sizeof(double)) != hipSuccess)
                return rec_->fail(HM_ERR_DEVICE, "adjoint buffer clear");
            for (int p = 0; p < np_; p++) {
                AdjLevelPatch& A = items[p];
                const size_t lstride = up256(sizeof(double) * size_t(neq_) * gvol(p)) / sizeof(double);
                for (int m = 0; m <= kMaxRK; m++) {
                    A.lam[m] = base + off_lam[p] + size_t(m) * lstride;
                    lam_base_[size_t(p) * (kMaxRK + 1) + m] = A.lam[m];
                }
                A.qbar = base + off_q[p];
                cells.items[p] = A;
            }
            char* tdev = reinterpret_cast<char*>(base + tot);
            std::vector<char> h(tb, 0);
            size_t o = 0;
            adj_cells_ = cells.write(h.data() + o, tdev + o);
            adj_cell_items_ = cells.total;
            o += cells.bytes();
            if (!rev.empty()) std::memcpy(h.data() + o, rev.data(), sizeof(HaloItem) * rev.size());
            rev_dev_ = reinterpret_cast<const HaloItem*>(tdev + o);
            o += up256(sizeof(HaloItem) * (rev.size() + 1));
            std::memcpy(h.data() + o, soff.data(), sizeof(int64_t) * np_);
            slot_off_dev_ = reinterpret_cast<const int64_t*>(tdev + o);
            if (hipMemcpy(tdev, h.data(), tb, hipMemcpyHostToDevice) != hipSuccess)
                return rec_->fail(HM_ERR_DEVICE, "adjoint table upload");
            // The VJP plan (built once the trajectory store exists) needs these.
            adj_items_ = items;
            slot_off_host_ = soff;
            adj_.reset(new AdjointSingle(rec_));
            adj_nst_ = nst_;
        }
        if (adj_nst_ != nst_) return rec_->fail(HM_ERR_UNSUPPORTED, "the stage count changed after the adjoint was set up");
        if (nsteps > 0) {
            const int nckpt = (nsteps + every - 1) / every;
            const int nslots = 1 + nckpt + every * nst_;
            if (nslots > store_slots_) {
                if (store_) {
                    (void)hipDeviceSynchronize();
                    (void)hipFree(store_);
                    store_ = nullptr;
                    store_slots_ = 0;
                }
                if (hipMalloc(reinterpret_cast<void**>(&store_), size_t(nslots) * slot_doubles_ * sizeof(double)) !=
                    hipSuccess) {
                    store_ = nullptr;
                    return rec_->fail(HM_ERR_OUT_OF_MEMORY, "checkpoint storage");
                }
                store_slots_ = nslots;
            }
            nckpt_ = nckpt;
            // The VJP plan reads each stage input from its trajectory slot
            // (KParams::qoff selects the slot): its Q views are slot 0.
            if (vjp_store_ != store_) {
                std::vector<AdjLamSource> lam_src(np_);
                std::vector<hm_vjp_patch_view> vv(np_);
                for (int p = 0; p < np_; p++) {
                    const AdjLevelPatch& A = adj_items_[p];
                    lam_src[p] = AdjLamSource{A.lam[0], A.lam[1] - A.lam[0], g_};
                    hm_vjp_patch_view v{};
                    for (int a = 0; a < 3; a++) {
                        v.n[a] = a < d_ ? n_[3 * p + a] : 1;
                        v.ghost[a] = a < d_ ? g_ : 0;
                        v.dx[a] = dx_[a];
                    }
                    for (int e = 0; e < neq_; e++) {
                        v.Q[e] = store_ + slot_off_host_[p] + size_t(e) * gvol(p);
                        v.Q_bar[e] = A.qbar + size_t(e) * gvol(p);
                    }
                    // Not read: level runs form the face-flux adjoints from the
                    // stage adjoints (lam_src); the view only needs an address.
                    for (int a = 0; a < d_; a++) v.F_bar[a] = A.qbar;
                    vv[p] = v;
                }
                vjp_store_ = nullptr;
                if (int st = adj_->build(vv.data(), np_, vjp_plan_, wst_off_.data(), lam_src.data())) return st;
                vjp_store_ = store_;
            }
            // Face states of one segment's stages, when they fit the budget;
            // else the adjoint's flux half reconstructs them again.
            const size_t nfs = size_t(every) * nst_;
            const size_t want = nfs * wst_slot_doubles_ * sizeof(double);
            if (nfs > wst_slots_ || want == 0) {
                if (wst_) {
                    (void)hipDeviceSynchronize();
                    (void)hipFree(wst_);
                    wst_ = nullptr;
                    wst_slots_ = 0;
                }
                size_t budget = 0;
                if (face_state_budget_ < 0) {
                    size_t fr = 0, tot = 0;
                    if (hipMemGetInfo(&fr, &tot) == hipSuccess) budget = fr / 2;
                } else {
                    budget = size_t(face_state_budget_);
                }
                if (want > 0 && want <= budget) {
                    if (hipMalloc(reinterpret_cast<void**>(&wst_), want) == hipSuccess)
                        wst_slots_ = nfs;
                    else
                        wst_ = nullptr;  // not fatal: reconstruct instead
                }
            }
        }
        return HM_OK;
    }

    // Face-state slot i, or NULL without a store.
    double* face_states(int i) const {
        return (wst_ && i >= 0 && size_t(i) < wst_slots_) ? wst_ + size_t(i) * wst_slot_doubles_ : nullptr;
    }

   public:
    // Device memory the adjoint may use to keep face states between its
    // recompute and reverse passes: < 0 half of the free memory (default),
    // 0 none.  Takes effect at the next advanceRecorded.
    int setFaceStateBudget(int64_t bytes) {
        face_state_budget_ = bytes;
        if (wst_) {
            (void)hipDeviceSynchronize();
            (void)hipFree(wst_);
            wst_ = nullptr;
            wst_slots_ = 0;
        }
        return HM_OK;
    }
    bool faceStatesKept() const { return wst_ != nullptr; }

   private:

    void drop_graphs() {
        if (graphs_.empty()) return;
        (void)hipDeviceSynchronize();
        for (auto& kv : graphs_) (void)hipGraphExecDestroy(kv.second);
        graphs_.clear();
    }

    ConvectiveFluxReconstructorGPU* rec_;
    int d_ = 0, neq_ = 0, nq_ = 0, g_ = kGc, np_ = 0, nst_ = 3;
    int dlo_[3] = {}, dn_[3] = {};
    bool periodic_[3] = {};
    double dx_[3] = {};
    double alpha_[kMaxRK][kMaxRK] = {}, beta_[kMaxRK][kMaxRK] = {}, gamma_[kMaxRK][kMaxRK] = {};
    std::vector<int> lo_, n_;
    void* storage_ = nullptr;
    void* tables_ = nullptr;
    std::vector<LevelPatch> patch_;
    std::vector<double*> fl_;
    std::vector<std::vector<hm_patch_view>> views_;
    void* rhs_scratch_ = nullptr;
    DiffusiveFluxReconstructorGPU* diff_ = nullptr;
    std::vector<double*> fdl_;  // diffusive flux per patch, stage and axis (with diff_)
    void* diff_scratch_ = nullptr;
    DiffusiveFluxReconstructorGPU::OwnedPlan diff_plan_[kMaxRK];
    void* dt_mem_ = nullptr;  // stable-dt work table and accumulators
    SegTableT<DevPatch> dt_work_{};
    int64_t dt_items_ = 0;
    DtBuffers dt_buf_;
    ConvectiveFluxReconstructorGPU::OwnedPlan rhs_plan_[kMaxRK];
    SegTableT<HaloItem> halo_{};
    SegTableT<BcItem> bc_table_{};
    const double* bc_dir_ = nullptr;
    int64_t bc_items_ = 0;
    int bc_cond_[6] = {HM_BC_NONE, HM_BC_NONE, HM_BC_NONE, HM_BC_NONE, HM_BC_NONE, HM_BC_NONE};
    double bc_values_[6 * (HM_MAX_EQN + 1)] = {};
    bool has_bc_ = false;
    bool bc_bad_ = false;
    Schedule sched_{};  // schedule of the batched RHS plans
    SegTableT<LevelPatch> copy_{}, zero_{};
    SegTableT<RKPatch> rk_[kMaxRK] = {};
    SegTableT<RKPatch> acc_[kMaxRK] = {};
    int64_t acc_items_ = 0;
    void* acc_mem_ = nullptr;           // flux sums (enableFluxSums)
    size_t acc_bytes_ = 0;
    std::vector<double*> facc_, sacc_;  // per patch and axis / per patch
    std::vector<double*> fdacc_;        // diffusive flux sums per patch and axis (Navier-Stokes form)
    AccTerms step_terms_{};             // gamma terms of one step's flux sums
    RKParams rkp_[kMaxRK] = {};
    int64_t halo_items_ = 0, copy_items_ = 0, zero_items_ = 0, rk_items_ = 0;
    int src_first_ = 0;  // first source component k_zero_source clears
    size_t rk_off_ = 0;
    hipStream_t cap_stream_ = nullptr;
    hipEvent_t ev_a_ = nullptr, ev_b_ = nullptr;
    std::map<double, hipGraphExec_t> graphs_;
    // adjoint
    std::vector<HaloItem> halo_host_;
    void* adj_mem_ = nullptr;
    std::vector<double*> lam_base_;
    SegTableT<AdjLevelPatch> adj_cells_{};
    int64_t adj_cell_items_ = 0;
    const HaloItem* rev_dev_ = nullptr;
    StageAdj stage_adj_[kMaxRK] = {};
    bool zero_first_ = false, out_in_place_ = false;
    const int64_t* slot_off_dev_ = nullptr;
    std::unique_ptr<AdjointSingle> adj_;
    AdjointSingle::Owned vjp_plan_{};
    const double* vjp_store_ = nullptr;  // the store the VJP plan's Q views point into
    std::vector<AdjLevelPatch> adj_items_;
    std::vector<int64_t> slot_off_host_;
    int adj_nst_ = 0;
    double* store_ = nullptr;
    size_t slot_doubles_ = 0;
    double* wst_ = nullptr;              // face-state store: wst_slots_ slots of wst_slot_doubles_
    size_t wst_slots_ = 0, wst_slot_doubles_ = 0;
    std::vector<int64_t> wst_off_;       // per patch and axis: chunk offset inside a slot
    int64_t face_state_budget_ = -1;
    int store_slots_ = 0, nckpt_ = 0;
    bool recorded_ = false;
    double rec_dt_ = 0.0;
    int rec_nsteps_ = 0, every_ = 1;
};

// Stage update for caller-owned patch data (Euler::advanceSingleStepOnPatch).
int advance_single_step(ConvectiveFluxReconstructorGPU* rec, const hm_rk_patch_view* views, int np, int nc,
                        const double* alpha, const double* beta, const double* gamma, hipStream_t s) {
    if (np < 0 || (np > 0 && !views) || nc < 1 || nc > kMaxRK || !alpha || !beta)
        return rec->fail(HM_ERR_INVALID_ARGUMENT, "bad stage-update arguments");
    if (np == 0) return HM_OK;
    if (int st = rec->selectDevice()) return st;
    const hm_config& c = rec->config();
    const int d = c.dim, neq = rec->numEqn(), nq = rec->numConservativeArrays();
    RKParams r{};
    r.nc = nc;
    r.neq = neq;
    const bool five = c.model == HM_MODEL_FIVE_EQN_ALLAIRE;
    r.nz = five ? c.num_species - 1 : 0;
    r.zfirst = five ? c.num_species + d + 1 : 0;
    r.last_beta = -1;
    for (int m = 0; m < kMaxRK; m++) {
        r.alpha[m] = m < nc ? alpha[m] : 0.0;
        r.beta[m] = m < nc ? beta[m] : 0.0;
        r.gamma[m] = (m < nc && gamma) ? gamma[m] : 0.0;
        if (m < nc && beta[m] != 0.0) r.last_beta = m;
    }
    TableImage<RKPatch> cells, acc;
    cells.items.resize(np);
    acc.items.resize(np);
    bool any_acc = false;
    char msg[160];
    for (int p = 0; p < np; p++) {
        const hm_rk_patch_view& v = views[p];
        RKPatch& R = cells.items[p];
        std::memset(&R, 0, sizeof R);
        int64_t ivol = 1;
        for (int a = 0; a < 3; a++) {
            const bool used = a < d;
            R.n[a] = used ? v.n[a] : 1;
            R.gq[a] = used ? v.ghost[a] : 0;
            R.gi[a] = used ? v.ghost_intermediate[a] : 0;
            R.dx[a] = used ? v.dx[a] : 1.0;
            if (used && (v.n[a] < 1 || !(v.dx[a] > 0.0) || v.ghost[a] < 0 || v.ghost_intermediate[a] < 0)) {
                std::snprintf(msg, sizeof msg, "patch %d: bad extent, ghost width or dx on axis %d", p, a);
                return rec->fail(HM_ERR_INVALID_ARGUMENT, msg);
            }
            ivol *= R.n[a];
        }
        // Kernels number the items of a patch's work segment in 32 bits.
        double items = double(ivol);
        for (int a = 0; a < d; a++) items = std::max(items, double(ivol) / R.n[a] * (R.n[a] + 1));
        if (items >= 2147483648.0) {
            std::snprintf(msg, sizeof msg, "patch %d: 2^31 cells or faces or more; split the patch", p);
            return rec->fail(HM_ERR_INVALID_ARGUMENT, msg);
        }
        for (int e = 0; e < nq; e++) {
            if (!v.Q[e]) {
                std::snprintf(msg, sizeof msg, "patch %d: conservative array %d is NULL", p, e);
                return rec->fail(HM_ERR_INVALID_ARGUMENT, msg);
            }
            R.Q[e] = v.Q[e];
        }
        for (int m = 0; m < nc; m++) {
            const bool needs_flux = beta[m] != 0.0 || (gamma && gamma[m] != 0.0);
            for (int e = 0; e < neq; e++) {
                if (alpha[m] != 0.0 && !v.Q_intermediate[m][e]) {
                    std::snprintf(msg, sizeof msg, "patch %d: intermediate %d array %d is NULL", p, m, e);
                    return rec->fail(HM_ERR_INVALID_ARGUMENT, msg);
                }
                R.Qi[m][e] = v.Q_intermediate[m][e];
            }
            for (int a = 0; a < d; a++) {
                if (needs_flux && !v.F_intermediate[m][a]) {
                    std::snprintf(msg, sizeof msg, "patch %d: intermediate %d flux on axis %d is NULL", p, m, a);
                    return rec->fail(HM_ERR_INVALID_ARGUMENT, msg);
                }
                R.Fi[m][a] = v.F_intermediate[m][a];
            }
            if (needs_flux && !v.S_intermediate[m]) {
                std::snprintf(msg, sizeof msg, "patch %d: intermediate %d source is NULL", p, m);
                return rec->fail(HM_ERR_INVALID_ARGUMENT, msg);
            }
            R.Si[m] = v.S_intermediate[m];
            // Navier-Stokes form: diffusive fluxes on every axis or on none.
            int nd = 0;
            for (int a = 0; a < d; a++) nd += v.F_diffusive_intermediate[m][a] ? 1 : 0;
            if (nd != 0 && nd != d) {
                std::snprintf(msg, sizeof msg, "patch %d: intermediate %d diffusive flux given on some axes only", p, m);
                return rec->fail(HM_ERR_INVALID_ARGUMENT, msg);
            }
            for (int a = 0; a < 3; a++) R.Fdi[m][a] = (a < d && nd) ? v.F_diffusive_intermediate[m][a] : nullptr;
            if (nd) r.ns = 1;
        }
        for (int a = 0; a < 3; a++) R.Facc[a] = a < d ? v.F_accumulated[a] : nullptr;
        for (int a = 0; a < 3; a++) R.Fdacc[a] = a < d ? v.F_diffusive_accumulated[a] : nullptr;
        R.Sacc = v.S_accumulated;
        cells.add(p, 0, ivol);
        acc.items[p] = R;
        if (gamma) {
            for (int a = 0; a < d; a++)
                if (R.Facc[a] || R.Fdacc[a]) {
                    acc.add(p, a, int64_t(R.n[0] + (a == 0)) * (R.n[1] + (a == 1)) * (R.n[2] + (a == 2)));
                    any_acc = true;
                }
            if (R.Sacc) {
                acc.add(p, d, ivol);
                any_acc = true;
            }
        }
    }
    // One patch: descriptors by value, no table upload.  A single-patch
    // segment's axis is its position, so this needs every accumulated axis.
    if (np == 1 && (!any_acc || int(acc.cnt.size()) == d + 1)) {
        SegSingleT<RKPatch> wc{}, wa{};
        wc.p = cells.items[0];
        wc.nseg = int(cells.cnt.size());
        for (int i = 0; i < wc.nseg; i++) {
            wc.start[i] = cells.start[i];
            wc.cnt[i] = cells.cnt[i];
        }
        wc.start[wc.nseg] = cells.total;
        wa.p = acc.items[0];
        wa.nseg = int(acc.cnt.size());
        for (int i = 0; i < wa.nseg; i++) {
            wa.start[i] = acc.start[i];
            wa.cnt[i] = acc.cnt[i];
        }
        wa.start[wa.nseg] = acc.total;
        if (launch_rk_update<SegSingleT<RKPatch>>(d, wc, cells.total, any_acc ? &wa : nullptr, acc.total, r, s) !=
            hipSuccess)
            return rec->fail(HM_ERR_DEVICE, "stage update launch failed");
        return HM_OK;
    }
    const size_t bytes = cells.bytes() + (any_acc ? acc.bytes() : 0);
    void* dev = rec->stageTables(bytes);
    if (!dev) return rec->fail(HM_ERR_OUT_OF_MEMORY, "stage-update tables");
    std::vector<char> h(bytes, 0);
    SegTableT<RKPatch> tc = cells.write(h.data(), static_cast<char*>(dev));
    SegTableT<RKPatch> ta{};
    if (any_acc) ta = acc.write(h.data() + cells.bytes(), static_cast<char*>(dev) + cells.bytes());
    // Synchronous upload: the host image dies with this call.
    if (hipMemcpy(dev, h.data(), bytes, hipMemcpyHostToDevice) != hipSuccess)
        return rec->fail(HM_ERR_DEVICE, "stage-update table upload");
    if (launch_rk_update<SegTableT<RKPatch>>(d, tc, cells.total, any_acc ? &ta : nullptr, acc.total, r, s) !=
        hipSuccess)
        return rec->fail(HM_ERR_DEVICE, "stage update launch failed");
    return rec->stageTablesInUse(s);
}

// ----------------------------------------------------------------------------
// Two-level hierarchy (refinement ratio 2): the time-refined advance of
// SAMRAI's TimeRefinementIntegrator as RungeKuttaLevelIntegrator drives it --
// one coarse step, two fine steps of half the size, then
// RungeKuttaLevelIntegrator::synchronizeLevelWithCoarser
// (RungeKuttaLevelIntegrator.cpp:2131-2225).
//   * Coarse-fine ghost cells of the fine level (inside the domain, covered by
//     no fine patch) are filled before each fine step from the coarse level,
//     time-interpolated between its states at the start and end of the coarse
//     step (the advance fill with the coarser level, :1500-1548, linear time
//     interpolation) and refined by conservative linear interpolation (the
//     CONSERVATIVE_LINEAR_REFINE operator the flow models register,
//     FlowModelSingleSpecies.cpp:578-597): monotonized-central limited slopes.
//     The intermediate stages keep them (their fills are same-level only,
//     :2513-2520).
//   * Both levels accumulate gamma-weighted flux sums; the coarse faces on the
//     coarse-fine boundary take the area average of the fine flux sums over
//     the two fine steps (upfluxsum, algs_upfluxsum2d.m4, coarsened to the
//     coarse level), the coarse update is repeated from the coarse state at the
//     start of the step with these fluxes (Euler::synchronizeFluxes,
//     Euler.cpp:1682-1990: Q += (-(F_R - F_L)/dx ... + S)), and the coarse
//     cells under the fine level take the average of their fine cells
//     (CONSERVATIVE_COARSEN).
// SAMRAI's refine and coarsen operators are not part of the reference; the
// ones here are the published forms those operator names denote, and they
// are pinned by conservation, exactness and monotonicity checks.
// ----------------------------------------------------------------------------

struct CfItem {      // one coarse-fine ghost cell of a fine patch
    int fp, cq;      // fine patch, coarse patch holding the parent cell
    int64_t fidx;    // cell in the fine patch's ghost box
    int64_t cidx;    // parent in the coarse patch's ghost box
    int child;       // bit a: the cell is the upper child along axis a
};

struct RefluxItem {  // one coarse face on the coarse-fine boundary, in a coarse patch's face array
    int cq, axis, fp, nf;
    int64_t cface;
    int64_t fface[4];
};

struct CoarsenItem { // one coarse cell under the fine level
    int cq, fp;
    int64_t cidx;    // coarse ghost-box index
    int64_t fidx;    // ghost-box index of the lower-corner child in the fine patch
};

struct SyncPatch {   // coarse patch data the synchronization kernels use
    int n[3];
    int64_t gvol;
    double dx[3];
    const double* Qold;  // state at the start of the coarse step (nq comps over the ghost box)
    double* Q;           // state (nq comps over the ghost box)
    double* Facc[3];     // flux sums (neq comps over the faces of each axis)
    double* Fdacc[3];    // diffusive flux sums (Navier-Stokes form), else NULL
    const double* Sacc;  // source sums (neq comps over the interior)
};

// MC-limited slope of the coarse values vL, v0, vR (per unit coarse cell).
__device__ __forceinline__ double mc_slope(double vL, double v0, double vR) {
    const double dR = vR - v0, dL = v0 - vL;
    if (!(dR * dL > 0.0)) return 0.0;
    const double c = 0.5 * (dR + dL);
    const double b = 2.0 * fmin(fabs(dR), fabs(dL));
    return copysign(fmin(fabs(c), b), c);
}

template <int D>
__global__ __launch_bounds__(kBlock) void k_cf_fill(const CfItem* items, int64_t n_items, const LevelPatch* fine,
                                                    const LevelPatch* coarse, double alpha, int nq, int g) {
    const int64_t t = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (t >= n_items) return;
    const CfItem it = items[t];
    const LevelPatch& C = coarse[it.cq];
    const int64_t st[3] = {1, int64_t(C.n[0] + 2 * g), int64_t(C.n[0] + 2 * g) * ((D > 1) ? C.n[1] + 2 * g : 1)};
    const double* old = C.Q[1];  // the coarse step's first stage input: its start state, ghosts filled
    const double* nw = C.Q[0];
    const LevelPatch& F = fine[it.fp];
    for (int c = 0; c < nq; c++) {
        auto val = [&](int64_t i) {
            const double o = old[c * C.gvol + i];
            return o + alpha * (nw[c * C.gvol + i] - o);
        };
        const double v0 = val(it.cidx);
        double v = v0;
        for (int a = 0; a < D; a++) {
            const double sl = mc_slope(val(it.cidx - st[a]), v0, val(it.cidx + st[a]));
            v = v + sl * (((it.child >> a) & 1) ? 0.25 : -0.25);
        }
        F.Q[0][c * F.gvol + it.fidx] = v;
    }
}

// Flux sums of the fine faces on the coarse-fine boundary: after each fine step
// its gamma-weighted flux is added to a sum that starts at zero
// (RungeKuttaLevelIntegrator::postprocessFluxAndSourceData,
// RungeKuttaLevelIntegrator.cpp:2920-2967, 3060-3180; algs_upfluxsum2d.m4).
// fsum holds [item][fine face j < 4][equation].
__global__ __launch_bounds__(kBlock) void k_fsum_add(const RefluxItem* items, int64_t n_items,
                                                     double* const* fine_facc, const int64_t* fine_fvol, int neq,
                                                     double* fsum, int first) {
    const int64_t t = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (t >= n_items) return;
    const RefluxItem it = items[t];
    const double* ff = fine_facc[it.fp * 3 + it.axis];
    const int64_t fv = fine_fvol[it.fp * 3 + it.axis];
    for (int j = 0; j < it.nf; j++)
        for (int e = 0; e < neq; e++) {
            double* o = fsum + (t * 4 + j) * neq + e;
            const double v = ff[e * fv + it.fface[j]];
            *o = first ? 0.0 + v : *o + v;
        }
}

// A coarse face on the coarse-fine boundary takes the area average of the
// fine flux sums over it: their sum times 2^(1-d).
// kind 0: the convective flux sums (SyncPatch::Facc), 1: the diffusive ones (Fdacc).
__global__ __launch_bounds__(kBlock) void k_reflux(const RefluxItem* items, int64_t n_items, const SyncPatch* coarse,
                                                   const double* fsum, int neq, int kind) {
    const int64_t t = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (t >= n_items) return;
    const RefluxItem it = items[t];
    const SyncPatch& C = coarse[it.cq];
    int64_t cv = 1;
    for (int b = 0; b < 3; b++) cv *= C.n[b] + (b == it.axis);
    const double w = 1.0 / double(it.nf);
    for (int e = 0; e < neq; e++) {
        double sum = fsum[(t * 4) * neq + e];
        for (int j = 1; j < it.nf; j++) sum += fsum[(t * 4 + j) * neq + e];
        (kind ? C.Fdacc : C.Facc)[it.axis][e * cv + it.cface] = sum * w;
    }
}

template <int D>
__global__ __launch_bounds__(kBlock) void k_sync_update(SegTableT<SyncPatch> w, int neq, int g, int nz, int zfirst,
                                                        int ns) {
    const int64_t gid = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    const SyncPatch* P;
    int ax;
    int64_t loc;
    if (!w.find(gid, P, ax, loc)) return;
    const int n0 = P->n[0], n1 = (D > 1) ? P->n[1] : 1, n2 = (D > 2) ? P->n[2] : 1;
    const int64_t r = loc / n0;
    const int i = int(loc - r * n0);
    const int j = (D > 1) ? int(r % n1) : 0;
    const int k = (D > 2) ? int(r / n1) : 0;
    const int e0 = n0 + 2 * g, e1 = (D > 1) ? n1 + 2 * g : 1;
    const int64_t iq = (i + g) + int64_t(j + ((D > 1) ? g : 0)) * e0 + int64_t(k + ((D > 2) ? g : 0)) * e0 * e1;
    const int64_t fL0 = i + int64_t(j) * (n0 + 1) + int64_t(k) * (n0 + 1) * n1, fR0 = fL0 + 1;
    const int64_t fL1 = i + int64_t(j) * n0 + int64_t(k) * n0 * (n1 + 1), fR1 = fL1 + n0;
    const int64_t fL2 = loc, fR2 = loc + int64_t(n0) * n1;
    const int64_t fv0 = int64_t(n0 + 1) * n1 * n2, fv1 = int64_t(n0) * (n1 + 1) * n2, fv2 = int64_t(n0) * n1 * (n2 + 1);
    const int64_t sv = int64_t(n0) * n1 * n2;
    for (int e = 0; e < neq; e++) {
        // Euler::synchronizeFluxes (Euler.cpp:1682-1943): Q += (-(F_x,R - F_x,L)/dx_0 - ... + S);
        // NavierStokes::synchronizeFluxes (NavierStokes.cpp:2424-2700):
        // Q += (-(F_c,x,R - F_c,x,L + F_d,x,R - F_d,x,L)/dx_0 - ... + S)
        auto diff = [&](int a, int64_t fL, int64_t fR, int64_t fv) {
            double dF = P->Facc[a][e * fv + fR] - P->Facc[a][e * fv + fL];
            if (ns) dF = dF + P->Fdacc[a][e * fv + fR] - P->Fdacc[a][e * fv + fL];
            return dF;
        };
        double t = -diff(0, fL0, fR0, fv0) / P->dx[0];
        if constexpr (D > 1) t = t - diff(1, fL1, fR1, fv1) / P->dx[1];
        if constexpr (D > 2) t = t - diff(2, fL2, fR2, fv2) / P->dx[2];
        P->Q[e * P->gvol + iq] = P->Qold[e * P->gvol + iq] + (t + P->Sacc[e * sv + loc]);
    }
    // Five-equation model: the last volume fraction is 1 minus the others, in
    // species order (FlowModelFiveEqnAllaire::updateCellDataOfConservativeVariables,
    // FlowModelFiveEqnAllaire.cpp:1739-1890, called at Euler.cpp:1937-1939).
    if (nz > 0) {
        double zl = 1.0;
        for (int z = 0; z < nz; z++) zl -= P->Q[(zfirst + z) * P->gvol + iq];
        P->Q[neq * P->gvol + iq] = zl;
    }
}

template <int D>
__global__ __launch_bounds__(kBlock) void k_coarsen(const CoarsenItem* items, int64_t n_items, const LevelPatch* coarse,
                                                    const LevelPatch* fine, int nq, int g) {
    const int64_t t = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (t >= n_items) return;
    const CoarsenItem it = items[t];
    const LevelPatch& C = coarse[it.cq];
    const LevelPatch& F = fine[it.fp];
    const int64_t s1 = F.n[0] + 2 * g, s2 = s1 * ((D > 1) ? F.n[1] + 2 * g : 1);
    constexpr int NC = 1 << D;
    const double w = 1.0 / double(NC);
    for (int c = 0; c < nq; c++) {
        const double* f = F.Q[0] + c * F.gvol + it.fidx;
        double sum = f[0];
        for (int q = 1; q < NC; q++) sum += f[(q & 1) + ((q >> 1) & 1) * s1 + ((q >> 2) & 1) * s2];
        C.Q[0][c * C.gvol + it.cidx] = sum * w;
    }
}

class Hierarchy {
   public:
    Hierarchy(Level* coarse, Level* fine) : c_(coarse), f_(fine) {}
    ~Hierarchy() {
        if (mem_) (void)hipFree(mem_);
    }

    int init() {
        ConvectiveFluxReconstructorGPU* rec = c_->reconstructor();
        const int d = c_->dim();
        if (f_->reconstructor() != rec) return rec->fail(HM_ERR_INVALID_ARGUMENT, "levels of different contexts");
        if (f_->dim() != d || f_->ghost() != c_->ghost())
            return rec->fail(HM_ERR_INVALID_ARGUMENT, "levels of different dimension or ghost width");
        if (c_->diffusive() != f_->diffusive())
            return rec->fail(HM_ERR_INVALID_ARGUMENT, "one level in the Navier-Stokes form, the other not");
        ns_ = c_->diffusive();
        for (int a = 0; a < d; a++) {
            if (f_->domainLo()[a] != 2 * c_->domainLo()[a] || f_->domainN()[a] != 2 * c_->domainN()[a])
                return rec->fail(HM_ERR_INVALID_ARGUMENT, "the fine domain must be the coarse domain refined by 2");
            if (f_->periodic(a) != c_->periodic(a))
                return rec->fail(HM_ERR_INVALID_ARGUMENT, "the levels' periodic axes differ");
            if (std::fabs(c_->dx()[a] - 2.0 * f_->dx()[a]) > 1e-12 * c_->dx()[a])
                return rec->fail(HM_ERR_INVALID_ARGUMENT, "coarse dx must be twice the fine dx");
        }
        for (int p = 0; p < f_->numPatches(); p++)
            for (int a = 0; a < d; a++)
                if ((f_->patchLo(p)[a] & 1) || (f_->patchN(p)[a] & 1))
                    return rec->fail(HM_ERR_INVALID_ARGUMENT, "fine patch boxes must start and end on coarse cells");
        g_ = c_->ghost();
        if (int st = c_->enableFluxSums()) return st;
        if (int st = f_->enableFluxSums()) return st;
        return build();
    }

    // num_steps coarse steps of dt, each with two fine steps of dt / 2 and the
    // synchronization.
    int advance(double dt, int nsteps, hipStream_t s) {
        ConvectiveFluxReconstructorGPU* rec = c_->reconstructor();
        if (nsteps < 0 || !(dt > 0.0)) return rec->fail(HM_ERR_INVALID_ARGUMENT, "bad dt or step count");
        if (int st = rec->selectDevice()) return st;
        const int d = c_->dim();
        for (int it = 0; it < nsteps; it++) {
            if (int st = c_->clearFluxSums(s)) return st;
            if (int st = c_->stepOnce(dt, s)) return st;
            if (int st = c_->fillGhostsOf(0, s)) return st;  // the coarse state at the end of the step
            for (int k = 0; k < 2; k++) {
                if (int st = f_->clearFluxSums(s)) return st;  // the fine step's own flux sums
                // coarse-fine ghosts first: the fine step's own fill (same level,
                // then the physical boundary conditions, which may read them) follows
                if (ncf_ > 0) {
                    const double alpha = 0.5 * k;
                    switch (d) {
                        case 1: hipLaunchKernelGGL(k_cf_fill<1>, dim3(blocks_for(ncf_)), dim3(kBlock), 0, s, cf_, ncf_, f_->devPatches(), c_->devPatches(), alpha, f_->numArrays(), g_); break;
                        case 2: hipLaunchKernelGGL(k_cf_fill<2>, dim3(blocks_for(ncf_)), dim3(kBlock), 0, s, cf_, ncf_, f_->devPatches(), c_->devPatches(), alpha, f_->numArrays(), g_); break;
                        default: hipLaunchKernelGGL(k_cf_fill<3>, dim3(blocks_for(ncf_)), dim3(kBlock), 0, s, cf_, ncf_, f_->devPatches(), c_->devPatches(), alpha, f_->numArrays(), g_); break;
                    }
                }
                if (int st = f_->stepOnce(0.5 * dt, s)) return st;
                if (nrf_ > 0) {
                    hipLaunchKernelGGL(k_fsum_add, dim3(blocks_for(nrf_)), dim3(kBlock), 0, s, rf_, nrf_, ffacc_, ffvol_,
                                       c_->numEqn(), fsum_, int(k == 0));
                    if (ns_)
                        hipLaunchKernelGGL(k_fsum_add, dim3(blocks_for(nrf_)), dim3(kBlock), 0, s, rf_, nrf_, ffdacc_,
                                           ffvol_, c_->numEqn(), fsumd_, int(k == 0));
                }
            }
            if (nrf_ > 0) {
                hipLaunchKernelGGL(k_reflux, dim3(blocks_for(nrf_)), dim3(kBlock), 0, s, rf_, nrf_, sync_.patches, fsum_,
                                   c_->numEqn(), 0);
                if (ns_)
                    hipLaunchKernelGGL(k_reflux, dim3(blocks_for(nrf_)), dim3(kBlock), 0, s, rf_, nrf_, sync_.patches,
                                       fsumd_, c_->numEqn(), 1);
            }
            const hm_config& cfg = rec->config();
            const bool five = cfg.model == HM_MODEL_FIVE_EQN_ALLAIRE;
            const int nz = five ? cfg.num_species - 1 : 0, zfirst = five ? cfg.num_species + d + 1 : 0;
            switch (d) {
                case 1: hipLaunchKernelGGL(k_sync_update<1>, dim3(blocks_for(sync_items_)), dim3(kBlock), 0, s, sync_, c_->numEqn(), g_, nz, zfirst, int(ns_)); break;
                case 2: hipLaunchKernelGGL(k_sync_update<2>, dim3(blocks_for(sync_items_)), dim3(kBlock), 0, s, sync_, c_->numEqn(), g_, nz, zfirst, int(ns_)); break;
                default: hipLaunchKernelGGL(k_sync_update<3>, dim3(blocks_for(sync_items_)), dim3(kBlock), 0, s, sync_, c_->numEqn(), g_, nz, zfirst, int(ns_)); break;
            }
            if (ncs_ > 0) {
                switch (d) {
                    case 1: hipLaunchKernelGGL(k_coarsen<1>, dim3(blocks_for(ncs_)), dim3(kBlock), 0, s, cs_, ncs_, c_->devPatches(), f_->devPatches(), c_->numArrays(), g_); break;
                    case 2: hipLaunchKernelGGL(k_coarsen<2>, dim3(blocks_for(ncs_)), dim3(kBlock), 0, s, cs_, ncs_, c_->devPatches(), f_->devPatches(), c_->numArrays(), g_); break;
                    default: hipLaunchKernelGGL(k_coarsen<3>, dim3(blocks_for(ncs_)), dim3(kBlock), 0, s, cs_, ncs_, c_->devPatches(), f_->devPatches(), c_->numArrays(), g_); break;
                }
            }
            if (hipGetLastError() != hipSuccess) return rec->fail(HM_ERR_DEVICE, "hierarchy launch failed");
        }
        return HM_OK;
    }

   private:
    // Whether coarse cell x (global, wrapped on periodic axes) lies under a fine patch; its fine patch.
    int fine_over(const int* xc) const {
        const int d = c_->dim();
        int xf[3] = {0, 0, 0};
        for (int a = 0; a < d; a++) xf[a] = 2 * xc[a];
        return f_->patchOf(xf);
    }

    int build() {
        ConvectiveFluxReconstructorGPU* rec = c_->reconstructor();
        const int d = c_->dim(), g = g_;
        std::vector<CfItem> cf;
        std::vector<RefluxItem> rf;
        std::vector<CoarsenItem> cs;
        auto ghost_index = [&](const int* n, const int* x) {  // x in ghost-box coordinates
            int64_t i = 0, st = 1;
            for (int a = 0; a < d; a++) {
                i += x[a] * st;
                st *= n[a] + 2 * g;
            }
            return i;
        };
        auto wrap = [&](const Level* L, int a, int v) {
            if (!L->periodic(a)) return v;
            const int lo = L->domainLo()[a], n = L->domainN()[a];
            return lo + ((v - lo) % n + n) % n;
        };
        // coarse-fine ghost cells of every fine patch
        for (int p = 0; p < f_->numPatches(); p++) {
            const int* lo = f_->patchLo(p);
            const int* n = f_->patchN(p);
            int ext[3] = {1, 1, 1};
            for (int a = 0; a < d; a++) ext[a] = n[a] + 2 * g;
            for (int z = 0; z < ext[2]; z++)
                for (int y = 0; y < ext[1]; y++)
                    for (int x = 0; x < ext[0]; x++) {
                        const int gx[3] = {x, y, z};
                        int xg[3] = {0, 0, 0};  // global fine index
                        bool interior = true, outside = false;
                        for (int a = 0; a < d; a++) {
                            xg[a] = lo[a] - g + gx[a];
                            if (gx[a] < g || gx[a] >= g + n[a]) interior = false;
                            if (!f_->periodic(a) &&
                                (xg[a] < f_->domainLo()[a] || xg[a] >= f_->domainLo()[a] + f_->domainN()[a]))
                                outside = true;
                        }
                        if (interior || outside || f_->patchOf(xg) >= 0) continue;
                        int xw[3] = {0, 0, 0}, xc[3] = {0, 0, 0}, child = 0;
                        for (int a = 0; a < d; a++) {
                            xw[a] = wrap(f_, a, xg[a]);
                            xc[a] = (xw[a] >= 0) ? xw[a] / 2 : -((-xw[a] + 1) / 2);
                            if (xw[a] - 2 * xc[a]) child |= 1 << a;
                        }
                        const int q = c_->patchOf(xc);
                        if (q < 0)
                            return rec->fail(HM_ERR_INVALID_ARGUMENT,
                                             "a coarse-fine ghost cell has no coarse parent: the fine level is not nested");
                        int cgx[3] = {0, 0, 0};
                        for (int a = 0; a < d; a++) {
                            cgx[a] = wrap(c_, a, xc[a]) - c_->patchLo(q)[a] + g;
                            if (cgx[a] < 1 || cgx[a] > c_->patchN(q)[a] + 2 * g - 2)
                                return rec->fail(HM_ERR_INVALID_ARGUMENT, "coarse parent too close to its ghost-box edge");
                        }
                        cf.push_back(CfItem{p, q, ghost_index(n, gx), ghost_index(c_->patchN(q), cgx), child});
                    }
        }
        // coarse cells under the fine level, and coarse faces on the coarse-fine boundary
        for (int q = 0; q < c_->numPatches(); q++) {
            const int* lo = c_->patchLo(q);
            const int* n = c_->patchN(q);
            int ext[3] = {1, 1, 1};
            for (int a = 0; a < d; a++) ext[a] = n[a];
            for (int z = 0; z < ext[2]; z++)
                for (int y = 0; y < ext[1]; y++)
                    for (int x = 0; x < ext[0]; x++) {
                        const int ix[3] = {x, y, z};
                        int xc[3] = {0, 0, 0}, gx[3] = {0, 0, 0};
                        for (int a = 0; a < d; a++) {
                            xc[a] = lo[a] + ix[a];
                            gx[a] = ix[a] + g;
                        }
                        const int p = fine_over(xc);
                        if (p < 0) continue;
                        int fx[3] = {0, 0, 0};
                        for (int a = 0; a < d; a++) fx[a] = 2 * xc[a] - f_->patchLo(p)[a] + g;
                        cs.push_back(CoarsenItem{q, p, ghost_index(n, gx), ghost_index(f_->patchN(p), fx)});
                    }
            for (int a = 0; a < d; a++) {
                int fe[3] = {1, 1, 1};
                for (int b = 0; b < d; b++) fe[b] = n[b] + (b == a);
                for (int z = 0; z < fe[2]; z++)
                    for (int y = 0; y < fe[1]; y++)
                        for (int x = 0; x < fe[0]; x++) {
                            const int ix[3] = {x, y, z};
                            int R[3] = {0, 0, 0}, Lc[3] = {0, 0, 0};
                            for (int b = 0; b < d; b++) {
                                R[b] = lo[b] + ix[b];
                                Lc[b] = R[b] - (b == a);
                            }
                            int Rw[3] = {0, 0, 0}, Lw[3] = {0, 0, 0};
                            bool outL = false, outR = false;
                            for (int b = 0; b < d; b++) {
                                Rw[b] = wrap(c_, b, R[b]);
                                Lw[b] = wrap(c_, b, Lc[b]);
                                const int dlo = c_->domainLo()[b], dhi = dlo + c_->domainN()[b];
                                if (Rw[b] < dlo || Rw[b] >= dhi) outR = true;
                                if (Lw[b] < dlo || Lw[b] >= dhi) outL = true;
                            }
                            if (outL || outR) continue;  // a physical boundary face
                            const int pL = fine_over(Lw), pR = fine_over(Rw);
                            if ((pL >= 0) == (pR >= 0)) continue;
                            const int p = (pL >= 0) ? pL : pR;
                            RefluxItem it{};
                            it.cq = q;
                            it.axis = a;
                            it.fp = p;
                            int64_t cf_i = 0, st = 1;
                            for (int b = 0; b < d; b++) {
                                cf_i += ix[b] * st;
                                st *= fe[b];
                            }
                            it.cface = cf_i;
                            // the fine faces: along a at 2 R (the covered side's fine patch face), 2^(d-1) transverse children
                            const int* flo = f_->patchLo(p);
                            const int* fn = f_->patchN(p);
                            const int* Rref = (pL >= 0) ? Lw : Rw;  // a cell of the covered side
                            int ffe[3] = {1, 1, 1};
                            for (int b = 0; b < d; b++) ffe[b] = fn[b] + (b == a);
                            int nf = 0;
                            for (int t = 0; t < (1 << (d - 1)); t++) {
                                int fx[3] = {0, 0, 0}, bit = 0;
                                for (int b = 0; b < d; b++) {
                                    if (b == a) {
                                        // the face between the two sides, in the fine patch's face numbering
                                        const int face_global = (pL >= 0) ? 2 * (Lw[a] + 1) : 2 * Rw[a];
                                        fx[b] = face_global - flo[b];
                                    } else {
                                        fx[b] = 2 * Rref[b] + ((t >> bit) & 1) - flo[b];
                                        bit++;
                                    }
                                    if (fx[b] < 0 || fx[b] >= ffe[b])
                                        return rec->fail(HM_ERR_INVALID_ARGUMENT, "fine face outside its patch");
                                }
                                int64_t fi = 0, fst = 1;
                                for (int b = 0; b < d; b++) {
                                    fi += fx[b] * fst;
                                    fst *= ffe[b];
                                }
                                it.fface[nf++] = fi;
                            }
                            it.nf = nf;
                            rf.push_back(it);
                        }
            }
        }
        // sync tables: coarse patch descriptors and their cells
        const int ncp = c_->numPatches(), nfp = f_->numPatches();
        TableImage<SyncPatch> sync;
        sync.items.resize(ncp);
        for (int q = 0; q < ncp; q++) {
            SyncPatch& P = sync.items[q];
            std::memset(&P, 0, sizeof P);
            for (int a = 0; a < 3; a++) {
                P.n[a] = a < d ? c_->patchN(q)[a] : 1;
                P.dx[a] = a < d ? c_->dx()[a] : 1.0;
            }
            P.gvol = c_->ghostVolume(q);
            P.Qold = c_->hostPatch(q).Q[1];
            P.Q = c_->hostPatch(q).Q[0];
            for (int a = 0; a < d; a++) {
                P.Facc[a] = c_->fluxSum(q, a);
                P.Fdacc[a] = c_->diffusiveFluxSum(q, a);
            }
            P.Sacc = c_->sourceSum(q);
            int64_t v = 1;
            for (int a = 0; a < d; a++) v *= c_->patchN(q)[a];
            sync.add(q, 0, v);
        }
        std::vector<double*> ffacc(size_t(nfp) * 3, nullptr), ffdacc(size_t(nfp) * 3, nullptr);
        std::vector<int64_t> ffvol(size_t(nfp) * 3, 0);
        for (int p = 0; p < nfp; p++)
            for (int a = 0; a < d; a++) {
                ffacc[size_t(p) * 3 + a] = f_->fluxSum(p, a);
                ffdacc[size_t(p) * 3 + a] = f_->diffusiveFluxSum(p, a);
                int64_t v = 1;
                for (int b = 0; b < d; b++) v *= f_->patchN(p)[b] + (b == a);
                ffvol[size_t(p) * 3 + a] = v;
            }
        const size_t fsum_bytes = up256(sizeof(double) * rf.size() * 4 * size_t(c_->numEqn()));
        const size_t bytes = up256(sizeof(CfItem) * cf.size()) + up256(sizeof(RefluxItem) * rf.size()) +
                             up256(sizeof(CoarsenItem) * cs.size()) + sync.bytes() +
                             2 * up256(sizeof(double*) * ffacc.size()) + up256(sizeof(int64_t) * ffvol.size()) +
                             fsum_bytes * (ns_ ? 2 : 1);
        if (hipMalloc(&mem_, bytes) != hipSuccess) {
            mem_ = nullptr;
            return rec->fail(HM_ERR_OUT_OF_MEMORY, "hierarchy tables");
        }
        std::vector<char> h(bytes, 0);
        char* dv = static_cast<char*>(mem_);
        size_t o = 0;
        auto put = [&](const void* src, size_t n) {
            if (n) std::memcpy(h.data() + o, src, n);
            const size_t at = o;
            o += up256(n);
            return dv + at;
        };
        cf_ = reinterpret_cast<const CfItem*>(put(cf.data(), sizeof(CfItem) * cf.size()));
        rf_ = reinterpret_cast<const RefluxItem*>(put(rf.data(), sizeof(RefluxItem) * rf.size()));
        cs_ = reinterpret_cast<const CoarsenItem*>(put(cs.data(), sizeof(CoarsenItem) * cs.size()));
        sync_ = sync.write(h.data() + o, dv + o);
        sync_items_ = sync.total;
        o += sync.bytes();
        ffacc_ = reinterpret_cast<double* const*>(put(ffacc.data(), sizeof(double*) * ffacc.size()));
        ffdacc_ = reinterpret_cast<double* const*>(put(ffdacc.data(), sizeof(double*) * ffdacc.size()));
        ffvol_ = reinterpret_cast<const int64_t*>(put(ffvol.data(), sizeof(int64_t) * ffvol.size()));
        fsum_ = reinterpret_cast<double*>(dv + o);  // written by k_fsum_add before it is read
        o += fsum_bytes;
        if (ns_) {
            fsumd_ = reinterpret_cast<double*>(dv + o);
            o += fsum_bytes;
        }
        ncf_ = int64_t(cf.size());
        nrf_ = int64_t(rf.size());
        ncs_ = int64_t(cs.size());
        if (hipMemcpy(mem_, h.data(), bytes, hipMemcpyHostToDevice) != hipSuccess)
            return rec->fail(HM_ERR_DEVICE, "hierarchy table upload");
        return HM_OK;
    }

    Level* c_;
    Level* f_;
    int g_ = kGc;
    void* mem_ = nullptr;
    const CfItem* cf_ = nullptr;
    const RefluxItem* rf_ = nullptr;
    const CoarsenItem* cs_ = nullptr;
    SegTableT<SyncPatch> sync_{};
    double* const* ffacc_ = nullptr;
    double* const* ffdacc_ = nullptr;  // the fine diffusive flux sums (Navier-Stokes form)
    const int64_t* ffvol_ = nullptr;
    double* fsum_ = nullptr;
    double* fsumd_ = nullptr;  // the fine diffusive flux sums over the coarse step
    bool ns_ = false;          // both levels in the Navier-Stokes form
    int64_t ncf_ = 0, nrf_ = 0, ncs_ = 0, sync_items_ = 0;
};

}  // namespace hm

struct hm_level {
    std::unique_ptr<hm::Level> lvl;
    hm_ctx* ctx;
    int hierarchies = 0;  // live hierarchies built on this level
};

struct hm_hierarchy {
    std::unique_ptr<hm::Hierarchy> h;
    hm_level* coarse;
    hm_level* fine;
};

extern "C" {

int hm_level_create(hm_ctx* ctx, int num_patches, const int* patch_lo, const int* patch_n, const int* domain_lo,
                    const int* domain_n, const int* periodic, const double* dx, hm_level** out) {
    if (!ctx || !ctx->rec || !out || !patch_lo || !patch_n || !domain_lo || !domain_n || !periodic || !dx)
        return HM_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    if (int st = ctx->rec->selectDevice()) return st;
    hm_level* l = new (std::nothrow) hm_level();
    if (!l) return HM_ERR_OUT_OF_MEMORY;
    l->ctx = ctx;
    ctx->levels++;
    l->lvl.reset(new (std::nothrow) hm::Level(ctx->rec.get(), ctx->diff.get()));
    if (!l->lvl) {
        ctx->levels--;
        delete l;
        return HM_ERR_OUT_OF_MEMORY;
    }
    const int st = l->lvl->init(num_patches, patch_lo, patch_n, domain_lo, domain_n, periodic, dx);
    if (st != HM_OK) {
        l->lvl.reset();
        ctx->levels--;
        delete l;
        return st;
    }
    *out = l;
    return HM_OK;
}

int hm_hierarchy_create(hm_level* coarse, hm_level* fine, int ratio, hm_hierarchy** out) {
    if (!coarse || !fine || !out || coarse == fine) return HM_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    if (ratio != 2) return coarse->ctx->rec->fail(HM_ERR_UNSUPPORTED, "only refinement ratio 2 is built");
    if (coarse->ctx != fine->ctx) return coarse->ctx->rec->fail(HM_ERR_INVALID_ARGUMENT, "levels of different contexts");
    if (int st = coarse->ctx->rec->selectDevice()) return st;
    hm_hierarchy* h = new (std::nothrow) hm_hierarchy();
    if (!h) return HM_ERR_OUT_OF_MEMORY;
    h->coarse = coarse;
    h->fine = fine;
    h->h.reset(new (std::nothrow) hm::Hierarchy(coarse->lvl.get(), fine->lvl.get()));
    if (!h->h) {
        delete h;
        return HM_ERR_OUT_OF_MEMORY;
    }
    if (int st = h->h->init()) {
        delete h;
        return st;
    }
    coarse->hierarchies++;
    fine->hierarchies++;
    *out = h;
    return HM_OK;
}

int hm_hierarchy_destroy(hm_hierarchy* h) {
    if (!h) return HM_OK;
    h->coarse->hierarchies--;
    h->fine->hierarchies--;
    delete h;
    return HM_OK;
}

int hm_hierarchy_advance(hm_hierarchy* h, double dt_coarse, int num_steps, void* stream) {
    if (!h) return HM_ERR_INVALID_ARGUMENT;
    return h->h->advance(dt_coarse, num_steps, static_cast<hipStream_t>(stream));
}

int hm_level_destroy(hm_level* level) {
    if (!level) return HM_OK;
    if (level->hierarchies > 0) {
        level->ctx->rec->fail(HM_ERR_INVALID_ARGUMENT, "destroy the hierarchies of this level first");
        return HM_ERR_INVALID_ARGUMENT;
    }
    level->lvl.reset();
    level->ctx->levels--;
    delete level;
    return HM_OK;
}

int hm_level_set_rk_coefficients(hm_level* level, int num_stages, const double* alpha, const double* beta,
                                 const double* gamma) {
    if (!level || !alpha || !beta) return HM_ERR_INVALID_ARGUMENT;
    return level->lvl->setCoefficients(num_stages, alpha, beta, gamma);
}

int hm_level_write_patch(hm_level* level, int patch, const double* Q) {
    if (!level) return HM_ERR_INVALID_ARGUMENT;
    return level->lvl->writePatch(patch, Q);
}

int hm_level_read_patch(hm_level* level, int patch, double* Q) {
    if (!level) return HM_ERR_INVALID_ARGUMENT;
    return level->lvl->readPatch(patch, Q);
}

int hm_level_patch_array(hm_level* level, int patch, double** Q, int* ghost) {
    if (!level) return HM_ERR_INVALID_ARGUMENT;
    return level->lvl->patchArray(patch, Q, ghost);
}

int hm_level_set_boundary_conditions(hm_level* level, const int* conditions, const double* dirichlet) {
    if (!level || !conditions) return HM_ERR_INVALID_ARGUMENT;
    if (int st = level->ctx->rec->selectDevice()) return st;
    return level->lvl->setBoundaryConditions(conditions, dirichlet);
}

int hm_level_autotune(hm_level* level, double dt, int repetitions, hm_schedule* chosen, double* ms) {
    if (!level) return HM_ERR_INVALID_ARGUMENT;
    hm::Schedule best;
    double best_ms = 0.0, default_ms = 0.0;
    if (int st = level->lvl->autotune(dt, repetitions, best, best_ms, default_ms)) return st;
    if (chosen) {
        chosen->tile_rows = best.tile_rows;
        chosen->xcd_remap = best.xcd;
        chosen->order = best.order;
        chosen->persistent = best.persistent;
    }
    if (ms) {
        ms[0] = best_ms;
        ms[1] = default_ms;
    }
    return HM_OK;
}

int hm_level_set_schedule(hm_level* level, const hm_schedule* schedule) {
    if (!level || !schedule) return HM_ERR_INVALID_ARGUMENT;
    if (int st = level->ctx->rec->selectDevice()) return st;
    hm::Schedule sc;
    sc.tile_rows = schedule->tile_rows;
    sc.xcd = schedule->xcd_remap;
    sc.order = schedule->order;
    sc.persistent = schedule->persistent;
    return level->lvl->setSchedule(sc);
}

int hm_level_get_schedule(hm_level* level, hm_schedule* schedule) {
    if (!level || !schedule) return HM_ERR_INVALID_ARGUMENT;
    const hm::Schedule& sc = level->lvl->schedule();
    schedule->tile_rows = sc.tile_rows;
    schedule->xcd_remap = sc.xcd;
    schedule->order = sc.order;
    schedule->persistent = sc.persistent;
    return HM_OK;
}

int hm_level_fill_ghosts(hm_level* level, void* stream) {
    if (!level) return HM_ERR_INVALID_ARGUMENT;
    return level->lvl->fillGhosts(static_cast<hipStream_t>(stream));
}

int hm_level_advance(hm_level* level, double dt, int num_steps, int use_graph, void* stream) {
    if (!level) return HM_ERR_INVALID_ARGUMENT;
    return level->lvl->advance(dt, num_steps, use_graph != 0, static_cast<hipStream_t>(stream));
}

int hm_level_advance_cfl(hm_level* level, double cfl, int num_steps, int use_graph, void* stream) {
    if (!level) return HM_ERR_INVALID_ARGUMENT;
    return level->lvl->advanceCfl(cfl, num_steps, use_graph != 0, static_cast<hipStream_t>(stream));
}

int hm_level_dt_info(hm_level* level, double* info) {
    if (!level) return HM_ERR_INVALID_ARGUMENT;
    return level->lvl->dtInfo(info);
}

int hm_level_advance_recorded(hm_level* level, double dt, int num_steps, int checkpoint_every, int use_graph,
                              void* stream) {
    if (!level) return HM_ERR_INVALID_ARGUMENT;
    return level->lvl->advanceRecorded(dt, num_steps, checkpoint_every, use_graph != 0,
                                       static_cast<hipStream_t>(stream));
}

int hm_level_set_face_state_budget(hm_level* level, int64_t bytes) {
    if (!level) return HM_ERR_INVALID_ARGUMENT;
    return level->lvl->setFaceStateBudget(bytes);
}

int hm_level_face_states_kept(hm_level* level) {
    if (!level) return HM_ERR_INVALID_ARGUMENT;
    return level->lvl->faceStatesKept() ? 1 : 0;
}

int hm_level_adjoint(hm_level* level, void* stream) {
    if (!level) return HM_ERR_INVALID_ARGUMENT;
    return level->lvl->adjoint(static_cast<hipStream_t>(stream));
}

int hm_level_write_adjoint(hm_level* level, int patch, const double* lam) {
    if (!level) return HM_ERR_INVALID_ARGUMENT;
    return level->lvl->writeAdjoint(patch, lam);
}

int hm_level_read_adjoint(hm_level* level, int patch, double* lam) {
    if (!level) return HM_ERR_INVALID_ARGUMENT;
    return level->lvl->readAdjoint(patch, lam);
}

int hm_advance_single_step_on_patches(hm_ctx* ctx, const hm_rk_patch_view* patches, int num_patches,
                                      int num_coeffs, const double* alpha, const double* beta, const double* gamma,
                                      void* stream) {
    if (!ctx || !ctx->rec) return HM_ERR_INVALID_ARGUMENT;
    return hm::advance_single_step(ctx->rec.get(), patches, num_patches, num_coeffs, alpha, beta, gamma,
                                   static_cast<hipStream_t>(stream));
}

int hm_advance_single_step_on_patch(hm_ctx* ctx, const hm_rk_patch_view* patch, int num_coeffs,
                                    const double* alpha, const double* beta, const double* gamma, void* stream) {
    return hm_advance_single_step_on_patches(ctx, patch, 1, num_coeffs, alpha, beta, gamma, stream);
}

}  // extern "C"
