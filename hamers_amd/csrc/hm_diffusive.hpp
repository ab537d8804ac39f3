// hm_diffusive.hpp -- host layer of the GPU diffusive flux reconstructor.
//
// DiffusiveFluxReconstructorGPU mirrors the reference's
// DiffusiveFluxReconstructor interface (DiffusiveFluxReconstructor.hpp:73-79)
// for the node-based SIXTH_ORDER scheme (DiffusiveFluxReconstructorNode.cpp:32-1738,
// DiffusiveFluxReconstructorNodeSixthOrder.cpp) of the four-equation
// conservative model, with the constant transport mixing rules.  It shares the
// convective reconstructor's configuration (species gamma, R, device) and
// reports failures through it.
//
// Pipeline (per call, all patches at once on the batched path):
//   k_diff_cells  cells of the 6-ghost box: rho, Y_i, u, T, h_i, mu, mu_v, kappa
//   k_diff_nodes  node fluxes F_node[e] = sum_b sum_terms D * d(var)/dx_b on the
//                 nodes [-3, n_a + 3) along each axis a
//   k_diff_faces  face fluxes dt (a_r (F_L + F_R) + b_r (F_LL + F_RR) + c_r (F_LLL + F_RRR))
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "hm_reconstructor.hpp"
#include "hm_tables.hpp"

namespace hm {

constexpr int kGd = 6;  // diffusive ghost width (DiffusiveFluxReconstructorNodeSixthOrder.cpp:22)
constexpr int kDiffStages = 3;

struct DiffPatch {
    int n[3];                     // interior extents (1 on unused axes)
    int gq[3];                    // ghost width of Q (0 on unused axes)
    double dxinv[3];              // 1 / dx (DiffusiveFluxReconstructorNode.cpp:1773)
    const double* Q[HM_MAX_EQN];  // conservative variables over the Q ghost box
    double* F[3];                 // output diffusive flux per axis (component-major)
    double* cell;                 // scratch: derived cell data over the 6-ghost box
    double* fnode[3];             // scratch: node fluxes per axis, neq comps over (n_a+6) x n_b
};

struct DParams {
    double gamma[HM_MAX_SPECIES];
    double R[HM_MAX_SPECIES];
    double cp[HM_MAX_SPECIES];
    double cv[HM_MAX_SPECIES];
    double mu[HM_MAX_SPECIES];
    double mu_v[HM_MAX_SPECIES];
    double kappa[HM_MAX_SPECIES];
    double D[HM_MAX_SPECIES];
    double wfac[HM_MAX_SPECIES];  // 1 / sqrt(M_i), the mixing weights' factors
    int num_species;
    double dt;
    const double* dtp;  // non-NULL: dt in device memory (KParams::dtp)
};

template <class W>
struct DPlan {
    W work[kDiffStages];
    int64_t items[kDiffStages];
};

template <class W>
hipError_t launch_diffusive_four(int dim, const DPlan<W>& plan, const DParams& k, hipStream_t s, int first,
                                 int last);

class DiffusiveFluxReconstructorGPU {
   public:
    DiffusiveFluxReconstructorGPU(ConvectiveFluxReconstructorGPU& rec, const hm_transport& tr) : rec_(rec) {
        const hm_config& c = rec.config();
        const KParams& kp = rec.kernelParams();
        for (int i = 0; i < HM_MAX_SPECIES; i++) {
            k_.gamma[i] = kp.gamma[i];
            k_.R[i] = kp.R[i];
            k_.cp[i] = kp.cp[i];
            k_.cv[i] = kp.cv[i];
            const bool used = i < c.num_species;
            k_.mu[i] = used ? tr.species_mu[i] : 0.0;
            k_.mu_v[i] = used ? tr.species_mu_v[i] : 0.0;
            k_.kappa[i] = used ? tr.species_kappa[i] : 0.0;
            k_.D[i] = used ? tr.species_D[i] : 0.0;
            // EquationOfShearViscosityMixingRulesConstant.cpp:497
            k_.wfac[i] = used ? double(1) / (std::sqrt(tr.species_M[i])) : 0.0;
        }
        k_.num_species = c.num_species;
        k_.dt = 0.0;
        k_.dtp = nullptr;
        neq_ = rec.numEqn();
    }

    ~DiffusiveFluxReconstructorGPU() {
        (void)rec_.selectDevice();
        (void)hipDeviceSynchronize();
        for (auto& p : plans_)
            if (p.dev) (void)hipFree(p.dev);
        if (arena_) (void)hipFree(arena_);
    }

    int getDiffusiveFluxNumberOfGhostCells() const { return kGd; }

    int computeDiffusiveFluxOnPatch(const hm_patch_view& v, double time, double dt, int rk_step, hipStream_t s) {
        (void)time;
        (void)rk_step;
        if (int st = check_view(v, 0)) return st;
        if (int st = rec_.selectDevice()) return st;
        const Layout L = layout(v);
        if (order_.enter(s) != hipSuccess) return rec_.fail(HM_ERR_DEVICE, "hipStreamWaitEvent failed");
        if (int st = ensure_arena(L.total, s)) return st;
        const DiffPatch p = make_patch(v, arena_, L);
        DPlan<SegSingleT<DiffPatch>> plan{};
        const int d = rec_.config().dim;
        const Ext e = extents(v);
        for (int stage = 0; stage < kDiffStages; stage++) {
            SegSingleT<DiffPatch>& w = plan.work[stage];
            w.p = p;
            int64_t acc = 0;
            const int nseg = (stage == 0) ? 1 : d;
            for (int a = 0; a < nseg; a++) {
                w.start[a] = acc;
                w.cnt[a] = (stage == 0) ? e.cells : (stage == 1) ? e.nodes[a] : e.faces[a];
                acc += (w.cnt[a] + kSegAlign - 1) / kSegAlign * kSegAlign;
            }
            w.nseg = nseg;
            w.start[nseg] = acc;
            plan.items[stage] = acc;
        }
        DParams k = k_;
        k.dt = dt;
        return launch_ordered(plan, k, s);
    }

    int computeDiffusiveFluxOnPatches(const hm_patch_view* views, int np, double time, double dt, int rk_step,
                                      hipStream_t s) {
        (void)time;
        (void)rk_step;
        if (np < 0 || (np > 0 && !views)) return rec_.fail(HM_ERR_INVALID_ARGUMENT, "invalid patch list");
        if (np == 0) return HM_OK;
        for (int i = 0; i < np; i++)
            if (int st = check_view(views[i], i)) return st;
        if (int st = rec_.selectDevice()) return st;
        size_t total = 0;
        for (int i = 0; i < np; i++) total += layout(views[i]).total;
        if (order_.enter(s) != hipSuccess) return rec_.fail(HM_ERR_DEVICE, "hipStreamWaitEvent failed");
        if (int st = ensure_arena(total, s)) return st;
        CachedPlan* cp = nullptr;
        if (int st = get_plan(views, np, s, cp)) return st;
        DParams k = k_;
        k.dt = dt;
        return launch_ordered(cp->plan, k, s);
    }

    // Work tables of a fixed patch list over a caller-owned scratch region
    // (scratchDoubles), outside the plan cache: a captured graph can hold their
    // addresses for as long as the owner keeps them.
    struct OwnedPlan {
        void* dev = nullptr;
        DPlan<SegTableT<DiffPatch>> plan{};
    };

    size_t scratchDoubles(const hm_patch_view* views, int np) const {
        size_t total = 0;
        for (int i = 0; i < np; i++) total += layout(views[i]).total;
        return total;
    }

    int buildOwnedPlan(const hm_patch_view* views, int np, double* scratch, OwnedPlan& out) {
        for (int i = 0; i < np; i++)
            if (int st = check_view(views[i], i)) return st;
        if (int st = rec_.selectDevice()) return st;
        releaseOwnedPlan(out);
        return upload(views, np, scratch, out.dev, out.plan);
    }

    void releaseOwnedPlan(OwnedPlan& p) {
        if (p.dev) (void)hipFree(p.dev);
        p.dev = nullptr;
    }

    // No allocation, no events: graph-capture safe.  dtp: dt in device memory, or NULL.
    int runOwnedPlan(const OwnedPlan& p, double dt, hipStream_t s, const double* dtp = nullptr) {
        DParams k = k_;
        k.dt = dt;
        k.dtp = dtp;
        return launch(p.plan, k, s);
    }

   private:
    // launch() on the arena, then mark the arena in use by stream s.
    template <class W>
    int launch_ordered(const DPlan<W>& plan, const DParams& k, hipStream_t s) {
        if (int st = launch(plan, k, s)) return st;
        if (order_.leave(s) != hipSuccess) return rec_.fail(HM_ERR_DEVICE, "hipEventRecord failed");
        return HM_OK;
    }

    struct Ext {
        int64_t cells, nodes[3], faces[3];
    };
    struct Layout {
        size_t cell, fnode[3], total;
    };
    struct CachedPlan {
        std::vector<hm_patch_view> views;
        void* arena_at_build = nullptr;
        void* dev = nullptr;
        DPlan<SegTableT<DiffPatch>> plan{};
        uint64_t use = 0;
    };

    // Derived components per cell: rho, Y_i, u (dim), T, h_i, mu, mu_v, kappa.
    int cell_comps() const { return 2 * k_.num_species + rec_.config().dim + 5; }

    Ext extents(const hm_patch_view& v) const {
        const int d = rec_.config().dim;
        Ext e{};
        e.cells = 1;
        for (int a = 0; a < d; a++) e.cells *= v.n[a] + 2 * kGd;
        for (int a = 0; a < 3; a++) {
            e.nodes[a] = e.faces[a] = 0;
            if (a >= d) continue;
            int64_t m = 1, f = 1;
            for (int b = 0; b < d; b++) {
                m *= (b == a) ? v.n[b] + 6 : v.n[b];
                f *= (b == a) ? v.n[b] + 1 : v.n[b];
            }
            e.nodes[a] = m;
            e.faces[a] = f;
        }
        return e;
    }

    Layout layout(const hm_patch_view& v) const {
        const Ext e = extents(v);
        Layout L{};
        L.cell = align_up(size_t(cell_comps()) * size_t(e.cells), kAlignDoubles);
        L.total = L.cell;
        for (int a = 0; a < 3; a++) {
            L.fnode[a] = align_up(size_t(neq_) * size_t(e.nodes[a]), kAlignDoubles);
            L.total += L.fnode[a];
        }
        return L;
    }

    DiffPatch make_patch(const hm_patch_view& v, double* base, const Layout& L) const {
        DiffPatch p{};
        const int d = rec_.config().dim;
        for (int a = 0; a < 3; a++) {
            p.n[a] = (a < d) ? v.n[a] : 1;
            p.gq[a] = (a < d) ? v.ghost[a] : 0;
            p.dxinv[a] = (a < d) ? double(1) / v.dx[a] : 1.0;
            p.F[a] = (a < d) ? v.F[a] : nullptr;
        }
        for (int e = 0; e < HM_MAX_EQN; e++) p.Q[e] = (e < neq_) ? v.Q[e] : nullptr;
        double* q = base;
        p.cell = q;
        q += L.cell;
        for (int a = 0; a < 3; a++) {
            p.fnode[a] = L.fnode[a] ? q : nullptr;
            q += L.fnode[a];
        }
        return p;
    }

    int check_view(const hm_patch_view& v, int which) {
        const int d = rec_.config().dim;
        char buf[160];
        for (int a = 0; a < d; a++) {
            if (v.n[a] < 1) {
                snprintf(buf, sizeof buf, "patch %d: interior extent n[%d]=%d < 1", which, a, v.n[a]);
                return rec_.fail(HM_ERR_INVALID_ARGUMENT, buf);
            }
            if (v.ghost[a] < kGd) {
                snprintf(buf, sizeof buf,
                         "patch %d: ghost width %d on axis %d is below the 6 cells the sixth-order diffusive flux needs",
                         which, v.ghost[a], a);
                return rec_.fail(HM_ERR_INVALID_ARGUMENT, buf);
            }
            if (!(v.dx[a] > 0.0)) {
                snprintf(buf, sizeof buf, "patch %d: dx[%d] must be positive", which, a);
                return rec_.fail(HM_ERR_INVALID_ARGUMENT, buf);
            }
            if (!v.F[a]) {
                snprintf(buf, sizeof buf, "patch %d: diffusive flux array for axis %d is NULL", which, a);
                return rec_.fail(HM_ERR_INVALID_ARGUMENT, buf);
            }
        }
        // Kernels number the items of a patch's work segment in 32 bits.
        double box = 1.0;
        for (int a = 0; a < d; a++) box *= double(v.n[a]) + 2.0 * double(v.ghost[a]);
        if (box >= 2147483648.0) {
            snprintf(buf, sizeof buf, "patch %d: ghost box of 2^31 cells or more; split the patch", which);
            return rec_.fail(HM_ERR_INVALID_ARGUMENT, buf);
        }
        for (int e = 0; e < neq_; e++)
            if (!v.Q[e]) {
                snprintf(buf, sizeof buf, "patch %d: conservative variable %d is NULL", which, e);
                return rec_.fail(HM_ERR_INVALID_ARGUMENT, buf);
            }
        return HM_OK;
    }

    int ensure_arena(size_t doubles, hipStream_t s) {
        if (doubles <= arena_doubles_) return HM_OK;
        const size_t want = std::max(doubles, arena_doubles_ + arena_doubles_ / 2);
        if (arena_) {
            if (hipStreamSynchronize(s) != hipSuccess) return rec_.fail(HM_ERR_DEVICE, "stream synchronize failed");
            (void)hipFree(arena_);
            arena_ = nullptr;
            arena_doubles_ = 0;
        }
        for (auto& p : plans_)
            if (p.dev) (void)hipFree(p.dev);
        plans_.clear();
        if (hipMalloc(reinterpret_cast<void**>(&arena_), want * sizeof(double)) != hipSuccess) {
            arena_ = nullptr;
            return rec_.fail(HM_ERR_OUT_OF_MEMORY, "hipMalloc of the diffusive scratch arena failed");
        }
        arena_doubles_ = want;
        return HM_OK;
    }

    // Work tables of a patch list, cached per distinct list (the three RK
    // stages of a level reuse theirs).
    int get_plan(const hm_patch_view* views, int np, hipStream_t s, CachedPlan*& out) {
        use_clock_++;
        for (auto& p : plans_)
            if (int(p.views.size()) == np && p.arena_at_build == arena_ &&
                std::memcmp(p.views.data(), views, sizeof(hm_patch_view) * size_t(np)) == 0) {
                p.use = use_clock_;
                out = &p;
                return HM_OK;
            }
        CachedPlan* p = nullptr;
        if (plans_.size() < kMaxPlans) {
            plans_.emplace_back();
            p = &plans_.back();
        } else {
            p = &plans_[0];
            for (auto& c : plans_)
                if (c.use < p->use) p = &c;
            if (hipStreamSynchronize(s) != hipSuccess) return rec_.fail(HM_ERR_DEVICE, "stream synchronize failed");
            if (p->dev) (void)hipFree(p->dev);
            p->dev = nullptr;
        }
        p->views.clear();
        if (int st = upload(views, np, arena_, p->dev, p->plan)) return st;
        p->views.assign(views, views + np);
        p->arena_at_build = arena_;
        p->use = use_clock_;
        out = p;
        return HM_OK;
    }

    // Device work tables of a patch list whose scratch starts at `base`.
    int upload(const hm_patch_view* views, int np, double* base, void*& dev, DPlan<SegTableT<DiffPatch>>& plan) {
        const int d = rec_.config().dim;
        TableImage<DiffPatch> img[kDiffStages];
        for (auto& t : img) t.items.resize(np);
        for (int i = 0; i < np; i++) {
            const Layout L = layout(views[i]);
            const DiffPatch dp = make_patch(views[i], base, L);
            base += L.total;
            for (auto& t : img) t.items[i] = dp;
            const Ext e = extents(views[i]);
            img[0].add(i, 0, e.cells);
            for (int a = 0; a < d; a++) img[1].add(i, a, e.nodes[a]);
            for (int a = 0; a < d; a++) img[2].add(i, a, e.faces[a]);
        }
        size_t bytes = 0;
        for (auto& t : img) bytes += t.bytes();
        if (hipMalloc(&dev, bytes) != hipSuccess) {
            dev = nullptr;
            return rec_.fail(HM_ERR_OUT_OF_MEMORY, "hipMalloc of the diffusive work tables failed");
        }
        std::vector<char> h(bytes, 0);
        size_t o = 0;
        for (int st = 0; st < kDiffStages; st++) {
            plan.work[st] = img[st].write(h.data() + o, static_cast<char*>(dev) + o);
            plan.items[st] = img[st].total;
            o += img[st].bytes();
        }
        // built once per distinct patch list: a blocking copy is fine here
        if (hipMemcpy(dev, h.data(), bytes, hipMemcpyHostToDevice) != hipSuccess)
            return rec_.fail(HM_ERR_DEVICE, "upload of the diffusive work tables failed");
        return HM_OK;
    }

    template <class W>
    int launch(const DPlan<W>& plan, const DParams& k, hipStream_t s) {
        const hipError_t e = launch_diffusive_four<W>(rec_.config().dim, plan, k, s, 0, kDiffStages);
        if (e != hipSuccess)
            return rec_.fail(HM_ERR_DEVICE, std::string("diffusive kernel launch failed: ") + hipGetErrorString(e));
        return HM_OK;
    }

    static constexpr size_t kMaxPlans = 4;
    ConvectiveFluxReconstructorGPU& rec_;
    DParams k_{};
    int neq_ = 0;
    double* arena_ = nullptr;
    size_t arena_doubles_ = 0;
    StreamOrder order_;  // calls on different streams take the arena in turn
    std::vector<CachedPlan> plans_;
    uint64_t use_clock_ = 0;
};

}  // namespace hm
