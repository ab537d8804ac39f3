// hm_reconstructor.hpp -- host layer: the GPU convective flux reconstructor.
//
// ConvectiveFluxReconstructorGPU mirrors the reference's
// ConvectiveFluxReconstructor interface (ConvectiveFluxReconstructor.hpp:23-124)
// for the WCNS56 hybrid HLLC / HLLC-HLL schemes, and make_reconstructor
// mirrors ConvectiveFluxReconstructorManager (ConvectiveFluxReconstructorManager.cpp:14-153).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "hm_kernels.hpp"

namespace hm {

namespace {

constexpr size_t kAlignDoubles = 32;  // 256-byte alignment of every scratch array

size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }
int64_t align_items(int64_t v) { return (v + kSegAlign - 1) / kSegAlign * kSegAlign; }

uint64_t fnv1a(const void* data, size_t n, uint64_t h = 1469598103934665603ull) {
    const unsigned char* p = static_cast<const unsigned char*>(data);
    for (size_t i = 0; i < n; i++) {
        h ^= p[i];
        h *= 1099511628211ull;
    }
    return h;
}

}  // namespace

inline int num_equations(int model, int dim, int ns) {
    switch (model) {
        case HM_MODEL_SINGLE_SPECIES: return dim + 2;                 // FlowModelSingleSpecies
        case HM_MODEL_FOUR_EQN_CONSERVATIVE: return ns + dim + 1;     // FlowModelFourEqnConservative
        case HM_MODEL_FIVE_EQN_ALLAIRE: return dim + 2 * ns;          // FlowModelFiveEqnAllaire.cpp:30
        default: return -1;
    }
}

// Orders the calls that share one scratch region of a context across streams.
// Each call records an event on its stream after its launches; a call that
// arrives on another stream first makes that stream wait for the event, so
// two streams never write the shared scratch at the same time.  Calls on one
// stream are ordered by the stream itself.
class StreamOrder {
   public:
    StreamOrder() = default;
    StreamOrder(const StreamOrder&) = delete;
    StreamOrder& operator=(const StreamOrder&) = delete;
    ~StreamOrder() {
        if (ev_) (void)hipEventDestroy(ev_);
    }
    // Before the call's first launch on s.
    hipError_t enter(hipStream_t s) {
        if (recorded_ && s != last_) return hipStreamWaitEvent(s, ev_, 0);
        return hipSuccess;
    }
    // After the call's last launch on s.
    hipError_t leave(hipStream_t s) {
        if (!ev_) {
            const hipError_t e = hipEventCreateWithFlags(&ev_, hipEventDisableTiming);
            if (e != hipSuccess) {
                ev_ = nullptr;
                return e;
            }
        }
        const hipError_t e = hipEventRecord(ev_, s);
        if (e != hipSuccess) return e;
        last_ = s;
        recorded_ = true;
        return hipSuccess;
    }

   private:
    hipEvent_t ev_ = nullptr;
    hipStream_t last_ = nullptr;
    bool recorded_ = false;
};

// Sizes (in doubles) of one patch's scratch arrays.
struct ScratchLayout {
    size_t prim, sens, fmid[3], ustar[3], total;
};

class ConvectiveFluxReconstructorGPU {
   public:
    explicit ConvectiveFluxReconstructorGPU(const hm_config& cfg) : cfg_(cfg) {
        neq_ = num_equations(cfg.model, cfg.dim, cfg.num_species);
        k_.p = cfg.constant_p;
        k_.q = cfg.constant_q;
        k_.C = cfg.constant_C;
        k_.alpha_tau = cfg.constant_alpha_tau;
        for (int i = 0; i < HM_MAX_SPECIES; i++) {
            k_.gamma[i] = cfg.species_gamma[i];
            k_.R[i] = cfg.species_R[i];
            // EquationOfStateMixingRulesIdealGas.cpp:108-120
            const bool used = i < cfg.num_species;
            k_.cp[i] = used ? cfg.species_gamma[i] / (cfg.species_gamma[i] - double(1)) * cfg.species_R[i] : 0.0;
            k_.cv[i] = used ? double(1) / (cfg.species_gamma[i] - double(1)) * cfg.species_R[i] : 0.0;
            // EquationOfStateMixingRulesIdealGas.cpp:7404-7800 divides per cell; the
            // quotient is the same correctly rounded number computed once here.
            k_.igm1[i] = used ? double(1) / (cfg.species_gamma[i] - double(1)) : 0.0;
        }
        k_.num_species = cfg.num_species;
        k_.dt = 0.0;
        k_.dtp = nullptr;
        k_.wrec = nullptr;
        k_.lam_src = -1;
        k_.lam_scale = 0.0;
        k_.qoff = 0;
        k_.xcd = 0;
        k_.queue = nullptr;
        has_advective_ = (cfg.model == HM_MODEL_FIVE_EQN_ALLAIRE && cfg.num_species > 1);
        fused_ = (cfg.model == HM_MODEL_SINGLE_SPECIES);
        plans_.reserve(kMaxPlans);  // plan addresses stay valid while cached
    }

    ~ConvectiveFluxReconstructorGPU() {
        int prev = -1;
        (void)hipGetDevice(&prev);
        (void)hipSetDevice(cfg_.device);
        (void)hipDeviceSynchronize();
        for (auto& pl : plans_) free_plan(pl);
        for (auto e : prof_events_) (void)hipEventDestroy(e);
        if (arena_) (void)hipFree(arena_);
        if (stage_buf_) (void)hipFree(stage_buf_);
        if (stage_ev_) (void)hipEventDestroy(stage_ev_);
        if (queue_) (void)hipFree(queue_);
        if (prev >= 0) (void)hipSetDevice(prev);
    }

    int numEqn() const { return neq_; }

    // Conservative arrays a patch view must provide: the five-equation model
    // also hands over the last species' volume fraction (stored, not evolved).
    int numConservativeArrays() const { return cfg_.model == HM_MODEL_FIVE_EQN_ALLAIRE ? neq_ + 1 : neq_; }

    // Scratch components per cell of the 4-ghost box: V, c (and rho for the
    // multi-species models).
    int prim_comps() const { return cfg_.model == HM_MODEL_SINGLE_SPECIES ? neq_ + 1 : neq_ + 2; }

    // ConvectiveFluxReconstructor::getConvectiveFluxNumberOfGhostCells (4 for WCNS56).
    int getConvectiveFluxNumberOfGhostCells() const { return kGc; }

    int computeConvectiveFluxAndSourceOnPatch(const hm_patch_view& v, double time, double dt, int rk_step,
                                              hipStream_t s) {
        (void)time;
        (void)rk_step;
        if (int st = check_view(v, 0)) return st;
        if (int st = select_device()) return st;
        const ScratchLayout L = layout(v);
        // (wait first: growing the arena synchronizes only this stream)
        if (arena_order_.enter(s) != hipSuccess) return fail(HM_ERR_DEVICE, "hipStreamWaitEvent failed");
        if (int st = ensure_arena(L.total, s)) return st;
        const Plan<SegSingle> plan = patch_plan(v, arena_, L);
        KParams k = k_;
        k.dt = dt;
        return run_ordered(plan, k, s);
    }

    // The per-patch RHS with its scratch in a caller-given device region of
    // `cap` doubles starting on a 256-byte boundary (the allocation-free
    // entry point's static array): no arena, no tables, no profiling events,
    // no allocation.
    int computeOnPatchInScratch(const hm_patch_view& v, double dt, hipStream_t s, double* scratch, size_t cap) {
        if (int st = check_view(v, 0)) return st;
        if (int st = select_device()) return st;
        const ScratchLayout L = layout(v);
        if (L.total > cap) {
            char buf[200];
            snprintf(buf, sizeof buf,
                     "patch needs %zu doubles of scratch; the allocation-free entry point holds %zu "
                     "(use hm_compute_convective_flux_and_source_on_patch)",
                     L.total, cap);
            return fail(HM_ERR_INVALID_ARGUMENT, buf);
        }
        const Plan<SegSingle> plan = patch_plan(v, scratch, L);
        KParams k = k_;
        k.dt = dt;
        return run_stages(plan, k, s, 0, kStages);
    }

    int device() const { return cfg_.device; }

    int computeConvectiveFluxAndSourceOnPatches(const hm_patch_view* views, int np, double time, double dt,
                                                int rk_step, hipStream_t s) {
        (void)time;
        (void)rk_step;
        if (np < 0 || (np > 0 && !views)) return fail(HM_ERR_INVALID_ARGUMENT, "invalid patch list");
        if (np == 0) return HM_OK;
        for (int i = 0; i < np; i++)
            if (int st = check_view(views[i], i)) return st;
        if (int st = select_device()) return st;
        size_t total = 0;
        for (int i = 0; i < np; i++) total += layout(views[i]).total;
        if (arena_order_.enter(s) != hipSuccess) return fail(HM_ERR_DEVICE, "hipStreamWaitEvent failed");
        if (int st = ensure_arena(total, s)) return st;
        TablePlan* tp = nullptr;
        if (int st = get_table_plan(views, np, s, tp)) return st;
        KParams k = k_;
        k.dt = dt;
        return run_ordered(tp->plan, k, s);
    }

    // Work tables of a fixed patch list over a caller-owned scratch region,
    // outside the plan cache: a captured graph can hold their addresses for as
    // long as the owner keeps them.
    struct OwnedPlan {
        void* dev = nullptr;
        Plan<SegTable> plan{};
    };

    size_t scratchDoubles(const hm_patch_view* views, int np) const {
        size_t total = 0;
        for (int i = 0; i < np; i++) total += layout(views[i]).total;
        return total;
    }

    // wrec_off (optional, 3 per patch): the patches' chunks in a face-state
    // slot, for runs that record the face states (runOwnedPlan's wrec).
    int buildOwnedPlan(const hm_patch_view* views, int np, double* scratch, OwnedPlan& out,
                       const int64_t* wrec_off = nullptr, const Schedule* sched = nullptr) {
        for (int i = 0; i < np; i++)
            if (int st = check_view(views[i], i)) return st;
        if (int st = select_device()) return st;
        releaseOwnedPlan(out);
        return upload_tables(views, np, scratch, out.dev, out.plan, wrec_off, sched);
    }

    // The built-in schedule of a patch list: tile rows by the work the call
    // holds (tile_rows_for), XCD block numbering for the single-species
    // kernels (measured faster there) and hardware order for the multi-species
    // ones, segments in patch order.
    Schedule defaultSchedule(const hm_patch_view* views, int np) const {
        Schedule sc;
        sc.tile_rows = tile_rows_of(views, np);
        sc.xcd = fused_ ? 1 : 0;
        sc.order = 0;
        return sc;
    }

    // Candidate schedules of the autotuner for this model.
    std::vector<Schedule> scheduleCandidates() const {
        std::vector<Schedule> c;
        const int rows[3] = {4, 8, 24};
        for (int r = 0; r < (fused_ && cfg_.dim > 1 ? 3 : 1); r++)
            for (int x = 0; x < 2; x++)
                for (int o = 0; o < 2; o++)
                    for (int p = 0; p < (fused_ ? 1 : 2); p++) {
                        Schedule sc;
                        sc.tile_rows = (fused_ && cfg_.dim > 1) ? rows[r] : 4;
                        sc.xcd = x;
                        sc.order = o;
                        sc.persistent = p;
                        c.push_back(sc);
                    }
        return c;
    }

    // Whether sc is a schedule of this model ("" if so, else the reason).
    const char* scheduleError(const Schedule& sc) const {
        if (sc.tile_rows != 4 && sc.tile_rows != 8 && sc.tile_rows != 24) return "tile rows must be 4, 8 or 24";
        if ((sc.xcd != 0 && sc.xcd != 1) || (sc.order != 0 && sc.order != 1) ||
            (sc.persistent != 0 && sc.persistent != 1))
            return "xcd, order and persistent are 0 or 1";
        if (fused_ && sc.persistent) return "the persistent midpoint flux is a schedule of the multi-species models";
        return "";
    }

    // Times `run` (one batched RHS of a plan) for each candidate schedule,
    // built by `build` (plan for a schedule), repetitions times after one
    // warm-up, and returns the fastest (median of the repetitions).
    template <class Build, class Run>
    int tuneSchedules(hipStream_t s, int reps, Build&& build, Run&& run, Schedule& best, double& best_ms,
                      double& default_ms, const Schedule& dflt) {
        if (reps < 1) return fail(HM_ERR_INVALID_ARGUMENT, "autotune needs at least one repetition");
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
            return fail(HM_ERR_DEVICE, "hipEventCreate failed");
        best_ms = -1.0;
        default_ms = -1.0;
        int st = HM_OK;
        for (const Schedule& sc : scheduleCandidates()) {
            if ((st = build(sc))) break;
            if ((st = run())) break;  // warm-up
            std::vector<float> t(size_t(reps), 0.f);
            for (int r = 0; r < reps && st == HM_OK; r++) {
                if (hipEventRecord(e0, s) != hipSuccess) st = fail(HM_ERR_DEVICE, "hipEventRecord failed");
                if (st == HM_OK) st = run();
                if (st == HM_OK && (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                                    hipEventElapsedTime(&t[size_t(r)], e0, e1) != hipSuccess))
                    st = fail(HM_ERR_DEVICE, "event timing failed");
            }
            if (st) break;
            std::sort(t.begin(), t.end());
            const double ms = t[t.size() / 2];
            if (sc.tile_rows == dflt.tile_rows && sc.xcd == dflt.xcd && sc.order == dflt.order &&
                sc.persistent == dflt.persistent)
                default_ms = ms;
            if (best_ms < 0.0 || ms < best_ms) {
                best_ms = ms;
                best = sc;
            }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        return st;
    }

    // Autotuner of the batched RHS of a patch list: times every candidate
    // schedule on these views (their F is overwritten and their S accumulated
    // by every trial: call it on scratch outputs or before the real call) and
    // keeps the fastest for the list; later batched calls on the same list use
    // it.
    int autotuneOnPatches(const hm_patch_view* views, int np, double dt, int reps, hipStream_t s, Schedule& best,
                          double& best_ms, double& default_ms) {
        if (np < 1 || !views) return fail(HM_ERR_INVALID_ARGUMENT, "invalid patch list");
        for (int i = 0; i < np; i++)
            if (int st = check_view(views[i], i)) return st;
        if (int st = select_device()) return st;
        size_t total = 0;
        for (int i = 0; i < np; i++) total += layout(views[i]).total;
        if (arena_order_.enter(s) != hipSuccess) return fail(HM_ERR_DEVICE, "hipStreamWaitEvent failed");
        if (int st = ensure_arena(total, s)) return st;
        KParams k = k_;
        k.dt = dt;
        OwnedPlan trial;
        const Schedule dflt = defaultSchedule(views, np);
        int st = tuneSchedules(
            s, reps,
            [&](const Schedule& sc) {
                if (hipStreamSynchronize(s) != hipSuccess) return fail(HM_ERR_DEVICE, "stream synchronize failed");
                releaseOwnedPlan(trial);
                return upload_tables(views, np, arena_, trial.dev, trial.plan, nullptr, &sc);
            },
            [&]() { return run_stages(trial.plan, k, s, 0, kStages); }, best, best_ms, default_ms, dflt);
        if (hipStreamSynchronize(s) != hipSuccess && st == HM_OK) st = fail(HM_ERR_DEVICE, "stream synchronize failed");
        releaseOwnedPlan(trial);
        if (st) return st;
        (void)arena_order_.leave(s);
        return setSchedule(views, np, best);
    }

    // The persistent midpoint flux's queue counter, allocated with the first
    // plan that uses it (calls on a context are ordered, so one serves all).
    int ensure_queue() {
        if (queue_) return HM_OK;
        if (hipMalloc(reinterpret_cast<void**>(&queue_), 256) != hipSuccess) {
            queue_ = nullptr;
            return fail(HM_ERR_OUT_OF_MEMORY, "hipMalloc of the work-queue counter failed");
        }
        return HM_OK;
    }

    // Pins the schedule of a patch list's batched calls (the autotuner's
    // choice, or a caller's).
    int setSchedule(const hm_patch_view* views, int np, const Schedule& sc) {
        if (np < 1 || !views) return fail(HM_ERR_INVALID_ARGUMENT, "invalid patch list");
        if (*scheduleError(sc)) return fail(HM_ERR_INVALID_ARGUMENT, scheduleError(sc));
        const uint64_t key = fnv1a(views, sizeof(hm_patch_view) * size_t(np), fnv1a(&np, sizeof np));
        bool found = false;
        for (auto& t : tuned_)
            if (t.key == key && t.views.size() == size_t(np) &&
                std::memcmp(t.views.data(), views, sizeof(hm_patch_view) * size_t(np)) == 0) {
                t.sched = sc;
                found = true;
            }
        if (!found) {
            if (tuned_.size() >= kMaxPlans) tuned_.erase(tuned_.begin());
            tuned_.push_back(TunedSchedule{key, std::vector<hm_patch_view>(views, views + np), sc});
        }
        // cached tables of this list were built with the previous schedule
        if (hipDeviceSynchronize() != hipSuccess) return fail(HM_ERR_DEVICE, "device synchronize failed");
        for (auto& pl : plans_)
            if (pl.key == key) {
                free_plan(pl);
                pl.key = 0;
            }
        return HM_OK;
    }

    // The schedule a batched call on this list uses.
    Schedule scheduleOf(const hm_patch_view* views, int np) const {
        const uint64_t key = fnv1a(views, sizeof(hm_patch_view) * size_t(np), fnv1a(&np, sizeof np));
        for (const auto& t : tuned_)
            if (t.key == key && t.views.size() == size_t(np) &&
                std::memcmp(t.views.data(), views, sizeof(hm_patch_view) * size_t(np)) == 0)
                return t.sched;
        return defaultSchedule(views, np);
    }

    void releaseOwnedPlan(OwnedPlan& p) {
        if (p.dev) (void)hipFree(p.dev);
        p.dev = nullptr;
    }

    // All five stages of an owned plan; no profiling events (graph capture
    // safe).  wrec: face-state slot to record into (single-species model,
    // plans built with wrec_off), NULL for none.
    int runOwnedPlan(const OwnedPlan& p, double dt, hipStream_t s, double* wrec = nullptr,
                     const double* dtp = nullptr) {
        return runOwnedPlanStages(p, dt, 0, kStages, s, wrec, 0, dtp);
    }

    // Pipeline stages [first, last) of an owned plan.
    int runOwnedPlanStages(const OwnedPlan& p, double dt, int first, int last, hipStream_t s, double* wrec = nullptr,
                           int64_t qoff = 0, const double* dtp = nullptr) {
        KParams k = k_;
        k.dt = dt;
        k.dtp = dtp;
        k.wrec = wrec;
        k.qoff = qoff;
        return run_stages(p.plan, k, s, first, last);
    }

    // Device descriptor of a patch whose scratch starts at `base` (the layout
    // of scratchDoubles / buildOwnedPlan, patch after patch).
    DevPatch devPatch(const hm_patch_view& v, double* base, const int64_t* wrec_off = nullptr) const {
        DevPatch p = make_dev_patch(v, base, layout(v));
        if (wrec_off)
            for (int a = 0; a < 3; a++) p.wrec[a] = wrec_off[a];
        return p;
    }
    size_t patchScratchDoubles(const hm_patch_view& v) const { return layout(v).total; }
    const KParams& kernelParams() const { return k_; }

    // Device buffer for per-call tables of the batched stage update.  The
    // previous user's launches are waited for before the buffer is handed out
    // again, so calls on different streams never overwrite tables in use.
    void* stageTables(size_t bytes) {
        if (stage_ev_ && hipEventSynchronize(stage_ev_) != hipSuccess) return nullptr;
        if (bytes > stage_bytes_) {
            if (stage_buf_) (void)hipFree(stage_buf_);
            stage_buf_ = nullptr;
            stage_bytes_ = 0;
            if (hipMalloc(&stage_buf_, bytes) != hipSuccess) {
                stage_buf_ = nullptr;
                return nullptr;
            }
            stage_bytes_ = bytes;
        }
        return stage_buf_;
    }

    int stageTablesInUse(hipStream_t s) {
        if (!stage_ev_ && hipEventCreateWithFlags(&stage_ev_, hipEventDisableTiming) != hipSuccess)
            return fail(HM_ERR_DEVICE, "hipEventCreate failed");
        if (hipEventRecord(stage_ev_, s) != hipSuccess) return fail(HM_ERR_DEVICE, "hipEventRecord failed");
        return HM_OK;
    }

    const std::string& lastError() const { return err_; }
    const hm_config& config() const { return cfg_; }
    int selectDevice() { return select_device(); }

    // Stage timing with HIP events recorded on the launch stream around every
    // pipeline stage (0 derived, 1 sensor, 2 midpoint flux, 3 face flux, 4 source).
    int setProfiling(int on) {
        if (int st = select_device()) return st;
        if (on && prof_events_.empty()) {
            for (int i = 0; i < kProfSlots; i++) {
                hipEvent_t e;
                if (hipEventCreate(&e) != hipSuccess) return fail(HM_ERR_DEVICE, "hipEventCreate failed");
                prof_events_.push_back(e);
            }
        }
        profiling_ = on != 0;
        prof_used_ = 0;
        for (int i = 0; i < kStages; i++) {
            prof_ms_[i] = 0.0;
            prof_n_[i] = 0;
        }
        return HM_OK;
    }

    // Fold recorded event pairs into per-stage totals (waits for them).
    int readProfile(double* ms, int64_t* launches) {
        if (int st = collect_profile()) return st;
        for (int i = 0; i < kStages; i++) {
            ms[i] = prof_ms_[i];
            launches[i] = prof_n_[i];
        }
        return HM_OK;
    }

    int fail(int code, const std::string& msg) {
        err_ = msg;
        return code;
    }

   private:
    // Work of one patch's RHS with its scratch at `base` (layout L).
    Plan<SegSingle> patch_plan(const hm_patch_view& v, double* base, const ScratchLayout& L) const {
        const DevPatch dp = make_dev_patch(v, base, L);
        Plan<SegSingle> plan{};
        const int d = cfg_.dim;
        plan.tile_rows = tile_rows_of(&v, 1);
        plan.xcd = fused_ ? 1 : 0;
        plan.persistent = 0;
        plan.queue = nullptr;
        const Extents e = extents(v, plan.tile_rows, true);
        // derived cells
        plan.work[0].p = dp;
        plan.work[0].nseg = 1;
        plan.work[0].start[0] = 0;
        plan.work[0].cnt[0] = e.vol4;
        plan.work[0].start[1] = align_items(e.vol4);
        plan.items[0] = plan.work[0].start[1];
        // sensor cells
        plan.work[1].p = dp;
        plan.work[1].nseg = 1;
        plan.work[1].start[0] = 0;
        plan.work[1].cnt[0] = 0;  // the derived-data kernels compute the sensor (stage 1 is empty)
        plan.work[1].start[1] = align_items(plan.work[1].cnt[0]);
        plan.items[1] = plan.work[1].start[1];
        // midpoint faces and output faces, one segment per axis
        for (int stage = 2; stage < 4; stage++) {
            SegSingle& w = plan.work[stage];
            w.p = dp;
            w.nseg = d;
            int64_t acc = 0;
            for (int a = 0; a < d; a++) {
                w.start[a] = acc;
                w.cnt[a] = (stage == 2) ? e.work2[a] : e.work3[a];
                acc += align_items(w.cnt[a]);
            }
            w.start[d] = acc;
            plan.items[stage] = acc;
        }
        if (fused_) {
            // The single-species model runs both flux stages in one launch on
            // one patch: stage 2 takes axis 0's chunks, padded to whole blocks,
            // and the tiles of axes 1 and 2; stage 3 is empty.
            SegSingle& w = plan.work[2];
            int64_t acc = align_up(size_t(e.work2[0]), kBlock);
            for (int a = 1; a < d; a++) {
                w.start[a] = acc;
                w.cnt[a] = e.work3[a];
                acc += e.work3[a];
            }
            w.start[d] = acc;
            plan.items[2] = acc;
            plan.items[3] = 0;
        }
        // advective source cells
        plan.work[4].p = dp;
        plan.work[4].nseg = 1;
        plan.work[4].start[0] = 0;
        plan.work[4].cnt[0] = has_advective_ ? e.vol0 : 0;
        plan.work[4].start[1] = align_items(plan.work[4].cnt[0]);
        plan.items[4] = plan.work[4].start[1];
        return plan;
    }

    struct Extents {
        int64_t vol4, vol2, vol0, mid[3], face[3];
        int64_t work2[3], work3[3];  // items of the midpoint-flux and face-flux stages per axis
    };

    struct TablePlan {
        uint64_t key = 0;
        std::vector<hm_patch_view> views;
        void* dev = nullptr;   // device copy of the tables
        Plan<SegTable> plan{};
        void* arena_at_build = nullptr;
        uint64_t use = 0;
    };

    // Tile rows of the single-species model's tiled flux stage for a call over
    // these patches (4 when the model does not tile).
    int tile_rows_of(const hm_patch_view* views, int np) const {
        if (!fused_ || cfg_.dim < 2) return 4;
        int64_t b24 = 0, b8 = 0;
        for (int i = 0; i < np; i++) {
            const int nn[3] = {views[i].n[0], cfg_.dim > 1 ? views[i].n[1] : 1, cfg_.dim > 2 ? views[i].n[2] : 1};
            for (int a = 1; a < cfg_.dim; a++) {
                b24 += tile_items(a, nn, cfg_.dim, 24) / 256;
                b8 += tile_items(a, nn, cfg_.dim, 8) / 256;
            }
        }
        return tile_rows_for(b24, b8);
    }

    // rows: the call's tile rows (the work-item counts of the tiled stage
    // depend on it; the scratch layout does not).  mtiles: a multi-species
    // call with the fused flux (every schedule but the persistent one).
    Extents extents(const hm_patch_view& v, int rows = 24, bool mtiles = false) const {
        const int d = cfg_.dim;
        Extents e{};
        int64_t n[3] = {1, 1, 1};
        for (int a = 0; a < d; a++) n[a] = v.n[a];
        auto grown = [&](int g) {
            int64_t r = 1;
            for (int a = 0; a < 3; a++) r *= (a < d) ? n[a] + 2 * g : 1;
            return r;
        };
        e.vol4 = grown(kGc);
        e.vol2 = grown(kGs);
        e.vol0 = grown(0);
        for (int a = 0; a < 3; a++) {
            e.mid[a] = 0;
            e.face[a] = 0;
            if (a >= d) continue;
            int64_t m = 1, f = 1;
            for (int b = 0; b < 3; b++) {
                m *= (b == a) ? n[b] + 3 : n[b];
                f *= (b == a) ? n[b] + 1 : n[b];
            }
            e.mid[a] = m;
            e.face[a] = f;
            // The single-species model fuses the midpoint and face flux: axis 0
            // in stage 2, axes 1 and 2 (tiles) in stage 3.
            const int nn[3] = {int(n[0]), int(n[1]), int(n[2])};
            if (mtiles && !fused_) {
                // multi-species: midpoint and face flux fused over tiles in stage 2
                e.work2[a] = mtile_items(a, nn, d);
                e.work3[a] = 0;
            } else if (!fused_) {
                e.work2[a] = m;
                e.work3[a] = f;
            } else if (a == 0) {
                e.work2[a] = fused_items(0, nn, d);
                e.work3[a] = 0;
            } else {
                e.work2[a] = 0;
                e.work3[a] = tile_items(a, nn, d, rows);
            }
        }
        return e;
    }

    ScratchLayout layout(const hm_patch_view& v) const {
        const Extents e = extents(v);
        ScratchLayout L{};
        L.prim = align_up(size_t(prim_comps()) * e.vol4, kAlignDoubles);
        L.sens = (cfg_.dim > 1) ? align_up(size_t(2) * e.vol2, kAlignDoubles) : 0;
        L.total = L.prim + L.sens;
        for (int a = 0; a < 3; a++) {
            // (the fused flux keeps its midpoint fluxes in registers or LDS)
            L.fmid[a] = fused_ ? 0 : align_up(size_t(neq_) * e.mid[a], kAlignDoubles);
            L.ustar[a] = has_advective_ ? align_up(size_t(e.mid[a]), kAlignDoubles) : 0;
            L.total += L.fmid[a] + L.ustar[a];
        }
        return L;
    }

    DevPatch make_dev_patch(const hm_patch_view& v, double* base, const ScratchLayout& L) const {
        DevPatch p{};
        const int d = cfg_.dim;
        for (int a = 0; a < 3; a++) {
            p.wrec[a] = -1;
            p.n[a] = (a < d) ? v.n[a] : 1;
            p.gq[a] = (a < d) ? v.ghost[a] : 0;
            p.dx[a] = (a < d) ? v.dx[a] : 1.0;
        }
        const int nq = numConservativeArrays();
        for (int e = 0; e < HM_MAX_EQN; e++) p.Q[e] = (e < nq) ? v.Q[e] : nullptr;
        for (int a = 0; a < 3; a++) p.F[a] = (a < d) ? v.F[a] : nullptr;
        p.S = v.S;
        double* q = base;
        p.prim = q;
        q += L.prim;
        p.sens = L.sens ? q : nullptr;
        q += L.sens;
        for (int a = 0; a < 3; a++) {
            p.fmid[a] = L.fmid[a] ? q : nullptr;
            q += L.fmid[a];
            p.ustar[a] = L.ustar[a] ? q : nullptr;
            q += L.ustar[a];
        }
        return p;
    }

    int check_view(const hm_patch_view& v, int which) {
        const int d = cfg_.dim;
        char buf[160];
        for (int a = 0; a < d; a++) {
            if (v.n[a] < 1) {
                snprintf(buf, sizeof buf, "patch %d: interior extent n[%d]=%d < 1", which, a, v.n[a]);
                return fail(HM_ERR_INVALID_ARGUMENT, buf);
            }
            if (v.ghost[a] < kGc) {
                snprintf(buf, sizeof buf, "patch %d: ghost width %d on axis %d is below the 4 cells WCNS56 needs",
                         which, v.ghost[a], a);
                return fail(HM_ERR_INVALID_ARGUMENT, buf);
            }
            if (!(v.dx[a] > 0.0)) {
                snprintf(buf, sizeof buf, "patch %d: dx[%d] must be positive", which, a);
                return fail(HM_ERR_INVALID_ARGUMENT, buf);
            }
            if (!v.F[a]) {
                snprintf(buf, sizeof buf, "patch %d: flux array for axis %d is NULL", which, a);
                return fail(HM_ERR_INVALID_ARGUMENT, buf);
            }
        }
        // Kernels number the items of a patch's work segment in 32 bits and
        // address one component of a patch array at a 32-bit byte offset.
        double box = 1.0;
        for (int a = 0; a < d; a++) box *= double(v.n[a]) + 2.0 * double(v.ghost[a] > 6 ? v.ghost[a] : 6);
        if (box >= 536870912.0) {
            snprintf(buf, sizeof buf, "patch %d: ghost box of 2^29 cells or more; split the patch", which);
            return fail(HM_ERR_INVALID_ARGUMENT, buf);
        }
        for (int e = 0; e < numConservativeArrays(); e++)
            if (!v.Q[e]) {
                snprintf(buf, sizeof buf, "patch %d: conservative variable %d is NULL", which, e);
                return fail(HM_ERR_INVALID_ARGUMENT, buf);
            }
        if (has_advective_ && !v.S) {
            snprintf(buf, sizeof buf, "patch %d: source array is NULL but the model has advective equations", which);
            return fail(HM_ERR_INVALID_ARGUMENT, buf);
        }
        return HM_OK;
    }

    int select_device() {
        int cur = -1;
        if (hipGetDevice(&cur) != hipSuccess) return fail(HM_ERR_DEVICE, "hipGetDevice failed");
        if (cur != cfg_.device && hipSetDevice(cfg_.device) != hipSuccess)
            return fail(HM_ERR_DEVICE, "hipSetDevice failed");
        return HM_OK;
    }

    // Grow the scratch arena to at least `doubles` entries.  Growth waits for
    // work already queued on `s` (which may still read the old arena) and
    // invalidates cached batch plans, whose tables hold arena addresses.
    int ensure_arena(size_t doubles, hipStream_t s) {
        if (doubles <= arena_doubles_) return HM_OK;
        size_t want = std::max(doubles, arena_doubles_ + arena_doubles_ / 2);
        if (arena_) {
            if (hipStreamSynchronize(s) != hipSuccess) return fail(HM_ERR_DEVICE, "stream synchronize failed");
            (void)hipFree(arena_);
            arena_ = nullptr;
            arena_doubles_ = 0;
        }
        for (auto& pl : plans_) free_plan(pl);
        plans_.clear();
        if (hipMalloc(reinterpret_cast<void**>(&arena_), want * sizeof(double)) != hipSuccess) {
            arena_ = nullptr;
            return fail(HM_ERR_OUT_OF_MEMORY, "hipMalloc of the scratch arena failed");
        }
        arena_doubles_ = want;
        return HM_OK;
    }

    void free_plan(TablePlan& pl) {
        if (pl.dev) (void)hipFree(pl.dev);
        pl.dev = nullptr;
    }

    // Build (or reuse) the device work tables of a patch list.
    int get_table_plan(const hm_patch_view* views, int np, hipStream_t s, TablePlan*& out) {
        const uint64_t key = fnv1a(views, sizeof(hm_patch_view) * size_t(np), fnv1a(&np, sizeof np));
        use_clock_++;
        for (auto& pl : plans_) {
            if (pl.key == key && int(pl.views.size()) == np && pl.arena_at_build == arena_ &&
                std::memcmp(pl.views.data(), views, sizeof(hm_patch_view) * size_t(np)) == 0) {
                pl.use = use_clock_;
                out = &pl;
                return HM_OK;
            }
        }
        // Evict the least recently used plan when the cache is full.
        TablePlan* pl = nullptr;
        if (plans_.size() < kMaxPlans) {
            plans_.emplace_back();
            pl = &plans_.back();
        } else {
            pl = &plans_[0];
            for (auto& c : plans_)
                if (c.use < pl->use) pl = &c;
            if (hipStreamSynchronize(s) != hipSuccess) return fail(HM_ERR_DEVICE, "stream synchronize failed");
            free_plan(*pl);
        }
        pl->key = 0;
        const Schedule* sc = nullptr;
        for (const auto& t : tuned_)
            if (t.key == key && t.views.size() == size_t(np) &&
                std::memcmp(t.views.data(), views, sizeof(hm_patch_view) * size_t(np)) == 0)
                sc = &t.sched;
        if (int st = upload_tables(views, np, arena_, pl->dev, pl->plan, nullptr, sc)) return st;
        pl->key = key;
        pl->views.assign(views, views + np);
        pl->arena_at_build = arena_;
        pl->use = use_clock_;
        out = pl;
        return HM_OK;
    }

    // Device work tables of a patch list whose scratch starts at `base`:
    // DevPatch[np], then per stage start/cnt/seg, in one allocation.
    int upload_tables(const hm_patch_view* views, int np, double* base, void*& dev, Plan<SegTable>& plan,
                      const int64_t* wrec_off = nullptr, const Schedule* sched = nullptr) {
        const Schedule sc = sched ? *sched : defaultSchedule(views, np);
        const int d = cfg_.dim;
        std::vector<DevPatch> patches(np);
        std::vector<int64_t> start[kStages], cnt[kStages];
        std::vector<int> seg[kStages];
        for (int i = 0; i < np; i++) {
            const ScratchLayout L = layout(views[i]);
            patches[i] = make_dev_patch(views[i], base, L);
            if (wrec_off)
                for (int a = 0; a < 3; a++) patches[i].wrec[a] = wrec_off[size_t(i) * 3 + a];
            base += L.total;
        }
        int64_t acc[kStages] = {0, 0, 0, 0, 0};
        plan.tile_rows = fused_ ? sc.tile_rows : 4;
        plan.xcd = sc.xcd;
        plan.persistent = (!fused_ && sc.persistent) ? 1 : 0;
        plan.queue = nullptr;
        if (plan.persistent) {
            if (int st = ensure_queue()) return st;
            plan.queue = queue_;
        }
        // (count, patch * 4 + axis) per stage, in patch order
        std::vector<std::pair<int64_t, int>> segs[kStages];
        for (int i = 0; i < np; i++) {
            const Extents e = extents(views[i], plan.tile_rows, !plan.persistent);
            auto add = [&](int stage, int axis, int64_t c) {
                if (c > 0) segs[stage].push_back({c, i * 4 + axis});
            };
            add(0, 0, e.vol4);
            // (stage 1 stays empty: the derived-data kernels compute the sensor)
            for (int a = 0; a < d; a++) add(2, a, e.work2[a]);
            for (int a = 0; a < d; a++) add(3, a, e.work3[a]);
            if (has_advective_) add(4, 0, e.vol0);
        }
        if (sc.order == 1)
            for (int st = 2; st < 4; st++)
                std::stable_sort(segs[st].begin(), segs[st].end(),
                                 [](const std::pair<int64_t, int>& x, const std::pair<int64_t, int>& y) {
                                     return x.first > y.first;
                                 });
        for (int st = 0; st < kStages; st++) {
            for (const auto& sg : segs[st]) {
                start[st].push_back(acc[st]);
                cnt[st].push_back(sg.first);
                seg[st].push_back(sg.second);
                acc[st] += align_items(sg.first);
            }
            start[st].push_back(acc[st]);
        }
        size_t bytes = align_up(sizeof(DevPatch) * size_t(np), 256);
        size_t off_start[kStages], off_cnt[kStages], off_seg[kStages];
        for (int st = 0; st < kStages; st++) {
            off_start[st] = bytes;
            bytes = align_up(bytes + sizeof(int64_t) * start[st].size(), 256);
            off_cnt[st] = bytes;
            bytes = align_up(bytes + sizeof(int64_t) * cnt[st].size(), 256);
            off_seg[st] = bytes;
            bytes = align_up(bytes + sizeof(int) * seg[st].size(), 256);
        }
        std::vector<char> img(bytes, 0);
        std::memcpy(img.data(), patches.data(), sizeof(DevPatch) * size_t(np));
        for (int st = 0; st < kStages; st++) {
            if (!start[st].empty()) std::memcpy(img.data() + off_start[st], start[st].data(), sizeof(int64_t) * start[st].size());
            if (!cnt[st].empty()) std::memcpy(img.data() + off_cnt[st], cnt[st].data(), sizeof(int64_t) * cnt[st].size());
            if (!seg[st].empty()) std::memcpy(img.data() + off_seg[st], seg[st].data(), sizeof(int) * seg[st].size());
        }
        dev = nullptr;
        if (hipMalloc(&dev, bytes) != hipSuccess) {
            dev = nullptr;
            return fail(HM_ERR_OUT_OF_MEMORY, "hipMalloc of the work tables failed");
        }
        // The tables are built once per distinct patch list, so a blocking copy is fine here.
        if (hipMemcpy(dev, img.data(), bytes, hipMemcpyHostToDevice) != hipSuccess)
            return fail(HM_ERR_DEVICE, "upload of the work tables failed");
        char* dv = static_cast<char*>(dev);
        for (int st = 0; st < kStages; st++) {
            SegTable& w = plan.work[st];
            w.patches = reinterpret_cast<const DevPatch*>(dv);
            w.start = reinterpret_cast<const int64_t*>(dv + off_start[st]);
            w.cnt = reinterpret_cast<const int64_t*>(dv + off_cnt[st]);
            w.seg = reinterpret_cast<const int*>(dv + off_seg[st]);
            w.nseg = int(cnt[st].size());
            plan.items[st] = acc[st];
        }
        return HM_OK;
    }

    // run() on the context's arena, then mark the arena in use by stream s.
    template <class W>
    int run_ordered(const Plan<W>& plan, const KParams& k, hipStream_t s) {
        if (int st = run(plan, k, s)) return st;
        if (arena_order_.leave(s) != hipSuccess) return fail(HM_ERR_DEVICE, "hipEventRecord failed");
        return HM_OK;
    }

    template <class W>
    int run(const Plan<W>& plan, const KParams& k, hipStream_t s) {
        if (!profiling_) return run_stages(plan, k, s, 0, kStages);
        // One launch per stage, bracketed by events.
        for (int stage = 0; stage < kStages; stage++) {
            if (plan.items[stage] == 0) continue;
            if (prof_used_ + 2 > int(prof_events_.size()))
                if (int st = collect_profile()) return st;
            hipEvent_t e0 = prof_events_[prof_used_], e1 = prof_events_[prof_used_ + 1];
            if (hipEventRecord(e0, s) != hipSuccess) return fail(HM_ERR_DEVICE, "hipEventRecord failed");
            if (int st = run_stages(plan, k, s, stage, stage + 1)) return st;
            if (hipEventRecord(e1, s) != hipSuccess) return fail(HM_ERR_DEVICE, "hipEventRecord failed");
            prof_stage_[prof_used_ / 2] = stage;
            prof_used_ += 2;
        }
        return HM_OK;
    }

    template <class W>
    int run_stages(const Plan<W>& plan, const KParams& k0, hipStream_t s, int first, int last) {
        Plan<W> sub = plan;
        for (int st = 0; st < kStages; st++)
            if (st < first || st >= last) sub.items[st] = 0;
        KParams k = k0;
        k.xcd = plan.xcd;
        k.queue = plan.queue;
        hipError_t e = hipSuccess;
        switch (cfg_.model) {
            case HM_MODEL_SINGLE_SPECIES: e = launch_rhs_single<W>(cfg_.dim, cfg_.scheme, sub, k, s); break;
            case HM_MODEL_FOUR_EQN_CONSERVATIVE: e = launch_rhs_four<W>(cfg_.dim, cfg_.scheme, sub, k, s); break;
            case HM_MODEL_FIVE_EQN_ALLAIRE: e = launch_rhs_five<W>(cfg_.dim, cfg_.scheme, sub, k, s); break;
            default: return fail(HM_ERR_UNSUPPORTED, "flow model not available in this build");
        }
        if (e != hipSuccess) return fail(HM_ERR_DEVICE, std::string("kernel launch failed: ") + hipGetErrorString(e));
        return HM_OK;
    }

    int collect_profile() {
        for (int i = 0; i + 1 < prof_used_; i += 2) {
            if (hipEventSynchronize(prof_events_[i + 1]) != hipSuccess)
                return fail(HM_ERR_DEVICE, "hipEventSynchronize failed");
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, prof_events_[i], prof_events_[i + 1]) != hipSuccess)
                return fail(HM_ERR_DEVICE, "hipEventElapsedTime failed");
            prof_ms_[prof_stage_[i / 2]] += ms;
            prof_n_[prof_stage_[i / 2]] += 1;
        }
        prof_used_ = 0;
        return HM_OK;
    }

    static constexpr size_t kMaxPlans = 8;
    static constexpr int kProfSlots = 512;
    bool profiling_ = false;
    std::vector<hipEvent_t> prof_events_;
    int prof_used_ = 0;
    int prof_stage_[kProfSlots / 2] = {};
    double prof_ms_[kStages] = {};
    int64_t prof_n_[kStages] = {};
    hm_config cfg_;
    int neq_ = 0;
    bool has_advective_ = false;
    bool fused_ = false;  // single-species model: midpoint and face flux fused (axis 0 in stage 2, axes 1, 2 in stage 3)
    KParams k_{};
    double* arena_ = nullptr;
    size_t arena_doubles_ = 0;
    StreamOrder arena_order_;  // calls on different streams take the arena in turn
    std::vector<TablePlan> plans_;
    struct TunedSchedule {
        uint64_t key;
        std::vector<hm_patch_view> views;
        Schedule sched;
    };
    std::vector<TunedSchedule> tuned_;  // schedules pinned per patch list (autotuner or caller)
    uint64_t use_clock_ = 0;
    void* stage_buf_ = nullptr;
    unsigned* queue_ = nullptr;  // work-queue counter of the persistent midpoint flux (allocated on first use)
    size_t stage_bytes_ = 0;
    hipEvent_t stage_ev_ = nullptr;
    std::string err_;
};

// ConvectiveFluxReconstructorManager (ConvectiveFluxReconstructorManager.cpp:14-153):
// validates the configuration the deck would give and builds the reconstructor.
inline int make_reconstructor(const hm_config& c, std::unique_ptr<ConvectiveFluxReconstructorGPU>& out,
                              std::string& err) {
    if (c.dim < 1 || c.dim > 3) {
        err = "dim must be 1, 2 or 3";
        return HM_ERR_INVALID_ARGUMENT;
    }
    if (c.scheme != HM_SCHEME_WCNS5_JS_HLLC_HLL && c.scheme != HM_SCHEME_WCNS6_LD_HLLC_HLL) {
        err = "unknown convective flux reconstructor";  // ConvectiveFluxReconstructorManager.cpp:146
        return HM_ERR_UNSUPPORTED;
    }
    if (c.model != HM_MODEL_SINGLE_SPECIES && c.model != HM_MODEL_FOUR_EQN_CONSERVATIVE &&
        c.model != HM_MODEL_FIVE_EQN_ALLAIRE) {
        err = "unknown flow model";
        return HM_ERR_UNSUPPORTED;
    }
    const int ns = (c.model == HM_MODEL_SINGLE_SPECIES) ? 1 : c.num_species;
    if (ns < 1 || ns > HM_MAX_SPECIES || num_equations(c.model, c.dim, ns) + 1 > HM_MAX_EQN) {
        err = "number of species out of range";
        return HM_ERR_INVALID_ARGUMENT;
    }
    if (c.model == HM_MODEL_FIVE_EQN_ALLAIRE && (ns < 2 || ns > 3)) {
        err = "the five-equation model is built for 2 or 3 species";
        return HM_ERR_UNSUPPORTED;
    }
    if (c.model == HM_MODEL_FOUR_EQN_CONSERVATIVE && ns > 3) {
        err = "the four-equation model is built for 1 to 3 species";
        return HM_ERR_UNSUPPORTED;
    }
    if (c.constant_p < 1 || (c.scheme == HM_SCHEME_WCNS6_LD_HLLC_HLL && c.constant_q < 1)) {
        err = "WCNS exponents must be positive";
        return HM_ERR_INVALID_ARGUMENT;
    }
    for (int i = 0; i < ns; i++) {
        if (!(c.species_gamma[i] > 1.0)) {
            err = "species_gamma must exceed 1";
            return HM_ERR_INVALID_ARGUMENT;
        }
        if (c.model == HM_MODEL_FOUR_EQN_CONSERVATIVE && !(c.species_R[i] > 0.0)) {
            err = "species_R must be positive";
            return HM_ERR_INVALID_ARGUMENT;
        }
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        err = "no HIP device available";
        return HM_ERR_DEVICE;
    }
    if (c.device < 0 || c.device >= ndev) {
        err = "device ordinal out of range";
        return HM_ERR_INVALID_ARGUMENT;
    }
    hm_config cc = c;
    if (c.model == HM_MODEL_SINGLE_SPECIES) cc.num_species = 1;
    try {
        out.reset(new ConvectiveFluxReconstructorGPU(cc));
    } catch (const std::bad_alloc&) {
        err = "out of host memory";
        return HM_ERR_OUT_OF_MEMORY;
    }
    return HM_OK;
}

}  // namespace hm

