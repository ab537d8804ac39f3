// hm_host.hip -- C-ABI of the GPU convective flux reconstructor
// (include/hamers_convective.h).
#include "hm_ctx.hpp"

extern "C" {

int hm_ctx_create(const hm_config* cfg, hm_ctx** out) {
    if (!cfg || !out) return HM_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    hm_ctx* c = new (std::nothrow) hm_ctx();
    if (!c) return HM_ERR_OUT_OF_MEMORY;
    std::string err;
    int st = hm::make_reconstructor(*cfg, c->rec, err);
    if (st == HM_OK) st = hm::noalloc_setup(cfg->device, err);
    if (st != HM_OK) {
        // The failed context is still returned so that hm_last_error can explain the failure.
        c->rec.reset();
        c->err = err;
        *out = c;
        return st;
    }
    *out = c;
    return HM_OK;
}

int hm_ctx_destroy(hm_ctx* ctx) {
    if (ctx && ctx->levels > 0) {
        if (ctx->rec) ctx->rec->fail(HM_ERR_INVALID_ARGUMENT, "destroy the levels of this context first");
        return HM_ERR_INVALID_ARGUMENT;
    }
    delete ctx;
    return HM_OK;
}

int hm_get_num_ghosts(const hm_ctx* ctx, int* num_ghosts) {
    if (!ctx || !ctx->rec || !num_ghosts) return HM_ERR_INVALID_ARGUMENT;
    *num_ghosts = ctx->rec->getConvectiveFluxNumberOfGhostCells();
    return HM_OK;
}

int hm_get_num_eqn(const hm_ctx* ctx, int* num_eqn) {
    if (!ctx || !ctx->rec || !num_eqn) return HM_ERR_INVALID_ARGUMENT;
    *num_eqn = ctx->rec->numEqn();
    return HM_OK;
}

const char* hm_last_error(const hm_ctx* ctx) {
    if (!ctx) return "NULL context";
    if (!ctx->rec) return ctx->err.c_str();
    return ctx->rec->lastError().c_str();
}

int hm_compute_convective_flux_and_source_on_patch(hm_ctx* ctx, const hm_patch_view* patch, double time, double dt,
                                                    int rk_step, void* stream) {
    if (!ctx || !ctx->rec) return HM_ERR_INVALID_ARGUMENT;
    if (!patch) return ctx->rec->fail(HM_ERR_INVALID_ARGUMENT, "NULL patch view");
    return ctx->rec->computeConvectiveFluxAndSourceOnPatch(*patch, time, dt, rk_step,
                                                           static_cast<hipStream_t>(stream));
}

int hm_compute_convective_flux_and_source_on_patches(hm_ctx* ctx, const hm_patch_view* patches, int num_patches,
                                                      double time, double dt, int rk_step, void* stream) {
    if (!ctx || !ctx->rec) return HM_ERR_INVALID_ARGUMENT;
    return ctx->rec->computeConvectiveFluxAndSourceOnPatches(patches, num_patches, time, dt, rk_step,
                                                             static_cast<hipStream_t>(stream));
}

namespace {
hm::Schedule to_sched(const hm_schedule& c) {
    hm::Schedule s;
    s.tile_rows = c.tile_rows;
    s.xcd = c.xcd_remap;
    s.order = c.order;
    s.persistent = c.persistent;
    return s;
}
hm_schedule from_sched(const hm::Schedule& s) {
    hm_schedule c;
    c.tile_rows = s.tile_rows;
    c.xcd_remap = s.xcd;
    c.order = s.order;
    c.persistent = s.persistent;
    return c;
}
}  // namespace

int hm_autotune_on_patches(hm_ctx* ctx, const hm_patch_view* patches, int num_patches, double dt, int repetitions,
                           void* stream, hm_schedule* chosen, double* ms) {
    if (!ctx || !ctx->rec) return HM_ERR_INVALID_ARGUMENT;
    hm::Schedule best;
    double best_ms = 0.0, default_ms = 0.0;
    const int st = ctx->rec->autotuneOnPatches(patches, num_patches, dt, repetitions, static_cast<hipStream_t>(stream),
                                               best, best_ms, default_ms);
    if (st) return st;
    if (chosen) *chosen = from_sched(best);
    if (ms) {
        ms[0] = best_ms;
        ms[1] = default_ms;
    }
    return HM_OK;
}

int hm_set_schedule_on_patches(hm_ctx* ctx, const hm_patch_view* patches, int num_patches,
                               const hm_schedule* schedule) {
    if (!ctx || !ctx->rec || !schedule) return HM_ERR_INVALID_ARGUMENT;
    if (int st = ctx->rec->selectDevice()) return st;
    return ctx->rec->setSchedule(patches, num_patches, to_sched(*schedule));
}

int hm_get_schedule_on_patches(hm_ctx* ctx, const hm_patch_view* patches, int num_patches, hm_schedule* schedule) {
    if (!ctx || !ctx->rec || !schedule || !patches || num_patches < 1) return HM_ERR_INVALID_ARGUMENT;
    *schedule = from_sched(ctx->rec->scheduleOf(patches, num_patches));
    return HM_OK;
}

int hm_profile_enable(hm_ctx* ctx, int enable) {
    if (!ctx || !ctx->rec) return HM_ERR_INVALID_ARGUMENT;
    return ctx->rec->setProfiling(enable);
}

int hm_get_num_conservative_arrays(const hm_ctx* ctx, int* num_arrays) {
    if (!ctx || !ctx->rec || !num_arrays) return HM_ERR_INVALID_ARGUMENT;
    *num_arrays = ctx->rec->numConservativeArrays();
    return HM_OK;
}

int hm_profile_read(hm_ctx* ctx, double* stage_ms, int64_t* stage_launches) {
    if (!ctx || !ctx->rec || !stage_ms || !stage_launches) return HM_ERR_INVALID_ARGUMENT;
    return ctx->rec->readProfile(stage_ms, stage_launches);
}

int hm_synchronize(hm_ctx* ctx, void* stream) {
    if (!ctx || !ctx->rec) return HM_ERR_INVALID_ARGUMENT;
    if (hipStreamSynchronize(static_cast<hipStream_t>(stream)) != hipSuccess)
        return ctx->rec->fail(HM_ERR_DEVICE, "hipStreamSynchronize failed");
    return HM_OK;
}

}  // extern "C"
