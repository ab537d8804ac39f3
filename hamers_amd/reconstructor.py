"""Python mirror of the reference's convective-flux reconstructor interface.

Mirrors ``ConvectiveFluxReconstructor`` (include/flow/convective_flux_reconstructors/
ConvectiveFluxReconstructor.hpp:23-124) and ``ConvectiveFluxReconstructorManager``
(src/flow/convective_flux_reconstructors/ConvectiveFluxReconstructorManager.cpp:14-153):
the same deck keys, the same method name and argument meaning, and the same
output semantics (flux overwritten, source accumulated, both scaled by dt).
Patch data are torch tensors on the GPU; torch only provides device memory
and streams, every computation runs in libhamers_amd.so.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

from . import capi

_SCHEMES = {
    "WCNS5_JS_HLLC_HLL": capi.HM_SCHEME_WCNS5_JS_HLLC_HLL,  # ConvectiveFluxReconstructorManager.cpp:40
    "WCNS6_LD_HLLC_HLL": capi.HM_SCHEME_WCNS6_LD_HLLC_HLL,  # ConvectiveFluxReconstructorManager.cpp:79
}
_MODELS = {
    "SINGLE_SPECIES": capi.HM_MODEL_SINGLE_SPECIES,
    "FOUR_EQN_CONSERVATIVE": capi.HM_MODEL_FOUR_EQN_CONSERVATIVE,
    "FIVE_EQN_ALLAIRE": capi.HM_MODEL_FIVE_EQN_ALLAIRE,
}


def num_equations(model: str, dim: int, num_species: int) -> int:
    if model == "SINGLE_SPECIES":
        return dim + 2
    if model == "FIVE_EQN_ALLAIRE":
        return dim + 2 * num_species
    if model == "FOUR_EQN_CONSERVATIVE":
        return num_species + dim + 1
    raise ValueError(f"unknown flow model {model!r}")


@dataclass
class PatchData:
    """The patch data one reconstructor call reads and writes.

    ``Q``: list of per-equation tensors over the ghost box, or one tensor of
    shape (neq, *ghost box reversed); ``F``: per-axis flux tensors of shape
    (neq, *face box reversed); ``S``: source tensor (neq, *interior reversed).
    ``n``, ``ghost`` and ``dx`` are in (x, y, z) order.
    """
    n: tuple
    ghost: tuple
    dx: tuple
    Q: object
    F: list
    S: object = None
    _keep: list = field(default_factory=list, repr=False)

    def view(self, neq: int) -> capi.hm_patch_view:
        v = capi.hm_patch_view()
        d = len(self.n)
        for a in range(3):
            v.n[a] = int(self.n[a]) if a < d else 1
            v.ghost[a] = int(self.ghost[a]) if a < d else 0
            v.dx[a] = float(self.dx[a]) if a < d else 1.0
        qs = list(self.Q) if isinstance(self.Q, (list, tuple)) else [self.Q[e] for e in range(self.Q.shape[0])]
        if len(qs) < neq:
            raise ValueError(f"patch has {len(qs)} conservative arrays, the model needs {neq}")
        for e in range(neq):
            t = qs[e]
            _check_tensor(t)
            v.Q[e] = t.data_ptr()
        for a in range(d):
            _check_tensor(self.F[a])
            v.F[a] = self.F[a].data_ptr()
        if self.S is not None:
            _check_tensor(self.S)
            v.S = self.S.data_ptr()
        return v


def _check_tensor(t):
    import torch

    if not (t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()):
        raise ValueError("patch arrays must be contiguous float64 tensors on the GPU")


class PatchList:
    """A level's patch list with its C views built once (the arrays' device
    addresses are captured, so the tensors must stay alive and in place)."""

    def __init__(self, patches, neq: int):
        self.patches = list(patches)
        self.views = (capi.hm_patch_view * len(self.patches))()
        for i, p in enumerate(self.patches):
            self.views[i] = p.view(neq)

    def __len__(self):
        return len(self.patches)


class ConvectiveFluxReconstructor:
    """A WCNS56 hybrid HLLC / HLLC-HLL convective flux reconstructor on the GPU."""

    def __init__(self, object_name: str, dim: int, flow_model: str, scheme: str, num_species: int = 1,
                 species_gamma=(1.4,), species_R=(1.0,), constant_p: int = 2, constant_q: int = 4,
                 constant_C: float = 1.0e9, constant_alpha_tau: float = 35.0, device: int = 0):
        self.object_name = object_name
        self.dim = dim
        self.flow_model = flow_model
        self.scheme = scheme
        self.num_species = 1 if flow_model == "SINGLE_SPECIES" else num_species
        self.num_eqn = num_equations(flow_model, dim, self.num_species)
        # arrays a patch hands over: the five-equation model adds the last volume fraction
        self.num_arrays = self.num_eqn + (1 if flow_model == "FIVE_EQN_ALLAIRE" else 0)
        cfg = capi.hm_config()
        cfg.dim = dim
        cfg.model = _MODELS[flow_model]
        cfg.num_species = self.num_species
        cfg.scheme = _SCHEMES[scheme]
        cfg.constant_p = constant_p
        cfg.constant_q = constant_q
        cfg.constant_C = constant_C
        cfg.constant_alpha_tau = constant_alpha_tau
        for i, gv in enumerate(species_gamma):
            cfg.species_gamma[i] = gv
        for i, rv in enumerate(species_R):
            cfg.species_R[i] = rv
        cfg.device = device
        self._ctx = ctypes.c_void_p()
        st = capi.lib.hm_ctx_create(ctypes.byref(cfg), ctypes.byref(self._ctx))
        if st != capi.HM_OK:
            msg = capi.lib.hm_last_error(self._ctx)
            capi.lib.hm_ctx_destroy(self._ctx)
            self._ctx = None
            raise capi.HamersError(st, msg.decode() if msg else "")

    def __del__(self):
        if getattr(self, "_ctx", None):
            capi.lib.hm_ctx_destroy(self._ctx)
            self._ctx = None

    def getConvectiveFluxNumberOfGhostCells(self) -> int:
        g = ctypes.c_int()
        capi.check(self._ctx, capi.lib.hm_get_num_ghosts(self._ctx, ctypes.byref(g)))
        return g.value

    @staticmethod
    def _stream_ptr(stream):
        if stream is None:
            import torch

            stream = torch.cuda.current_stream()
        return ctypes.c_void_p(stream.cuda_stream)

    def computeConvectiveFluxAndSourceOnPatch(self, patch: PatchData, time: float, dt: float,
                                              RK_step_number: int = 0, stream=None):
        v = patch.view(self.num_arrays)
        st = capi.lib.hm_compute_convective_flux_and_source_on_patch(
            self._ctx, ctypes.byref(v), time, dt, RK_step_number, self._stream_ptr(stream))
        capi.check(self._ctx, st)

    STAGES = ("derived", "sensor", "midflux", "faceflux", "source")

    def enable_profiling(self, on: bool = True):
        """Time each pipeline stage with HIP events on the launch stream."""
        capi.check(self._ctx, capi.lib.hm_profile_enable(self._ctx, int(on)))

    def read_profile(self) -> dict:
        ms = (ctypes.c_double * capi.HM_NUM_STAGES)()
        n = (ctypes.c_int64 * capi.HM_NUM_STAGES)()
        capi.check(self._ctx, capi.lib.hm_profile_read(self._ctx, ms, n))
        return {name: (ms[i], n[i]) for i, name in enumerate(self.STAGES)}

    def computeConvectiveFluxAndSourceOnPatches(self, patches, time: float, dt: float, RK_step_number: int = 0,
                                                stream=None):
        plist = patches if isinstance(patches, PatchList) else PatchList(patches, self.num_arrays)
        st = capi.lib.hm_compute_convective_flux_and_source_on_patches(
            self._ctx, plist.views, len(plist), time, dt, RK_step_number, self._stream_ptr(stream))
        capi.check(self._ctx, st)


def _schedule(sched):
    if isinstance(sched, capi.hm_schedule):
        return sched
    c = capi.hm_schedule()
    c.tile_rows, c.xcd_remap, c.order = int(sched["tile_rows"]), int(sched["xcd_remap"]), int(sched["order"])
    c.persistent = int(sched.get("persistent", 0))
    return c


def autotune(rec: "ConvectiveFluxReconstructor", patches, dt: float, repetitions: int = 3, stream=None) -> dict:
    """Autotuner of the batched RHS of a patch list (hm_autotune_on_patches):
    every candidate schedule is timed on these patches (their F is overwritten,
    their S accumulated), the fastest is pinned for the list.  Returns the
    schedule with its time and the built-in schedule's (ms)."""
    plist = patches if isinstance(patches, PatchList) else PatchList(patches, rec.num_arrays)
    sc = capi.hm_schedule()
    ms = (ctypes.c_double * 2)()
    capi.check(rec._ctx, capi.lib.hm_autotune_on_patches(rec._ctx, plist.views, len(plist), float(dt),
                                                          int(repetitions), rec._stream_ptr(stream),
                                                          ctypes.byref(sc), ms))
    return dict(sc.as_dict(), ms=ms[0], default_ms=ms[1])


def set_schedule(rec: "ConvectiveFluxReconstructor", patches, sched):
    plist = patches if isinstance(patches, PatchList) else PatchList(patches, rec.num_arrays)
    capi.check(rec._ctx, capi.lib.hm_set_schedule_on_patches(rec._ctx, plist.views, len(plist),
                                                              ctypes.byref(_schedule(sched))))


def get_schedule(rec: "ConvectiveFluxReconstructor", patches) -> dict:
    plist = patches if isinstance(patches, PatchList) else PatchList(patches, rec.num_arrays)
    sc = capi.hm_schedule()
    capi.check(rec._ctx, capi.lib.hm_get_schedule_on_patches(rec._ctx, plist.views, len(plist), ctypes.byref(sc)))
    return sc.as_dict()


def computeSpectralRadiusesAndStableDtOnPatches(rec: "ConvectiveFluxReconstructor", patches, stream=None):
    """Euler::computeSpectralRadiusesAndStableDtOnPatch over the interior cells
    of every patch, reduced over the list as RungeKuttaLevelIntegrator::getLevelDt
    does before its CFL factor: [s_0 .. s_{d-1}, min over patches of the
    patch dt, the level dt].  Only the patches' conservative arrays are read."""
    plist = patches if isinstance(patches, PatchList) else PatchList(patches, rec.num_arrays)
    out = (ctypes.c_double * (rec.dim + 2))()
    capi.check(rec._ctx, capi.lib.hm_compute_spectral_radiuses_and_stable_dt_on_patches(
        rec._ctx, plist.views, len(plist), out, rec._stream_ptr(stream)))
    return list(out)


def _vjp_view(pd: PatchData, F_bar, Q_bar, num_arrays: int, neq: int) -> capi.hm_vjp_patch_view:
    v = capi.hm_vjp_patch_view()
    d = len(pd.n)
    for a in range(3):
        v.n[a] = int(pd.n[a]) if a < d else 1
        v.ghost[a] = int(pd.ghost[a]) if a < d else 0
        v.dx[a] = float(pd.dx[a]) if a < d else 1.0
    qs = list(pd.Q) if isinstance(pd.Q, (list, tuple)) else [pd.Q[e] for e in range(pd.Q.shape[0])]
    for e in range(num_arrays):
        _check_tensor(qs[e])
        v.Q[e] = qs[e].data_ptr()
    for a in range(d):
        _check_tensor(F_bar[a])
        v.F_bar[a] = F_bar[a].data_ptr()
    qb = list(Q_bar) if isinstance(Q_bar, (list, tuple)) else [Q_bar[e] for e in range(Q_bar.shape[0])]
    for e in range(neq):
        _check_tensor(qb[e])
        v.Q_bar[e] = qb[e].data_ptr()
    return v


def convective_flux_vjp(rec: ConvectiveFluxReconstructor, patches, F_bar, Q_bar, dt: float, stream=None):
    """Vector-Jacobian product of computeConvectiveFluxAndSourceOnPatch(es):
    for each patch, Q_bar[i] <- (dF/dQ)^T F_bar[i] over the ghost box of Q.
    ``F_bar[i]`` is the list of per-axis face-flux adjoints (layout of F)."""
    views = (capi.hm_vjp_patch_view * len(patches))()
    for i, pd in enumerate(patches):
        views[i] = _vjp_view(pd, F_bar[i], Q_bar[i], rec.num_arrays, rec.num_eqn)
    st = capi.lib.hm_convective_flux_vjp_on_patches(rec._ctx, views, len(patches), 0.0, float(dt), 0,
                                                    rec._stream_ptr(stream))
    capi.check(rec._ctx, st)


class DiffusiveFluxReconstructor:
    """The node-based SIXTH_ORDER diffusive flux reconstructor of the
    four-equation model (DiffusiveFluxReconstructor.hpp:73-79,
    DiffusiveFluxReconstructorManager.cpp:13-73) with constant transport
    mixing rules.  It shares the convective reconstructor's context (species
    gamma and R, device), as NavierStokes holds both reconstructors of one
    flow model.  Transport values are per species."""

    def __init__(self, convective: ConvectiveFluxReconstructor, species_mu, species_mu_v, species_kappa,
                 species_D, species_M):
        self.rec = convective
        tr = capi.hm_transport()
        for key, vals in (("species_mu", species_mu), ("species_mu_v", species_mu_v),
                          ("species_kappa", species_kappa), ("species_D", species_D), ("species_M", species_M)):
            arr = getattr(tr, key)
            for i, val in enumerate(vals):
                arr[i] = float(val)
        capi.check(convective._ctx, capi.lib.hm_set_diffusive_transport(convective._ctx, ctypes.byref(tr)))

    def getDiffusiveFluxNumberOfGhostCells(self) -> int:
        g = ctypes.c_int()
        capi.check(self.rec._ctx, capi.lib.hm_get_diffusive_num_ghosts(self.rec._ctx, ctypes.byref(g)))
        return g.value

    def computeDiffusiveFluxOnPatch(self, patch: PatchData, time: float, dt: float, RK_step_number: int = 0,
                                    stream=None):
        """``patch.F`` are the diffusive flux arrays (overwritten); Q needs 6 ghosts."""
        v = patch.view(self.rec.num_eqn)
        st = capi.lib.hm_compute_diffusive_flux_on_patch(self.rec._ctx, ctypes.byref(v), time, dt, RK_step_number,
                                                         self.rec._stream_ptr(stream))
        capi.check(self.rec._ctx, st)

    def computeDiffusiveFluxOnPatches(self, plist: PatchList, time: float, dt: float, RK_step_number: int = 0,
                                      stream=None):
        st = capi.lib.hm_compute_diffusive_flux_on_patches(self.rec._ctx, plist.views, len(plist), time, dt,
                                                           RK_step_number, self.rec._stream_ptr(stream))
        capi.check(self.rec._ctx, st)


class ConvectiveFluxReconstructorManager:
    """Builds the reconstructor named by the deck key, as
    ConvectiveFluxReconstructorManager.cpp:14-153 does; unknown keys raise."""

    def __init__(self, object_name: str, dim: int, flow_model: str, convective_flux_reconstructor: str, **db):
        if convective_flux_reconstructor not in _SCHEMES:
            raise ValueError(f"Unknown convective_flux_reconstructor string {convective_flux_reconstructor!r}")
        if flow_model not in _MODELS:
            raise ValueError(f"Unknown flow_model string {flow_model!r}")
        self.reconstructor = ConvectiveFluxReconstructor(object_name, dim, flow_model,
                                                         convective_flux_reconstructor, **db)

    def getConvectiveFluxReconstructor(self) -> ConvectiveFluxReconstructor:
        return self.reconstructor
