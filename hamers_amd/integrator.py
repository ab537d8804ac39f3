"""Python mirror of the Runge-Kutta stage update and of one level's stage loop.

``advanceSingleStepOnPatch`` mirrors ``Euler::advanceSingleStepOnPatch``
(src/apps/Euler/Euler.cpp:1003-1680; with diffusive fluxes
``NavierStokes::advanceSingleStepOnPatch``, NavierStokes.cpp:1003-2400; the zero fill of
``RungeKuttaLevelIntegrator::advanceLevel`` included) and ``Level`` runs the
stage loop of ``RungeKuttaLevelIntegrator::advanceLevel``
(src/algs/integrator/RungeKuttaLevelIntegrator.cpp:1457-1939) for one level of
patches held on the GPU.  Every computation runs in libhamers_amd.so.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np

from . import capi
from .reconstructor import ConvectiveFluxReconstructor, _check_tensor

# SSP-RK3 (RungeKuttaLevelIntegrator.cpp:3907-3924)
SSP_RK3_ALPHA = ((1.0,), (0.75, 0.25), (1.0 / 3.0, 0.0, 2.0 / 3.0))
SSP_RK3_BETA = ((1.0,), (0.0, 0.25), (0.0, 0.0, 2.0 / 3.0))
SSP_RK3_GAMMA = ((1.0 / 6.0,), (0.0, 1.0 / 6.0), (0.0, 0.0, 2.0 / 3.0))


def _arrays(Q, count):
    """A list of tensors, or one (count, ...) tensor -> list of per-component tensors."""
    if Q is None:
        return [None] * count
    if isinstance(Q, (list, tuple)):
        return list(Q)
    return [Q[c] for c in range(Q.shape[0])]


def _ptr(t):
    if t is None:
        return None
    _check_tensor(t)
    return t.data_ptr()


@dataclass
class RKPatchData:
    """One patch of a stage update: scratch arrays ``Q`` (updated), and per
    coefficient n the intermediate ``Q_intermediate[n]``, ``F_intermediate[n]``
    (per axis) and ``S_intermediate[n]``; optional ``F_accumulated`` /
    ``S_accumulated`` receive the gamma-weighted sums."""
    n: tuple
    ghost: tuple
    ghost_intermediate: tuple
    dx: tuple
    Q: object
    Q_intermediate: list
    F_intermediate: list
    S_intermediate: list
    F_accumulated: list = None
    S_accumulated: object = None
    Fd_intermediate: list = None   # Navier-Stokes: per coefficient, per-axis diffusive fluxes
    Fd_accumulated: list = None
    _keep: list = field(default_factory=list, repr=False)

    def view(self, num_arrays: int, num_eqn: int) -> capi.hm_rk_patch_view:
        v = capi.hm_rk_patch_view()
        d = len(self.n)
        for a in range(3):
            v.n[a] = int(self.n[a]) if a < d else 1
            v.ghost[a] = int(self.ghost[a]) if a < d else 0
            v.ghost_intermediate[a] = int(self.ghost_intermediate[a]) if a < d else 0
            v.dx[a] = float(self.dx[a]) if a < d else 1.0
        Qs = _arrays(self.Q, num_arrays)
        if len(Qs) != num_arrays:
            raise ValueError(f"need {num_arrays} conservative arrays, got {len(Qs)}")
        self._keep = list(Qs)
        for e, t in enumerate(Qs):
            v.Q[e] = _ptr(t)
        for m, Qm in enumerate(self.Q_intermediate):
            for e, t in enumerate(_arrays(Qm, num_eqn)[:num_eqn]):
                v.Q_intermediate[m][e] = _ptr(t)
                self._keep.append(t)
        for m, Fm in enumerate(self.F_intermediate):
            for a, t in enumerate(Fm or []):
                v.F_intermediate[m][a] = _ptr(t)
        for m, t in enumerate(self.S_intermediate):
            v.S_intermediate[m] = _ptr(t)
        for a, t in enumerate(self.F_accumulated or []):
            v.F_accumulated[a] = _ptr(t)
        v.S_accumulated = _ptr(self.S_accumulated)
        for m, Fm in enumerate(self.Fd_intermediate or []):
            for a, t in enumerate(Fm or []):
                v.F_diffusive_intermediate[m][a] = _ptr(t)
        for a, t in enumerate(self.Fd_accumulated or []):
            v.F_diffusive_accumulated[a] = _ptr(t)
        return v


def _coeffs(values):
    arr = (ctypes.c_double * capi.HM_MAX_RK_COEFFS)()
    for i, x in enumerate(values):
        arr[i] = float(x)
    return arr


def advanceSingleStepOnPatch(rec: ConvectiveFluxReconstructor, patch, alpha, beta, gamma=None, stream=None):
    """Stage update of one patch (``RKPatchData``) or of a list of patches (one launch)."""
    if not (1 <= len(alpha) <= capi.HM_MAX_RK_COEFFS) or len(beta) != len(alpha):
        raise ValueError("alpha and beta need the same length, 1..5")
    if gamma is not None and len(gamma) != len(alpha):
        raise ValueError("gamma needs the length of alpha")
    a, b = _coeffs(alpha), _coeffs(beta)
    c = _coeffs(gamma) if gamma is not None else None
    s = rec._stream_ptr(stream)
    if isinstance(patch, RKPatchData):
        v = patch.view(rec.num_arrays, rec.num_eqn)
        st = capi.lib.hm_advance_single_step_on_patch(rec._ctx, ctypes.byref(v), len(alpha), a, b, c, s)
    else:
        patches = list(patch)
        views = (capi.hm_rk_patch_view * len(patches))()
        for i, p in enumerate(patches):
            views[i] = p.view(rec.num_arrays, rec.num_eqn)
        st = capi.lib.hm_advance_single_step_on_patches(rec._ctx, views, len(patches), len(alpha), a, b, c, s)
    capi.check(rec._ctx, st)


class Level:
    """Patches of one level on the GPU, advanced by the Runge-Kutta stage loop.

    ``boxes`` is a list of (lo, n) pairs in level index space, (x, y, z)
    ordered; ``domain`` is (lo, n) of the level domain; ``periodic`` per axis.
    """

    def __init__(self, rec: ConvectiveFluxReconstructor, boxes, domain, periodic, dx):
        self.rec = rec  # the context must outlive the level
        d = rec.dim
        self.dim = d
        self.boxes = [(tuple(lo), tuple(n)) for lo, n in boxes]
        np_ = len(self.boxes)
        lo = (ctypes.c_int * (3 * np_))()
        nn = (ctypes.c_int * (3 * np_))()
        for i, (bl, bn) in enumerate(self.boxes):
            for a in range(d):
                lo[3 * i + a] = int(bl[a])
                nn[3 * i + a] = int(bn[a])
        dlo = (ctypes.c_int * 3)(*([int(x) for x in domain[0]] + [0] * (3 - d)))
        dn = (ctypes.c_int * 3)(*([int(x) for x in domain[1]] + [1] * (3 - d)))
        per = (ctypes.c_int * 3)(*([int(bool(x)) for x in periodic] + [0] * (3 - d)))
        ddx = (ctypes.c_double * 3)(*([float(x) for x in dx] + [1.0] * (3 - d)))
        self._lvl = ctypes.c_void_p()
        capi.check(rec._ctx, capi.lib.hm_level_create(rec._ctx, np_, lo, nn, dlo, dn, per, ddx,
                                                      ctypes.byref(self._lvl)))

    def __del__(self):
        if getattr(self, "_lvl", None):
            capi.lib.hm_level_destroy(self._lvl)
            self._lvl = None

    def __len__(self):
        return len(self.boxes)

    def set_rk_coefficients(self, alpha, beta, gamma=None):
        """Rows of per-stage coefficients, e.g. SSP_RK3_ALPHA."""
        ns = len(alpha)

        def table(rows):
            t = (ctypes.c_double * (ns * ns))()
            for s_, row in enumerate(rows):
                for m, x in enumerate(row):
                    t[s_ * ns + m] = float(x)
            return t

        c = table(gamma) if gamma is not None else None
        capi.check(self.rec._ctx, capi.lib.hm_level_set_rk_coefficients(self._lvl, ns, table(alpha), table(beta), c))

    def write_patch(self, i: int, Q: np.ndarray):
        """Interior values (num_arrays, *reversed(n)) of patch i (float64, host)."""
        bn = self.boxes[i][1]
        shape = (self.rec.num_arrays,) + tuple(reversed(bn))
        Q = np.ascontiguousarray(Q, dtype=np.float64)
        if Q.shape != shape:
            raise ValueError(f"patch {i}: expected shape {shape}, got {Q.shape}")
        capi.check(self.rec._ctx, capi.lib.hm_level_write_patch(
            self._lvl, i, Q.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))

    def read_patch(self, i: int) -> np.ndarray:
        bn = self.boxes[i][1]
        Q = np.empty((self.rec.num_arrays,) + tuple(reversed(bn)))
        capi.check(self.rec._ctx, capi.lib.hm_level_read_patch(
            self._lvl, i, Q.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return Q

    def write_global(self, Q: np.ndarray):
        """Scatter a whole-domain interior array over the patches."""
        for i, (lo, n) in enumerate(self.boxes):
            sl = tuple(slice(lo[a], lo[a] + n[a]) for a in reversed(range(self.dim)))
            self.write_patch(i, Q[(slice(None),) + sl])

    def read_global(self, shape) -> np.ndarray:
        Q = np.full(shape, np.nan)
        for i, (lo, n) in enumerate(self.boxes):
            sl = tuple(slice(lo[a], lo[a] + n[a]) for a in reversed(range(self.dim)))
            Q[(slice(None),) + sl] = self.read_patch(i)
        return Q

    def patch_array(self, i: int):
        """(device address, ghost width) of patch i's scratch arrays."""
        ptr = ctypes.c_void_p()
        g = ctypes.c_int()
        capi.check(self.rec._ctx, capi.lib.hm_level_patch_array(self._lvl, i, ctypes.byref(ptr), ctypes.byref(g)))
        return ptr.value, g.value

    # Physical boundary conditions (include/hamers_convective.h HM_BC_*).
    BC = {"NONE": 0, "FLOW": 1, "REFLECT": 2, "SYMMETRY": 3, "DIRICHLET": 4}

    def set_boundary_conditions(self, conditions, dirichlet=None):
        """Conditions of the non-periodic sides, (xlo, xhi, ylo, yhi, zlo, zhi)
        as HM_BC_* codes or names; ``dirichlet`` [2 dim][num_arrays] values of
        the DIRICHLET sides (EulerBoundaryConditions::setPhysicalBoundaryConditions)."""
        d = self.dim
        codes = [self.BC[c] if isinstance(c, str) else int(c) for c in conditions]
        if len(codes) != 2 * d:
            raise ValueError(f"{2 * d} side conditions expected")
        cc = (ctypes.c_int * (2 * d))(*codes)
        dv = None
        if dirichlet is not None:
            arr = np.ascontiguousarray(dirichlet, dtype=np.float64).reshape(2 * d, self.rec.num_arrays)
            self._dirichlet = arr
            dv = arr.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        capi.check(self.rec._ctx, capi.lib.hm_level_set_boundary_conditions(self._lvl, cc, dv))

    def autotune(self, dt: float, repetitions: int = 3) -> dict:
        """hm_level_autotune: the fastest schedule of the level's batched RHS
        (timed on the current state), with its time and the built-in one's."""
        sc = capi.hm_schedule()
        ms = (ctypes.c_double * 2)()
        capi.check(self.rec._ctx, capi.lib.hm_level_autotune(self._lvl, float(dt), int(repetitions),
                                                             ctypes.byref(sc), ms))
        return dict(sc.as_dict(), ms=ms[0], default_ms=ms[1])

    def set_schedule(self, sched: dict):
        sc = capi.hm_schedule()
        sc.tile_rows, sc.xcd_remap, sc.order = int(sched["tile_rows"]), int(sched["xcd_remap"]), int(sched["order"])
        sc.persistent = int(sched.get("persistent", 0))
        capi.check(self.rec._ctx, capi.lib.hm_level_set_schedule(self._lvl, ctypes.byref(sc)))

    def schedule(self) -> dict:
        sc = capi.hm_schedule()
        capi.check(self.rec._ctx, capi.lib.hm_level_get_schedule(self._lvl, ctypes.byref(sc)))
        return sc.as_dict()

    def fill_ghosts(self, stream=None):
        capi.check(self.rec._ctx, capi.lib.hm_level_fill_ghosts(self._lvl, self.rec._stream_ptr(stream)))

    def advance(self, dt: float, num_steps: int = 1, use_graph: bool = False, stream=None):
        capi.check(self.rec._ctx, capi.lib.hm_level_advance(self._lvl, float(dt), int(num_steps), int(use_graph),
                                                            self.rec._stream_ptr(stream)))

    def advance_cfl(self, cfl: float, num_steps: int = 1, use_graph: bool = False, stream=None):
        """Steps of dt = cfl * the level's stable dt, computed on the device at
        the start of every step (no host round trip)."""
        capi.check(self.rec._ctx, capi.lib.hm_level_advance_cfl(self._lvl, float(cfl), int(num_steps),
                                                                int(use_graph), self.rec._stream_ptr(stream)))

    def dt_info(self) -> dict:
        """The last advance_cfl step's reduction."""
        info = (ctypes.c_double * capi.HM_DT_INFO_SLOTS)()
        capi.check(self.rec._ctx, capi.lib.hm_level_dt_info(self._lvl, info))
        return {"spectral_radiuses": list(info[0:self.dim]), "dt_patches": info[3], "dt_level": info[4],
                "dt": info[5], "time": info[6]}

    # -- discrete adjoint ---------------------------------------------------

    def advance_recorded(self, dt: float, num_steps: int, checkpoint_every: int, use_graph: bool = False,
                         stream=None):
        """Advance and keep a checkpoint every ``checkpoint_every`` steps for adjoint()."""
        capi.check(self.rec._ctx, capi.lib.hm_level_advance_recorded(
            self._lvl, float(dt), int(num_steps), int(checkpoint_every), int(use_graph), self.rec._stream_ptr(stream)))

    def adjoint(self, stream=None):
        """Turn the adjoint of the final state into that of the initial state."""
        capi.check(self.rec._ctx, capi.lib.hm_level_adjoint(self._lvl, self.rec._stream_ptr(stream)))

    def set_face_state_budget(self, nbytes: int):
        """Bytes of device memory adjoint() may keep face states in between its
        recompute and reverse passes (< 0: half the free memory, 0: none)."""
        capi.check(self.rec._ctx, capi.lib.hm_level_set_face_state_budget(self._lvl, int(nbytes)))

    @property
    def face_states_kept(self) -> bool:
        r = capi.lib.hm_level_face_states_kept(self._lvl)
        capi.check(self.rec._ctx, min(r, 0))
        return r == 1

    def write_adjoint(self, i: int, lam: np.ndarray):
        bn = self.boxes[i][1]
        shape = (self.rec.num_eqn,) + tuple(reversed(bn))
        lam = np.ascontiguousarray(lam, dtype=np.float64)
        if lam.shape != shape:
            raise ValueError(f"patch {i}: expected shape {shape}, got {lam.shape}")
        capi.check(self.rec._ctx, capi.lib.hm_level_write_adjoint(
            self._lvl, i, lam.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))

    def read_adjoint(self, i: int) -> np.ndarray:
        bn = self.boxes[i][1]
        lam = np.empty((self.rec.num_eqn,) + tuple(reversed(bn)))
        capi.check(self.rec._ctx, capi.lib.hm_level_read_adjoint(
            self._lvl, i, lam.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return lam

    def write_adjoint_global(self, lam: np.ndarray):
        for i, (lo, n) in enumerate(self.boxes):
            sl = tuple(slice(lo[a], lo[a] + n[a]) for a in reversed(range(self.dim)))
            self.write_adjoint(i, lam[(slice(None),) + sl])

    def read_adjoint_global(self, shape) -> np.ndarray:
        lam = np.full(shape, np.nan)
        for i, (lo, n) in enumerate(self.boxes):
            sl = tuple(slice(lo[a], lo[a] + n[a]) for a in reversed(range(self.dim)))
            lam[(slice(None),) + sl] = self.read_adjoint(i)
        return lam


class Hierarchy:
    """Two levels, refinement ratio 2 (hm_hierarchy_create): ``advance`` runs
    coarse steps of dt, each with two fine steps of dt / 2, the coarse-fine
    ghost fill and the synchronization (refluxing, then the coarse cells under
    the fine level averaged from it)."""

    def __init__(self, coarse: Level, fine: Level):
        self.coarse, self.fine = coarse, fine  # the levels must outlive the hierarchy
        self._h = ctypes.c_void_p()
        capi.check(coarse.rec._ctx, capi.lib.hm_hierarchy_create(coarse._lvl, fine._lvl, 2, ctypes.byref(self._h)))

    def __del__(self):
        if getattr(self, "_h", None):
            capi.lib.hm_hierarchy_destroy(self._h)
            self._h = None

    def advance(self, dt: float, num_steps: int = 1, stream=None):
        capi.check(self.coarse.rec._ctx, capi.lib.hm_hierarchy_advance(self._h, float(dt), int(num_steps),
                                                                        self.coarse.rec._stream_ptr(stream)))
