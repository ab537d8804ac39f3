"""ctypes wrapper around the CPU oracle (``libhamers_oracle.so``).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()``
and the CPU-baseline leg of ``bench.py``, never by the product package.

Array convention (numpy, C order): a d-dimensional component-major SAMRAI
array is a numpy array of shape ``(ncomp, n_{d-1}, ..., n_0)`` -- the last axis
is x, so x is the fastest-varying index as in SAMRAI CellData/SideData.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libhamers_oracle.so")

SINGLE_SPECIES, FOUR_EQN_CONSERVATIVE, FIVE_EQN_ALLAIRE = 0, 1, 2
WCNS5_JS_HLLC_HLL, WCNS6_LD_HLLC_HLL = 0, 1


class _Params(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int),
        ("model", ctypes.c_int),
        ("num_species", ctypes.c_int),
        ("scheme", ctypes.c_int),
        ("p", ctypes.c_int),
        ("q", ctypes.c_int),
        ("C", ctypes.c_double),
        ("alpha_tau", ctypes.c_double),
        ("gamma", ctypes.c_double * 8),
        ("R", ctypes.c_double * 8),
        ("mutation", ctypes.c_int),
    ]


class _Patch(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int * 3),
        ("g", ctypes.c_int),
        ("dx", ctypes.c_double * 3),
        ("Q", ctypes.c_void_p),
        ("F", ctypes.c_void_p * 3),
        ("S", ctypes.c_void_p),
    ]


@dataclass
class Params:
    dim: int = 2
    model: int = SINGLE_SPECIES
    num_species: int = 1
    scheme: int = WCNS6_LD_HLLC_HLL
    p: int = 2
    q: int = 4
    C: float = 1.0e9
    alpha_tau: float = 35.0
    gamma: list = field(default_factory=lambda: [1.4])
    R: list = field(default_factory=lambda: [1.0])
    mutation: int = 0

    def to_c(self) -> _Params:
        c = _Params()
        c.dim, c.model, c.num_species, c.scheme = self.dim, self.model, self.num_species, self.scheme
        c.p, c.q, c.C, c.alpha_tau = self.p, self.q, self.C, self.alpha_tau
        for i, gv in enumerate(self.gamma[:8]):
            c.gamma[i] = gv
        for i in range(8):  # gas constants not given default to 1
            c.R[i] = self.R[i] if i < len(self.R) else 1.0
        c.mutation = self.mutation
        return c

    @property
    def nq(self) -> int:
        """Number of conservative arrays handed over: the five-equation model
        also passes the last species' volume fraction (stored, not evolved)."""
        return self.neq + 1 if self.model == FIVE_EQN_ALLAIRE else self.neq

    @property
    def neq(self) -> int:
        if self.model == SINGLE_SPECIES:
            return self.dim + 2
        if self.model == FIVE_EQN_ALLAIRE:
            return self.dim + 2 * self.num_species
        return self.num_species + self.dim + 1


def build() -> None:
    """Compile the oracle with its Makefile (idempotent)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.hmo_rhs_patch.restype = ctypes.c_int
        L.hmo_rhs_patch.argtypes = [
            ctypes.POINTER(_Params), ctypes.POINTER(ctypes.c_int), ctypes.c_int,
            ctypes.POINTER(ctypes.c_double), ctypes.c_double, ctypes.c_void_p,
            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
        ]
        L.hmo_rhs_batch.restype = ctypes.c_int
        L.hmo_rhs_batch.argtypes = [
            ctypes.POINTER(_Params), ctypes.c_int, ctypes.POINTER(_Patch), ctypes.c_double, ctypes.c_int,
        ]
        L.hmo_max_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _ptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


def face_shape(neq: int, n: tuple, axis: int) -> tuple:
    """numpy shape of the axis-``axis`` flux SideData (0 ghosts)."""
    ext = [n[a] + (1 if a == axis else 0) for a in range(len(n))]
    return (neq,) + tuple(reversed(ext))


def rhs_patch(prm: Params, Q: np.ndarray, n: tuple, g: int, dx: tuple, dt: float, S: np.ndarray | None = None):
    """computeConvectiveFluxAndSourceOnPatch on one patch.

    ``Q`` has shape ``(neq, *(n[a]+2g reversed))``.  Returns ``(F, S)`` where F
    is a list of per-axis face arrays (overwritten) and S the source (accumulated).
    """
    d = prm.dim
    neq = prm.neq
    n = tuple(int(v) for v in n)
    Q = np.ascontiguousarray(Q, dtype=np.float64)
    assert Q.shape == (prm.nq,) + tuple(reversed([v + 2 * g for v in n])), Q.shape
    F = [np.zeros(face_shape(neq, n, a)) for a in range(d)]
    if S is None:
        S = np.zeros((neq,) + tuple(reversed(n)))
    nn = (ctypes.c_int * 3)(*(list(n) + [1] * (3 - d)))
    dd = (ctypes.c_double * 3)(*(list(dx) + [1.0] * (3 - d)))
    cp = prm.to_c()
    st = lib().hmo_rhs_patch(ctypes.byref(cp), nn, g, dd, dt, _ptr(Q),
                             _ptr(F[0]), _ptr(F[1]) if d > 1 else None, _ptr(F[2]) if d > 2 else None, _ptr(S))
    if st != 0:
        raise RuntimeError(f"oracle hmo_rhs_patch failed with status {st}")
    return F, S


def rhs_batch(prm: Params, patches: list, dt: float, nthreads: int = 0):
    """Many patches in parallel (OpenMP).  ``patches`` is a list of dicts with
    keys Q, n, g, dx, F (list), S -- arrays are written in place."""
    arr = (_Patch * len(patches))()
    keep = []
    d = prm.dim
    for i, P in enumerate(patches):
        c = arr[i]
        n = list(P["n"]) + [1] * (3 - d)
        for a in range(3):
            c.n[a] = n[a]
        c.g = P["g"]
        dx = list(P["dx"]) + [1.0] * (3 - d)
        for a in range(3):
            c.dx[a] = dx[a]
        c.Q = _ptr(P["Q"])
        for a in range(3):
            c.F[a] = _ptr(P["F"][a]) if a < d else None
        c.S = _ptr(P["S"])
        keep.append(P)
    cp = prm.to_c()
    st = lib().hmo_rhs_batch(ctypes.byref(cp), len(patches), arr, dt, nthreads)
    if st != 0:
        raise RuntimeError(f"oracle hmo_rhs_batch failed with status {st}")


def max_threads() -> int:
    return lib().hmo_max_threads()


# ---- pointwise entry points (pinning tests only) ----

def _arr(x, n=None):
    a = np.ascontiguousarray(x, dtype=np.float64)
    if n is not None:
        assert a.size == n
    return a


def weno_pair(prm: Params, U) -> tuple:
    L = lib()
    L.hmo_weno_pair.argtypes = [ctypes.POINTER(_Params), ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    U = _arr(U, 6)
    um, up = ctypes.c_double(), ctypes.c_double()
    cp = prm.to_c()
    L.hmo_weno_pair(ctypes.byref(cp), U.ctypes.data, ctypes.byref(um), ctypes.byref(up))
    return um.value, up.value


def smoothness(U) -> np.ndarray:
    L = lib()
    L.hmo_smoothness.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    U = _arr(U, 6)
    b = np.zeros(4)
    L.hmo_smoothness(U.ctypes.data, b.ctypes.data)
    return b


def char_roundtrip_single(dim: int, axis: int, rho_avg: float, c_avg: float, V):
    L = lib()
    L.hmo_char_roundtrip_single.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    V = _arr(V, dim + 2)
    Wc = np.zeros(dim + 2)
    Vb = np.zeros(dim + 2)
    L.hmo_char_roundtrip_single(dim, axis, rho_avg, c_avg, V.ctypes.data, Wc.ctypes.data, Vb.ctypes.data)
    return Wc, Vb


def riemann_multi(prm: Params, axis: int, use_hll: bool, VL, VR):
    """Midpoint flux of one face, four- or five-equation model."""
    L = lib()
    L.hmo_riemann_multi.argtypes = [ctypes.POINTER(_Params), ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                    ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    L.hmo_riemann_multi.restype = ctypes.c_int
    VL = _arr(VL, prm.neq)
    VR = _arr(VR, prm.neq)
    F = np.zeros(prm.neq)
    us = ctypes.c_double()
    st = L.hmo_riemann_multi(ctypes.byref(prm.to_c()), axis, int(use_hll), VL.ctypes.data, VR.ctypes.data,
                             F.ctypes.data, ctypes.byref(us))
    if st != 0:
        raise ValueError(f"oracle riemann_multi status {st}")
    return F, us.value


def riemann_single(dim: int, axis: int, gamma: float, use_hll: bool, VL, VR, mutation: int = 0):
    """Midpoint flux of one face, single-species model; ``mutation`` plants a
    known mistake (8: HLLC-HLL blend weights alpha_1, alpha_2 swapped)."""
    L = lib()
    L.hmo_riemann_single_mutated.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                             ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                             ctypes.c_int]
    VL = _arr(VL, dim + 2)
    VR = _arr(VR, dim + 2)
    F = np.zeros(dim + 2)
    us = ctypes.c_double()
    L.hmo_riemann_single_mutated(dim, axis, gamma, int(use_hll), VL.ctypes.data, VR.ctypes.data, F.ctypes.data,
                                 ctypes.byref(us), int(mutation))
    return F, us.value


def midpoint_probe_single(prm: Params, Q: np.ndarray, n: tuple, g: int, dx: tuple, axis: int):
    """Sensor value and midpoint flux of every midpoint face of one axis of a
    single-species patch, from the pass ``rhs_patch`` runs.  Returns
    (sensor, fmid) shaped like the face box (i_a in [-1, n_a+1], others
    interior; numpy order z, y, x) and (neq,) + that."""
    L = lib()
    L.hmo_midpoint_probe_single.restype = ctypes.c_int
    L.hmo_midpoint_probe_single.argtypes = [ctypes.POINTER(_Params), ctypes.POINTER(ctypes.c_int), ctypes.c_int,
                                            ctypes.POINTER(ctypes.c_double), ctypes.c_void_p, ctypes.c_int,
                                            ctypes.c_void_p, ctypes.c_void_p]
    d = prm.dim
    ext = [n[b] + 3 if b == axis else n[b] for b in range(d)]
    shape = tuple(reversed(ext))
    Q = np.ascontiguousarray(Q, dtype=np.float64)
    sensor = np.zeros(shape)
    fmid = np.zeros((prm.neq,) + shape)
    nn = (ctypes.c_int * 3)(*(list(n) + [1] * (3 - d)))
    dd = (ctypes.c_double * 3)(*(list(dx) + [1.0] * (3 - d)))
    st = L.hmo_midpoint_probe_single(ctypes.byref(prm.to_c()), nn, g, dd, Q.ctypes.data, axis, sensor.ctypes.data,
                                     fmid.ctypes.data)
    if st != 0:
        raise ValueError(f"oracle midpoint_probe_single status {st}")
    return sensor, fmid


def spectral_radiuses_and_dt(prm: Params, Q: np.ndarray, n: tuple, g: int, dx: tuple) -> np.ndarray:
    """Euler::computeSpectralRadiusesAndStableDtOnPatch of one patch (interior
    cells): [max (|u_a| + c) / dx_a for each axis a, 1 / max of their sum]."""
    L = lib()
    L.hmo_spectral_radiuses_and_dt.restype = ctypes.c_int
    L.hmo_spectral_radiuses_and_dt.argtypes = [ctypes.POINTER(_Params), ctypes.POINTER(ctypes.c_int), ctypes.c_int,
                                               ctypes.POINTER(ctypes.c_double), ctypes.c_void_p, ctypes.c_void_p]
    d = prm.dim
    Q = np.ascontiguousarray(Q, dtype=np.float64)
    out = np.zeros(d + 1)
    nn = (ctypes.c_int * 3)(*(list(n) + [1] * (3 - d)))
    dd = (ctypes.c_double * 3)(*(list(dx) + [1.0] * (3 - d)))
    st = L.hmo_spectral_radiuses_and_dt(ctypes.byref(prm.to_c()), nn, g, dd, Q.ctypes.data, out.ctypes.data)
    if st != 0:
        raise ValueError(f"oracle spectral_radiuses_and_dt status {st}")
    return out


def level_dt(patch_results, cfl: float = 1.0) -> float:
    """RungeKuttaLevelIntegrator::getLevelDt (RungeKuttaLevelIntegrator.cpp:1175-1346)
    from the patches' [s_0 .. s_{d-1}, dt_patch]: the per-axis maxima over the
    patches, dt = min over patches of dt_patch, then min(dt, 1 / sum of the
    maxima) (tbox Min: a < b ? a : b), times cfl = min(cfl_init, cfl)."""
    d = len(patch_results[0]) - 1
    s = [0.0] * d
    dt = float(np.finfo(np.float64).max)
    for r in patch_results:
        for a in range(d):
            s[a] = s[a] if s[a] >= r[a] else float(r[a])
        dt = dt if dt < r[d] else float(r[d])
    ssum = 0.0
    for a in range(d):
        ssum += s[a]
    dt = dt if dt < 1.0 / ssum else 1.0 / ssum
    return dt * cfl


def bounded_multi(prm: Params, axis: int, V) -> bool:
    """Bounds check of one side of a face normal to ``axis`` (four- or
    five-equation model)."""
    L = lib()
    L.hmo_bounded_multi.restype = ctypes.c_int
    L.hmo_bounded_multi.argtypes = [ctypes.POINTER(_Params), ctypes.c_int, ctypes.c_void_p]
    V = _arr(V, prm.neq)
    st = L.hmo_bounded_multi(ctypes.byref(prm.to_c()), axis, V.ctypes.data)
    if st < 0:
        raise ValueError("oracle bounded_multi: bad arguments")
    return bool(st)


def branch_tape(mode: int) -> int:
    """Branch tape of the single-species path: 1 records every branch decision
    (clearing the tape), 2 replays the recorded decisions from the start, 0
    switches the tape off.  Returns the tape length."""
    L = lib()
    L.hmo_branch_tape.restype = ctypes.c_int64
    L.hmo_branch_tape.argtypes = [ctypes.c_int]
    return int(L.hmo_branch_tape(int(mode)))


def branch_tape_status() -> tuple:
    """(length, replay position, overrun) of the branch tape."""
    L = lib()
    L.hmo_branch_tape_status.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    n, pos, ov = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int()
    L.hmo_branch_tape_status(ctypes.byref(n), ctypes.byref(pos), ctypes.byref(ov))
    return n.value, pos.value, bool(ov.value)


# Categories of branch decisions (bit c of branch_tape_free_mask's mask).
TAPE_RTAU, TAPE_SIGMA, TAPE_WAVE_SPEED, TAPE_SPM, TAPE_SSTAR, TAPE_HLL, TAPE_FALLBACK, TAPE_SENSOR = range(1, 9)


def branch_tape_free_mask(mask: int) -> None:
    """Categories (bit c = TAPE_* c) that a replay decides afresh, their tape
    entries still consumed; 0 replays every decision."""
    L = lib()
    L.hmo_branch_tape_free_mask.argtypes = [ctypes.c_int]
    L.hmo_branch_tape_free_mask(int(mask))


class frozen_branches:
    """Context manager: record the branch decisions of the first oracle run
    inside it, replay them for every later run (``replay()`` rewinds)."""

    def __enter__(self):
        branch_tape(1)
        return self

    def replay(self):
        n, pos, ov = branch_tape_status()
        branch_tape(2)
        return n

    def consumed(self) -> bool:
        n, pos, ov = branch_tape_status()
        return pos == n and not ov

    def __exit__(self, *exc):
        branch_tape(0)
        return False


class _Transport(ctypes.Structure):
    _fields_ = [(k, ctypes.c_double * 8) for k in ("mu", "mu_v", "kappa", "D", "M")]


@dataclass
class Transport:
    """Constant per-species transport properties (shear and bulk viscosity,
    thermal conductivity, mass diffusivity) and molecular weights."""
    mu: list = field(default_factory=lambda: [1e-3, 2e-3])
    mu_v: list = field(default_factory=lambda: [0.0, 0.0])
    kappa: list = field(default_factory=lambda: [1e-3, 1.5e-3])
    D: list = field(default_factory=lambda: [1e-3, 2e-3])
    M: list = field(default_factory=lambda: [1.0, 4.0])

    def to_c(self) -> _Transport:
        c = _Transport()
        for k in ("mu", "mu_v", "kappa", "D", "M"):
            vals = getattr(self, k)
            arr = getattr(c, k)
            for i in range(8):
                arr[i] = vals[i] if i < len(vals) else (1.0 if k == "M" else 0.0)
        return c


def diffusive_patch(prm: Params, tr: Transport, Q: np.ndarray, n: tuple, g: int, dx: tuple, dt: float):
    """computeDiffusiveFluxOnPatch (node SIXTH_ORDER) of the four-equation
    model on one patch; ``Q`` has ghost width ``g`` >= 6.  Returns the per-axis
    face flux arrays (dt-scaled)."""
    L = lib()
    L.hmo_diffusive_patch.restype = ctypes.c_int
    L.hmo_diffusive_patch.argtypes = [ctypes.POINTER(_Params), ctypes.POINTER(_Transport),
                                      ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                      ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                      ctypes.c_void_p]
    d = prm.dim
    neq = prm.neq
    n = tuple(int(v) for v in n)
    Q = np.ascontiguousarray(Q, dtype=np.float64)
    assert Q.shape == (prm.nq,) + tuple(reversed([v + 2 * g for v in n])), Q.shape
    F = [np.zeros(face_shape(neq, n, a)) for a in range(d)]
    nn = (ctypes.c_int * 3)(*(list(n) + [1] * (3 - d)))
    dd = (ctypes.c_double * 3)(*(list(dx) + [1.0] * (3 - d)))
    cp, ct = prm.to_c(), tr.to_c()
    st = L.hmo_diffusive_patch(ctypes.byref(cp), ctypes.byref(ct), nn, g, dd, dt, _ptr(Q), _ptr(F[0]),
                               _ptr(F[1]) if d > 1 else None, _ptr(F[2]) if d > 2 else None)
    if st != 0:
        raise RuntimeError(f"oracle hmo_diffusive_patch failed with status {st}")
    return F


# Physical boundary conditions (codes of include/hamers_convective.h).
BC_NONE, BC_FLOW, BC_REFLECT, BC_SYMMETRY, BC_DIRICHLET = range(5)


def fill_boundary(Qi: np.ndarray, g: int, periodic, conds, dirichlet=None, mom0: int = 1) -> np.ndarray:
    """Ghosted array of a one-patch domain (ncomp, *rev(n)) -> (ncomp, *rev(n + 2g))
    with the basic physical boundary conditions of
    BasicCartesianBoundaryUtilities2.cpp:215-440 (edges) / 590-900 (nodes) and
    BasicCartesianBoundaryUtilities3.cpp, applied to every conservative
    variable as EulerBoundaryConditions::setPhysicalBoundaryConditions does
    (EulerBoundaryConditions.cpp:749-1100):
      FLOW      the boundary cell (pivot interior_box_lo / hi, :285-310);
      REFLECT   the mirror cell lo + (fill_hi - i) / hi - (i - fill_lo), the
                vector variable's component normal to the side negated (:335-375);
                scalar variables take SYMMETRY (EulerBoundaryConditions.cpp:196-345);
      SYMMETRY  the mirror cell (:383-420);
      DIRICHLET the side's values (:262-276).
    conds[2a + hi] per side; periodic axes wrap.  Momentum components are
    mom0 .. mom0 + d - 1.  A cell outside the domain along several axes
    composes the axes' rules (the x-named node and edge conditions of the
    reference, whose pivot lies in the next axis's ghost box); with DIRICHLET
    among them the first such axis (x, y, z order) gives the values."""
    d = Qi.ndim - 1
    n = [Qi.shape[d - a] for a in range(d)]  # x first
    src, flip, side = [], [], []
    for a in range(d):
        x = np.arange(-g, n[a] + g)
        s = x.copy()
        f = np.zeros(x.size, dtype=bool)
        sd = np.full(x.size, -1)
        if periodic[a]:
            s = np.mod(x, n[a])
        else:
            for hi in (0, 1):
                c = conds[2 * a + hi]
                out = (x < 0) if hi == 0 else (x >= n[a])
                if c == BC_NONE or not np.any(out):
                    continue
                if c == BC_FLOW:
                    s[out] = 0 if hi == 0 else n[a] - 1
                elif c in (BC_REFLECT, BC_SYMMETRY):
                    s[out] = (-1 - x[out]) if hi == 0 else (2 * n[a] - 1 - x[out])
                    f[out] = c == BC_REFLECT
                elif c == BC_DIRICHLET:
                    s[out] = 0 if hi == 0 else n[a] - 1  # any interior cell: overwritten below
                    sd[out] = 2 * a + hi
                else:
                    raise ValueError(f"unknown boundary condition {c}")
            if np.any((s < 0) | (s >= n[a])):
                raise ValueError("cells outside the domain on a side without a condition")
        src.append(s)
        flip.append(f)
        side.append(sd)
    # gather: axis a of the array is numpy axis d - a
    idx = np.ix_(*[src[a] for a in reversed(range(d))])
    Q = np.stack([Qi[c][idx] for c in range(Qi.shape[0])])
    for a in range(d):
        if not np.any(flip[a]):
            continue
        shp = [1] * d
        shp[d - 1 - a] = flip[a].size
        sign = np.where(flip[a], -1.0, 1.0).reshape(shp)
        for b in range(d):
            if a == b:
                Q[mom0 + b] = Q[mom0 + b] * sign
    if any(np.any(sd >= 0) for sd in side):
        dirichlet = np.asarray(dirichlet, dtype=np.float64)
        first = np.full(Q.shape[1:], -1)
        for a in range(d):  # the first DIRICHLET axis in x, y, z order wins
            shp = [1] * d
            shp[d - 1 - a] = side[a].size
            sd = np.broadcast_to(side[a].reshape(shp), first.shape)
            first = np.where((first < 0) & (sd >= 0), sd, first)
        for sdv in np.unique(first[first >= 0]):
            m = first == sdv
            for c in range(Q.shape[0]):
                Q[c][m] = dirichlet[sdv][c]
    return Q
