"""Two-level (refinement ratio 2) advance of the oracle: the restatement the
GPU hierarchy (hm_hierarchy_advance) is checked against.  Test
infrastructure, not product code.

Follows RungeKuttaLevelIntegrator's time-refined advance
(RungeKuttaLevelIntegrator.cpp:1457-1939 per level, synchronizeLevelWithCoarser
:2131-2225): the coarse level (one patch holding the whole periodic domain)
takes one SSP-RK3 step of dt with gamma-weighted flux sums
(Euler.cpp:1230-1270); the fine level (a list of patches in the refined index
space) takes two steps of dt/2, its coarse-fine ghost cells filled before each
step from the coarse level -- linear in time between the coarse states at the
start and end of the step, refined by conservative linear interpolation with
monotonized-central slopes (the published form of SAMRAI's
CONSERVATIVE_LINEAR_REFINE, which the flow models register,
FlowModelSingleSpecies.cpp:578-597; SAMRAI itself is not part of the
reference) -- and kept through the intermediate stages (same-level fills only,
:2513-2520); then the coarse faces on the coarse-fine boundary take the area
average of the fine flux sums -- each fine step's gamma-weighted flux added to
a sum that starts at zero (RungeKuttaLevelIntegrator.cpp:2920-2967, 3060-3180,
algs_upfluxsum2d.m4) -- as the sum of the 2^(d-1) fine faces times 2^(1-d)
(SAMRAI's outerface coarsening is not part of the reference; this is its
area average written out), the
coarse update is repeated from the start state with the flux sums
(Euler::synchronizeFluxes, Euler.cpp:1682-1943) and coarse cells under fine
patches take the average of their 2^d fine cells (CONSERVATIVE_COARSEN).
With transport properties (four-equation Navier-Stokes form) both levels add
the diffusive flux to every stage (NavierStokes::advanceSingleStepOnPatch), the
diffusive flux is a second flux variable with its own sums, refluxed the same
way (it is registered as FLUX with CONSERVATIVE_COARSEN, NavierStokes.cpp:
393-402), and the coarse update of the synchronization is the Navier-Stokes
form Q += (-(F_c,R - F_c,L + F_d,R - F_d,L)/dx_0 - ... + S)
(NavierStokes::synchronizeFluxes, NavierStokes.cpp:2424-2700).
Every arithmetic step keeps the operation order of the GPU kernels it checks.
"""
from __future__ import annotations

import numpy as np

from . import oracle as O

RK_ALPHA = [[1.0], [0.75, 0.25], [1.0 / 3.0, 0.0, 2.0 / 3.0]]
RK_BETA = [[1.0], [0.0, 0.25], [0.0, 0.0, 2.0 / 3.0]]
RK_GAMMA = [[1.0 / 6.0], [0.0, 1.0 / 6.0], [0.0, 0.0, 2.0 / 3.0]]


def mc_slope(vL, v0, vR):
    """Monotonized-central limited slope per unit cell: 0 at extrema, else the
    central difference limited to twice the smaller one-sided difference."""
    dR = vR - v0
    dL = v0 - vL
    c = 0.5 * (dR + dL)
    b = 2.0 * np.minimum(np.abs(dR), np.abs(dL))
    return np.where(dR * dL > 0.0, np.copysign(np.minimum(np.abs(c), b), c), 0.0)


def _divergence(F, S, dx, Fd=None):
    """-(sum_a (F_a,R - F_a,L)/dx_a) + S in the stage update's order; with
    diffusive fluxes Fd the Navier-Stokes form -(sum_a (F_a,R - F_a,L + Fd_a,R -
    Fd_a,L)/dx_a) + S."""
    d = len(F)
    acc = None
    for a in range(d):
        Fa = F[a]
        ax = Fa.ndim - 1 - a
        R = [slice(None)] * Fa.ndim
        L = [slice(None)] * Fa.ndim
        R[ax] = slice(1, None)
        L[ax] = slice(0, -1)
        dF = Fa[tuple(R)] - Fa[tuple(L)]
        if Fd is not None:
            dF = dF + Fd[a][tuple(R)] - Fd[a][tuple(L)]
        t = dF / dx[a]
        acc = -t if acc is None else acc - t
    return acc + S


class TwoLevel:
    """prm: oracle Params; coarse domain n (x first); fine patch boxes (lo, n)
    in the refined index space (even lo and n).  bc = (periodic per axis, side
    conditions, DIRICHLET values) as oracle.fill_boundary takes them; None:
    periodic on every axis.  transport: oracle Transport for the four-equation
    Navier-Stokes form (6 ghost cells), None for the Euler form."""

    def __init__(self, prm, n, dx, fine_boxes, g=4, bc=None, transport=None):
        self.tr = transport
        if transport is not None:
            g = max(g, 6)
        self.prm, self.d, self.g = prm, len(n), g
        self.bc = bc
        self.mom0 = 1 if prm.model == O.SINGLE_SPECIES else prm.num_species
        self.n = tuple(n)
        self.dx = tuple(dx)
        self.nf = tuple(2 * v for v in n)
        self.dxf = tuple(0.5 * v for v in dx)
        self.boxes = [(tuple(lo), tuple(m)) for lo, m in fine_boxes]
        self.nq = prm.nq
        self.neq = prm.neq
        # coverage of the fine domain: patch index or -1
        self.cover = np.full(tuple(reversed(self.nf)), -1, dtype=int)
        for p, (lo, m) in enumerate(self.boxes):
            self.cover[self._sl(lo, m)] = p

    @staticmethod
    def _sl(lo, m):
        return tuple(slice(lo[a], lo[a] + m[a]) for a in reversed(range(len(lo))))

    def _gather(self, G, idx_per_axis):
        """G[(comp,) + wrapped indices] for index lists per axis (x first)."""
        return G[(slice(None),) + np.ix_(*[idx_per_axis[a] for a in reversed(range(self.d))])]

    # ---- ghost fills ------------------------------------------------------
    def fill(self, Q):
        """Ghosted array of a whole-level array: periodic images, or the
        physical boundary conditions (oracle.fill_boundary)."""
        g, d = self.g, self.d
        if self.bc is None:
            return np.pad(Q, [(0, 0)] + [(g, g)] * d, mode="wrap")
        periodic, conds, dirichlet = self.bc
        return O.fill_boundary(Q, g, periodic, conds, dirichlet, self.mom0)

    def fine_ghosted_all(self, pts, CF):
        """Ghost boxes of every fine patch: interiors from the patches, cells
        under no fine patch from CF (the coarse-fine values of this fine step,
        on the whole fine domain), then the level's periodic images or physical
        boundary conditions.  A cell inside the domain has one value in every
        ghost box that holds it (same-level ghosts copy interiors, coarse-fine
        ghosts depend on the position only), so one whole-domain array gives
        them all; the boundary fill reads only cells inside the domain."""
        g, d = self.g, self.d
        G = CF.copy()
        for (lo, m), q in zip(self.boxes, pts):
            G[(slice(None),) + self._sl(lo, m)] = q
        Gg = self.fill(G)
        return [Gg[(slice(None),) + self._sl(lo, [v + 2 * g for v in m])].copy() for lo, m in self.boxes]

    def refine(self, Cold_g, Cnew_g, alpha, xw):
        """Refined values at the fine cells xw (wrapped fine indices per axis)
        from the coarse ghosted arrays at time fraction alpha."""
        g, d = self.g, self.d
        par = [xw[a] // 2 for a in range(d)]
        child = [xw[a] - 2 * par[a] for a in range(d)]

        def val(shift_axis, sh):
            idx = [par[a] + g + (sh if a == shift_axis else 0) for a in range(d)]
            o = self._gather(Cold_g, idx)
            return o + alpha * (self._gather(Cnew_g, idx) - o)

        v0 = val(-1, 0)
        v = v0.copy()
        for a in range(d):
            sl = mc_slope(val(a, -1), v0, val(a, 1))
            shp = [1] * d
            shp[d - 1 - a] = child[a].size
            off = np.where(child[a] == 1, 0.25, -0.25).reshape(shp)
            v = v + sl * off
        return v

    def refine_patches(self, Qc):
        """Fine patch interiors refined from the coarse state (the fine level's
        initial data in the tests)."""
        Cg = self.fill(Qc)
        return [self.refine(Cg, Cg, 0.0, [np.arange(lo[a], lo[a] + m[a]) for a in range(self.d)])
                for lo, m in self.boxes]

    def composite_sum(self, Qc, pts, ncomp):
        """Sum of the first ncomp components over the composite grid (coarse
        cells not under the fine level, then the fine cells), times the cell volumes."""
        cov = self.covered_coarse()
        s = Qc[:ncomp][:, ~cov].sum(axis=1) * float(np.prod(self.dx))
        for q in pts:
            s = s + q[:ncomp].reshape(ncomp, -1).sum(axis=1) * float(np.prod(self.dxf))
        return s

    # ---- one level step ----------------------------------------------------
    def _step_level(self, patches, dt, dxl, fill, acc):
        """SSP-RK3 step of a list of patch states; fill(sn, interiors, ghosted
        stage-0 inputs) -> ghosted inputs of stage sn; acc: per patch the flux,
        source and diffusive-flux sums (added to; the last is None in the Euler
        form)."""
        nst = 3
        Qint = []
        R = []
        prev = None
        for sn in range(nst):
            Qg = fill(sn, patches, prev)
            prev = Qg
            Qint.append([q[(slice(None),) + self._interior_sl(q)] for q in Qg])
            Rs = []
            for i, q in enumerate(Qg):
                n = tuple(reversed(q.shape[1:]))
                n = tuple(v - 2 * self.g for v in n)
                qc = np.ascontiguousarray(q)
                F, S = O.rhs_patch(self.prm, qc, n, self.g, dxl, dt)
                Fd = O.diffusive_patch(self.prm, self.tr, qc, n, self.g, dxl, dt) if self.tr is not None else None
                Rs.append(_divergence(F, S[:self.neq], dxl, Fd))
                Facc, Sacc, Fdacc = acc[i]
                for a in range(self.d):
                    if RK_GAMMA[sn][sn] != 0.0:
                        Facc[a] += RK_GAMMA[sn][sn] * F[a]
                        if Fd is not None:
                            Fdacc[a] += RK_GAMMA[sn][sn] * Fd[a]
                if RK_GAMMA[sn][sn] != 0.0:
                    Sacc += RK_GAMMA[sn][sn] * S[:self.neq]
            R.append(Rs)
            new = []
            for i in range(len(patches)):
                q = np.zeros_like(patches[i])
                for mm in range(sn + 1):
                    if RK_ALPHA[sn][mm] != 0.0:
                        q[:self.neq] = q[:self.neq] + RK_ALPHA[sn][mm] * Qint[mm][i][:self.neq]
                    if RK_BETA[sn][mm] != 0.0:
                        q[:self.neq] = q[:self.neq] + RK_BETA[sn][mm] * R[mm][i]
                if self.nq > self.neq:
                    ns = self.prm.num_species
                    q[-1] = 1.0
                    for si in range(ns - 1):
                        q[-1] -= q[self.neq - (ns - 1) + si]
                new.append(q)
            patches = new
        return patches

    def _interior_sl(self, q):
        return tuple(slice(self.g, s - self.g) for s in q.shape[1:])

    # ---- two-level advance --------------------------------------------------
    def _sums(self, n):
        """Zero flux, source and (Navier-Stokes form) diffusive-flux sums of a patch of extents n."""
        d = self.d
        F = [np.zeros(O.face_shape(self.neq, n, a)) for a in range(d)]
        Fd = [np.zeros(O.face_shape(self.neq, n, a)) for a in range(d)] if self.tr is not None else None
        return (F, np.zeros((self.neq,) + tuple(reversed(n))), Fd)

    def advance(self, Qc, Qf_patches, dt, reflux=True, reflux_diffusive=True):
        """One coarse step (with two fine steps): Qc (nq, coarse domain),
        Qf_patches: interior arrays of the fine patches.  Returns the new pair."""
        g, d = self.g, self.d
        fill_c = self.fill
        Cold_g = fill_c(Qc)
        facc_c = self._sums(self.n)

        def fill_coarse(sn, pts, prev):
            # the whole level as one patch: periodic images or the boundary
            # conditions set every ghost cell
            return [fill_c(pts[0])]

        Qc_new = self._step_level([Qc], dt, self.dx, fill_coarse, [facc_c])[0]
        Cnew_g = fill_c(Qc_new)
        fsum = None  # fine flux sums: (0 + F of the first fine step) + F of the second
        fsumd = None  # the same of the diffusive flux (Navier-Stokes form)
        pts = [q.copy() for q in Qf_patches]
        for k in range(2):
            facc_f = [self._sums(m) for _, m in self.boxes]
            alpha = 0.5 * k

            # the coarse-fine values of this fine step on the whole fine domain,
            # kept through its intermediate stages
            CF = self.refine(Cold_g, Cnew_g, alpha, [np.arange(self.nf[a]) for a in range(d)])

            def fill_fine(sn, ps, prev):
                return self.fine_ghosted_all(ps, CF)

            pts = self._step_level(pts, 0.5 * dt, self.dxf, fill_fine, facc_f)
            if fsum is None:
                fsum = [[0.0 + Fa for Fa in F] for F, _, _ in facc_f]
            else:
                fsum = [[fs + Fa for fs, Fa in zip(fsp, F)] for fsp, (F, _, _) in zip(fsum, facc_f)]
            if self.tr is not None:
                if fsumd is None:
                    fsumd = [[0.0 + Fa for Fa in Fd] for _, _, Fd in facc_f]
                else:
                    fsumd = [[fs + Fa for fs, Fa in zip(fsp, Fd)] for fsp, (_, _, Fd) in zip(fsumd, facc_f)]
        # synchronization
        Fc, Sc, Fdc = facc_c
        if reflux:
            self._reflux(Fc, fsum)
            if self.tr is not None and reflux_diffusive:
                self._reflux(Fdc, fsumd)
        Qs = Qc.copy()
        Qs[:self.neq] = Qc[:self.neq] + _divergence(Fc, Sc, self.dx, Fdc)
        if self.nq > self.neq:
            # last volume fraction = 1 - the others, in species order
            # (FlowModelFiveEqnAllaire::updateCellDataOfConservativeVariables, Euler.cpp:1937-1939)
            ns = self.prm.num_species
            Qs[-1] = 1.0
            for si in range(ns - 1):
                Qs[-1] -= Qs[self.neq - (ns - 1) + si]
        self._coarsen(Qs, pts)
        return Qs, pts

    def covered_coarse(self):
        """Coarse cells under a fine patch (x last)."""
        return self.cover[tuple(slice(0, None, 2) for _ in range(self.d))] >= 0

    def _reflux(self, Fc, fsum):
        d = self.d
        covc = self.covered_coarse()
        for a in range(d):
            ax = d - 1 - a
            F = Fc[a]
            nfaces = F.shape[1:]
            it = np.ndindex(*nfaces)
            for idx in it:  # idx in numpy order (z, y, x) of the face array
                fidx = list(reversed(idx))  # x first
                R = list(fidx)
                L = list(fidx)
                L[a] -= 1
                if self.bc is not None and not self.bc[0][a] and (L[a] < 0 or R[a] >= self.n[a]):
                    continue  # a face on the physical boundary
                Rw = [R[b] % self.n[b] for b in range(d)]
                Lw = [L[b] % self.n[b] for b in range(d)]
                cR = covc[tuple(reversed(Rw))]
                cL = covc[tuple(reversed(Lw))]
                if cR == cL:
                    continue
                cell = Lw if cL else Rw
                p = self.cover[tuple(reversed([2 * v for v in cell]))]
                lo, m = self.boxes[p]
                face_global = 2 * (Lw[a] + 1) if cL else 2 * Rw[a]
                trans = [b for b in range(d) if b != a]
                vals = []
                for t in range(1 << (d - 1)):
                    fx = [0] * d
                    for bit, b in enumerate(trans):
                        fx[b] = 2 * cell[b] + ((t >> bit) & 1) - lo[b]
                    fx[a] = face_global - lo[a]
                    vals.append(fsum[p][a][(slice(None),) + tuple(reversed(fx))])
                s = vals[0]
                for v in vals[1:]:
                    s = s + v
                F[(slice(None),) + idx] = s * (1.0 / len(vals))

    def _coarsen(self, Qs, pts):
        d = self.d
        for (lo, m), q in zip(self.boxes, pts):
            kids = []
            for c in range(1 << d):
                sl = tuple(slice((c >> a) & 1, None, 2) for a in reversed(range(d)))
                kids.append(q[(slice(None),) + sl])
            s = kids[0]
            for k in kids[1:]:
                s = s + k
            clo = [v // 2 for v in lo]
            cm = [v // 2 for v in m]
            Qs[(slice(None),) + self._sl(clo, cm)] = s * (1.0 / (1 << d))
