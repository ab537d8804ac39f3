"""Shared helpers of the test suite (oracle-side reference computations)."""
from __future__ import annotations

import numpy as np

from oracle import oracle as O
from hamers_amd import workloads as W

# SSP-RK3 weights of RungeKuttaLevelIntegrator.cpp:3907-3924.
RK_ALPHA = [[1.0], [0.75, 0.25], [1.0 / 3.0, 0.0, 2.0 / 3.0]]
RK_BETA = [[1.0], [0.0, 0.25], [0.0, 0.0, 2.0 / 3.0]]


def flux_divergence(F, S, dx, Fd=None):
    """-(sum_a (F_a,R - F_a,L)/dx_a) + S, with the reference's association
    (Euler.cpp:1212, 1344-1347, 1544-1548); with diffusive fluxes Fd the
    Navier-Stokes form -(sum_a (F_a,R - F_a,L + Fd_a,R - Fd_a,L)/dx_a) + S
    (NavierStokes.cpp:1743-1748, 2085-2092)."""
    d = len(F)
    acc = None
    for a in range(d):
        Fa = F[a]
        ax = Fa.ndim - 1 - a
        R = [slice(None)] * Fa.ndim
        Lf = [slice(None)] * Fa.ndim
        R[ax] = slice(1, None)
        Lf[ax] = slice(0, -1)
        dF = Fa[tuple(R)] - Fa[tuple(Lf)]
        if Fd is not None:
            dF = dF + Fd[a][tuple(R)] - Fd[a][tuple(Lf)]
        t = dF / dx[a]
        acc = -t if acc is None else acc - t
    return acc + S


def ghost_fill(Qi, g, prm, bc=None):
    """Ghosted one-patch domain: periodic, or with the physical boundary
    conditions bc = (periodic per axis, side conditions, dirichlet values)."""
    if bc is None:
        return W.fill_periodic(Qi, g)
    periodic, conds, dirichlet = bc
    mom0 = 1 if prm.model == O.SINGLE_SPECIES else prm.num_species
    return O.fill_boundary(Qi, g, periodic, conds, dirichlet, mom0)


def oracle_rhs_periodic(prm, Qi, n, dx, dt, g=4, transport=None, bc=None):
    """Oracle RHS (flux divergence, dt-scaled) on a periodic domain treated as
    one patch (or with physical boundary conditions ``bc``, see ghost_fill);
    for the five-equation model the source is part of the RHS and the last
    volume fraction's row is zero.  With ``transport`` (four-equation model)
    the diffusive flux is added in the Navier-Stokes form."""
    if transport is not None:
        g = max(g, 6)
    Qg = ghost_fill(Qi, g, prm, bc)
    F, S = O.rhs_patch(prm, Qg, n, g, dx, dt)
    Fd = O.diffusive_patch(prm, transport, Qg, n, g, dx, dt) if transport is not None else None
    R = flux_divergence(F, S, dx, Fd)
    if prm.nq > prm.neq:
        R = np.concatenate([R, np.zeros_like(R[:1])], axis=0)
    return R, F, S


def oracle_rk3(prm, Qi, n, dx, dt, nsteps, alpha=RK_ALPHA, beta=RK_BETA, transport=None, bc=None):
    """SSP-RK3 steps on a periodic domain with the oracle RHS; the stage
    update follows Euler::advanceSingleStepOnPatch (Euler.cpp:1003-1680), or
    with ``transport`` NavierStokes::advanceSingleStepOnPatch."""
    Q = Qi.copy()
    for _ in range(nsteps):
        Qint, R = [], []
        # (the zero-fill of Euler::advanceSingleStepOnPatch covers all stored arrays)
        for sn in range(len(alpha)):
            Qint.append(Q.copy())
            R.append(oracle_rhs_periodic(prm, Q, n, dx, dt, transport=transport, bc=bc)[0])
            Qn = np.zeros_like(Q)
            for m in range(sn + 1):
                if alpha[sn][m] != 0.0:
                    Qn = Qn + alpha[sn][m] * Qint[m]
                if beta[sn][m] != 0.0:
                    Qn = Qn + beta[sn][m] * R[m]
            if prm.nq > prm.neq:
                # last volume fraction = 1 - sum of the others (FlowModelFiveEqnAllaire.cpp:1739-1890)
                ns = prm.num_species
                Qn[-1] = 1.0
                for si in range(ns - 1):
                    Qn[-1] -= Qn[prm.neq - (ns - 1) + si]
            Q = Qn
    return Q


def max_rel_err(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    scale = max(np.max(np.abs(b)), 1e-300)
    return float(np.max(np.abs(a - b)) / scale)
