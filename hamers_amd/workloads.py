"""Synthetic workloads for the convective-flux RHS (numpy, host side).

Arrays are component-major SAMRAI layouts in numpy C order: a d-dimensional
field with ``ncomp`` components is ``(ncomp, n_{d-1}, ..., n_0)`` (x last).

Initial conditions follow the reference's problem files where one exists
(``problems/Euler/initial_conditions/ConvergenceSingleSpecies.cpp:77-101,
150-177``) and otherwise the workload descriptions the benchmark is quoted on.
"""
from __future__ import annotations

import numpy as np


def ghost_shape(ncomp: int, n: tuple, g: int) -> tuple:
    return (ncomp,) + tuple(reversed([v + 2 * g for v in n]))


def fill_periodic(Qi: np.ndarray, g: int) -> np.ndarray:
    """Interior array (ncomp, *rev(n)) -> ghosted array with periodic halo."""
    d = Qi.ndim - 1
    pad = [(0, 0)] + [(g, g)] * d
    return np.pad(Qi, pad, mode="wrap")


def conservative_from_primitive_single(rho, vel, p, gamma):
    """(rho, [u, v, w], p) -> (rho, rho*u.., E) for the single-species model
    (E = p/(gamma-1) + 1/2 rho |u|^2)."""
    d = len(vel)
    Q = np.empty((d + 2,) + np.shape(rho))
    Q[0] = rho
    ke = np.zeros_like(rho)
    for b in range(d):
        Q[1 + b] = rho * vel[b]
        ke = ke + vel[b] * vel[b]
    Q[d + 1] = p / (gamma - 1.0) + 0.5 * rho * ke
    return Q


def cell_centres(n: tuple, xlo: tuple, dx: tuple):
    """Meshgrid of cell-centre coordinates, returned in reversed (numpy) order
    as a list [x, y, (z)] of arrays of shape rev(n)."""
    axes = [xlo[a] + (np.arange(n[a]) + 0.5) * dx[a] for a in range(len(n))]
    grids = np.meshgrid(*reversed(axes), indexing="ij")
    return list(reversed(grids))


def density_wave_convergence(N: int, dim: int = 2, gamma: float = 1.4):
    """Periodic density wave of the reference's convergence test on [-1,1]^d:
    rho = 1 + 1/2 sin(pi sum x), u = 1, p = 1
    (problems/Euler/initial_conditions/ConvergenceSingleSpecies.cpp:77-101)."""
    n = (N,) * dim
    dx = (2.0 / N,) * dim
    X = cell_centres(n, (-1.0,) * dim, dx)
    s = sum(X)
    rho = 1.0 + 0.5 * np.sin(np.pi * s)
    vel = [np.ones_like(rho) for _ in range(dim)]
    p = np.ones_like(rho)
    return conservative_from_primitive_single(rho, vel, p, gamma), n, dx


def density_wave_unit(N: int = 128, dim: int = 2, gamma: float = 1.4):
    """Smooth advected density wave on [0,1]^d: rho = 1 + 0.2 sin(2 pi sum x),
    u = 1, p = 1 (single 128^2 patch workload)."""
    n = (N,) * dim
    dx = (1.0 / N,) * dim
    X = cell_centres(n, (0.0,) * dim, dx)
    rho = 1.0 + 0.2 * np.sin(2.0 * np.pi * sum(X))
    vel = [np.ones_like(rho) for _ in range(dim)]
    p = np.ones_like(rho)
    return conservative_from_primitive_single(rho, vel, p, gamma), n, dx


def four_quadrant_riemann(N: int = 2048, seed: int = 1234, gamma: float = 1.4, noise: float = 1.0e-3):
    """2D four-quadrant Riemann data on [0,1]^2 split at (1/2,1/2): per quadrant
    (rho,u,v,p) with rho,p ~ U[0.1,1.5], u,v ~ U[-0.8,0.8]; then additive white
    noise noise*U[-1,1] on every primitive variable."""
    rng = np.random.default_rng(seed)
    quad = {}
    for q in range(4):
        rho = rng.uniform(0.1, 1.5)
        u = rng.uniform(-0.8, 0.8)
        v = rng.uniform(-0.8, 0.8)
        p = rng.uniform(0.1, 1.5)
        quad[q] = (rho, u, v, p)
    n = (N, N)
    dx = (1.0 / N, 1.0 / N)
    X, Y = cell_centres(n, (0.0, 0.0), dx)
    qi = (X >= 0.5).astype(int) + 2 * (Y >= 0.5).astype(int)
    prim = np.empty((4, N, N))
    for q in range(4):
        for c in range(4):
            prim[c][qi == q] = quad[q][c]
    prim += noise * rng.uniform(-1.0, 1.0, size=prim.shape)
    return conservative_from_primitive_single(prim[0], [prim[1], prim[2]], prim[3], gamma), n, dx


def kelvin_helmholtz(N: int = 512, amplitudes=None, seed: int = 7, gamma: float = 1.4):
    """Periodic Kelvin-Helmholtz double shear layer on [0,1]^2 with 8 sinusoidal
    transverse-velocity perturbation modes whose amplitudes are the parameters.
    u = U0 tanh((y-1/4)/delta) - U0 tanh((y-3/4)/delta) - U0 (periodic pair),
    v = sum_k a_k sin(2 pi k x + phi_k) * envelope(y), rho = 1 (+ 1 inside), p = 2.5."""
    rng = np.random.default_rng(seed)
    default_amplitudes = 1.0e-2 * rng.uniform(0.5, 1.0, size=8)  # drawn always: the phases follow
    if amplitudes is None:
        amplitudes = default_amplitudes
    phases = rng.uniform(0.0, 2.0 * np.pi, size=8)
    n = (N, N)
    dx = (1.0 / N, 1.0 / N)
    X, Y = cell_centres(n, (0.0, 0.0), dx)
    delta = 0.02
    U0 = 0.5
    band = 0.5 * (np.tanh((Y - 0.25) / delta) - np.tanh((Y - 0.75) / delta))
    u = U0 * (2.0 * band - 1.0)
    rho = 1.0 + band
    env = np.exp(-((Y - 0.25) / 0.05) ** 2) + np.exp(-((Y - 0.75) / 0.05) ** 2)
    v = np.zeros_like(X)
    for k in range(8):
        v = v + amplitudes[k] * np.sin(2.0 * np.pi * (k + 1) * X + phases[k]) * env
    p = 2.5 * np.ones_like(X)
    return conservative_from_primitive_single(rho, [u, v], p, gamma), n, dx, np.asarray(amplitudes), phases


def kelvin_helmholtz_mode_derivatives(N: int, Q0: np.ndarray, seed: int = 7):
    """dQ0/da_k of kelvin_helmholtz: v = sum_k a_k s_k(x, y) enters rho v and
    E = p/(gamma-1) + rho (u^2 + v^2)/2, so d(rho v)/da_k = rho s_k and
    dE/da_k = rho v s_k.  Returns an array (8, 4, N, N)."""
    rng = np.random.default_rng(seed)
    rng.uniform(0.5, 1.0, size=8)  # the default amplitudes drawn first
    phases = rng.uniform(0.0, 2.0 * np.pi, size=8)
    n = (N, N)
    dx = (1.0 / N, 1.0 / N)
    X, Y = cell_centres(n, (0.0, 0.0), dx)
    env = np.exp(-((Y - 0.25) / 0.05) ** 2) + np.exp(-((Y - 0.75) / 0.05) ** 2)
    rho = Q0[0]
    v = Q0[2] / rho
    out = np.zeros((8,) + Q0.shape)
    for k in range(8):
        sk = np.sin(2.0 * np.pi * (k + 1) * X + phases[k]) * env
        out[k, 2] = rho * sk
        out[k, 3] = rho * v * sk
    return out


def _smooth_mode_shapes(N: int, seed: int):
    """Phases and the 8 transverse-velocity mode shapes of ``smooth_modes``."""
    rng = np.random.default_rng(seed)
    default_amplitudes = 0.05 * rng.uniform(0.5, 1.0, size=8)  # drawn always: the phases follow
    phases = rng.uniform(0.0, 2.0 * np.pi, size=8)
    X, Y = cell_centres((N, N), (0.0, 0.0), (1.0 / N, 1.0 / N))
    shapes = [np.sin(2.0 * np.pi * ((k % 4 + 1) * X + (k // 4 + 1) * Y) + phases[k]) for k in range(8)]
    return default_amplitudes, phases, X, Y, shapes


def smooth_modes(N: int = 64, amplitudes=None, seed: int = 11, gamma: float = 1.4):
    """Smooth periodic flow on [0,1]^2 in which every primitive variable varies
    along both axes at O(0.1) amplitude, plus 8 transverse-velocity modes
    sin(2 pi (kx x + ky y) + phi_k) whose amplitudes are the parameters.
    No region is near-uniform, so the scheme's smoothness indicators stay far
    above rounding level and the final state is a smooth function of the
    amplitudes (a gradient check by differences resolves ~1e-9)."""
    default_amplitudes, phases, X, Y, shapes = _smooth_mode_shapes(N, seed)
    if amplitudes is None:
        amplitudes = default_amplitudes
    tp = 2.0 * np.pi
    rho = 1.0 + 0.2 * np.sin(tp * (X + 2.0 * Y))
    u = 0.3 + 0.2 * np.sin(tp * (2.0 * X + Y) + 1.0)
    v = 0.2 + 0.1 * np.cos(tp * (X - Y) + 2.0)
    for k in range(8):
        v = v + amplitudes[k] * shapes[k]
    p = 1.0 + 0.1 * np.cos(tp * (3.0 * X + Y))
    return (conservative_from_primitive_single(rho, [u, v], p, gamma), (N, N), (1.0 / N, 1.0 / N),
            np.asarray(amplitudes), phases)


def smooth_modes_derivatives(N: int, Q0: np.ndarray, seed: int = 11):
    """dQ0/da_k of ``smooth_modes`` (v is linear in a; rho v and E = p/(gamma-1)
    + rho |u|^2/2 follow).  Returns an array (8, 4, N, N)."""
    _, _, _, _, shapes = _smooth_mode_shapes(N, seed)
    rho = Q0[0]
    v = Q0[2] / rho
    out = np.zeros((8,) + Q0.shape)
    for k in range(8):
        out[k, 2] = rho * shapes[k]
        out[k, 3] = rho * v * shapes[k]
    return out


def split_patches(Q_global_interior: np.ndarray, patch: tuple, g: int):
    """Split a periodic global interior field into patches of extent ``patch``
    (same rank), each returned with its ghost halo filled from its neighbours.

    Returns a list of (offset, Q_patch) with offset in (x, y, z) order."""
    d = Q_global_interior.ndim - 1
    N = tuple(reversed(Q_global_interior.shape[1:]))
    Qg = fill_periodic(Q_global_interior, g)
    out = []
    counts = [N[a] // patch[a] for a in range(d)]
    for idx in np.ndindex(*reversed(counts)):
        off = tuple(reversed([i * p for i, p in zip(idx, reversed(patch))]))
        sl = [slice(None)] + [slice(off[a], off[a] + patch[a] + 2 * g) for a in reversed(range(d))]
        out.append((off, np.ascontiguousarray(Qg[tuple(sl)])))
    return out


def conservative_from_primitive_five(Z_rho, vel, p, Z, gamma):
    """Five-equation Allaire model: (Z_i rho_i, u, p, Z_i (all ns)) ->
    (Z_i rho_i, rho u, E, Z_i (all ns)); E = p/(gamma_mix - 1) + 1/2 rho |u|^2
    with 1/(gamma_mix - 1) = sum Z_i / (gamma_i - 1) (isobaric mixing)."""
    ns = len(Z_rho)
    d = len(vel)
    rho = sum(Z_rho)
    xi = sum(Z[i] / (gamma[i] - 1.0) for i in range(ns))
    ke = sum(v * v for v in vel)
    Q = np.empty((2 * ns + d + 1,) + np.shape(rho))
    for i in range(ns):
        Q[i] = Z_rho[i]
    for b in range(d):
        Q[ns + b] = rho * vel[b]
    gamma_m = 1.0 / xi + 1.0
    Q[ns + d] = p / (gamma_m - 1.0) + 0.5 * rho * ke
    for i in range(ns):
        Q[ns + d + 1 + i] = Z[i]
    return Q


def interface_advection_convergence(N: int, dim: int = 2):
    """Periodic material-interface advection of the reference's five-equation
    convergence test on [-1,1]^d: Z_1 = 1/2 + 1/4 sin(pi sum x), rho_1 = 2,
    rho_2 = 1, gamma = (1.6, 1.4), u = 1, p = 1
    (problems/Euler/initial_conditions/ConvergenceFiveEqnAllaire.cpp:77-130)."""
    n = (N,) * dim
    dx = (2.0 / N,) * dim
    X = cell_centres(n, (-1.0,) * dim, dx)
    Z1 = 0.5 + 0.25 * np.sin(np.pi * sum(X))
    Z2 = 1.0 - Z1
    Q = conservative_from_primitive_five([Z1 * 2.0, Z2 * 1.0], [np.ones_like(Z1)] * dim, np.ones_like(Z1),
                                         [Z1, Z2], (1.6, 1.4))
    return Q, n, dx


def species_heat_capacities(gamma, R):
    """c_p,i = gamma_i/(gamma_i-1) R_i and c_v,i = 1/(gamma_i-1) R_i
    (EquationOfStateMixingRulesIdealGas.cpp:108-120)."""
    cp = [g / (g - 1.0) * r for g, r in zip(gamma, R)]
    cv = [1.0 / (g - 1.0) * r for g, r in zip(gamma, R)]
    return cp, cv


def mixture_gamma_mass(Y, gamma, R):
    """Mixture ratio of specific heats with mass fractions (ISOTHERMAL_AND_ISOBARIC
    closure): sum Y_i c_p,i / sum Y_i c_v,i."""
    cp, cv = species_heat_capacities(gamma, R)
    c_p = sum(Y[i] * cp[i] for i in range(len(gamma)))
    c_v = sum(Y[i] * cv[i] for i in range(len(gamma)))
    return c_p / c_v


def conservative_from_primitive_four(rho_Y, vel, p, gamma, R):
    """(rho Y_i, [u, v, w], p) -> (rho Y_i, rho u.., E) for the four-equation
    conservative model, E = p/(gamma_mix-1) + 1/2 rho |u|^2."""
    ns, d = len(rho_Y), len(vel)
    rho = sum(rho_Y)
    Y = [r / rho for r in rho_Y]
    gm = mixture_gamma_mass(Y, gamma, R)
    Q = np.empty((ns + d + 1,) + np.shape(rho))
    for i in range(ns):
        Q[i] = rho_Y[i]
    ke = np.zeros_like(rho)
    for b in range(d):
        Q[ns + b] = rho * vel[b]
        ke = ke + vel[b] * vel[b]
    Q[ns + d] = p / (gm - 1.0) + 0.5 * rho * ke
    return Q


FOUR_EQN_GAMMA = (1.6, 1.4)
FOUR_EQN_R = (1.0, 2.0)


def species_advection_convergence(N: int, dim: int = 2, gamma=FOUR_EQN_GAMMA, R=FOUR_EQN_R):
    """Two ideal gases of different gamma and R advected on [-1,1]^d at u = 1,
    p = 1: rho = 1 + 1/2 sin(pi sum x), Y_1 = 1/2 + 1/4 cos(pi sum x).  With
    uniform velocity and pressure the exact solution is a pure translation."""
    n = (N,) * dim
    dx = (2.0 / N,) * dim
    X = cell_centres(n, (-1.0,) * dim, dx)
    s = sum(X)
    rho = 1.0 + 0.5 * np.sin(np.pi * s)
    Y1 = 0.5 + 0.25 * np.cos(np.pi * s)
    Q = conservative_from_primitive_four([rho * Y1, rho * (1.0 - Y1)], [np.ones_like(rho)] * dim, np.ones_like(rho),
                                         gamma, R)
    return Q, n, dx
