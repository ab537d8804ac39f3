"""Pins of the oracle's stable time step: Euler::computeSpectralRadiusesAndStableDtOnPatch
(Euler.cpp:489-888) and the level reduction of RungeKuttaLevelIntegrator::getLevelDt
(RungeKuttaLevelIntegrator.cpp:1175-1346).

Closed forms: on a uniform state the spectral radius of axis a is
(|u_a| + c) / dx_a with c = sqrt(gamma p / rho) (single species; for the
mixtures c^2 = gamma_mix p / rho, since both mixing rules' c^2 = Gamma p / rho +
sum Y_i Psi_i reduce to that) and the patch dt is 1 / their sum.  On a state
whose fastest x- and y-waves sit in different cells the patch dt is
1 / max over cells of the sum, and the level dt 1 / (the sum of the per-axis
maxima) is strictly smaller; a planted mistake (the max of the axes' radiuses
instead of their sum) is caught."""
import numpy as np
import pytest

from oracle import oracle as O
from hamers_amd import workloads as W

G = 4


def _single(dim, shape, rho, vel, p):
    return W.conservative_from_primitive_single(np.full(shape, rho), [np.full(shape, v) for v in vel],
                                                np.full(shape, p), 1.4)


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_uniform_state_single(dim):
    n = (6, 5, 4)[:dim]
    dx = (0.1, 0.05, 0.2)[:dim]
    vel = (0.4, -0.7, 0.25)[:dim]
    rho, p = 1.3, 0.9
    Q = _single(dim, tuple(reversed([v + 2 * G for v in n])), rho, vel, p)
    out = O.spectral_radiuses_and_dt(O.Params(dim=dim, scheme=1), Q, n, G, dx)
    c = np.sqrt(1.4 * p / rho)
    s = [(abs(vel[a]) + c) / dx[a] for a in range(dim)]
    assert np.allclose(out[:dim], s, rtol=1e-14)
    assert out[dim] == pytest.approx(1.0 / sum(s), rel=1e-14)
    assert O.level_dt([out], cfl=0.4) == pytest.approx(0.4 / sum(s), rel=1e-14)


def _two_hot_cells(mutation=0):
    """Fluid at rest (c = 1) with one cell moving at u = 2 and another at v = 3."""
    n, dx = (8, 8), (0.1, 0.2)
    shape = (8 + 2 * G, 8 + 2 * G)
    rho = np.ones(shape)
    u = np.zeros(shape)
    v = np.zeros(shape)
    p = np.full(shape, 1.0 / 1.4)  # c = 1
    u[G + 2, G + 5] = 2.0          # cell (x=5, y=2)
    v[G + 6, G + 1] = 3.0          # cell (x=1, y=6)
    Q = W.conservative_from_primitive_single(rho, [u, v], p, 1.4)
    return O.spectral_radiuses_and_dt(O.Params(dim=2, scheme=1, mutation=mutation), Q, n, G, dx), dx


def test_per_axis_maxima_in_different_cells():
    out, dx = _two_hot_cells()
    sx, sy = (2.0 + 1.0) / dx[0], (3.0 + 1.0) / dx[1]
    assert out[0] == pytest.approx(sx, rel=1e-14) and out[1] == pytest.approx(sy, rel=1e-14)
    # patch dt: the largest per-cell sum (the u cell: 30 + 5, the v cell: 10 + 20)
    assert out[2] == pytest.approx(1.0 / max(sx + 1.0 / dx[1], 1.0 / dx[0] + sy), rel=1e-14)
    # level dt: 1 / (sum of the axis maxima), smaller than the patch dt
    lvl = O.level_dt([out])
    assert lvl == pytest.approx(1.0 / (sx + sy), rel=1e-14)
    assert lvl < out[2]


def test_planted_max_instead_of_sum_is_caught():
    out, dx = _two_hot_cells(mutation=9)
    sx, sy = (2.0 + 1.0) / dx[0], (3.0 + 1.0) / dx[1]
    assert out[2] != pytest.approx(1.0 / max(sx + 1.0 / dx[1], 1.0 / dx[0] + sy), rel=1e-6)


def test_level_reduction_over_patches():
    """The level's axis maxima come from different patches."""
    rs = [np.array([3.0, 1.0, 0.2]), np.array([1.5, 4.0, 0.3])]
    assert O.level_dt(rs, cfl=0.5) == pytest.approx(0.5 / (3.0 + 4.0), rel=1e-15)
    rs = [np.array([3.0, 1.0, 0.1])]  # a patch dt below 1 / sum wins
    assert O.level_dt(rs) == 0.1


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_uniform_state_five_eqn(dim):
    gamma = (1.6, 1.4)
    n = (5,) * dim
    dx = (0.1, 0.07, 0.3)[:dim]
    shape = tuple([5 + 2 * G] * dim)
    Z1 = 0.3
    vel = (0.2, -0.5, 0.1)[:dim]
    Zr = [Z1 * 2.0, (1 - Z1) * 1.0]
    p = 1.1
    Q = W.conservative_from_primitive_five([np.full(shape, z) for z in Zr], [np.full(shape, v) for v in vel],
                                           np.full(shape, p), [np.full(shape, Z1), np.full(shape, 1 - Z1)], gamma)
    prm = O.Params(dim=dim, model=O.FIVE_EQN_ALLAIRE, num_species=2, scheme=1, gamma=list(gamma))
    out = O.spectral_radiuses_and_dt(prm, Q, n, G, dx)
    gmix = 1.0 / (Z1 / (gamma[0] - 1) + (1 - Z1) / (gamma[1] - 1)) + 1.0
    c = np.sqrt(gmix * p / sum(Zr))
    s = [(abs(vel[a]) + c) / dx[a] for a in range(dim)]
    assert np.allclose(out[:dim], s, rtol=1e-13)
    assert out[dim] == pytest.approx(1.0 / sum(s), rel=1e-13)


@pytest.mark.parametrize("dim", [2, 3])
def test_uniform_state_four_eqn(dim):
    gamma, R = (1.4, 1.67), (1.0, 0.25)
    n = (5,) * dim
    dx = (0.1, 0.07, 0.3)[:dim]
    shape = tuple([5 + 2 * G] * dim)
    Y1, rho, p = 0.35, 1.7, 0.8
    vel = (0.2, -0.5, 0.1)[:dim]
    gm = W.mixture_gamma_mass([Y1, 1 - Y1], gamma, R)
    ke = sum(v * v for v in vel)
    Q = np.empty((2 + dim + 1,) + shape)
    Q[0], Q[1] = rho * Y1, rho * (1 - Y1)
    for b in range(dim):
        Q[2 + b] = rho * vel[b]
    Q[2 + dim] = p / (gm - 1.0) + 0.5 * rho * ke
    prm = O.Params(dim=dim, model=O.FOUR_EQN_CONSERVATIVE, num_species=2, scheme=1, gamma=list(gamma), R=list(R))
    out = O.spectral_radiuses_and_dt(prm, Q, n, G, dx)
    c = np.sqrt(gm * p / rho)
    s = [(abs(vel[a]) + c) / dx[a] for a in range(dim)]
    assert np.allclose(out[:dim], s, rtol=1e-13)
    assert out[dim] == pytest.approx(1.0 / sum(s), rel=1e-13)
