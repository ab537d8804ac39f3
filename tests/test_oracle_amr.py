"""Pins of the oracle's two-level advance (oracle/amr.py), the restatement the
GPU hierarchy is checked against.

- Conservative linear refinement: the children of a coarse cell average to it,
  a linear field is refined exactly, and refined values stay within the range
  of the coarse cell and its neighbours (the MC limiter).
- A fine level covering the whole domain steps exactly as the single-level
  oracle RK3 on the refined grid with dt / 2 (no coarse-fine ghosts), and the
  coarse level is then the average of it.
- After one coarse step, coarse cells that are neither under the fine level
  nor next to a coarse-fine face equal the single-level coarse RK3 (the
  synchronization recomputes them from the same fluxes), while the cells next
  to the coarse-fine faces take the refluxed correction.
- The composite sums of the conserved variables stay constant to rounding
  over several coarse steps on a periodic domain (2D single-species, 3D
  five-equation), and a planted mistake -- no refluxing -- breaks that.
- Navier-Stokes form (four-equation model with transport, 6 ghost cells): a
  fully covered fine level is the single-level Navier-Stokes RK3 bit for bit;
  the composite sums stay constant to rounding, and a planted mistake --
  refluxing the convective flux but not the diffusive one -- breaks that (2D;
  3D conservation too); the diffusive flux is in the step (the result differs
  from the Euler form).
"""
import numpy as np
import pytest

from oracle import amr
from oracle import oracle as O
from hamers_amd import workloads as W
from tests.util import max_rel_err, oracle_rk3

G = 4


def _two_level_2d(boxes, N=16):
    Qi, n, dx = W.density_wave_convergence(N, dim=2)
    prm = O.Params(dim=2, model=O.SINGLE_SPECIES, scheme=1)
    return Qi, n, dx, prm, amr.TwoLevel(prm, n, dx, boxes)


def test_refine_conserves_and_is_exact_on_linear_fields():
    rng = np.random.default_rng(3)
    n = (10, 8)
    T = amr.TwoLevel(O.Params(dim=2, model=O.SINGLE_SPECIES, scheme=1), n, (0.1, 0.1), [((0, 0), (20, 16))])
    # random field: children average to the parent
    Qc = rng.uniform(0.5, 1.5, size=(4, n[1], n[0]))
    (Qf,) = T.refine_patches(Qc)
    mean = 0.25 * (Qf[:, 0::2, 0::2] + Qf[:, 0::2, 1::2] + Qf[:, 1::2, 0::2] + Qf[:, 1::2, 1::2])
    assert np.max(np.abs(mean - Qc)) < 1e-15
    # monotone: within the range of the parent and its axis neighbours
    Cg = np.pad(Qc, [(0, 0), (1, 1), (1, 1)], mode="wrap")
    lo = np.minimum.reduce([Cg[:, 1:-1, 1:-1], Cg[:, :-2, 1:-1], Cg[:, 2:, 1:-1], Cg[:, 1:-1, :-2], Cg[:, 1:-1, 2:]])
    hi = np.maximum.reduce([Cg[:, 1:-1, 1:-1], Cg[:, :-2, 1:-1], Cg[:, 2:, 1:-1], Cg[:, 1:-1, :-2], Cg[:, 1:-1, 2:]])
    for cy in range(2):
        for cx in range(2):
            ch = Qf[:, cy::2, cx::2]
            assert np.all(ch >= lo - 1e-15) and np.all(ch <= hi + 1e-15)
    # linear field (away from the periodic seam): refined exactly
    x = np.arange(n[0]) + 0.5
    y = np.arange(n[1]) + 0.5
    lin = np.broadcast_to(2.0 + 0.3 * x[None, :] - 0.2 * y[:, None], (4, n[1], n[0])).copy()
    (Qf,) = T.refine_patches(lin)
    xf = (np.arange(2 * n[0]) + 0.5) / 2
    yf = (np.arange(2 * n[1]) + 0.5) / 2
    exact = 2.0 + 0.3 * xf[None, :] - 0.2 * yf[:, None]
    inner = (slice(2, -2), slice(2, -2))
    assert np.max(np.abs(Qf[0][inner] - exact[inner])) < 1e-14
    # at an extremum the slope is zero: the children copy the parent
    spike = np.ones((4, n[1], n[0]))
    spike[:, 4, 5] = 3.0
    (Qf,) = T.refine_patches(spike)
    assert np.all(Qf[:, 8:10, 10:12] == 3.0)


def test_fully_covered_fine_level_is_the_single_level_rk3():
    Qi, n, dx, prm, T = _two_level_2d([((0, 0), (16, 32)), ((16, 0), (16, 32))])
    fine = T.refine_patches(Qi)
    Qf_glob = np.concatenate(fine, axis=2)
    dt = 0.3 * dx[0]
    Qc, pts = T.advance(Qi, fine, dt)
    ref = oracle_rk3(prm, Qf_glob, T.nf, T.dxf, 0.5 * dt, 2)
    assert np.array_equal(np.concatenate(pts, axis=2), ref)
    avg = 0.25 * (ref[:, 0::2, 0::2] + ref[:, 0::2, 1::2] + ref[:, 1::2, 0::2] + ref[:, 1::2, 1::2])
    assert max_rel_err(Qc, avg) < 1e-15


def test_coarse_cells_away_from_the_fine_level_are_the_single_level_step():
    Qi, n, dx, prm, T = _two_level_2d([((8, 8), (12, 10))])
    fine = T.refine_patches(Qi)
    dt = 0.3 * dx[0]
    Qc, _ = T.advance(Qi, fine, dt)
    ref = oracle_rk3(prm, Qi, n, dx, dt, 1)
    cov = T.covered_coarse()
    near = np.zeros_like(cov)
    for sh in ((0, 1), (0, -1), (1, 1), (1, -1)):
        near |= np.roll(cov, sh[1], axis=sh[0])
    near &= ~cov
    far = ~cov & ~near
    scale = np.max(np.abs(ref), axis=(1, 2))[:, None]
    assert np.max(np.abs(Qc[:, far] - ref[:, far]) / scale) < 1e-14
    # next to the coarse-fine faces the refluxed fluxes change the update
    assert np.max(np.abs(Qc[:, near] - ref[:, near]) / scale) > 1e-9


@pytest.mark.parametrize("reflux", [True, False], ids=["refluxed", "planted-no-reflux"])
def test_composite_conservation_2d(reflux):
    Qi, n, dx, prm, T = _two_level_2d([((8, 8), (12, 16)), ((20, 8), (4, 8))])
    fine = T.refine_patches(Qi)
    m0 = T.composite_sum(Qi, fine, 4)
    Qc, pts = Qi, fine
    for _ in range(3):
        Qc, pts = T.advance(Qc, pts, 0.3 * dx[0], reflux=reflux)
    drift = np.max(np.abs(T.composite_sum(Qc, pts, 4) - m0) / np.abs(m0))
    if reflux:
        assert drift < 1e-14
    else:
        assert drift > 1e-5


def test_composite_conservation_3d_five_eqn():
    Qi, n, dx = W.interface_advection_convergence(8, dim=3)
    prm = O.Params(dim=3, model=O.FIVE_EQN_ALLAIRE, num_species=2, scheme=1, gamma=[1.6, 1.4])
    T = amr.TwoLevel(prm, n, dx, [((4, 4, 2), (6, 8, 10)), ((10, 4, 2), (4, 4, 10))])
    fine = T.refine_patches(Qi)
    nc = prm.neq - (prm.num_species - 1)  # partial densities, momentum, energy
    m0 = T.composite_sum(Qi, fine, nc)
    Qc, pts = T.advance(Qi, fine, 0.3 * dx[0])
    assert np.max(np.abs(T.composite_sum(Qc, pts, nc) - m0) / np.abs(m0)) < 1e-14
    # the last volume fraction is 1 minus the others everywhere
    assert np.max(np.abs(Qc[-1] - (1.0 - Qc[-2]))) < 1e-15
    for q in pts:
        assert np.max(np.abs(q[-1] - (1.0 - q[-2]))) < 1e-15


def test_time_interpolation_is_linear():
    """Coarse-fine values at time fraction alpha of a field linear in space
    and time: the exact field at the fine centres at that time (the MC limiter
    leaves linear data alone)."""
    n = (10, 8)
    T = amr.TwoLevel(O.Params(dim=2, model=O.SINGLE_SPECIES, scheme=1), n, (0.1, 0.1), [((4, 4), (8, 6))])
    x = np.arange(n[0]) + 0.5
    y = np.arange(n[1]) + 0.5
    old = np.broadcast_to(1.0 + 0.25 * x[None, :] + 0.5 * y[:, None], (4, n[1], n[0])).copy()
    new = np.broadcast_to(3.0 - 0.5 * x[None, :] + 0.75 * y[:, None], (4, n[1], n[0])).copy()
    Cold, Cnew = T.fill(old), T.fill(new)
    xs = [np.arange(4, 16), np.arange(4, 12)]  # away from the periodic seam
    xf = (xs[0] + 0.5) / 2
    yf = (xs[1] + 0.5) / 2
    for alpha in (0.0, 0.25, 0.5):
        R = T.refine(Cold, Cnew, alpha, xs)
        want = (1 - alpha) * (1.0 + 0.25 * xf[None, :] + 0.5 * yf[:, None]) + \
            alpha * (3.0 - 0.5 * xf[None, :] + 0.75 * yf[:, None])
        assert np.max(np.abs(R[0] - want)) < 1e-14


def _reflect_x_case():
    """A smooth state whose mirror image about x = 0 (and x = 1) with the x
    momentum negated is its own smooth continuation: rho, v, p even in x, u odd."""
    n = (16, 16)
    dx = (1.0 / 16, 1.0 / 16)
    X = W.cell_centres(n, (0.0, 0.0), dx)
    x, y = X[0], X[1]
    rho = 1.0 + 0.2 * np.cos(2 * np.pi * x) * np.cos(2 * np.pi * y)
    u = 0.3 * np.sin(2 * np.pi * x) * (1.0 + 0.3 * np.sin(2 * np.pi * y))
    v = 0.2 * np.cos(2 * np.pi * y) + 0.1 * np.cos(4 * np.pi * x)
    p = 1.0 + 0.1 * np.cos(2 * np.pi * x) + 0.05 * np.sin(2 * np.pi * y)
    Qi = W.conservative_from_primitive_single(rho, [u, v], p, 1.4)
    prm = O.Params(dim=2, model=O.SINGLE_SPECIES, scheme=1)
    return np.ascontiguousarray(Qi), n, dx, prm


def test_fully_covered_fine_level_with_walls_is_the_single_level_rk3():
    Qi, n, dx, prm = _reflect_x_case()
    bc = ((0, 1), [O.BC_REFLECT, O.BC_FLOW, 0, 0], None)
    T = amr.TwoLevel(prm, n, dx, [((0, 0), (16, 32)), ((16, 0), (16, 32))], bc=bc)
    fine = T.refine_patches(Qi)
    Qf = np.concatenate(fine, axis=2)
    dt = 0.2 * dx[0]
    Qc, pts = T.advance(Qi, fine, dt)
    ref = oracle_rk3(prm, Qf, T.nf, T.dxf, 0.5 * dt, 2, bc=bc)
    assert np.array_equal(np.concatenate(pts, axis=2), ref)


def test_reflect_wall_two_level_equals_mirrored_periodic_domain():
    """A REFLECT wall at x = 0 (and x = 1) with a fine patch against it: the
    two-level step equals, on its own cells, the periodic two-level step of
    the domain twice as wide holding the data and its mirror image (x momentum
    negated), fine patches mirrored too.  The mirrored half rounds on its own,
    so the match is to rounding after the first stage (on a smooth state, so
    that no branch sits at a knife edge)."""
    Qi, n, dx, prm = _reflect_x_case()
    boxes = [((0, 8), (8, 12)), ((12, 2), (10, 10))]
    bc = ((0, 1), [O.BC_REFLECT, O.BC_REFLECT, 0, 0], None)
    T = amr.TwoLevel(prm, n, dx, boxes, bc=bc)
    fine = T.refine_patches(Qi)
    dt = 0.2 * dx[0]
    Qc, pts = T.advance(Qi, fine, dt)
    mirror = Qi[:, :, ::-1].copy()
    mirror[1] = -mirror[1]
    Qd = np.concatenate([Qi, mirror], axis=2)
    nfx = 2 * n[0]
    mboxes = boxes + [((2 * nfx - lo[0] - m[0], lo[1]), m) for lo, m in boxes]
    Td = amr.TwoLevel(prm, (2 * n[0], n[1]), dx, mboxes)
    fined = Td.refine_patches(Qd)
    Qcd, ptsd = Td.advance(Qd, fined, dt)
    scale = np.max(np.abs(Qc), axis=(1, 2))[:, None, None]
    assert np.max(np.abs(Qc - Qcd[:, :, :n[0]]) / scale) < 1e-12
    for a, b in zip(pts, ptsd[:len(boxes)]):
        assert np.max(np.abs(a - b) / scale) < 1e-12
    # a planted mistake: the wall treated as FLOW breaks the match
    Tb = amr.TwoLevel(prm, n, dx, boxes, bc=((0, 1), [O.BC_FLOW, O.BC_REFLECT, 0, 0], None))
    Qb, _ = Tb.advance(Qi, fine, dt)
    assert np.max(np.abs(Qb - Qcd[:, :, :n[0]]) / scale) > 1e-6


# Four-equation Navier-Stokes form: constant transport properties of two species.
NS_GAM, NS_R = (1.6, 1.4), (1.0, 2.0)
NS_TR = O.Transport(mu=[2e-2, 1e-2], mu_v=[5e-3, 0.0], kappa=[3e-2, 2e-2], D=[2e-2, 4e-2], M=[1.0, 4.0])


def ns_state(N):
    """Smooth periodic two-species state on [0,1]^2 in which every primitive
    variable varies along both axes (so every diffusive term is nonzero)."""
    n, dx = (N, N), (1.0 / N, 1.0 / N)
    X = W.cell_centres(n, (0.0, 0.0), dx)
    ph = lambda k: 2 * np.pi * ((k + 1) * X[0] + (k + 2) * X[1]) + k  # noqa: E731
    rho = 1.0 + 0.2 * np.sin(ph(0))
    Y0 = 0.5 + 0.3 * np.sin(ph(1))
    vel = [0.3 * np.sin(ph(2)), 0.2 * np.cos(ph(3))]
    p = 1.0 + 0.1 * np.cos(ph(4))
    Q = W.conservative_from_primitive_four([rho * Y0, rho * (1.0 - Y0)], vel, p, NS_GAM, NS_R)
    prm = O.Params(dim=2, model=O.FOUR_EQN_CONSERVATIVE, num_species=2, scheme=1, gamma=list(NS_GAM),
                   R=list(NS_R))
    return Q, n, dx, prm


def test_navier_stokes_fully_covered_fine_level_is_the_single_level_rk3():
    Qi, n, dx, prm = ns_state(8)
    T = amr.TwoLevel(prm, n, dx, [((0, 0), (16, 6)), ((0, 6), (16, 10))], transport=NS_TR)
    assert T.g == 6
    fine = T.refine_patches(Qi)
    Qf_glob = np.concatenate(fine, axis=1)
    dt = 0.2 * dx[0]
    Qc, pts = T.advance(Qi, fine, dt)
    ref = oracle_rk3(prm, Qf_glob, T.nf, T.dxf, 0.5 * dt, 2, transport=NS_TR)
    assert np.array_equal(np.concatenate(pts, axis=1), ref)
    avg = 0.25 * (ref[:, 0::2, 0::2] + ref[:, 0::2, 1::2] + ref[:, 1::2, 0::2] + ref[:, 1::2, 1::2])
    assert max_rel_err(Qc, avg) < 1e-15
    euler = oracle_rk3(prm, Qf_glob, T.nf, T.dxf, 0.5 * dt, 2)
    assert max_rel_err(euler, ref) > 1e-8  # the diffusive flux is in the step


@pytest.mark.parametrize("reflux_diffusive", [True, False], ids=["refluxed", "planted-no-diffusive-reflux"])
def test_navier_stokes_composite_conservation(reflux_diffusive):
    Qi, n, dx, prm = ns_state(12)
    T = amr.TwoLevel(prm, n, dx, [((6, 6), (10, 12)), ((16, 8), (4, 6))], transport=NS_TR)
    fine = T.refine_patches(Qi)
    m0 = T.composite_sum(Qi, fine, prm.neq)
    Qc, pts = Qi, fine
    for _ in range(2):
        Qc, pts = T.advance(Qc, pts, 0.2 * dx[0], reflux_diffusive=reflux_diffusive)
    drift = np.max(np.abs(T.composite_sum(Qc, pts, prm.neq) - m0) / np.maximum(np.abs(m0), 1e-3))
    if reflux_diffusive:
        assert drift < 1e-14
    else:
        assert drift > 1e-9


def ns_state_3d(N):
    """The 3D counterpart of ns_state: every primitive variable varies along all three axes."""
    n, dx = (N, N, N), (1.0 / N,) * 3
    X = W.cell_centres(n, (0.0, 0.0, 0.0), dx)
    ph = lambda k: 2 * np.pi * ((k + 1) * X[0] + (k + 2) * X[1] + (3 - k % 2) * X[2]) + k  # noqa: E731
    rho = 1.0 + 0.2 * np.sin(ph(0))
    Y0 = 0.5 + 0.3 * np.sin(ph(1))
    vel = [0.3 * np.sin(ph(2)), 0.2 * np.cos(ph(3)), 0.1 * np.sin(ph(4))]
    p = 1.0 + 0.1 * np.cos(ph(5))
    Q = W.conservative_from_primitive_four([rho * Y0, rho * (1.0 - Y0)], vel, p, NS_GAM, NS_R)
    prm = O.Params(dim=3, model=O.FOUR_EQN_CONSERVATIVE, num_species=2, scheme=1, gamma=list(NS_GAM),
                   R=list(NS_R))
    return Q, n, dx, prm


def test_navier_stokes_3d_composite_conservation():
    Qi, n, dx, prm = ns_state_3d(8)
    T = amr.TwoLevel(prm, n, dx, [((4, 4, 2), (6, 8, 10)), ((10, 4, 2), (4, 4, 10))], transport=NS_TR)
    fine = T.refine_patches(Qi)
    m0 = T.composite_sum(Qi, fine, prm.neq)
    Qc, pts = T.advance(Qi, fine, 0.2 * dx[0])
    drift = np.max(np.abs(T.composite_sum(Qc, pts, prm.neq) - m0) / np.maximum(np.abs(m0), 1e-3))
    assert drift < 1e-14
