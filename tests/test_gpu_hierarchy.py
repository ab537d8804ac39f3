"""GPU two-level advance (hm_hierarchy_advance) against the oracle's
restatement (oracle/amr.py, pinned in test_oracle_amr.py).

- 2D single-species (WCNS6-LD), a fine level of two adjacent patches inside a
  four-patch coarse level, three coarse steps; and a fine level whose two
  patches meet across the periodic boundary.
- 3D five-equation (WCNS5-JS and WCNS6-LD), two fine patches, one coarse step.
- Physical boundaries: 2D single-species with REFLECT and FLOW x sides and a
  fine patch against the wall, and 3D five-equation with SYMMETRY and DIRICHLET
  z sides, against the restatement with the same conditions.
- Navier-Stokes form: four-equation model with transport (6 ghost cells), the
  diffusive flux refluxed beside the convective one: 2D, two coarse steps; 3D,
  two fine patches, one coarse step; 2D with REFLECT / FLOW x sides.
Coarse and fine states: max-norm relative error <= 1e-12 (FP64; the
restatement keeps the kernels' operation order).  The GPU result keeps the
composite sums of the conserved variables to rounding.  Errors: misaligned
fine boxes, a ratio other than 2, and destroying a level that a hierarchy uses.
"""
import ctypes
import sys

import numpy as np
import pytest

from oracle import amr
from oracle import oracle as O
from hamers_amd import workloads as W
from tests.util import max_rel_err

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def torch_dev():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a visible GPU"
    return torch


def _rec(dim, model="SINGLE_SPECIES", scheme="WCNS6_LD_HLLC_HLL", ns=1, gamma=(1.4,), R=None):
    from hamers_amd import ConvectiveFluxReconstructor

    return ConvectiveFluxReconstructor("amr", dim, model, scheme, num_species=ns, species_gamma=gamma,
                                       species_R=(1.0,) * len(gamma) if R is None else R)


def _levels(rec, n, dx, coarse_boxes, fine_boxes, T, bc=None):
    from hamers_amd import Level

    d = len(n)
    periodic = (1,) * d if bc is None else tuple(bc[0])
    C = Level(rec, coarse_boxes, ((0,) * d, n), periodic, dx)
    F = Level(rec, fine_boxes, ((0,) * d, T.nf), periodic, T.dxf)
    if bc is not None:
        for lvl in (C, F):
            lvl.set_boundary_conditions(bc[1], bc[2])
    return C, F


def _case(torch, rec, prm, Qi, n, dx, coarse_boxes, fine_boxes, dt, nsteps, nconserved, label, bc=None,
          transport=None):
    from hamers_amd import Hierarchy

    _say(f"case: {label}")
    T = amr.TwoLevel(prm, n, dx, fine_boxes, bc=bc, transport=transport)
    fine0 = T.refine_patches(Qi)
    C, F = _levels(rec, n, dx, coarse_boxes, fine_boxes, T, bc)
    C.write_global(Qi)
    for i, q in enumerate(fine0):
        F.write_patch(i, q)
    H = Hierarchy(C, F)
    H.advance(dt, nsteps)
    torch.cuda.synchronize()
    Qc = C.read_global(Qi.shape)
    pts = [F.read_patch(i) for i in range(len(fine_boxes))]
    del H
    ref_c, ref_p = Qi, fine0
    for _ in range(nsteps):
        ref_c, ref_p = T.advance(ref_c, ref_p, dt)
    err = max([max_rel_err(Qc, ref_c)] + [max_rel_err(a, b) for a, b in zip(pts, ref_p)])
    same = np.array_equal(Qc, ref_c) and all(np.array_equal(a, b) for a, b in zip(pts, ref_p))
    _say(f"  {label}: max rel err {err:.3e} (bit-identical: {same})")
    assert err <= TOL
    if nconserved == 0:
        return
    m0 = T.composite_sum(Qi, fine0, nconserved)
    # (sums near zero, such as the momentum of a periodic wave, against the largest sum's scale)
    scale = np.maximum(np.abs(m0), 1e-3 * np.max(np.abs(m0)))
    drift = np.max(np.abs(T.composite_sum(Qc, pts, nconserved) - m0) / scale)
    _say(f"  composite conservation drift {drift:.2e}")
    assert drift < 1e-13


def test_two_level_2d_single_species(torch_dev):
    Qi, n, dx = W.density_wave_convergence(16, dim=2)
    coarse = [((i, j), (8, 8)) for j in (0, 8) for i in (0, 8)]
    _case(torch_dev, _rec(2), O.Params(dim=2, model=O.SINGLE_SPECIES, scheme=1), Qi, n, dx, coarse,
          [((8, 8), (12, 16)), ((20, 8), (4, 8))], 0.3 * dx[0], 3, 4, "2D WCNS6-LD, two adjacent fine patches")


def test_two_level_2d_across_the_periodic_boundary(torch_dev):
    Qi, n, dx = W.density_wave_convergence(16, dim=2)
    coarse = [((0, 0), (16, 6)), ((0, 6), (16, 10))]
    _case(torch_dev, _rec(2, scheme="WCNS5_JS_HLLC_HLL"), O.Params(dim=2, model=O.SINGLE_SPECIES, scheme=0), Qi, n,
          dx, coarse, [((26, 4), (6, 10)), ((0, 4), (4, 10))], 0.3 * dx[0], 2, 4,
          "2D WCNS5-JS, fine patches meeting across the periodic boundary")


@pytest.mark.parametrize("scheme", [0, 1], ids=["WCNS5_JS", "WCNS6_LD"])
def test_two_level_3d_five_eqn(torch_dev, scheme):
    Qi, n, dx = W.interface_advection_convergence(8, dim=3)
    key = "WCNS6_LD_HLLC_HLL" if scheme == 1 else "WCNS5_JS_HLLC_HLL"
    rec = _rec(3, "FIVE_EQN_ALLAIRE", key, 2, (1.6, 1.4))
    prm = O.Params(dim=3, model=O.FIVE_EQN_ALLAIRE, num_species=2, scheme=scheme, gamma=[1.6, 1.4])
    coarse = [((0, 0, 0), (8, 8, 4)), ((0, 0, 4), (8, 8, 4))]
    _case(torch_dev, rec, prm, Qi, n, dx, coarse, [((4, 4, 2), (6, 8, 10)), ((10, 4, 2), (4, 4, 10))],
          0.3 * dx[0], 1, prm.neq - 1, f"3D five-eqn scheme={scheme}, two fine patches")


def test_two_level_2d_walls(torch_dev):
    from tests.test_oracle_amr import _reflect_x_case

    Qi, n, dx, prm = _reflect_x_case()
    bc = ((0, 1), [O.BC_REFLECT, O.BC_FLOW, 0, 0], None)
    coarse = [((0, 0), (8, 16)), ((8, 0), (8, 16))]
    _case(torch_dev, _rec(2), prm, Qi, n, dx, coarse, [((0, 8), (8, 12)), ((12, 2), (10, 10)), ((26, 0), (6, 32))],
          0.2 * dx[0], 2, 0, "2D WCNS6-LD, REFLECT / FLOW x sides, fine patches against both", bc=bc)


def test_two_level_3d_five_eqn_walls(torch_dev):
    Qi, n, dx = W.interface_advection_convergence(8, dim=3)
    rec = _rec(3, "FIVE_EQN_ALLAIRE", "WCNS6_LD_HLLC_HLL", 2, (1.6, 1.4))
    prm = O.Params(dim=3, model=O.FIVE_EQN_ALLAIRE, num_species=2, scheme=1, gamma=[1.6, 1.4])
    dirichlet = np.zeros((6, prm.nq))
    dirichlet[5] = Qi[:, -1, 0, 0]  # z-hi: the state of one boundary cell
    bc = ((1, 1, 0), [0, 0, 0, 0, O.BC_SYMMETRY, O.BC_DIRICHLET], dirichlet)
    coarse = [((0, 0, 0), (8, 8, 4)), ((0, 0, 4), (8, 8, 4))]
    _case(torch_dev, rec, prm, Qi, n, dx, coarse, [((4, 4, 0), (6, 8, 10)), ((10, 4, 8), (4, 4, 8))],
          0.3 * dx[0], 1, 0, "3D five-eqn, SYMMETRY / DIRICHLET z sides, fine patches against both", bc=bc)


def test_hierarchy_errors(torch_dev):
    from hamers_amd import Hierarchy, capi

    Qi, n, dx = W.density_wave_convergence(16, dim=2)
    rec = _rec(2)
    prm = O.Params(dim=2, model=O.SINGLE_SPECIES, scheme=1)
    T = amr.TwoLevel(prm, n, dx, [((8, 8), (8, 8))])
    C, F = _levels(rec, n, dx, [((0, 0), (16, 16))], [((8, 8), (8, 8))], T)
    h = ctypes.c_void_p()
    assert capi.lib.hm_hierarchy_create(C._lvl, F._lvl, 3, ctypes.byref(h)) == capi.HM_ERR_UNSUPPORTED
    # misaligned fine box (odd lower corner)
    from hamers_amd import Level

    Fb = Level(rec, [((7, 8), (8, 8))], ((0, 0), T.nf), (1, 1), T.dxf)
    with pytest.raises(capi.HamersError):
        Hierarchy(C, Fb)
    # a fine domain that is not the coarse one refined by 2
    Fw = Level(rec, [((8, 8), (8, 8))], ((0, 0), (30, 32)), (1, 1), T.dxf)
    with pytest.raises(capi.HamersError):
        Hierarchy(C, Fw)
    # a level in use by a hierarchy is not destroyed
    H = Hierarchy(C, F)
    assert capi.lib.hm_level_destroy(F._lvl) == capi.HM_ERR_INVALID_ARGUMENT
    del H


def test_two_level_2d_navier_stokes(torch_dev):
    """Four-equation Navier-Stokes form on both levels: a context with
    transport properties, whose levels keep 6 ghost cells; the restatement
    refluxes the diffusive flux sums beside the convective ones."""
    from hamers_amd import DiffusiveFluxReconstructor
    from tests.test_oracle_amr import NS_TR, ns_state

    Qi, n, dx, prm = ns_state(16)
    rec = _rec(2, "FOUR_EQN_CONSERVATIVE", "WCNS6_LD_HLLC_HLL", 2, tuple(prm.gamma), tuple(prm.R))
    diff = DiffusiveFluxReconstructor(rec, NS_TR.mu, NS_TR.mu_v, NS_TR.kappa, NS_TR.D, NS_TR.M)
    coarse = [((i, j), (8, 8)) for j in (0, 8) for i in (0, 8)]
    _case(torch_dev, rec, prm, Qi, n, dx, coarse, [((8, 8), (12, 14)), ((20, 10), (6, 8))], 0.2 * dx[0], 2, prm.neq,
          "2D four-equation Navier-Stokes, two fine patches", transport=NS_TR)
    del diff


def _ns_context(prm):
    from hamers_amd import DiffusiveFluxReconstructor
    from tests.test_oracle_amr import NS_TR

    rec = _rec(prm.dim, "FOUR_EQN_CONSERVATIVE", "WCNS6_LD_HLLC_HLL", 2, tuple(prm.gamma), tuple(prm.R))
    diff = DiffusiveFluxReconstructor(rec, NS_TR.mu, NS_TR.mu_v, NS_TR.kappa, NS_TR.D, NS_TR.M)
    return rec, diff


def test_two_level_3d_navier_stokes(torch_dev):
    from tests.test_oracle_amr import NS_TR, ns_state_3d

    Qi, n, dx, prm = ns_state_3d(8)
    rec, diff = _ns_context(prm)
    _case(torch_dev, rec, prm, Qi, n, dx, [((0, 0, 0), (8, 8, 8))], [((4, 4, 2), (6, 8, 10)), ((10, 4, 2), (4, 4, 10))],
          0.2 * dx[0], 1, prm.neq, "3D four-equation Navier-Stokes, two fine patches", transport=NS_TR)
    del diff


def test_two_level_2d_navier_stokes_walls(torch_dev):
    from tests.test_oracle_amr import NS_TR, ns_state

    Qi, n, dx, prm = ns_state(16)
    rec, diff = _ns_context(prm)
    bc = ((0, 1), [O.BC_REFLECT, O.BC_FLOW, 0, 0], None)
    coarse = [((0, 0), (8, 16)), ((8, 0), (8, 16))]
    _case(torch_dev, rec, prm, Qi, n, dx, coarse, [((0, 8), (8, 12)), ((12, 2), (10, 10)), ((26, 0), (6, 32))],
          0.2 * dx[0], 2, 0, "2D four-equation Navier-Stokes, REFLECT / FLOW x sides, fine patches against both",
          bc=bc, transport=NS_TR)
    del diff
