"""GPU physical boundary conditions of the level integrator
(hm_level_set_boundary_conditions) against the oracle (oracle.fill_boundary,
pinned in test_oracle_bc.py): every patch's ghost box after a fill is
bit-identical to the oracle's ghosted domain, and Runge-Kutta steps of levels
of uneven patches with mixed FLOW / REFLECT / SYMMETRY / DIRICHLET sides (and
periodic axes) are bit-identical to the oracle's steps on the whole domain:
2D single species, 3D five-equation, 3D four-equation Navier-Stokes (6 ghost
cells).  Bad conditions and the unsupported adjoint are reported as status
codes."""
import ctypes
import os
import sys

import numpy as np
import pytest

from oracle import oracle as O
from hamers_amd import workloads as W
from util import max_rel_err, oracle_rk3

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


def _hip(torch):
    """torch's HIP runtime (the one the library is bound to), for device-to-host copies."""
    lib = ctypes.CDLL(os.path.join(os.path.dirname(torch.__file__), "lib", "libamdhip64.so"))
    lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    lib.hipMemcpy.restype = ctypes.c_int
    return lib


def _ghost_box(torch, lvl, i):
    """Patch i's scratch state over its ghost box (num_arrays, *rev(n + 2g))."""
    ptr, g = lvl.patch_array(i)
    n = lvl.boxes[i][1]
    shape = (lvl.rec.num_arrays,) + tuple(v + 2 * g for v in reversed(n))
    out = np.empty(shape)
    torch.cuda.synchronize()
    assert _hip(torch).hipMemcpy(out.ctypes.data, ptr, out.nbytes, 2) == 0
    return out, g


def _check_ghosts(torch, lvl, Qg_oracle, g):
    for i, (lo, n) in enumerate(lvl.boxes):
        got, gg = _ghost_box(torch, lvl, i)
        assert gg == g
        sl = tuple(slice(lo[a], lo[a] + n[a] + 2 * g) for a in reversed(range(lvl.dim)))
        want = Qg_oracle[(slice(None),) + sl]
        assert np.array_equal(got, want), f"patch {i}: ghost box differs (max {np.nanmax(np.abs(got - want)):.3e})"


def _level(rec, boxes, n, periodic, dx, conds, dirichlet):
    from hamers_amd import Level

    lvl = Level(rec, boxes, ((0,) * len(n), n), periodic, dx)
    lvl.set_boundary_conditions(conds, dirichlet)
    return lvl


def _run(torch, tag, rec, prm, Qi, n, dx, boxes, periodic, conds, dirichlet, dt, nsteps, transport=None):
    lvl = _level(rec, boxes, n, periodic, dx, conds, dirichlet)
    lvl.write_global(Qi)
    _say(f"case {tag}: ghost fill")
    lvl.fill_ghosts()
    g = 6 if transport is not None else 4
    mom0 = 1 if prm.model == O.SINGLE_SPECIES else prm.num_species
    _check_ghosts(torch, lvl, O.fill_boundary(Qi, g, periodic, conds, dirichlet, mom0), g)
    _say(f"{tag}: every ghost box bit-identical to the oracle's")
    lvl.write_global(Qi)
    for use_graph in (False, True):
        _say(f"case {tag}: {nsteps} RK3 steps, graph {use_graph}")
        lvl.write_global(Qi)
        lvl.advance(dt, nsteps, use_graph=use_graph)
        torch.cuda.synchronize()
        got = lvl.read_global(Qi.shape)
        want = oracle_rk3(prm, Qi, n, dx, dt, nsteps, transport=transport, bc=(periodic, conds, dirichlet))
        err = max_rel_err(got, want)
        same = int(np.sum(got == want))
        _say(f"{tag}: max rel err {err:.3e}, bit-identical {same}/{got.size}")
        assert np.all(np.isfinite(got))
        assert err <= TOL, f"{tag}: {err:.3e}"


def _rec(dim, model, scheme, ns, gamma, R=None):
    from hamers_amd import ConvectiveFluxReconstructor

    return ConvectiveFluxReconstructor("bc", dim, model, scheme, num_species=ns, species_gamma=gamma,
                                       species_R=R or (1.0,) * ns)


def test_bc_single_species_2d_mixed_sides(torch_dev):
    Qn, _, _ = W.four_quadrant_riemann(48, seed=11)
    n = (40, 36)
    Qi = np.ascontiguousarray(Qn[:, :n[1], :n[0]])
    dx = (1.0 / n[0], 1.0 / n[1])
    boxes = [((0, 0), (16, 12)), ((16, 0), (24, 12)), ((0, 12), (24, 24)), ((24, 12), (16, 24))]
    state = np.array([0.8, 0.1, -0.2, 2.0])
    dirichlet = np.zeros((4, 4))
    dirichlet[3] = state
    rec = _rec(2, "SINGLE_SPECIES", "WCNS5_JS_HLLC_HLL", 1, (1.4,))
    prm = O.Params(dim=2, model=O.SINGLE_SPECIES, scheme=0)
    _run(torch_dev, "2D single-species REFLECT / FLOW / SYMMETRY / DIRICHLET", rec, prm, Qi, n, dx, boxes, (0, 0),
         [O.BC_REFLECT, O.BC_FLOW, O.BC_SYMMETRY, O.BC_DIRICHLET], dirichlet, 0.2 * dx[0], 2)


def test_bc_five_eqn_3d_with_periodic_axis(torch_dev):
    from test_gpu_parity_five import shock_interface

    n, g = (24, 16, 20), 4
    Qg0 = shock_interface(n, g, 12)
    Qi = np.ascontiguousarray(Qg0[(slice(None),) + (slice(g, -g),) * 3])
    dx = tuple(1.0 / v for v in n)
    boxes = [((0, 0, 0), (8, 16, 12)), ((8, 0, 0), (16, 8, 12)), ((8, 8, 0), (16, 8, 12)),
             ((0, 0, 12), (24, 16, 8))]
    rec = _rec(3, "FIVE_EQN_ALLAIRE", "WCNS6_LD_HLLC_HLL", 2, (1.4, 1.67))
    prm = O.Params(dim=3, model=O.FIVE_EQN_ALLAIRE, num_species=2, scheme=1, gamma=[1.4, 1.67])
    dirichlet = np.zeros((6, rec.num_arrays))
    dirichlet[1] = Qi[:, 0, 0, -1]  # x-hi: the state of one boundary cell
    _run(torch_dev, "3D five-eqn FLOW / DIRICHLET x, periodic y, REFLECT / SYMMETRY z", rec, prm, Qi, n, dx, boxes,
         (0, 1, 0), [O.BC_FLOW, O.BC_DIRICHLET, 0, 0, O.BC_REFLECT, O.BC_SYMMETRY], dirichlet, 2e-4, 2)


def test_bc_navier_stokes_3d_walls(torch_dev):
    from test_gpu_navier_stokes import GAM, RR, TR, _recs, _smooth_state_3d

    Qi, n, dx = _smooth_state_3d(16)
    boxes = [((0, 0, 0), (16, 8, 16)), ((0, 8, 0), (10, 8, 16)), ((10, 8, 0), (6, 8, 16))]
    rec, diff = _recs(3)
    prm = O.Params(dim=3, model=O.FOUR_EQN_CONSERVATIVE, num_species=2, gamma=list(GAM), R=list(RR))
    _run(torch_dev, "3D four-eqn Navier-Stokes REFLECT walls in y, SYMMETRY / FLOW z, periodic x", rec, prm, Qi,
         n, dx, boxes, (1, 0, 0), [0, 0, O.BC_REFLECT, O.BC_REFLECT, O.BC_SYMMETRY, O.BC_FLOW], None, 2e-3, 1,
         transport=TR)


def test_bc_errors_are_status_codes(torch_dev):
    from hamers_amd import Level, capi

    rec = _rec(2, "SINGLE_SPECIES", "WCNS6_LD_HLLC_HLL", 1, (1.4,))
    lvl = Level(rec, [((0, 0), (16, 16))], ((0, 0), (16, 16)), (1, 0), (1 / 16, 1 / 16))
    with pytest.raises(capi.HamersError, match="periodic"):
        lvl.set_boundary_conditions([O.BC_FLOW, O.BC_FLOW, 0, 0])
    with pytest.raises(capi.HamersError, match="DIRICHLET"):
        lvl.set_boundary_conditions([0, 0, O.BC_DIRICHLET, O.BC_FLOW])
    lvl.set_boundary_conditions([0, 0, O.BC_REFLECT, O.BC_FLOW])
    lvl.write_global(np.ones((4, 16, 16)))
    with pytest.raises(capi.HamersError) as e:
        lvl.advance_recorded(1e-3, 2, 1)
    assert e.value.status == capi.HM_ERR_UNSUPPORTED
    _say("boundary-condition errors reported as status codes")
