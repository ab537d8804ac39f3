"""GPU Runge-Kutta stage update and level integrator.

- The stage update entry points (per patch and batched) against numpy with the
  reference's association order (Euler.cpp:1003-1680): bit-identical.
- Whole SSP-RK3 steps of a multi-patch periodic level (halo fill, batched RHS,
  stage update) against the oracle's RK3 on the same domain as one patch:
  max-norm relative error <= 1e-12 (FP64).
- The hipGraph-replayed step against the eagerly launched one: bit-identical,
  checked after every replay with eager work in between.
"""
import sys

import numpy as np
import pytest

from oracle import oracle as O
from hamers_amd import workloads as W
from tests.util import max_rel_err, oracle_rk3

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


def _rec(dim, model="SINGLE_SPECIES", scheme="WCNS6_LD_HLLC_HLL", ns=1, gamma=(1.4,)):
    from hamers_amd import ConvectiveFluxReconstructor

    return ConvectiveFluxReconstructor("rk", dim, model, scheme, num_species=ns, species_gamma=gamma,
                                       species_R=(1.0,) * len(gamma))


# ----------------------------------------------------------------------------
# stage update
# ----------------------------------------------------------------------------

def _interior(a, g, d):
    return a[(slice(None),) + (slice(g, -g),) * d] if g > 0 else a


def _numpy_stage(Qi, Fi, Si, dx, alpha, beta, gq, gi, d, neq, nq, zfirst, nz):
    """Expected interior of Q after the stage update (Euler.cpp:1003-1680)."""
    shape = _interior(Qi[0], gi, d).shape[1:]
    q = np.zeros((nq,) + shape)
    for m in range(len(alpha)):
        if alpha[m] != 0.0:
            q[:neq] = q[:neq] + alpha[m] * _interior(Qi[m], gi, d)[:neq]
        if beta[m] != 0.0:
            t = None
            for a in range(d):
                F = Fi[m][a]
                ax = F.ndim - 1 - a
                R = [slice(None)] * F.ndim
                L = [slice(None)] * F.ndim
                R[ax] = slice(1, None)
                L[ax] = slice(0, -1)
                div = (F[tuple(R)] - F[tuple(L)]) / dx[a]
                t = -div if t is None else t - div
            q[:neq] = q[:neq] + beta[m] * (t + Si[m])
        if beta[m] != 0.0 and nz > 0:
            q[neq] = 1.0
            for z in range(nz):
                q[neq] = q[neq] - q[zfirst + z]
    return q


def _rk_case(torch, rng, n, gq, gi, neq, nq, alpha, beta, gamma, acc_axes):
    d = len(n)
    dev = "cuda:0"
    gshape = lambda c, g: (c,) + tuple(reversed([v + 2 * g for v in n]))  # noqa: E731
    ishape = (neq,) + tuple(reversed(n))
    fshape = [O.face_shape(neq, n, a) for a in range(d)]
    nc = len(alpha)
    Qi = [rng.uniform(0.5, 1.5, gshape(neq, gi)) for _ in range(nc)]
    Fi = [[rng.normal(size=fshape[a]) for a in range(d)] for _ in range(nc)]
    Si = [rng.normal(size=ishape) for _ in range(nc)]
    Q0 = rng.normal(size=gshape(nq, gq))
    Facc = [rng.normal(size=fshape[a]) for a in range(d)]
    Sacc = rng.normal(size=ishape)
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
    from hamers_amd import RKPatchData

    pd = RKPatchData(n, (gq,) * d, (gi,) * d, tuple(0.1 + 0.05 * a for a in range(d)), T(Q0),
                     [T(q) for q in Qi], [[T(f) for f in fm] for fm in Fi], [T(s) for s in Si],
                     [T(Facc[a]) if a in acc_axes else None for a in range(d)],
                     T(Sacc) if d in acc_axes else None)
    exp = dict(Qi=Qi, Fi=Fi, Si=Si, Q0=Q0, Facc=Facc, Sacc=Sacc)
    return pd, exp


def _check_rk(pd, exp, alpha, beta, gamma, neq, nq, zfirst, nz, acc_axes, tag):
    d = len(pd.n)
    gq = pd.ghost[0]
    gi = pd.ghost_intermediate[0]
    want = _numpy_stage(exp["Qi"], exp["Fi"], exp["Si"], pd.dx, alpha, beta, gq, gi, d, neq, nq, zfirst, nz)
    got = pd.Q.cpu().numpy()
    same = np.array_equal(_interior(got, gq, d), want)
    # ghost cells untouched
    mask = np.ones(got.shape[1:], bool)
    mask[(slice(gq, -gq),) * d] = False
    ghosts_kept = np.array_equal(got[:, mask], exp["Q0"][:, mask])
    _say(f"{tag}: interior bit-identical {same}, ghosts kept {ghosts_kept}")
    assert ghosts_kept, tag
    assert same, f"{tag}: max rel err {max_rel_err(_interior(got, gq, d), want):.3e}"
    if gamma is not None:
        for a in range(d):
            if a in acc_axes:
                w = exp["Facc"][a].copy()
                for m in range(len(gamma)):
                    if gamma[m] != 0.0:
                        w = w + gamma[m] * exp["Fi"][m][a]
                assert np.array_equal(pd.F_accumulated[a].cpu().numpy(), w), f"{tag}: F_acc axis {a}"
        if d in acc_axes:
            w = exp["Sacc"].copy()
            for m in range(len(gamma)):
                if gamma[m] != 0.0:
                    w = w + gamma[m] * exp["Si"][m]
            assert np.array_equal(pd.S_accumulated.cpu().numpy(), w), f"{tag}: S_acc"


@pytest.mark.parametrize("dim", [1, 2, 3])
@pytest.mark.parametrize("stage", [0, 1, 2])
def test_stage_update_single_patch(torch_dev, dim, stage):
    from hamers_amd import SSP_RK3_ALPHA, SSP_RK3_BETA, SSP_RK3_GAMMA, advanceSingleStepOnPatch

    rec = _rec(dim)
    rng = np.random.default_rng(10 * dim + stage)
    n = (13, 9, 6)[:dim]
    alpha, beta, gamma = SSP_RK3_ALPHA[stage], SSP_RK3_BETA[stage], SSP_RK3_GAMMA[stage]
    acc = set(range(dim + 1))
    pd, exp = _rk_case(torch_dev, rng, n, 4, 2, rec.num_eqn, rec.num_arrays, alpha, beta, gamma, acc)
    _say(f"case stage update {dim}D stage {stage}")
    advanceSingleStepOnPatch(rec, pd, alpha, beta, gamma)
    torch_dev.cuda.synchronize()
    _check_rk(pd, exp, alpha, beta, gamma, rec.num_eqn, rec.num_arrays, 0, 0, acc, f"{dim}D stage {stage}")


def test_stage_update_batched_uneven_with_partial_accumulation(torch_dev):
    from hamers_amd import advanceSingleStepOnPatch

    rec = _rec(2)
    rng = np.random.default_rng(3)
    alpha, beta, gamma = (0.3, 0.0, 0.5, 0.2), (0.0, 0.7, 0.0, 0.4), (0.1, 0.0, 0.3, 0.25)
    cases = []
    for k, (n, acc) in enumerate([((20, 7), {0, 1, 2}), ((5, 31), {2}), ((16, 16), set())]):
        cases.append((_rk_case(torch_dev, rng, n, 4, 3, rec.num_eqn, rec.num_arrays, alpha, beta, gamma, acc), acc))
    _say("case batched stage update, 3 uneven patches")
    advanceSingleStepOnPatch(rec, [c[0][0] for c in cases], alpha, beta, gamma)
    torch_dev.cuda.synchronize()
    for k, ((pd, exp), acc) in enumerate(cases):
        _check_rk(pd, exp, alpha, beta, gamma, rec.num_eqn, rec.num_arrays, 0, 0, acc, f"batched patch {k}")
    # one patch with only the source accumulated takes the table path
    (pd, exp) = _rk_case(torch_dev, rng, (11, 6), 4, 4, rec.num_eqn, rec.num_arrays, alpha, beta, gamma, {2})
    advanceSingleStepOnPatch(rec, pd, alpha, beta, gamma)
    torch_dev.cuda.synchronize()
    _check_rk(pd, exp, alpha, beta, gamma, rec.num_eqn, rec.num_arrays, 0, 0, {2}, "single patch, source only")


@pytest.mark.parametrize("ns", [2, 3])
def test_stage_update_five_eqn_last_volume_fraction(torch_dev, ns):
    from hamers_amd import SSP_RK3_ALPHA, SSP_RK3_BETA, advanceSingleStepOnPatch

    gam = (1.4, 1.67, 1.3)[:ns]
    rec = _rec(2, "FIVE_EQN_ALLAIRE", "WCNS5_JS_HLLC_HLL", ns, gam)
    rng = np.random.default_rng(ns)
    for stage in range(3):
        alpha, beta = SSP_RK3_ALPHA[stage], SSP_RK3_BETA[stage]
        pd, exp = _rk_case(torch_dev, rng, (12, 10), 4, 4, rec.num_eqn, rec.num_arrays, alpha, beta, None, set())
        advanceSingleStepOnPatch(rec, pd, alpha, beta)
        torch_dev.cuda.synchronize()
        _check_rk(pd, exp, alpha, beta, None, rec.num_eqn, rec.num_arrays, ns + 2 + 1, ns - 1, set(),
                  f"five-eqn ns={ns} stage {stage}")
    # no nonzero beta: the zero-filled last volume fraction stays zero
    pd, exp = _rk_case(torch_dev, rng, (6, 6), 4, 4, rec.num_eqn, rec.num_arrays, (1.0,), (0.0,), None, set())
    advanceSingleStepOnPatch(rec, pd, (1.0,), (0.0,))
    torch_dev.cuda.synchronize()
    assert np.all(_interior(pd.Q.cpu().numpy(), 4, 2)[-1] == 0.0)


# ----------------------------------------------------------------------------
# level integrator
# ----------------------------------------------------------------------------

def _level_vs_oracle(torch, rec, prm, Qi, n, dx, boxes, dt, nsteps, tag, use_graph=False):
    from hamers_amd import Level

    lvl = Level(rec, boxes, ((0,) * len(n), n), (1,) * len(n), dx)
    lvl.write_global(Qi)
    _say(f"case {tag}")
    lvl.advance(dt, nsteps, use_graph=use_graph)
    torch.cuda.synchronize()
    got = lvl.read_global(Qi.shape)
    want = oracle_rk3(prm, Qi, n, dx, dt, nsteps)
    err = max_rel_err(got, want)
    same = int(np.sum(got == want))
    _say(f"{tag}: max rel err {err:.3e}, bit-identical {same}/{got.size}")
    assert np.all(np.isfinite(got))
    assert err <= TOL, f"{tag}: {err:.3e}"
    return got


def test_level_rk3_single_species_2d_uneven(torch_dev):
    Qi, n, dx = W.density_wave_convergence(32, dim=2)
    boxes = [((0, 0), (20, 12)), ((20, 0), (12, 12)), ((0, 12), (8, 20)), ((8, 12), (24, 20))]
    rec = _rec(2)
    prm = O.Params(dim=2, model=O.SINGLE_SPECIES, scheme=1)
    _level_vs_oracle(torch_dev, rec, prm, Qi, n, dx, boxes, 0.3 * dx[0], 3, "level RK3 2D WCNS6-LD, 4 uneven patches")


def test_level_rk3_single_species_3d(torch_dev):
    Qi, n, dx = W.density_wave_convergence(16, dim=3)
    boxes = [((i, j, k), (8, 8, 8)) for k in (0, 8) for j in (0, 8) for i in (0, 8)]
    rec = _rec(3, scheme="WCNS5_JS_HLLC_HLL")
    prm = O.Params(dim=3, model=O.SINGLE_SPECIES, scheme=0)
    _level_vs_oracle(torch_dev, rec, prm, Qi, n, dx, boxes, 0.25 * dx[0], 2, "level RK3 3D WCNS5-JS, 8 patches",
                     use_graph=True)


def test_level_rk3_single_patch_1d(torch_dev):
    Qi, n, dx = W.density_wave_convergence(24, dim=1)
    rec = _rec(1)
    prm = O.Params(dim=1, model=O.SINGLE_SPECIES, scheme=1)
    _level_vs_oracle(torch_dev, rec, prm, Qi, n, dx, [((0,), (24,))], 0.3 * dx[0], 4,
                     "level RK3 1D, one patch with its own periodic images")


def test_level_rk3_four_quadrant_shocks(torch_dev):
    Qi, n, dx = W.four_quadrant_riemann(48, seed=5)
    boxes = [((i, j), (16, 16)) for j in (0, 16, 32) for i in (0, 16, 32)]
    rec = _rec(2, scheme="WCNS5_JS_HLLC_HLL")
    prm = O.Params(dim=2, model=O.SINGLE_SPECIES, scheme=0)
    _level_vs_oracle(torch_dev, rec, prm, Qi, n, dx, boxes, 0.2 * dx[0], 3, "level RK3 four-quadrant, 9 patches")


@pytest.mark.parametrize("scheme", [1, 0], ids=["WCNS6_LD", "WCNS5_JS"])
def test_level_rk3_five_eqn(torch_dev, scheme):
    Qi, n, dx = W.interface_advection_convergence(24, dim=2)
    boxes = [((0, 0), (24, 10)), ((0, 10), (13, 14)), ((13, 10), (11, 14))]
    key = "WCNS6_LD_HLLC_HLL" if scheme == 1 else "WCNS5_JS_HLLC_HLL"
    rec = _rec(2, "FIVE_EQN_ALLAIRE", key, 2, (1.6, 1.4))
    prm = O.Params(dim=2, model=O.FIVE_EQN_ALLAIRE, num_species=2, scheme=scheme, gamma=[1.6, 1.4])
    _level_vs_oracle(torch_dev, rec, prm, Qi, n, dx, boxes, 0.3 * dx[0], 2,
                     f"level RK3 five-eqn scheme={scheme}, 3 patches")


def test_graph_replay_equals_eager(torch_dev):
    from hamers_amd import Level

    torch = torch_dev
    Qi, n, dx = W.four_quadrant_riemann(64, seed=9)
    boxes = [((i, j), (16, 16)) for j in range(0, 64, 16) for i in range(0, 64, 16)]
    rec = _rec(2)
    eager = Level(rec, boxes, ((0, 0), n), (1, 1), dx)
    graph = Level(rec, boxes, ((0, 0), n), (1, 1), dx)
    eager.write_global(Qi)
    graph.write_global(Qi)
    dt = 0.2 * dx[0]
    for it in range(4):
        _say(f"case graph replay {it}")
        eager.advance(dt, 1, use_graph=False)
        graph.advance(dt, 1, use_graph=True)
        torch.cuda.synchronize()
        a = eager.read_global(Qi.shape)
        b = graph.read_global(Qi.shape)
        assert np.array_equal(a, b), f"replay {it} differs"
    # several replays in one call, and a second captured dt
    eager.advance(dt, 3)
    graph.advance(dt, 3, use_graph=True)
    eager.advance(0.5 * dt, 2)
    graph.advance(0.5 * dt, 2, use_graph=True)
    torch.cuda.synchronize()
    assert np.array_equal(eager.read_global(Qi.shape), graph.read_global(Qi.shape))
    _say("graph replays bit-identical to eager steps")


def test_level_rk3_convergence_rate(torch_dev):
    """Advected density wave, dt proportional to dx: SSP-RK3 with WCNS6-LD
    converges at third order overall (the time error dominates)."""
    from hamers_amd import Level

    errs = []
    T = 0.5
    for N in (16, 32, 64):
        Qi, n, dx = W.density_wave_convergence(N, dim=1)
        rec = _rec(1)
        lvl = Level(rec, [((0,), (N // 2,)), ((N // 2,), (N // 2,))], ((0,), n), (1,), dx)
        lvl.write_global(Qi)
        nsteps = int(round(T / (0.2 * dx[0])))
        lvl.advance(T / nsteps, nsteps, use_graph=True)
        torch_dev.cuda.synchronize()
        rho = lvl.read_global(Qi.shape)[0]
        x = -1.0 + (np.arange(N) + 0.5) * dx[0]
        exact = 1.0 + 0.5 * np.sin(np.pi * (x - T))
        errs.append(float(np.max(np.abs(rho - exact))))
    rates = [np.log2(errs[i] / errs[i + 1]) for i in range(2)]
    _say(f"RK3 level errors {errs}, rates {rates}")
    assert min(rates) > 2.7
