"""GPU Navier-Stokes stage (four-equation model): the stage update with
diffusive fluxes against numpy in the reference's association order
(NavierStokes.cpp:1743-1748, 2085-2092, gamma accumulation 1803-1900):
bit-identical; and one SSP-RK3 step of a periodic 2D patch (convective +
diffusive flux and stage update on the GPU, periodic halos refreshed on the
host between stages) against the oracle's Navier-Stokes RK3: max-norm
relative error <= 1e-12."""
import sys

import numpy as np
import pytest

from oracle import oracle as O
from hamers_amd import workloads as W
from tests.util import flux_divergence, max_rel_err, oracle_rk3, RK_ALPHA, RK_BETA

pytestmark = pytest.mark.gpu
GAM, RR = (1.6, 1.4), (1.0, 2.0)
TR = O.Transport(mu=[2e-2, 1e-2], mu_v=[5e-3, 0.0], kappa=[3e-2, 2e-2], D=[2e-2, 4e-2], M=[1.0, 4.0])


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def torch_dev():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a visible GPU"
    return torch


def _recs(dim):
    from hamers_amd import ConvectiveFluxReconstructor, DiffusiveFluxReconstructor

    rec = ConvectiveFluxReconstructor("ns", dim, "FOUR_EQN_CONSERVATIVE", "WCNS6_LD_HLLC_HLL", num_species=2,
                                      species_gamma=GAM, species_R=RR)
    diff = DiffusiveFluxReconstructor(rec, TR.mu, TR.mu_v, TR.kappa, TR.D, TR.M)
    return rec, diff


def test_stage_update_navier_stokes_form(torch_dev):
    torch = torch_dev
    from hamers_amd import RKPatchData, advanceSingleStepOnPatch

    rng = np.random.default_rng(17)
    dim, n, g = 3, (12, 10, 8), 6
    neq = 2 + dim + 1
    dx = (0.1, 0.07, 0.05)
    rec, _ = _recs(dim)
    gshape = (neq,) + tuple(v + 2 * g for v in reversed(n))
    fsh = [O.face_shape(neq, n, a) for a in range(dim)]
    alpha, beta, gamma = (1.0 / 3.0, 0.0, 2.0 / 3.0), (0.0, 0.0, 2.0 / 3.0), (0.0, 0.0, 2.0 / 3.0)
    Qi = [rng.normal(size=gshape) for _ in range(3)]
    Fc = [[rng.normal(size=fsh[a]) for a in range(dim)] for _ in range(3)]
    Fd = [[rng.normal(size=fsh[a]) for a in range(dim)] for _ in range(3)]
    S = [rng.normal(size=(neq,) + tuple(reversed(n))) for _ in range(3)]
    Fc_acc0 = [rng.normal(size=fsh[a]) for a in range(dim)]
    Fd_acc0 = [rng.normal(size=fsh[a]) for a in range(dim)]
    dev = "cuda:0"
    T = lambda x: torch.from_numpy(np.ascontiguousarray(x)).to(dev)  # noqa: E731
    Q = T(np.full(gshape, np.nan))
    pd = RKPatchData(n, (g,) * dim, (g,) * dim, dx, Q, [T(q) for q in Qi], [[T(f) for f in F] for F in Fc],
                     [T(s) for s in S], F_accumulated=[T(f) for f in Fc_acc0],
                     Fd_intermediate=[[T(f) for f in F] for F in Fd], Fd_accumulated=[T(f) for f in Fd_acc0])
    _say("case Navier-Stokes stage update, 3D, 3 coefficients")
    advanceSingleStepOnPatch(rec, pd, alpha, beta, gamma)
    torch.cuda.synchronize()
    sl = (slice(None),) + (slice(g, -g),) * dim
    q = np.zeros((neq,) + tuple(reversed(n)))
    for m in range(3):
        if alpha[m] != 0.0:
            q = q + alpha[m] * Qi[m][sl]
        if beta[m] != 0.0:
            q = q + beta[m] * flux_divergence(Fc[m], S[m], dx, Fd[m])
    got = pd.Q.cpu().numpy()[sl]
    assert np.array_equal(got, q), float(np.max(np.abs(got - q)))
    for a in range(dim):
        ec, ed = Fc_acc0[a].copy(), Fd_acc0[a].copy()
        for m in range(3):
            if gamma[m] != 0.0:
                ec = ec + gamma[m] * Fc[m][a]
                ed = ed + gamma[m] * Fd[m][a]
        assert np.array_equal(pd.F_accumulated[a].cpu().numpy(), ec)
        assert np.array_equal(pd.Fd_accumulated[a].cpu().numpy(), ed)


def _smooth_state(N):
    n, dx = (N, N), (1.0 / N, 1.0 / N)
    X = W.cell_centres(n, (0.0, 0.0), dx)
    ph = lambda k: 2 * np.pi * ((k + 1) * X[0] + (k + 2) * X[1]) + k  # noqa: E731
    rho = 1.0 + 0.2 * np.sin(ph(0))
    Y0 = 0.5 + 0.3 * np.sin(ph(1))
    vel = [0.3 * np.sin(ph(2)), 0.2 * np.cos(ph(3))]
    p = 1.0 + 0.1 * np.cos(ph(4))
    return W.conservative_from_primitive_four([rho * Y0, rho * (1.0 - Y0)], vel, p, GAM, RR), n, dx


def test_navier_stokes_rk3_step_matches_oracle(torch_dev):
    torch = torch_dev
    from hamers_amd import PatchData, RKPatchData, advanceSingleStepOnPatch

    Q0, n, dx = _smooth_state(32)
    g, neq, dim = 6, 5, 2
    dt = 2e-3
    rec, diff = _recs(dim)
    dev = "cuda:0"
    T = lambda x: torch.from_numpy(np.ascontiguousarray(x)).to(dev)  # noqa: E731
    fsh = [O.face_shape(neq, n, a) for a in range(dim)]
    _say("case Navier-Stokes SSP-RK3 step, 2D periodic 32^2, two species")
    stages_Q, stages_F, stages_Fd, stages_S = [], [], [], []
    Qcur = Q0.copy()
    for sn in range(3):
        Qg = T(W.fill_periodic(Qcur, g))
        F = [torch.empty(fsh[a], dtype=torch.float64, device=dev) for a in range(dim)]
        Fd = [torch.empty(fsh[a], dtype=torch.float64, device=dev) for a in range(dim)]
        S = torch.zeros((neq,) + tuple(reversed(n)), dtype=torch.float64, device=dev)
        rec.computeConvectiveFluxAndSourceOnPatch(PatchData(n, (g, g), dx, Qg, F, S), 0.0, dt)
        diff.computeDiffusiveFluxOnPatch(PatchData(n, (g, g), dx, Qg, Fd, None), 0.0, dt)
        stages_Q.append(Qg)
        stages_F.append(F)
        stages_Fd.append(Fd)
        stages_S.append(S)
        Qout = torch.full_like(Qg, float("nan"))
        pd = RKPatchData(n, (g, g), (g, g), dx, Qout, list(stages_Q), list(stages_F), list(stages_S),
                         Fd_intermediate=list(stages_Fd))
        advanceSingleStepOnPatch(rec, pd, RK_ALPHA[sn], RK_BETA[sn])
        torch.cuda.synchronize()
        Qcur = Qout.cpu().numpy()[:, g:-g, g:-g]
    prm = O.Params(dim=2, model=O.FOUR_EQN_CONSERVATIVE, num_species=2, gamma=list(GAM), R=list(RR))
    ref = oracle_rk3(prm, Q0, n, dx, dt, 1, transport=TR)
    err = max_rel_err(Qcur, ref)
    same = int(np.sum(Qcur == ref))
    _say(f"Navier-Stokes RK3 step: max rel err {err:.3e}, bit-identical {same}/{ref.size}")
    assert err <= 1e-12
    # the diffusive flux matters at this viscosity: the Euler step differs
    euler = oracle_rk3(prm, Q0, n, dx, dt, 1)
    assert max_rel_err(euler, ref) > 1e-8


def _smooth_state_3d(N):
    n, dx = (N, N, N), (1.0 / N,) * 3
    X = W.cell_centres(n, (0.0, 0.0, 0.0), dx)
    ph = lambda k: 2 * np.pi * ((k + 1) * X[0] + (k + 2) * X[1] + (3 - k % 2) * X[2]) + k  # noqa: E731
    rho = 1.0 + 0.2 * np.sin(ph(0))
    Y0 = 0.5 + 0.3 * np.sin(ph(1))
    vel = [0.3 * np.sin(ph(2)), 0.2 * np.cos(ph(3)), 0.1 * np.sin(ph(4))]
    p = 1.0 + 0.1 * np.cos(ph(5))
    return W.conservative_from_primitive_four([rho * Y0, rho * (1.0 - Y0)], vel, p, GAM, RR), n, dx


@pytest.mark.parametrize("dim,use_graph", [(2, False), (2, True), (3, True)])
def test_navier_stokes_level_rk3_matches_oracle(torch_dev, dim, use_graph):
    """A level built on a context with transport properties keeps 6 ghost cells
    and adds the diffusive flux to every stage (Navier-Stokes form): SSP-RK3
    steps of an uneven periodic level against the oracle's Navier-Stokes RK3."""
    from hamers_amd import Level

    torch = torch_dev
    if dim == 2:
        Q0, n, dx = _smooth_state(32)
        boxes = [((0, 0), (20, 12)), ((20, 0), (12, 12)), ((0, 12), (8, 20)), ((8, 12), (24, 20))]
        dt, nsteps = 2e-3, 2
    else:
        Q0, n, dx = _smooth_state_3d(16)
        boxes = [((i, j, k), (8, 8, 8)) for k in (0, 8) for j in (0, 8) for i in (0, 8)]
        dt, nsteps = 2e-3, 1
    rec, diff = _recs(dim)
    lvl = Level(rec, boxes, ((0,) * dim, n), (1,) * dim, dx)
    lvl.write_global(Q0)
    _say(f"case Navier-Stokes level RK3, {dim}D, {len(boxes)} patches, {nsteps} steps, graph {use_graph}")
    lvl.advance(dt, nsteps, use_graph=use_graph)
    torch.cuda.synchronize()
    got = lvl.read_global(Q0.shape)
    prm = O.Params(dim=dim, model=O.FOUR_EQN_CONSERVATIVE, num_species=2, gamma=list(GAM), R=list(RR))
    ref = oracle_rk3(prm, Q0, n, dx, dt, nsteps, transport=TR)
    err = max_rel_err(got, ref)
    same = int(np.sum(got == ref))
    _say(f"Navier-Stokes level: max rel err {err:.3e}, bit-identical {same}/{ref.size}")
    assert np.all(np.isfinite(got))
    assert err <= 1e-12
    euler = oracle_rk3(prm, Q0, n, dx, dt, nsteps)
    assert max_rel_err(euler, ref) > 1e-8  # the diffusive flux is in the step


def test_navier_stokes_level_refuses_stable_dt(torch_dev):
    """The stable dt of the Navier-Stokes form needs the maximum diffusivity,
    which is not built: hm_level_advance_cfl says so instead of using the
    Euler dt."""
    from hamers_amd import Level
    from hamers_amd.capi import HamersError

    Q0, n, dx = _smooth_state(16)
    rec, diff = _recs(2)
    lvl = Level(rec, [((0, 0), (16, 16))], ((0, 0), n), (1, 1), dx)
    lvl.write_global(Q0)
    with pytest.raises(HamersError, match="Navier-Stokes"):
        lvl.advance_cfl(0.4, 1)
