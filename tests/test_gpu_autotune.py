"""The batched RHS's launch schedules and their autotuner
(hm_autotune_on_patches, hm_level_autotune): every candidate schedule gives
the same fluxes and sources bit for bit (the schedule only lays the work out
over the GPU), the autotuner returns a candidate no slower than the built-in
schedule and pins it for the list, and a level advanced with its tuned
schedule (eager and graph-captured) is bit-identical to one advanced with the
built-in schedule."""
import sys

import numpy as np
import pytest

from hamers_amd import workloads as W

pytestmark = pytest.mark.gpu

CANDIDATES_SINGLE = [{"tile_rows": r, "xcd_remap": x, "order": o, "persistent": 0}
                     for r in (4, 8, 24) for x in (0, 1) for o in (0, 1)]
CANDIDATES_MULTI = [{"tile_rows": 4, "xcd_remap": x, "order": o, "persistent": p}
                    for x in (0, 1) for o in (0, 1) for p in (0, 1)]
KEYS = ("tile_rows", "xcd_remap", "order", "persistent")


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def torch_dev():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a visible GPU"
    return torch


def _face_shape(neq, n, a):
    ext = [v + (1 if b == a else 0) for b, v in enumerate(n)]
    return (neq,) + tuple(reversed(ext))


def _patches(torch, rec, Qs, ns, g, dx):
    from hamers_amd import PatchData

    pds = []
    for Q, n in zip(Qs, ns):
        F = [torch.full(_face_shape(rec.num_eqn, n, a), np.nan, dtype=torch.float64, device="cuda:0")
             for a in range(len(n))]
        S = torch.zeros((rec.num_eqn,) + tuple(reversed(n)), dtype=torch.float64, device="cuda:0")
        pds.append(PatchData(tuple(n), (g,) * len(n), tuple(dx), torch.from_numpy(np.ascontiguousarray(Q)).to("cuda:0"),
                             F, S))
    return pds


def _rhs(torch, rec, plist, pds, dt):
    for pd in pds:
        for f in pd.F:
            f.fill_(np.nan)
        pd.S.zero_()
    rec.computeConvectiveFluxAndSourceOnPatches(plist, 0.0, dt)
    torch.cuda.synchronize()
    return [([f.cpu().numpy() for f in pd.F], pd.S.cpu().numpy()) for pd in pds]


def _same(a, b, tag):
    for (Fa, Sa), (Fb, Sb) in zip(a, b):
        for x, y in zip(Fa, Fb):
            assert np.all(np.isfinite(x)), f"{tag}: non-finite flux"
            assert np.array_equal(x, y), f"{tag}: flux differs"
        assert np.array_equal(Sa, Sb), f"{tag}: source differs"


def _check_list(torch, rec, pds, dt, candidates, tag):
    from hamers_amd import PatchList
    from hamers_amd.reconstructor import autotune, get_schedule, set_schedule

    plist = PatchList(pds, rec.num_arrays)
    _say(f"case schedules, {tag}: built-in {get_schedule(rec, plist)}")
    ref = _rhs(torch, rec, plist, pds, dt)
    for sc in candidates:
        set_schedule(rec, plist, sc)
        assert get_schedule(rec, plist) == sc
        _same(_rhs(torch, rec, plist, pds, dt), ref, f"{tag} {sc}")
    _say(f"{tag}: {len(candidates)} schedules bit-identical")
    best = autotune(rec, plist, dt, repetitions=3)
    _say(f"{tag}: autotuned {best}")
    assert {k: best[k] for k in KEYS} in candidates
    assert 0.0 < best["ms"] <= best["default_ms"]
    assert get_schedule(rec, plist) == {k: best[k] for k in KEYS}
    _same(_rhs(torch, rec, plist, pds, dt), ref, f"{tag} after autotune")


def test_schedules_single_species_2d_uneven(torch_dev):
    from hamers_amd import ConvectiveFluxReconstructor

    rec = ConvectiveFluxReconstructor("sched", 2, "SINGLE_SPECIES", "WCNS5_JS_HLLC_HLL", species_gamma=(1.4,))
    g = 4
    ns = [(64, 40), (24, 72), (128, 8), (8, 8), (96, 96), (40, 64)]
    Qs = []
    for i, n in enumerate(ns):
        Qn, _, dx = W.four_quadrant_riemann(128, seed=20 + i)
        Qs.append(W.fill_periodic(np.ascontiguousarray(Qn[:, :n[1], :n[0]]), g))
    pds = _patches(torch_dev, rec, Qs, ns, g, (1 / 128, 1 / 128))
    _check_list(torch_dev, rec, pds, 1e-3, CANDIDATES_SINGLE, "2D single-species, 6 uneven patches")


def test_schedules_five_eqn_3d_uneven(torch_dev):
    from hamers_amd import ConvectiveFluxReconstructor
    from test_gpu_parity_five import shock_interface

    rec = ConvectiveFluxReconstructor("sched5", 3, "FIVE_EQN_ALLAIRE", "WCNS6_LD_HLLC_HLL", num_species=2,
                                      species_gamma=(1.4, 1.67), species_R=(1.0, 1.0))
    g = 4
    ns = [(8, 16, 24), (32, 8, 8), (16, 16, 16), (24, 8, 32)]
    Qs = [shock_interface(n, g, 30 + i) for i, n in enumerate(ns)]
    pds = _patches(torch_dev, rec, Qs, ns, g, (1 / 32,) * 3)
    _check_list(torch_dev, rec, pds, 1e-4, CANDIDATES_MULTI, "3D five-equation, 4 uneven patches")


@pytest.mark.parametrize("use_graph", [False, True])
def test_level_autotune_keeps_results(torch_dev, use_graph):
    from hamers_amd import ConvectiveFluxReconstructor, Level

    torch = torch_dev
    rec = ConvectiveFluxReconstructor("lvl_sched", 2, "SINGLE_SPECIES", "WCNS6_LD_HLLC_HLL", species_gamma=(1.4,))
    Qi, n, dx = W.four_quadrant_riemann(96, seed=40)
    boxes = [((0, 0), (64, 40)), ((64, 0), (32, 40)), ((0, 40), (24, 56)), ((24, 40), (72, 56))]
    dt = 0.2 * dx[0]
    ref = Level(rec, boxes, ((0, 0), n), (1, 1), dx)
    ref.write_global(Qi)
    ref.advance(dt, 2, use_graph=use_graph)
    torch.cuda.synchronize()
    want = ref.read_global(Qi.shape)
    lvl = Level(rec, boxes, ((0, 0), n), (1, 1), dx)
    lvl.write_global(Qi)
    _say(f"case level autotune, 4 uneven patches, graph {use_graph}: built-in {lvl.schedule()}")
    best = lvl.autotune(dt, repetitions=3)
    _say(f"level autotuned {best}")
    assert 0.0 < best["ms"] <= best["default_ms"]
    assert lvl.schedule() == {k: best[k] for k in KEYS}
    lvl.write_global(Qi)
    lvl.advance(dt, 2, use_graph=use_graph)
    torch.cuda.synchronize()
    assert np.array_equal(lvl.read_global(Qi.shape), want)
    for sc in CANDIDATES_SINGLE[::3]:
        lvl.set_schedule(sc)
        lvl.write_global(Qi)
        lvl.advance(dt, 2, use_graph=use_graph)
        torch.cuda.synchronize()
        assert np.array_equal(lvl.read_global(Qi.shape), want), sc
    _say("level steps bit-identical under every schedule tried")


def test_bad_schedule_is_rejected(torch_dev):
    from hamers_amd import ConvectiveFluxReconstructor, Level, capi

    rec = ConvectiveFluxReconstructor("bad", 2, "SINGLE_SPECIES", "WCNS6_LD_HLLC_HLL", species_gamma=(1.4,))
    lvl = Level(rec, [((0, 0), (16, 16))], ((0, 0), (16, 16)), (1, 1), (1 / 16, 1 / 16))
    with pytest.raises(capi.HamersError, match="tile rows"):
        lvl.set_schedule({"tile_rows": 5, "xcd_remap": 0, "order": 0})
    with pytest.raises(capi.HamersError, match="persistent"):
        lvl.set_schedule({"tile_rows": 4, "xcd_remap": 0, "order": 0, "persistent": 1})


@pytest.mark.parametrize("use_graph", [False, True])
def test_level_five_eqn_every_schedule(torch_dev, use_graph):
    """A five-equation level stepped under each multi-species schedule (the
    persistent midpoint flux included, its queue cleared by a kernel inside the
    captured step) is bit-identical to the built-in schedule's steps."""
    from hamers_amd import ConvectiveFluxReconstructor, Level

    torch = torch_dev
    rec = ConvectiveFluxReconstructor("lvl5_sched", 2, "FIVE_EQN_ALLAIRE", "WCNS6_LD_HLLC_HLL", num_species=2,
                                      species_gamma=(1.6, 1.4), species_R=(1.0, 1.0))
    Qi, n, dx = W.interface_advection_convergence(32, dim=2)
    boxes = [((0, 0), (32, 12)), ((0, 12), (20, 20)), ((20, 12), (12, 20))]
    dt = 0.3 * dx[0]
    ref = Level(rec, boxes, ((0, 0), n), (1, 1), dx)
    ref.write_global(Qi)
    ref.advance(dt, 2, use_graph=use_graph)
    torch.cuda.synchronize()
    want = ref.read_global(Qi.shape)
    lvl = Level(rec, boxes, ((0, 0), n), (1, 1), dx)
    for sc in CANDIDATES_MULTI:
        lvl.set_schedule(sc)
        assert lvl.schedule() == sc
        lvl.write_global(Qi)
        lvl.advance(dt, 2, use_graph=use_graph)
        torch.cuda.synchronize()
        assert np.array_equal(lvl.read_global(Qi.shape), want), sc
    _say(f"five-equation level, graph {use_graph}: {len(CANDIDATES_MULTI)} schedules bit-identical")
