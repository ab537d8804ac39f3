"""GPU parity of the five-equation (Allaire) model against the CPU oracle:
flux and advective source, per-patch and batched, 1D-3D, 2 and 3 species.
Tolerance: max-norm relative error <= 1e-12 per array (FP64)."""
import sys

import numpy as np
import pytest

from oracle import oracle as O
from hamers_amd import workloads as W

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


def _prm(dim, scheme, ns=2, gamma=(1.4, 1.67)):
    return O.Params(dim=dim, model=O.FIVE_EQN_ALLAIRE, num_species=ns, scheme=scheme, gamma=list(gamma))


def _rec(dim, scheme, ns=2, gamma=(1.4, 1.67)):
    from hamers_amd import ConvectiveFluxReconstructor

    key = "WCNS6_LD_HLLC_HLL" if scheme == 1 else "WCNS5_JS_HLLC_HLL"
    return ConvectiveFluxReconstructor("five", dim, "FIVE_EQN_ALLAIRE", key, num_species=ns, species_gamma=gamma,
                                       species_R=(1.0,) * ns)


def _patchdata(torch, Q, n, g, dx, neq):
    from hamers_amd import PatchData

    d = len(n)
    Qd = torch.from_numpy(np.ascontiguousarray(Q)).to("cuda:0")
    F = [torch.full(O.face_shape(neq, n, a), np.nan, dtype=torch.float64, device="cuda:0") for a in range(d)]
    S = torch.zeros((neq,) + tuple(reversed(n)), dtype=torch.float64, device="cuda:0")
    return PatchData(tuple(n), (g,) * d, tuple(dx), Qd, F, S)


def compare(tag, Fg, Fr, Sg, Sr):
    worst = 0.0
    same = total = 0
    for fg, fr in zip(Fg, Fr):
        assert np.all(np.isfinite(fg)), f"{tag}: non-finite GPU flux"
        worst = max(worst, float(np.max(np.abs(fg - fr)) / max(np.max(np.abs(fr)), 1e-300)))
        same += int(np.sum(fg == fr))
        total += fg.size
    serr = float(np.max(np.abs(Sg - Sr)) / max(np.max(np.abs(Sr)), 1e-300)) if np.any(Sr) else float(np.max(np.abs(Sg)))
    same += int(np.sum(Sg == Sr))
    total += Sg.size
    _say(f"{tag}: flux max rel err {worst:.3e}, source max rel err {serr:.3e}, bit-identical {same}/{total}")
    assert worst <= TOL and serr <= TOL, f"{tag}: {worst:.3e} / {serr:.3e}"


def gpu_vs_oracle(torch, tag, Q, n, g, dx, dt, dim, scheme, ns=2, gamma=(1.4, 1.67)):
    rec = _rec(dim, scheme, ns, gamma)
    pd = _patchdata(torch, Q, n, g, dx, rec.num_eqn)
    _say(f"case {tag}")
    rec.computeConvectiveFluxAndSourceOnPatch(pd, 0.0, dt)
    torch.cuda.synchronize()
    Fr, Sr = O.rhs_patch(_prm(dim, scheme, ns, gamma), Q, n, g, dx, dt)
    compare(tag, [f.cpu().numpy() for f in pd.F], Fr, pd.S.cpu().numpy(), Sr)


def shock_interface(n, g, seed, ns=2, gamma=(1.4, 1.67)):
    """A Mach-1.5 shock in species 1 approaching a perturbed material interface."""
    rng = np.random.default_rng(seed)
    d = len(n)
    shp = tuple(reversed([v + 2 * g for v in n]))
    axes = [(np.arange(n[a] + 2 * g) - g + 0.5) / n[a] for a in range(d)]
    X = list(reversed(np.meshgrid(*reversed(axes), indexing="ij")))
    x = X[0]
    pert = sum(0.05 * np.cos(2 * np.pi * (k + 2) * X[1 + (k % (d - 1))] + rng.uniform(0, 6)) for k in range(3)) \
        if d > 1 else 0.0
    phi = 0.5 * (1.0 + np.tanh((x - (0.6 + pert)) / 0.03))  # 0 in species 1, 1 in species 2
    Z = [np.clip(1.0 - phi, 1e-6, 1 - 1e-6)]
    if ns == 3:
        Z = [Z[0] * 0.7, Z[0] * 0.3]
    Z.append(1.0 - sum(Z))
    rho_s = [1.0, 0.2, 3.0][:ns]
    shock = x < 0.3
    M, g1 = 1.5, gamma[0]
    p2 = 1.0 + 2 * g1 / (g1 + 1) * (M * M - 1)
    r2 = (g1 + 1) * M * M / ((g1 - 1) * M * M + 2)
    u2 = M * np.sqrt(g1) * (1 - 1 / r2)
    Zr = [Z[i] * rho_s[i] * np.where(shock, r2, 1.0) for i in range(ns)]
    vel = [np.where(shock, u2, 0.0) + 0.01 * rng.uniform(-1, 1, shp)] + \
          [0.01 * rng.uniform(-1, 1, shp) for _ in range(d - 1)]
    p = np.where(shock, p2, 1.0)
    return W.conservative_from_primitive_five(Zr, vel, p, Z, gamma)


@pytest.mark.parametrize("scheme", [1, 0], ids=["WCNS6_LD", "WCNS5_JS"])
def test_five_eqn_2d_interface_advection(torch_dev, scheme):
    Qi, n, dx = W.interface_advection_convergence(48, dim=2)
    Q = W.fill_periodic(Qi, 4)
    gpu_vs_oracle(torch_dev, f"five-eqn 2D interface scheme={scheme}", Q, n, 4, dx, 1e-3, 2, scheme,
                  gamma=(1.6, 1.4))


@pytest.mark.parametrize("scheme", [1, 0], ids=["WCNS6_LD", "WCNS5_JS"])
def test_five_eqn_3d_shock_interface(torch_dev, scheme):
    n = (16, 12, 10)
    Q = shock_interface(n, 4, seed=3)
    gpu_vs_oracle(torch_dev, f"five-eqn 3D shock-interface scheme={scheme}", Q, n, 4, tuple(1.0 / v for v in n),
                  1e-3, 3, scheme)


def test_five_eqn_2d_three_species(torch_dev):
    n = (24, 20)
    Q = shock_interface(n, 4, seed=5, ns=3, gamma=(1.4, 1.67, 1.3))
    gpu_vs_oracle(torch_dev, "five-eqn 2D three species", Q, n, 4, (1 / 24, 1 / 20), 1e-3, 2, 1, ns=3,
                  gamma=(1.4, 1.67, 1.3))


@pytest.mark.parametrize("scheme", [1, 0], ids=["WCNS6_LD", "WCNS5_JS"])
def test_five_eqn_1d(torch_dev, scheme):
    n = (40,)
    Q = shock_interface(n, 4, seed=7)
    gpu_vs_oracle(torch_dev, f"five-eqn 1D scheme={scheme}", Q, n, 4, (1 / 40,), 1e-3, 1, scheme)


def test_five_eqn_positivity_fallback(torch_dev):
    N, g = 20, 4
    rng = np.random.default_rng(4)
    Zi = rng.uniform(0.0, 1.0, (N, N))
    Qi = W.conservative_from_primitive_five(
        [Zi * 10 ** rng.uniform(-4, 0, (N, N)) + 1e-12, (1 - Zi) * 10 ** rng.uniform(-4, 0, (N, N)) + 1e-12],
        [rng.normal(size=(N, N)), rng.normal(size=(N, N))], 10 ** rng.uniform(-4, 0, (N, N)), [Zi, 1 - Zi],
        (1.4, 1.67))
    Q = W.fill_periodic(Qi, g)
    gpu_vs_oracle(torch_dev, "five-eqn positivity", Q, (N, N), g, (1 / N, 1 / N), 1e-3, 2, 1)


def test_five_eqn_batched_uneven_3d(torch_dev):
    torch = torch_dev
    sizes = [(8, 8, 8), (16, 8, 24), (24, 16, 8), (8, 32, 16), (12, 12, 12)]
    rec = _rec(3, 1)
    pds, refs = [], []
    for i, n in enumerate(sizes):
        Q = shock_interface(n, 4, seed=20 + i)
        dx = tuple(1.0 / v for v in n)
        pds.append(_patchdata(torch, Q, n, 4, dx, rec.num_eqn))
        refs.append(O.rhs_patch(_prm(3, 1), Q, n, 4, dx, 1e-3))
    _say("case five-eqn batched uneven 3D")
    rec.computeConvectiveFluxAndSourceOnPatches(pds, 0.0, 1e-3)
    torch.cuda.synchronize()
    batched = [([f.cpu().numpy().copy() for f in pd.F], pd.S.cpu().numpy().copy()) for pd in pds]
    for pd in pds:
        for f in pd.F:
            f.fill_(np.nan)
        pd.S.zero_()
        rec.computeConvectiveFluxAndSourceOnPatch(pd, 0.0, 1e-3)
    torch.cuda.synchronize()
    for i, pd in enumerate(pds):
        for a in range(3):
            assert np.array_equal(batched[i][0][a], pd.F[a].cpu().numpy()), f"patch {i} axis {a}"
        assert np.array_equal(batched[i][1], pd.S.cpu().numpy())
        compare(f"five-eqn batched patch {i} {sizes[i]}", batched[i][0], refs[i][0], batched[i][1], refs[i][1])


def test_five_eqn_needs_last_volume_fraction(torch_dev):
    from hamers_amd.capi import HamersError

    torch = torch_dev
    n = (8, 8)
    Q = shock_interface(n, 4, seed=1)
    rec = _rec(2, 1)
    pd = _patchdata(torch, Q[:-1].copy(), n, 4, (0.1, 0.1), rec.num_eqn)
    with pytest.raises(ValueError):
        rec.computeConvectiveFluxAndSourceOnPatch(pd, 0.0, 1e-3)


@pytest.mark.parametrize("dim", [2, 3])
def test_five_eqn_bounds_by_axis_gpu(torch_dev, dim):
    """A patch in the Gamma in (-1, 0) window: the reference's bounds check passes
    on x-faces ((Gamma + 1) p / rho > 0) and fails on y- and z-faces
    (Gamma p / rho < 0), which then take first-order states."""
    from test_oracle_five import negative_gamma_patch

    n, g = (12,) * dim, 4
    Q, dx = negative_gamma_patch(n, g)
    gpu_vs_oracle(torch_dev, f"five-eqn {dim}D Gamma<0 bounds by axis", Q, n, g, dx, 1e-3, dim, 1)


def test_five_eqn_null_last_volume_fraction_status(torch_dev):
    """Through the C-ABI itself: a five-equation view without the last volume
    fraction (Q[num_eqn] = NULL) is refused with HM_ERR_INVALID_ARGUMENT and a
    message naming the array, before any launch."""
    import ctypes

    from hamers_amd import capi

    torch = torch_dev
    n = (8, 8)
    Q = shock_interface(n, 4, seed=2)
    rec = _rec(2, 1)
    pd = _patchdata(torch, Q, n, 4, (0.1, 0.1), rec.num_eqn)
    v = pd.view(rec.num_arrays)
    neq = rec.num_eqn
    assert rec.num_arrays == neq + 1
    v.Q[neq] = None
    st = capi.lib.hm_compute_convective_flux_and_source_on_patch(rec._ctx, ctypes.byref(v), 0.0, 1e-3, 0, None)
    assert st == capi.HM_ERR_INVALID_ARGUMENT
    msg = capi.lib.hm_last_error(rec._ctx).decode()
    assert f"conservative variable {neq} is NULL" in msg, msg
    # the batched entry point refuses it the same way
    arr = (capi.hm_patch_view * 1)(v)
    st = capi.lib.hm_compute_convective_flux_and_source_on_patches(rec._ctx, arr, 1, 0.0, 1e-3, 0, None)
    assert st == capi.HM_ERR_INVALID_ARGUMENT


@pytest.mark.parametrize("dim", [2, 3])
def test_five_eqn_tile_shapes(torch_dev, dim):
    """Extents that exercise the fused flux's tiles of whole columns: columns
    of two and three tiles (n_A >= 62 faces: 62, 63, 64, 65, 125, 130 cells),
    columns that do not fill a tile's width, single-cell axes; batched and per
    patch, against the oracle."""
    torch = torch_dev
    if dim == 2:
        sizes = [(62, 5), (63, 3), (64, 7), (125, 4), (3, 130), (1, 70), (65, 1), (9, 13)]
    else:
        sizes = [(64, 6, 5), (6, 65, 4), (5, 4, 62), (1, 3, 64), (12, 9, 7)]
    rec = _rec(dim, 1)
    pds, refs = [], []
    for i, n in enumerate(sizes):
        Q = shock_interface(n, 4, seed=40 + i)
        dx = tuple(1.0 / max(v, 8) for v in n)
        pds.append(_patchdata(torch, Q, n, 4, dx, rec.num_eqn))
        refs.append(O.rhs_patch(_prm(dim, 1), Q, n, 4, dx, 1e-3))
    _say(f"case five-eqn {dim}D tile shapes, batched")
    rec.computeConvectiveFluxAndSourceOnPatches(pds, 0.0, 1e-3)
    torch.cuda.synchronize()
    for i, pd in enumerate(pds):
        compare(f"five-eqn {dim}D tile shapes batched {sizes[i]}", [f.cpu().numpy() for f in pd.F], refs[i][0],
                pd.S.cpu().numpy(), refs[i][1])
    _say(f"case five-eqn {dim}D tile shapes, per patch")
    for i, pd in enumerate(pds):
        for f in pd.F:
            f.fill_(np.nan)
        pd.S.zero_()
        rec.computeConvectiveFluxAndSourceOnPatch(pd, 0.0, 1e-3)
        torch.cuda.synchronize()
        compare(f"five-eqn {dim}D tile shapes per patch {sizes[i]}", [f.cpu().numpy() for f in pd.F], refs[i][0],
                pd.S.cpu().numpy(), refs[i][1])
