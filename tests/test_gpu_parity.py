"""GPU parity: the HIP path through the C-ABI against the CPU oracle on the
same inputs.

Tolerance (floating point, FP64): max-norm relative error <= 1e-12 per array.
Both sides are built without multiply-add contraction and keep the
reference's expression trees, so most values agree bit for bit; each test
also reports the bit-identical fraction.
"""
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


def _rec(dim, scheme, **kw):
    from hamers_amd import ConvectiveFluxReconstructor

    key = "WCNS6_LD_HLLC_HLL" if scheme == 1 else "WCNS5_JS_HLLC_HLL"
    return ConvectiveFluxReconstructor("test", dim, "SINGLE_SPECIES", key, species_gamma=(1.4,), **kw)


def _to_dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


def _patchdata(torch, Q, n, g, dx, neq):
    from hamers_amd import PatchData

    d = len(n)
    Qd = _to_dev(torch, Q)
    F = [torch.full(O.face_shape(neq, n, a), np.nan, dtype=torch.float64, device="cuda:0") for a in range(d)]
    S = torch.zeros((neq,) + tuple(reversed(n)), dtype=torch.float64, device="cuda:0")
    return PatchData(tuple(n), (g,) * d, tuple(dx), Qd, F, S)


def gpu_rhs(torch, rec, Q, n, g, dx, dt):
    pd = _patchdata(torch, Q, n, g, dx, rec.num_eqn)
    rec.computeConvectiveFluxAndSourceOnPatch(pd, 0.0, dt)
    torch.cuda.synchronize()
    return [f.cpu().numpy() for f in pd.F], pd.S.cpu().numpy()


def compare(tag, F_gpu, F_ref, S_gpu=None, S_ref=None):
    worst = 0.0
    same = total = 0
    for a, (fg, fr) in enumerate(zip(F_gpu, F_ref)):
        assert fg.shape == fr.shape
        assert np.all(np.isfinite(fg)), f"{tag}: non-finite GPU flux on axis {a}"
        scale = max(np.max(np.abs(fr)), 1e-300)
        err = float(np.max(np.abs(fg - fr)) / scale)
        worst = max(worst, err)
        same += int(np.sum(fg == fr))
        total += fg.size
    if S_gpu is not None:
        scale = max(np.max(np.abs(S_ref)), 1e-300)
        worst = max(worst, float(np.max(np.abs(S_gpu - S_ref)) / scale) if np.any(S_ref) else float(np.max(np.abs(S_gpu))))
    _say(f"{tag}: max rel err {worst:.3e}, bit-identical {same}/{total}")
    assert worst <= TOL, f"{tag}: max rel err {worst:.3e} > {TOL}"
    return worst, same / max(total, 1)


def _random_smooth_3d(n, g, seed):
    rng = np.random.default_rng(seed)
    shp = tuple(reversed([v + 2 * g for v in n]))
    z, y, x = np.meshgrid(*[np.linspace(0, 1, s) for s in shp], indexing="ij")
    rho = 1.0 + 0.3 * np.sin(2 * np.pi * (x + 2 * y + 0.5 * z)) + 0.05 * rng.uniform(-1, 1, shp)
    u = 0.4 * np.cos(2 * np.pi * (y - z)) + 0.05 * rng.uniform(-1, 1, shp)
    v = -0.3 * np.sin(2 * np.pi * (x + z))
    w = 0.2 * np.cos(2 * np.pi * x)
    p = 1.0 + 0.2 * np.cos(2 * np.pi * (x - y)) + np.where(x > 0.5, 1.5, 0.0)
    return W.conservative_from_primitive_single(rho, [u, v, w], p, 1.4)


# ---------------------------------------------------------------------------

@pytest.mark.parametrize("scheme", [1, 0], ids=["WCNS6_LD", "WCNS5_JS"])
def test_density_wave_single_patch_128(torch_dev, scheme):
    torch = torch_dev
    Qi, n, dx = W.density_wave_unit(128)
    Q = W.fill_periodic(Qi, 4)
    dt = 0.4 * dx[0] / (1.0 + np.sqrt(1.4))
    _say(f"case density wave 128^2 scheme={scheme}")
    Fg, Sg = gpu_rhs(torch, _rec(2, scheme), Q, n, 4, dx, dt)
    Fr, Sr = O.rhs_patch(O.Params(dim=2, scheme=scheme), Q, n, 4, dx, dt)
    compare(f"density wave scheme={scheme}", Fg, Fr, Sg, Sr)


@pytest.mark.parametrize("scheme", [1, 0], ids=["WCNS6_LD", "WCNS5_JS"])
def test_four_quadrant_riemann(torch_dev, scheme):
    torch = torch_dev
    Qi, n, dx = W.four_quadrant_riemann(N=96, seed=11)
    Q = W.fill_periodic(Qi, 4)
    _say(f"case four-quadrant 96^2 scheme={scheme}")
    Fg, Sg = gpu_rhs(torch, _rec(2, scheme), Q, n, 4, dx, 1e-3)
    Fr, Sr = O.rhs_patch(O.Params(dim=2, scheme=scheme), Q, n, 4, dx, 1e-3)
    compare(f"four-quadrant scheme={scheme}", Fg, Fr, Sg, Sr)


@pytest.mark.parametrize("scheme", [1, 0], ids=["WCNS6_LD", "WCNS5_JS"])
def test_three_dimensional_patch(torch_dev, scheme):
    torch = torch_dev
    n, g = (12, 10, 8), 4
    Q = _random_smooth_3d(n, g, seed=4)
    dx = (1 / 12, 1 / 10, 1 / 8)
    _say(f"case 3D 12x10x8 scheme={scheme}")
    Fg, Sg = gpu_rhs(torch, _rec(3, scheme), Q, n, g, dx, 2e-3)
    Fr, Sr = O.rhs_patch(O.Params(dim=3, scheme=scheme), Q, n, g, dx, 2e-3)
    compare(f"3D scheme={scheme}", Fg, Fr, Sg, Sr)


@pytest.mark.parametrize("scheme", [1, 0], ids=["WCNS6_LD", "WCNS5_JS"])
def test_one_dimensional_shock_tube(torch_dev, scheme):
    torch = torch_dev
    N, g = 40, 4
    x = (np.arange(N + 2 * g) - g + 0.5) / N
    rho = np.where(x < 0.5, 1.0, 0.125)
    p = np.where(x < 0.5, 1.0, 0.1)
    u = np.where(x < 0.5, 0.75, 0.0)
    Q = W.conservative_from_primitive_single(rho, [u], p, 1.4)
    _say(f"case 1D Sod-like N=40 scheme={scheme}")
    Fg, Sg = gpu_rhs(torch, _rec(1, scheme), Q, (N,), g, (1 / N,), 1e-3)
    Fr, Sr = O.rhs_patch(O.Params(dim=1, scheme=scheme), Q, (N,), g, (1 / N,), 1e-3)
    compare(f"1D scheme={scheme}", Fg, Fr, Sg, Sr)


@pytest.mark.parametrize("scheme", [1, 0], ids=["WCNS6_LD", "WCNS5_JS"])
def test_positivity_fallback_faces(torch_dev, scheme):
    torch = torch_dev
    N = 24
    rng = np.random.default_rng(3)
    rho = 10 ** rng.uniform(-6, 0, (N, N))
    p = 10 ** rng.uniform(-6, 0, (N, N))
    u = rng.normal(size=(N, N))
    Qi = W.conservative_from_primitive_single(rho, [u, np.zeros_like(rho)], p, 1.4)
    Q = W.fill_periodic(Qi, 4)
    _say(f"case positivity fallback scheme={scheme}")
    Fg, Sg = gpu_rhs(torch, _rec(2, scheme), Q, (N, N), 4, (1 / N, 1 / N), 1e-3)
    Fr, Sr = O.rhs_patch(O.Params(dim=2, scheme=scheme), Q, (N, N), 4, (1 / N, 1 / N), 1e-3)
    compare(f"positivity scheme={scheme}", Fg, Fr, Sg, Sr)


def test_sensor_selects_hllc_hll_under_compression(torch_dev):
    torch = torch_dev
    N, g = 32, 4
    x = (np.arange(N + 2 * g) - g + 0.5) / N
    X, Y = np.meshgrid(x, x, indexing="xy")
    rho = 1.0 + 0.1 * np.sin(2 * np.pi * Y)
    u = -2.0 * (X - 0.5)          # strong compression: theta < 0, sensor -> 1
    v = 0.3 * np.sin(2 * np.pi * X)
    p = np.ones_like(X)
    Q = W.conservative_from_primitive_single(rho, [u, v], p, 1.4)
    _say("case compression (HLLC-HLL selected)")
    Fg, Sg = gpu_rhs(torch, _rec(2, 1), Q, (N, N), g, (1 / N, 1 / N), 1e-3)
    Fr, Sr = O.rhs_patch(O.Params(dim=2, scheme=1), Q, (N, N), g, (1 / N, 1 / N), 1e-3)
    compare("compression", Fg, Fr, Sg, Sr)


def test_wider_ghost_and_nonsquare_patch(torch_dev):
    torch = torch_dev
    n, g = (17, 9), 6
    rng = np.random.default_rng(8)
    shp = tuple(reversed([v + 2 * g for v in n]))
    rho = 1.0 + 0.2 * rng.uniform(-1, 1, shp)
    u = 0.3 * rng.uniform(-1, 1, shp)
    v = 0.3 * rng.uniform(-1, 1, shp)
    p = 1.0 + 0.2 * rng.uniform(-1, 1, shp)
    Q = W.conservative_from_primitive_single(rho, [u, v], p, 1.4)
    _say("case ghost 6, 17x9")
    Fg, Sg = gpu_rhs(torch, _rec(2, 1), Q, n, g, (0.1, 0.2), 1e-3)
    Fr, Sr = O.rhs_patch(O.Params(dim=2, scheme=1), Q, n, g, (0.1, 0.2), 1e-3)
    compare("ghost 6", Fg, Fr, Sg, Sr)


@pytest.mark.parametrize("dim", [2, 3])
def test_batched_uneven_patches_match_per_patch_and_oracle(torch_dev, dim):
    """Uneven patch sizes through the batched entry point: bit-identical to the
    per-patch entry point, and within tolerance of the oracle."""
    torch = torch_dev
    rng = np.random.default_rng(dim)
    g = 4
    sizes = [(8, 8, 8), (16, 8, 24), (24, 16, 8), (8, 32, 16), (12, 12, 12)] if dim == 3 else \
        [(8, 8), (16, 40), (64, 8), (24, 24), (128, 16), (8, 72)]
    rec = _rec(dim, 1)
    pds, refs = [], []
    for i, nn in enumerate(sizes):
        n = nn[:dim]
        Q = _random_smooth_3d(n, g, seed=10 + i) if dim == 3 else None
        if dim == 2:
            shp = tuple(reversed([v + 2 * g for v in n]))
            rho = 10 ** rng.uniform(-1, 0, shp)
            p = 10 ** rng.uniform(-1, 0, shp)
            Q = W.conservative_from_primitive_single(rho, [rng.normal(size=shp), rng.normal(size=shp)], p, 1.4)
        dx = tuple(1.0 / v for v in n)
        pds.append(_patchdata(torch, Q, n, g, dx, rec.num_eqn))
        refs.append(O.rhs_patch(O.Params(dim=dim, scheme=1), Q, n, g, dx, 1e-3))
    _say(f"case batched uneven dim={dim} ({len(sizes)} patches)")
    rec.computeConvectiveFluxAndSourceOnPatches(pds, 0.0, 1e-3)
    torch.cuda.synchronize()
    batched = [[f.cpu().numpy().copy() for f in pd.F] for pd in pds]
    for i, pd in enumerate(pds):
        for f in pd.F:
            f.fill_(np.nan)
        rec.computeConvectiveFluxAndSourceOnPatch(pd, 0.0, 1e-3)
    torch.cuda.synchronize()
    for i, pd in enumerate(pds):
        single = [f.cpu().numpy() for f in pd.F]
        for a in range(dim):
            assert np.array_equal(batched[i][a], single[a]), f"patch {i} axis {a}: batched != per-patch"
        compare(f"batched patch {i} {sizes[i][:dim]}", batched[i], refs[i][0])


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_tiny_and_odd_extents(torch_dev, dim):
    """Patches of 1..7 cells per axis (batched and per patch): the axis-0
    fused flux packs rows of n0+3 midpoint faces into wavefronts of 64 with
    62 output lanes, so short rows, rows split between wavefronts and the
    last partial wavefront all occur."""
    torch = torch_dev
    g = 4
    sizes = {1: [(1,), (2,), (3,), (5,), (61,), (62,), (63,), (124,)],
             2: [(1, 1), (1, 5), (2, 3), (3, 7), (7, 2), (59, 3), (60, 1)],
             3: [(1, 1, 1), (1, 2, 3), (3, 1, 2), (2, 5, 1), (7, 3, 2)]}[dim]
    rec = _rec(dim, 1)
    pds, refs = [], []
    for i, n in enumerate(sizes):
        Q = _random_smooth_3d(n, g, seed=40 + i) if dim == 3 else None
        if dim < 3:
            rng = np.random.default_rng(100 + i)
            shp = tuple(reversed([v + 2 * g for v in n]))
            rho = 10 ** rng.uniform(-1, 0, shp)
            p = 10 ** rng.uniform(-1, 0, shp)
            vel = [rng.normal(size=shp) for _ in range(dim)]
            Q = W.conservative_from_primitive_single(rho, vel, p, 1.4)
        dx = tuple(1.0 / (v + 1) for v in n)
        pds.append(_patchdata(torch, Q, n, g, dx, rec.num_eqn))
        refs.append(O.rhs_patch(O.Params(dim=dim, scheme=1), Q, n, g, dx, 1e-3))
    _say(f"case tiny and odd extents dim={dim} ({len(sizes)} patches)")
    rec.computeConvectiveFluxAndSourceOnPatches(pds, 0.0, 1e-3)
    torch.cuda.synchronize()
    batched = [[f.cpu().numpy().copy() for f in pd.F] for pd in pds]
    for pd in pds:
        for f in pd.F:
            f.fill_(np.nan)
        rec.computeConvectiveFluxAndSourceOnPatch(pd, 0.0, 1e-3)
    torch.cuda.synchronize()
    for i, pd in enumerate(pds):
        single = [f.cpu().numpy() for f in pd.F]
        for a in range(dim):
            assert np.array_equal(batched[i][a], single[a]), f"patch {i} axis {a}: batched != per-patch"
        compare(f"tiny patch {i} {sizes[i]}", batched[i], refs[i][0])


def _tile_rows(shapes):
    """Rows per tile of the single-species kernel's axis-1/2 stage for a call
    over patches of these interior shapes: 24 when 24-row tiles give 2048
    blocks, else 8 when 8-row tiles give 512, else 4."""
    def blocks(rows):
        b = 0
        for n in shapes:
            for a in range(1, len(n)):
                trans = int(np.prod([n[c] for c in range(1, len(n)) if c != a]))
                b += -(-n[0] // 64) * trans * -(-(n[a] + 1) // (rows - 2))
        return b
    return 24 if blocks(24) >= 2048 else (8 if blocks(8) >= 512 else 4)


@pytest.mark.parametrize("case", ["2d_level_8_rows", "3d_level_8_rows", "2d_uneven_24_rows"])
def test_tile_row_bands_agree(torch_dev, case):
    """The axis-1/2 tiles come in 24, 8 and 4 rows by the size of the call:
    a level runs batched on 8-row (or 24-row) tiles and per patch on 8- or
    4-row tiles; all must equal each other bit for bit and the oracle."""
    torch = torch_dev
    g = 4
    if case == "2d_level_8_rows":
        Qi, _, dx = W.four_quadrant_riemann(N=512, seed=77)
        patches = [Qp for _, Qp in W.split_patches(Qi, (128, 128), g)]
        shapes = [(128, 128)] * len(patches)
        rows, single_rows = 8, 4
    elif case == "3d_level_8_rows":
        shapes = [(40, 36, 30), (64, 20, 24), (33, 50, 17), (70, 40, 28)]
        patches = [_random_smooth_3d(n, g, seed=300 + i) for i, n in enumerate(shapes)]
        dx = (0.02, 0.025, 0.03)
        rows, single_rows = 8, 4
    else:
        # x extents off the 64-lane tile width, rows off the 22-face tile height
        shapes = [(1000, 1000), (777, 1203), (1500, 600), (333, 1111), (65, 1)]
        patches = []
        for i, n in enumerate(shapes):
            rng = np.random.default_rng(500 + i)
            shp = tuple(reversed([v + 2 * g for v in n]))
            rho = 10 ** rng.uniform(-1, 0, shp)
            p = 10 ** rng.uniform(-1, 0, shp)
            patches.append(W.conservative_from_primitive_single(rho, [rng.normal(size=shp) for _ in range(2)], p, 1.4))
        dx = (1.0e-3, 1.0e-3)
        rows, single_rows = 24, 8
    dim = len(shapes[0])
    assert _tile_rows(shapes) == rows and _tile_rows(shapes[:1]) == single_rows
    rec = _rec(dim, 0)
    pds = [_patchdata(torch, Qp, n, g, dx, rec.num_eqn) for Qp, n in zip(patches, shapes)]
    _say(f"case tile-row bands {case} ({len(pds)} patches)")
    rec.computeConvectiveFluxAndSourceOnPatches(pds, 0.0, 1e-4)
    torch.cuda.synchronize()
    batched = [[f.cpu().numpy().copy() for f in pd.F] for pd in pds]
    for pd in pds:
        for f in pd.F:
            f.fill_(np.nan)
        rec.computeConvectiveFluxAndSourceOnPatch(pd, 0.0, 1e-4)
    torch.cuda.synchronize()
    for i, pd in enumerate(pds):
        for a in range(dim):
            assert np.array_equal(batched[i][a], pd.F[a].cpu().numpy()), f"patch {i} axis {a}: batched != per-patch"
    for i in (0, len(pds) - 1):
        Fr, _ = O.rhs_patch(O.Params(dim=dim, scheme=0), patches[i], shapes[i], g, dx, 1e-4)
        compare(f"tile-row bands patch {i}", batched[i], Fr)


def test_invalid_ghost_width_is_an_error_not_a_fault(torch_dev):
    from hamers_amd.capi import HamersError

    torch = torch_dev
    n, g = (8, 8), 3
    Q = np.ones((4, 14, 14))
    Q[3] = 2.5
    pd = _patchdata(torch, Q, n, g, (0.1, 0.1), 4)
    with pytest.raises(HamersError):
        _rec(2, 1).computeConvectiveFluxAndSourceOnPatch(pd, 0.0, 1e-3)


def test_full_size_2048_four_quadrant_batched(torch_dev):
    """The 2048^2 / 256-patch workload: sampled patches against the oracle,
    plus the periodic conservation property of the whole level."""
    torch = torch_dev
    Qi, n, dx = W.four_quadrant_riemann(N=2048, seed=1234)
    g = 4
    patches = W.split_patches(Qi, (128, 128), g)
    rec = _rec(2, 0)
    pds = [_patchdata(torch, Qp, (128, 128), g, dx, 4) for _, Qp in patches]
    _say("case 2048^2 four-quadrant batched (256 patches)")
    rec.computeConvectiveFluxAndSourceOnPatches(pds, 0.0, 1e-4)
    torch.cuda.synchronize()
    # conservation: with periodic data the summed flux divergence vanishes
    tot = np.zeros(4)
    mag = np.zeros(4)
    for pd in pds:
        Fx, Fy = pd.F[0].cpu().numpy(), pd.F[1].cpu().numpy()
        div = (Fx[:, :, 1:] - Fx[:, :, :-1]) / dx[0] + (Fy[:, 1:, :] - Fy[:, :-1, :]) / dx[1]
        tot += div.sum(axis=(1, 2))
        mag += np.abs(div).sum(axis=(1, 2))
    assert np.all(np.abs(tot) <= 1e-11 * mag)
    for idx in (0, 17, 100, 255):
        Fr, _ = O.rhs_patch(O.Params(dim=2, scheme=0), patches[idx][1], (128, 128), g, dx, 1e-4)
        compare(f"2048^2 patch {idx}", [f.cpu().numpy() for f in pds[idx].F], Fr)
