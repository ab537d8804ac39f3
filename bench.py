"""Benchmark of the convective-flux RHS on one MI355X.

Workload (one "step" = one RHS evaluation of a whole AMR level): 2D
single-species Euler (gamma 1.4), FP64, a 2048^2 periodic domain split into
256 patches of 128^2 with 4-cell halos, four-quadrant Riemann data (seeded)
with 1e-3 white noise, WCNS5-JS + hybrid HLLC/HLLC-HLL.  Inputs are resident
in HBM before the timed region.  The batched path (one work-list launch per
pipeline stage for the whole level) is the headline; the per-patch path (one
launch sequence per patch, the plain drop-in) is timed beside it, and the
CPU oracle is timed on a bounded sample of patches on the host cores.

A second workload is timed beside it and reported under "uneven_3d": 3D
five-equation (Allaire) two-gas Euler, a Mach-1.5 shock in gas 1 (gamma 1.4)
approaching a perturbed interface with gas 2 (gamma 1.67), WCNS6-LD + hybrid
HLLC/HLLC-HLL, on a level tiled by uneven patches whose extents per axis are
drawn (seeded) from {8, 16, 24, 32, 48, 64}; batched RHS against the
per-patch path.  Its fields are generated on the GPU with torch (synthetic
input only; every RHS runs in libhamers_amd.so).

A third leg, "navier_stokes_3d", times the convective + diffusive RHS of a
3D four-equation two-species Navier-Stokes level of uneven patches (multimode
interface, constant transport properties), batched against per-patch calls.

A fourth leg, reported under "adjoint", times the discrete adjoint of 200
SSP-RK3 steps on a 512^2 Kelvin-Helmholtz level (16 patches of 128^2,
checkpoints every 20 steps) against the forward run, and checks its gradient
w.r.t. 8 mode amplitudes against central differences.

Prints one JSON line (rank 0).
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E
# Kernels of the single-species pipeline's two flux stages.
STAGE_KERNEL = {"midflux": "k_flux_single", "faceflux": "k_flux_tile_single"}
# DRAM bytes per launch (FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0) of each flux
# kernel, 2048^2 in 128^2 patches, WCNS5-JS.
FLUX_TRAFFIC_2048_WCNS5 = {"midflux": 354.92e6 + 135.27e6, "faceflux": 506.47e6 + 135.27e6}
FP64_PEAK_TFLOPS = 78.6        # MI355X FP64 vector


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    # The GPU's clocks ramp over the first few milliseconds of load; ten
    # untimed RHS evaluations (about 8 ms) keep the ramp out of the timing.
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--N", type=int, default=2048)
    ap.add_argument("--patch", type=int, default=128)
    ap.add_argument("--scheme", default="WCNS5_JS_HLLC_HLL")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="target CPU-baseline work")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-per-patch", action="store_true")
    ap.add_argument("--uneven-seg", type=int, default=14, help="patches per axis of the uneven 3D level (0: skip)")
    ap.add_argument("--uneven-steps", type=int, default=5)
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="only copy this many MiB device to device 5 times (calibrates HBM byte counters) and exit")
    ap.add_argument("--ns-seg", type=int, default=8, help="patches per axis of the uneven 3D Navier-Stokes level (0: skip)")
    ap.add_argument("--ns-steps", type=int, default=4)
    ap.add_argument("--adjoint-steps", type=int, default=200, help="RK3 steps of the adjoint leg (0: skip)")
    ap.add_argument("--adjoint-N", type=int, default=512)
    ap.add_argument("--adjoint-patch", type=int, default=128)
    ap.add_argument("--adjoint-every", type=int, default=20)
    return ap.parse_args()


UNEVEN_CHOICES = (8, 16, 24, 32, 48, 64)


def shock_interface_patch_torch(torch, lo, n, g, dx, modes, gamma=(1.4, 1.67), rho_gas=(1.0, 0.2)):
    """Conservative five-equation arrays (8, n2+2g, n1+2g, n0+2g) of one patch of the
    shock / perturbed-interface field, halos from the same analytic field."""
    dev = "cuda:0"
    ax = [(torch.arange(n[a] + 2 * g, dtype=torch.float64, device=dev) + (lo[a] - g) + 0.5) * dx[a]
          for a in range(3)]
    z, y, x = torch.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
    xi = torch.full_like(y, 0.6)
    for k, (amp, ph_y, ph_z) in zip(range(2, 9), modes):
        xi = xi + 0.05 * amp * 0.5 * (torch.cos(2 * torch.pi * k * y + ph_y) + torch.cos(2 * torch.pi * k * z + ph_z))
    phi = 0.5 * (1.0 + torch.tanh((x - xi) / (2.0 * dx[0])))
    Z1 = torch.clamp(1.0 - phi, 1e-6, 1.0 - 1e-6)
    Z2 = 1.0 - Z1
    M, g1 = 1.5, gamma[0]
    p2 = 1.0 + 2 * g1 / (g1 + 1) * (M * M - 1)
    r2 = (g1 + 1) * M * M / ((g1 - 1) * M * M + 2)
    u2 = M * (g1 ** 0.5) * (1 - 1 / r2)
    shock = x < 0.3
    comp = torch.where(shock, torch.full_like(x, r2), torch.ones_like(x))
    Zr1 = Z1 * rho_gas[0] * comp
    Zr2 = Z2 * rho_gas[1] * comp
    rho = Zr1 + Zr2
    u = torch.where(shock, torch.full_like(x, u2), torch.zeros_like(x))
    p = torch.where(shock, torch.full_like(x, p2), torch.ones_like(x))
    E = p * (Z1 / (gamma[0] - 1.0) + Z2 / (gamma[1] - 1.0)) + 0.5 * rho * u * u
    zero = torch.zeros_like(x)
    return torch.stack([Zr1, Zr2, rho * u, zero, zero, E, Z1, Z2]).contiguous()


def bench_uneven_3d(args, torch, np, time):
    """Batched vs per-patch RHS of the uneven-patch 3D five-equation level."""
    import ctypes

    from hamers_amd import ConvectiveFluxReconstructor, PatchData, PatchList, capi

    rng = np.random.default_rng(2024)
    segs = [rng.choice(UNEVEN_CHOICES, size=args.uneven_seg) for _ in range(3)]
    starts = [np.concatenate([[0], np.cumsum(s)[:-1]]) for s in segs]
    Ng = [int(s.sum()) for s in segs]
    dx = tuple(1.0 / v for v in Ng)
    modes = [(rng.uniform(0.0, 1.0), rng.uniform(0, 2 * np.pi), rng.uniform(0, 2 * np.pi)) for _ in range(7)]
    tot = sum(m[0] for m in modes)
    modes = [(m[0] / tot, m[1], m[2]) for m in modes]
    g, neq = 4, 7
    rec = ConvectiveFluxReconstructor("uneven", 3, "FIVE_EQN_ALLAIRE", "WCNS6_LD_HLLC_HLL", num_species=2,
                                      species_gamma=(1.4, 1.67), species_R=(1.0, 1.0))
    dev = "cuda:0"
    pds, boxes = [], []
    for k in range(args.uneven_seg):
        for j in range(args.uneven_seg):
            for i in range(args.uneven_seg):
                lo = (int(starts[0][i]), int(starts[1][j]), int(starts[2][k]))
                n = (int(segs[0][i]), int(segs[1][j]), int(segs[2][k]))
                Q = shock_interface_patch_torch(torch, lo, n, g, dx, modes)
                F = [torch.empty((neq, n[2] + (a == 2), n[1] + (a == 1), n[0] + (a == 0)), dtype=torch.float64,
                                 device=dev) for a in range(3)]
                S = torch.zeros((neq, n[2], n[1], n[0]), dtype=torch.float64, device=dev)
                pds.append(PatchData(n, (g, g, g), dx, Q, F, S))
                boxes.append((lo, n))
    cells = sum(b[1][0] * b[1][1] * b[1][2] for b in boxes)
    plist = PatchList(pds, rec.num_arrays)
    dt = 1.0e-4
    stream = torch.cuda.current_stream()
    sp = ctypes.c_void_p(stream.cuda_stream)
    fn = capi.lib.hm_compute_convective_flux_and_source_on_patch
    views = [pd.view(rec.num_arrays) for pd in pds]

    def batched():
        rec.computeConvectiveFluxAndSourceOnPatches(plist, 0.0, dt, 0, stream)

    def per_patch():
        for v in views:
            st = fn(rec._ctx, ctypes.byref(v), 0.0, dt, 0, sp)
            if st:
                capi.check(rec._ctx, st)

    def timed(f, steps):
        f()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(steps):
            f()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / steps

    ms_b = timed(batched, args.uneven_steps)
    rec.enable_profiling(True)
    for _ in range(args.uneven_steps):
        batched()
    prof = rec.read_profile()
    rec.enable_profiling(False)
    stage = {k: (v[0] / v[1] if v[1] else 0.0) for k, v in prof.items()}
    ms_p = timed(per_patch, max(2, args.uneven_steps // 2)) if not args.no_per_patch else None
    bytes_per_cell = 8 * (neq + 3 * neq + 2 * 1)  # Q read, F on 3 axes, S read-modify-write (1 advective eq)
    ext = [np.bincount(np.concatenate(segs), minlength=65)[list(UNEVEN_CHOICES)].tolist()]
    return {
        "workload": "3D five-equation Allaire two-gas Euler, Mach-1.5 shock and perturbed interface, "
                    "WCNS6_LD_HLLC_HLL, uneven patches (extents per axis from {8,16,24,32,48,64})",
        "patches": len(pds), "cells": int(cells), "domain": Ng, "extent_counts": ext[0],
        "ms_per_rhs_batched": ms_b, "value_batched": cells / (ms_b * 1e-3),
        "ms_per_rhs_per_patch": ms_p, "value_per_patch": (cells / (ms_p * 1e-3)) if ms_p else None,
        "batched_speedup": (ms_p / ms_b) if ms_p else None,
        "stage_ms": stage, "bytes_per_cell": bytes_per_cell,
        "rhs_hbm_frac": bytes_per_cell * cells / (ms_b * 1e-3) / 1e9 / HBM_PEAK_GBS,
    }


def rtm_interface_patch_torch(torch, lo, n, g, dx, modes, gamma=(1.4, 1.67), R=(1.0, 0.25), rho_gas=(3.0, 1.0)):
    """Conservative four-equation arrays (6, n2+2g, n1+2g, n0+2g) of one patch of a
    multimode heavy/light interface x_I(y, z) = 0.5 + sum_k a_k (cos(2 pi k y + phi_k)
    + cos(2 pi k z + psi_k)) / 2 over k in [4, 32] with a_k ~ k^-2, smoothed over
    two cells, at rest at p = 1 (halos from the same analytic field)."""
    dev = "cuda:0"
    ax = [(torch.arange(n[a] + 2 * g, dtype=torch.float64, device=dev) + (lo[a] - g) + 0.5) * dx[a]
          for a in range(3)]
    z, y, x = torch.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
    xi = torch.full_like(y, 0.5)
    for k, (amp, ph_y, ph_z) in zip(range(4, 33), modes):
        xi = xi + amp * 0.5 * (torch.cos(2 * torch.pi * k * y + ph_y) + torch.cos(2 * torch.pi * k * z + ph_z))
    Y1 = 0.5 * (1.0 - torch.tanh((x - xi) / (2.0 * dx[0])))
    Y1 = torch.clamp(Y1, 1e-6, 1.0 - 1e-6)
    rho = 1.0 / (Y1 / rho_gas[0] + (1.0 - Y1) / rho_gas[1])
    cp = [gm / (gm - 1.0) * r for gm, r in zip(gamma, R)]
    cv = [1.0 / (gm - 1.0) * r for gm, r in zip(gamma, R)]
    gmix = (Y1 * cp[0] + (1 - Y1) * cp[1]) / (Y1 * cv[0] + (1 - Y1) * cv[1])
    E = 1.0 / (gmix - 1.0)
    zero = torch.zeros_like(x)
    return torch.stack([rho * Y1, rho * (1.0 - Y1), zero, zero, zero, E]).contiguous()


def bench_ns_3d(args, torch, np, time):
    """Convective + diffusive RHS of an uneven-patch 3D four-equation
    Navier-Stokes level, batched against per-patch calls."""
    import ctypes

    from hamers_amd import ConvectiveFluxReconstructor, DiffusiveFluxReconstructor, PatchData, PatchList, capi

    rng = np.random.default_rng(4096)
    segs = [rng.choice(UNEVEN_CHOICES, size=args.ns_seg) for _ in range(3)]
    starts = [np.concatenate([[0], np.cumsum(s)[:-1]]) for s in segs]
    Ng = [int(s.sum()) for s in segs]
    dx = tuple(1.0 / v for v in Ng)
    ks = np.arange(4, 33)
    amps = ks ** -2.0
    amps = 0.02 * amps / amps.sum()
    modes = [(float(a), rng.uniform(0, 2 * np.pi), rng.uniform(0, 2 * np.pi)) for a in amps]
    g, neq = 6, 6
    gamma, R = (1.4, 1.67), (1.0, 0.25)
    rec = ConvectiveFluxReconstructor("ns", 3, "FOUR_EQN_CONSERVATIVE", "WCNS6_LD_HLLC_HLL", num_species=2,
                                      species_gamma=gamma, species_R=R)
    diff = DiffusiveFluxReconstructor(rec, species_mu=(1e-3, 2e-3), species_mu_v=(0.0, 0.0),
                                      species_kappa=(1.5e-3, 2.5e-3), species_D=(1e-3, 1e-3), species_M=(1.0, 4.0))
    dev = "cuda:0"
    conv, dif, boxes = [], [], []
    for k in range(args.ns_seg):
        for j in range(args.ns_seg):
            for i in range(args.ns_seg):
                lo = (int(starts[0][i]), int(starts[1][j]), int(starts[2][k]))
                n = (int(segs[0][i]), int(segs[1][j]), int(segs[2][k]))
                Q = rtm_interface_patch_torch(torch, lo, n, g, dx, modes, gamma, R)
                shp = [(neq, n[2] + (a == 2), n[1] + (a == 1), n[0] + (a == 0)) for a in range(3)]
                Fc = [torch.empty(s_, dtype=torch.float64, device=dev) for s_ in shp]
                Fd = [torch.empty(s_, dtype=torch.float64, device=dev) for s_ in shp]
                S = torch.zeros((neq, n[2], n[1], n[0]), dtype=torch.float64, device=dev)
                conv.append(PatchData(n, (g, g, g), dx, Q, Fc, S))
                dif.append(PatchData(n, (g, g, g), dx, Q, Fd, None))
                boxes.append((lo, n))
    cells = sum(b[1][0] * b[1][1] * b[1][2] for b in boxes)
    pc, pdf = PatchList(conv, neq), PatchList(dif, neq)
    dt = 1.0e-4
    stream = torch.cuda.current_stream()
    sp = ctypes.c_void_p(stream.cuda_stream)
    fc, fd = capi.lib.hm_compute_convective_flux_and_source_on_patch, capi.lib.hm_compute_diffusive_flux_on_patch
    vc = [p.view(neq) for p in conv]
    vd = [p.view(neq) for p in dif]

    def batched_conv():
        rec.computeConvectiveFluxAndSourceOnPatches(pc, 0.0, dt, 0, stream)

    def batched_diff():
        diff.computeDiffusiveFluxOnPatches(pdf, 0.0, dt, 0, stream)

    def batched():
        batched_conv()
        batched_diff()

    def per_patch():
        for a, b in zip(vc, vd):
            st = fc(rec._ctx, ctypes.byref(a), 0.0, dt, 0, sp) or fd(rec._ctx, ctypes.byref(b), 0.0, dt, 0, sp)
            if st:
                capi.check(rec._ctx, st)

    def timed(f, steps):
        f()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(steps):
            f()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / steps

    ms_b = timed(batched, args.ns_steps)
    ms_bc = timed(batched_conv, args.ns_steps)
    ms_bd = timed(batched_diff, args.ns_steps)
    ms_p = timed(per_patch, max(2, args.ns_steps // 2)) if not args.no_per_patch else None
    return {
        "workload": "3D four-equation two-species Navier-Stokes (constant transport), multimode interface with "
                    "a k^-2 spectrum over k in [4, 32], WCNS6_LD_HLLC_HLL convective + sixth-order node diffusive "
                    "flux, one level of uneven patches (extents per axis from {8,16,24,32,48,64}), 6-cell halos",
        "patches": len(conv), "cells": int(cells), "domain": Ng,
        "ms_per_rhs_batched": ms_b, "ms_convective_batched": ms_bc, "ms_diffusive_batched": ms_bd,
        "value_batched": cells / (ms_b * 1e-3),
        "ms_per_rhs_per_patch": ms_p, "value_per_patch": (cells / (ms_p * 1e-3)) if ms_p else None,
        "batched_speedup": (ms_p / ms_b) if ms_p else None,
        "bytes_per_cell": 8 * (neq + 3 * neq + 3 * neq),
    }


def bench_adjoint(args, torch, np, time):
    """Discrete adjoint of SSP-RK3 on a Kelvin-Helmholtz level: 512^2 in 16
    patches of 128^2, WCNS6-LD + hybrid HLLC/HLLC-HLL, 200 steps, checkpoints
    every 20 steps with recompute; J = transverse kinetic energy at the final
    time, gradient w.r.t. the 8 mode amplitudes.  Times a plain (graph)
    forward, the recorded forward and the adjoint; checks the gradient against
    central differences of the GPU forward (bit-identical to the oracle's), at
    h = 1e-6 a_k and at h = 1e-3 a_k (the objective carries rounding noise near
    1e-7 relative, from the near-uniform tails of the shear layer), and on a
    smooth multi-mode flow of the same size, where differences resolve it."""
    from hamers_amd.adjoint import KHProblem, ModalLevelProblem, kinetic_energy_transverse_grad

    def wall(f):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        f()
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    def fd_error(prob, g, rel_h):
        a = prob.a0
        fd = np.zeros(8)
        for k in range(8):
            h = rel_h * a[k]
            ap, am = a.copy(), a.copy()
            ap[k] += h
            am[k] -= h
            fd[k] = (prob.objective(ap)[0] - prob.objective(am)[0]) / (2 * h)
        return float(np.max(np.abs(g - fd)) / np.max(np.abs(fd)))

    N, P, nsteps, every = args.adjoint_N, args.adjoint_patch, args.adjoint_steps, args.adjoint_every
    prob = KHProblem(N=N, P=P, nsteps=nsteps, checkpoint_every=every, use_graph=True)
    lvl = prob.level
    Q0 = prob.initial_state(prob.a0)
    J, g = prob.gradient(prob.a0)  # warm-up (graph capture, adjoint buffers)
    lvl.write_global(Q0)
    t_fwd = min(wall(lambda: lvl.advance(prob.dt, nsteps, use_graph=True)) for _ in range(2))
    lvl.write_global(Q0)
    t_rec = wall(lambda: lvl.advance_recorded(prob.dt, nsteps, every, use_graph=True))
    lvl.write_adjoint_global(kinetic_energy_transverse_grad(lvl.read_global(Q0.shape), prob.dx))
    t_adj = wall(lvl.adjoint)
    err6, err3 = fd_error(prob, g, 1e-6), fd_error(prob, g, 1e-3)
    smooth = ModalLevelProblem("smooth_modes", N=N, P=P, nsteps=nsteps, checkpoint_every=every, use_graph=True)
    _, gs = smooth.gradient(smooth.a0)
    err_smooth = fd_error(smooth, gs, 1e-6)
    cells = N * N
    return {
        "workload": f"Kelvin-Helmholtz {N}^2 in {(N // P) ** 2} patches of {P}^2, WCNS6_LD_HLLC_HLL, {nsteps} SSP-RK3 "
                    f"steps, checkpoint every {every} with recompute, gradient w.r.t. 8 mode amplitudes",
        "J": J, "gradient": [float(x) for x in g],
        "forward_s": t_fwd, "recorded_forward_s": t_rec, "adjoint_s": t_adj,
        "adjoint_over_forward": t_adj / t_fwd, "gradient_over_forward": (t_rec + t_adj) / t_fwd,
        "forward_cell_rhs_per_s": 3 * nsteps * cells / t_fwd,
        "fd_rel_err_h1e-6": err6, "fd_rel_err_h1e-3": err3,
        "smooth_modes_fd_rel_err_h1e-6": err_smooth,
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is design only "
                          "(see the sharding section of the design notes)", "n_gpus": args.gpus}))
        sys.exit(2)

    import numpy as np
    import torch

    if args.copy_probe_mb > 0:
        # A plain copy of known size: its FETCH_SIZE / WRITE_SIZE readings give
        # the counters' scale for the byte figures of the RHS kernels.
        src = torch.ones(args.copy_probe_mb * (1 << 20) // 8, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": src.numel() * 8, "copies": 5}), flush=True)
        return

    from hamers_amd import ConvectiveFluxReconstructor, PatchData, PatchList, capi
    from hamers_amd import workloads as W

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    dev = "cuda:0"
    N, P, g = args.N, args.patch, 4
    Qi, n_glob, dx = W.four_quadrant_riemann(N=N, seed=1234)
    patches = W.split_patches(Qi, (P, P), g)
    neq = 4
    rec = ConvectiveFluxReconstructor("bench", 2, "SINGLE_SPECIES", args.scheme, species_gamma=(1.4,))
    pds = []
    for _, Qp in patches:
        Qd = torch.from_numpy(Qp).to(dev)
        F = [torch.empty((neq, P, P + 1), dtype=torch.float64, device=dev),
             torch.empty((neq, P + 1, P), dtype=torch.float64, device=dev)]
        S = torch.zeros((neq, P, P), dtype=torch.float64, device=dev)
        pds.append(PatchData((P, P), (g, g), dx, Qd, F, S))
    cells = N * N
    dt = 1.0e-4
    stream = torch.cuda.current_stream()

    plist = PatchList(pds, neq)  # the level's patch list, built once as a C++ caller would hold it

    def step_batched():
        rec.computeConvectiveFluxAndSourceOnPatches(plist, 0.0, dt, 0, stream)

    # Per-patch path: views prebuilt so the loop is one C call per patch.
    views = [pd.view(neq) for pd in pds]
    sp = ctypes.c_void_p(stream.cuda_stream)
    fn = capi.lib.hm_compute_convective_flux_and_source_on_patch

    def step_per_patch():
        for v in views:
            st = fn(rec._ctx, ctypes.byref(v), 0.0, dt, 0, sp)
            if st:
                capi.check(rec._ctx, st)

    def timed(fn_step, steps, warmup):
        for _ in range(warmup):
            fn_step()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        t0 = time.perf_counter()
        for _ in range(steps):
            fn_step()
        e1.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        return e0.elapsed_time(e1) / steps, wall / steps * 1e3

    ms_batched, wall_batched = timed(step_batched, args.steps, args.warmup)

    # Stage timing of the batched path (events around each stage launch).
    rec.enable_profiling(True)
    for _ in range(args.steps):
        step_batched()
    prof = rec.read_profile()
    rec.enable_profiling(False)
    stage_avg = {k: (v[0] / v[1] if v[1] else 0.0) for k, v in prof.items()}
    dominant = max(stage_avg, key=stage_avg.get)

    ms_pp = None
    if not args.no_per_patch:
        ms_pp, _ = timed(step_per_patch, max(3, args.steps // 4), 1)

    # Algorithmic (compulsory) HBM bytes per cell per RHS: read Q (neq), write
    # F on d axes (d*neq): 8*(4 + 2*4) = 96 B for 2D single-species.
    bytes_per_cell = 8 * (neq + 2 * neq)
    # Each flux kernel covers one axis here (axis 0, axis 1): its compulsory
    # bytes are Q read for the node fluxes and F of its axis written,
    # 8*(4 + 4) = 64 B per cell per launch.
    kernel_bytes_per_cell = 8 * (neq + neq)
    achieved_gbs = kernel_bytes_per_cell * cells / (stage_avg[dominant] * 1e-3) / 1e9
    # DRAM bytes per launch of the dominant kernel, measured with rocprofv3
    # (FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0, factors from a copy of known size)
    # for the default workload only; other sizes report null.
    traffic, traffic_source = None, None
    if (args.N, args.patch, args.scheme) == (2048, 128, "WCNS5_JS_HLLC_HLL") and \
            FLUX_TRAFFIC_2048_WCNS5.get(dominant) is not None:
        traffic = FLUX_TRAFFIC_2048_WCNS5[dominant]
        traffic_source = "rocprofv3 FETCH_SIZE/WRITE_SIZE, profiles/flux_kernels_2048sq_counters.md"
    rhs_gbs = bytes_per_cell * cells / (ms_batched * 1e-3) / 1e9
    value = cells / (ms_batched * 1e-3)

    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O

        nthreads = min(16, os.cpu_count() or 1)
        prm = O.Params(dim=2, scheme=1 if args.scheme.startswith("WCNS6") else 0)

        cpu_list = [{"Q": Qp, "n": (P, P), "g": g, "dx": dx,
                  "F": [np.empty((neq, P, P + 1)), np.empty((neq, P + 1, P))],
                  "S": np.zeros((neq, P, P))} for _, Qp in patches]
        t0 = time.perf_counter()
        O.rhs_batch(prm, cpu_list, dt, nthreads)        # one pass over the whole level, to size the sample
        t1 = time.perf_counter() - t0
        reps = max(1, int(round(args.cpu_seconds / max(t1, 1e-3))))
        t0 = time.perf_counter()
        for _ in range(reps):
            O.rhs_batch(prm, cpu_list, dt, nthreads)
        tk = time.perf_counter() - t0
        cpu = {"value": reps * cells / tk, "unit": "cell-RHS-evals/s", "cores": nthreads, "kind": "port",
               "sample": f"the whole 256-patch level, {reps} RHS evaluations ({tk:.1f} s), oracle built -O3 "
                         f"-fopenmp -ffp-contract=off, {nthreads} OpenMP threads over patches"}

    uneven = None
    npatches = len(pds)
    if args.uneven_seg > 0:
        # free the 2D level first
        del plist, pds, views
        torch.cuda.empty_cache()
        uneven = bench_uneven_3d(args, torch, np, time)
        torch.cuda.empty_cache()

    ns3d = None
    if args.ns_seg > 0:
        ns3d = bench_ns_3d(args, torch, np, time)
        torch.cuda.empty_cache()

    adjoint = None
    if args.adjoint_steps > 0:
        adjoint = bench_adjoint(args, torch, np, time)

    out = {
        "metric": "FP64 cell-RHS-evals/s (batched convective RHS, WCNS5-JS + hybrid HLLC/HLLC-HLL)",
        "value": value,
        "unit": "cell-RHS-evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_batched,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "2D single-species Euler four-quadrant Riemann, 2048^2 in 256 patches of "
                               "128^2, 4-cell halos, WCNS5_JS_HLLC_HLL, batched RHS", "cells": cells,
                   "patches": npatches, "patch": [P, P], "scheme": args.scheme},
        "roofline": {"bound": "hbm", "achieved": achieved_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved_gbs / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": traffic_source,
                     "kernel": STAGE_KERNEL.get(dominant, dominant), "kernel_ms": stage_avg[dominant],
                     "bytes_per_cell": kernel_bytes_per_cell, "rhs_bytes_per_cell": bytes_per_cell},
        "rhs_hbm_frac": rhs_gbs / HBM_PEAK_GBS,
        "stage_ms": stage_avg,
        "per_patch": {"ms_per_step": ms_pp, "value": (cells / (ms_pp * 1e-3)) if ms_pp else None,
                      "batched_speedup": (ms_pp / ms_batched) if ms_pp else None},
        "host_wall_ms_per_step": wall_batched,
        "cpu_baseline": cpu,
        "uneven_3d": uneven,
        "navier_stokes_3d": ns3d,
        "adjoint": adjoint,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
