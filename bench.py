"""Benchmark of the convective-flux RHS on one MI355X.

Headline workload (one "step" = one RHS evaluation of a whole AMR level): 3D
five-equation (Allaire) two-gas Euler, FP64, a Mach-1.5 shock in gas 1
(gamma 1.4) approaching a sinusoidally perturbed interface with gas 2
(gamma 1.67), WCNS6-LD + hybrid HLLC/HLLC-HLL, on a level tiled by uneven
patches whose extents per axis are drawn (seeded) from {8, 16, 24, 32, 48, 64}:
14^3 = 2744 patches, 68.7 M cells, 4-cell halos from the analytic field.  The
fields are generated on the GPU with torch (synthetic input; every RHS runs
in libhamers_amd.so) and are resident in HBM before the timed region.  The
batched path (one work-list launch per pipeline stage for the whole level) is
the headline; the per-patch path (one launch sequence per patch, the plain
drop-in) is timed beside it; one SSP-RK3 step of the same level through the
level integrator with the step captured in a hipGraph is timed too; the CPU
oracle is timed on a bounded sample of the level's patches on the host cores.

Secondary legs: "rhs_2d" (2D single-species four-quadrant Riemann problem,
2048^2 in 256 patches of 128^2, WCNS5-JS, batched vs per-patch),
"navier_stokes_3d" (3D four-equation Navier-Stokes level of uneven patches,
convective + diffusive RHS, batched vs per-patch) and "adjoint" (discrete
adjoint of 200 SSP-RK3 steps on a 512^2 Kelvin-Helmholtz level, checkpoints
every 20 steps, gradient w.r.t. 8 mode amplitudes).

Prints one JSON line (rank 0).
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E
FP64_PEAK_TFLOPS = 78.6        # MI355X FP64 (vector and matrix)
# Counted per-launch figures of the headline's dominant kernel (rocprofv3 on
# the default headline workload; profiles/ holds the summaries and commands):
# FP64 operations per launch (64 lanes x (ADD + MUL + TRANS + 2 FMA) wave
# instructions of SQ_INSTS_VALU_*_F64), DRAM bytes per launch (FETCH_SIZE x 2.0
# + WRITE_SIZE x 1.0, factors from a copy of known size) and the fraction of
# the VALU issue peak its SQ_INSTS_VALU reaches.
HEADLINE_KERNEL_COUNTS = {
    "k_flux_multi": {
        "cells": 68726784,
        # 64 x (4 082 214 738 + 6 044 320 354 + 443 484 271 + 2 x 3 195 942 072)
        "fp64_ops": 64 * (4082214738 + 6044320354 + 443484271 + 2 * 3195942072),
        # FETCH_SIZE 54 973 956 KiB x 2.0 + WRITE_SIZE 26 786 422 KiB x 1.0 (includes the 72 B/lane spill)
        "dram_bytes": 54973956 * 1024 * 2 + 26786422 * 1024,
        "valu_issue_frac": 0.755,
        "source": "profiles/rhs_uneven3d_five_eqn_counters.md",
    },
}
# The same for the 2D single-species leg's flux kernels (2048^2 in 128^2
# patches, WCNS5-JS): DRAM bytes per launch.
FLUX_TRAFFIC_2048_WCNS5 = {"midflux": 354.92e6 + 135.27e6, "faceflux": 506.47e6 + 135.27e6}
STAGE_KERNEL_SINGLE = {"midflux": "k_flux_single", "faceflux": "k_flux_tile_single"}
# (the multi-species midpoint and face flux run fused in one kernel, timed as the midflux stage)
STAGE_KERNEL_MULTI = {"derived": "k_derived_multi", "midflux": "k_flux_multi", "faceflux": "k_faceflux_multi",
                      "source": "k_source_five"}
UNEVEN_CHOICES = (8, 16, 24, 32, 48, 64)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed RHS evaluations of the headline")
    # The GPU's clocks ramp over the first milliseconds of load: warm-up RHS
    # evaluations keep the ramp out of the timing.
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--seg", type=int, default=14, help="patches per axis of the headline's uneven 3D level")
    ap.add_argument("--rk-steps", type=int, default=3, help="timed SSP-RK3 steps of the headline level (0: skip)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="target CPU-baseline work (multi-thread run)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-per-patch", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="keep the built-in launch schedule")
    ap.add_argument("--headline-only", action="store_true", help="skip the secondary legs")
    ap.add_argument("--leg", choices=("2d", "ns", "adjoint", "two_level", "two_level_ns"), default=None,
                    help="run only this secondary leg and print its JSON (for profiling it)")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="only copy this many MiB device to device 5 times (calibrates HBM byte counters) and exit")
    ap.add_argument("--N", type=int, default=2048, help="2D leg: cells per axis")
    ap.add_argument("--patch", type=int, default=128, help="2D leg: patch size")
    ap.add_argument("--scheme", default="WCNS5_JS_HLLC_HLL", help="2D leg: scheme")
    ap.add_argument("--steps-2d", type=int, default=30)
    ap.add_argument("--ns-seg", type=int, default=8, help="patches per axis of the uneven 3D Navier-Stokes level (0: skip)")
    ap.add_argument("--ns-steps", type=int, default=4)
    ap.add_argument("--adjoint-steps", type=int, default=200, help="RK3 steps of the adjoint leg (0: skip)")
    ap.add_argument("--adjoint-N", type=int, default=512)
    ap.add_argument("--adjoint-patch", type=int, default=128)
    ap.add_argument("--adjoint-every", type=int, default=20)
    ap.add_argument("--two-level-ns-n", type=int, default=256,
                    help="cells per axis of the Navier-Stokes two-level leg's base level (0: skip)")
    return ap.parse_args()


def host_description():
    """nproc, the CPU model (from /proc/cpuinfo, as lscpu prints it) and the
    OpenMP thread count of this host."""
    model = platform.processor() or "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "cpu_model": model, "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}


def timed_events(torch, stream, f, steps, warmup):
    """Average ms per call of f over `steps` calls after `warmup`, from HIP
    events on `stream`, bracketed by device synchronizations."""
    for _ in range(warmup):
        f()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    t0 = time.perf_counter()
    for _ in range(steps):
        f()
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return e0.elapsed_time(e1) / steps, wall / steps * 1e3


def shock_interface_patch_torch(torch, lo, n, g, dx, modes, gamma=(1.4, 1.67), rho_gas=(1.0, 0.2)):
    """Conservative five-equation arrays (8, n2+2g, n1+2g, n0+2g) of the box of
    cells lo - g .. lo + n + g - 1 of the shock / perturbed-interface field
    (one patch with halos from the same analytic field, or the whole level
    grown by its halo)."""
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


class UnevenLevel:
    """The headline's level: 3D five-equation two-gas Euler, shock and perturbed
    interface, seg^3 uneven patches (extents per axis from UNEVEN_CHOICES,
    seeded), 4-cell halos from the analytic field, on the GPU."""

    def __init__(self, args, torch, np):
        from hamers_amd import ConvectiveFluxReconstructor, PatchData, PatchList

        rng = np.random.default_rng(2024)
        self.seg = args.seg
        self.segs = [rng.choice(UNEVEN_CHOICES, size=args.seg) for _ in range(3)]
        starts = [np.concatenate([[0], np.cumsum(s)[:-1]]) for s in self.segs]
        self.Ng = [int(s.sum()) for s in self.segs]
        self.dx = tuple(1.0 / v for v in self.Ng)
        modes = [(rng.uniform(0.0, 1.0), rng.uniform(0, 2 * np.pi), rng.uniform(0, 2 * np.pi)) for _ in range(7)]
        tot = sum(m[0] for m in modes)
        self.modes = [(m[0] / tot, m[1], m[2]) for m in modes]
        self.g, self.neq = 4, 7
        self.rec = ConvectiveFluxReconstructor("uneven", 3, "FIVE_EQN_ALLAIRE", "WCNS6_LD_HLLC_HLL", num_species=2,
                                               species_gamma=(1.4, 1.67), species_R=(1.0, 1.0))
        dev = "cuda:0"
        self.pds, self.boxes = [], []
        g, neq = self.g, self.neq
        # the field over the whole level grown by the halo, one evaluation; each
        # patch takes its ghost box out of it (the same values as evaluating
        # the field per patch: the coordinates are formed from the global index)
        whole = shock_interface_patch_torch(torch, (0, 0, 0), tuple(self.Ng), g, self.dx, self.modes)
        for k in range(args.seg):
            for j in range(args.seg):
                for i in range(args.seg):
                    lo = (int(starts[0][i]), int(starts[1][j]), int(starts[2][k]))
                    n = (int(self.segs[0][i]), int(self.segs[1][j]), int(self.segs[2][k]))
                    Q = whole[:, lo[2]:lo[2] + n[2] + 2 * g, lo[1]:lo[1] + n[1] + 2 * g,
                              lo[0]:lo[0] + n[0] + 2 * g].contiguous()
                    F = [torch.empty((neq, n[2] + (a == 2), n[1] + (a == 1), n[0] + (a == 0)), dtype=torch.float64,
                                     device=dev) for a in range(3)]
                    S = torch.zeros((neq, n[2], n[1], n[0]), dtype=torch.float64, device=dev)
                    self.pds.append(PatchData(n, (g, g, g), self.dx, Q, F, S))
                    self.boxes.append((lo, n))
        del whole
        torch.cuda.empty_cache()
        self.cells = int(sum(b[1][0] * b[1][1] * b[1][2] for b in self.boxes))
        self.plist = PatchList(self.pds, self.rec.num_arrays)
        self.dt = 1.0e-4

    def describe(self, np):
        ext = np.bincount(np.concatenate(self.segs), minlength=65)[list(UNEVEN_CHOICES)].tolist()
        return {"workload": "3D five-equation Allaire two-gas Euler, Mach-1.5 shock and perturbed interface, "
                            "WCNS6_LD_HLLC_HLL, uneven patches (extents per axis from {8,16,24,32,48,64}), "
                            "batched RHS of the whole level",
                "patches": len(self.pds), "cells": self.cells, "domain": self.Ng, "extent_counts": ext,
                "scheme": "WCNS6_LD_HLLC_HLL", "model": "FIVE_EQN_ALLAIRE", "num_species": 2}


def bench_headline(args, torch, np, L):
    """Batched RHS of the uneven level (headline), stage times with HIP events
    on the launch stream, the per-patch path beside it."""
    from hamers_amd import capi
    from hamers_amd.reconstructor import autotune

    rec = L.rec
    stream = torch.cuda.current_stream()
    sp = ctypes.c_void_p(stream.cuda_stream)
    fn = capi.lib.hm_compute_convective_flux_and_source_on_patch
    views = [pd.view(rec.num_arrays) for pd in L.pds]

    def batched():
        rec.computeConvectiveFluxAndSourceOnPatches(L.plist, 0.0, L.dt, 0, stream)

    def per_patch():
        for v in views:
            st = fn(rec._ctx, ctypes.byref(v), 0.0, L.dt, 0, sp)
            if st:
                capi.check(rec._ctx, st)

    # the autotuner picks the level's launch schedule (timed trials on this level)
    sched = autotune(rec, L.plist, L.dt, repetitions=3, stream=stream) if not args.no_autotune else None
    ms, wall = timed_events(torch, stream, batched, args.steps, args.warmup)
    rec.enable_profiling(True)
    for _ in range(args.steps):
        batched()
    prof = rec.read_profile()
    rec.enable_profiling(False)
    stage = {k: (v[0] / v[1] if v[1] else 0.0) for k, v in prof.items()}
    ms_p = None
    if not args.no_per_patch:
        ms_p, _ = timed_events(torch, stream, per_patch, max(2, args.steps // 4), 1)
    return ms, wall, stage, ms_p, sched


def headline_roofline(L, stage, ms):
    """Roofline of the dominant kernel of the headline RHS.  Its bound is FP64
    VALU issue (counters: profiles/), so the figure is its FP64 operation rate
    against MI355X's FP64 peak; the HBM figures are reported beside it: the
    compulsory bytes of the whole RHS (SURVEY-style count) per RHS time, and the
    kernel's measured DRAM traffic per launch."""
    neq, nadv = L.neq, 1
    rhs_bytes_per_cell = 8 * (neq + 3 * neq + 2 * nadv)  # Q read, F on 3 axes, S read-modify-write
    dominant = max(stage, key=stage.get)
    kernel = STAGE_KERNEL_MULTI.get(dominant, dominant)
    kms = stage[dominant]
    c = HEADLINE_KERNEL_COUNTS.get(kernel)
    rhs_gbs = rhs_bytes_per_cell * L.cells / (ms * 1e-3) / 1e9
    out = {"kernel": kernel, "kernel_ms": kms, "rhs_bytes_per_cell": rhs_bytes_per_cell,
           "rhs_hbm": {"achieved": rhs_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": rhs_gbs / HBM_PEAK_GBS}}
    if c is not None and (L.seg, L.cells) == (14, c["cells"]):
        tflops = c["fp64_ops"] / (kms * 1e-3) / 1e12
        out.update({"bound": "valu", "achieved": tflops, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                    "frac": tflops / FP64_PEAK_TFLOPS, "traffic": c["dram_bytes"],
                    "valu_issue_frac": c["valu_issue_frac"], "fp64_ops_per_launch": c["fp64_ops"],
                    "counts_source": c["source"],
                    "dram_gbs": c["dram_bytes"] / (kms * 1e-3) / 1e9})
    else:
        # without counted figures for this workload: the compulsory-byte view only
        out.update({"bound": "hbm", "achieved": rhs_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": rhs_gbs / HBM_PEAK_GBS, "traffic": None})
    return out


def bench_level_rk3(args, torch, np, L):
    """SSP-RK3 steps of the headline level through the level integrator
    (hm_level_advance), one step captured in a hipGraph and replayed."""
    from hamers_amd.integrator import Level

    g = L.g
    interiors = [pd.Q[:, g:-g, g:-g, g:-g].cpu().numpy() for pd in L.pds]
    # the RHS leg's arrays are no longer needed: free them before the level's storage
    L.plist = None
    L.pds = None
    torch.cuda.empty_cache()
    lvl = Level(L.rec, L.boxes, ((0, 0, 0), tuple(L.Ng)), (1, 1, 1), L.dx)
    for i, Qi in enumerate(interiors):
        lvl.write_patch(i, Qi)
    del interiors
    dt = 2.0e-5
    sched = lvl.autotune(dt, repetitions=2) if not args.no_autotune else None
    lvl.advance(dt, 1, use_graph=True)  # capture
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    lvl.advance(dt, args.rk_steps, use_graph=True)
    torch.cuda.synchronize()
    t = (time.perf_counter() - t0) / args.rk_steps
    ok = bool(np.all(np.isfinite(lvl.read_patch(0))))
    del lvl
    return {"ms_per_step": t * 1e3, "steps": args.rk_steps, "dt": dt, "graph": True, "schedule": sched,
            "cell_rhs_per_s": 3 * L.cells / t, "finite": ok,
            "note": "one SSP-RK3 step = ghost fill + 3 batched RHS + 3 stage updates, replayed from a hipGraph"}


def cpu_baseline_uneven(args, np, L):
    """The CPU oracle (C++ restatement, -O3 -fopenmp, OpenMP over patches) on a
    bounded sample of the headline level's patches: every 28th patch, on the
    host's OpenMP thread count, then a smaller sample on one thread."""
    from oracle import oracle as O

    prm = O.Params(dim=3, model=O.FIVE_EQN_ALLAIRE, num_species=2, scheme=O.WCNS6_LD_HLLC_HLL, gamma=[1.4, 1.67])

    def sample(stride):
        out = []
        for i in range(0, len(L.pds), stride):
            pd = L.pds[i]
            n = pd.n
            out.append({"Q": pd.Q.cpu().numpy(), "n": n, "g": L.g, "dx": L.dx,
                        "F": [np.empty(O.face_shape(L.neq, n, a)) for a in range(3)],
                        "S": np.zeros((L.neq,) + tuple(reversed(n)))})
        return out, sum(int(np.prod(p["n"])) for p in out)

    host = host_description()
    nthreads = O.max_threads()
    big, cells_big = sample(28)
    t0 = time.perf_counter()
    O.rhs_batch(prm, big, L.dt, nthreads)
    t1 = time.perf_counter() - t0
    reps = max(1, int(round(args.cpu_seconds / max(t1, 1e-3))))
    t0 = time.perf_counter()
    for _ in range(reps):
        O.rhs_batch(prm, big, L.dt, nthreads)
    tk = time.perf_counter() - t0
    small, cells_small = sample(280)
    t0 = time.perf_counter()
    O.rhs_batch(prm, small, L.dt, 1)
    t_one = time.perf_counter() - t0
    v_multi = reps * cells_big / tk
    v_one = cells_small / t_one
    return {"value": v_multi, "unit": "cell-RHS-evals/s", "cores": nthreads, "kind": "port",
            "sample": f"{len(big)} of the level's {len(L.pds)} patches (every 28th, {cells_big} cells), "
                      f"{reps} RHS evaluations in {tk:.1f} s; oracle built -O3 -fopenmp -ffp-contract=off, "
                      f"{nthreads} OpenMP threads over patches (the GPU box's CPU share); 1-thread run on every "
                      f"280th patch ({cells_small} cells)",
            "one_thread_value": v_one, "host": host,
            "full_host_estimate": v_one * (host["nproc"] or 1)}


def bench_2d(args, torch, np):
    """2D single-species four-quadrant Riemann level: batched vs per-patch RHS,
    stage times, the flux kernels' HBM roofline (measured traffic)."""
    from hamers_amd import ConvectiveFluxReconstructor, PatchData, PatchList, capi
    from hamers_amd import workloads as W

    dev = "cuda:0"
    N, P, g = args.N, args.patch, 4
    Qi, _, dx = W.four_quadrant_riemann(N=N, seed=1234)
    patches = W.split_patches(Qi, (P, P), g)
    neq = 4
    rec = ConvectiveFluxReconstructor("bench2d", 2, "SINGLE_SPECIES", args.scheme, species_gamma=(1.4,))
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
    plist = PatchList(pds, neq)
    views = [pd.view(neq) for pd in pds]
    sp = ctypes.c_void_p(stream.cuda_stream)
    fn = capi.lib.hm_compute_convective_flux_and_source_on_patch

    def batched():
        rec.computeConvectiveFluxAndSourceOnPatches(plist, 0.0, dt, 0, stream)

    def per_patch():
        for v in views:
            st = fn(rec._ctx, ctypes.byref(v), 0.0, dt, 0, sp)
            if st:
                capi.check(rec._ctx, st)

    from hamers_amd.reconstructor import autotune

    sched = autotune(rec, plist, dt, repetitions=3, stream=stream) if not args.no_autotune else None
    ms, _ = timed_events(torch, stream, batched, args.steps_2d, 10)
    rec.enable_profiling(True)
    for _ in range(args.steps_2d):
        batched()
    prof = rec.read_profile()
    rec.enable_profiling(False)
    stage = {k: (v[0] / v[1] if v[1] else 0.0) for k, v in prof.items()}
    dominant = max(stage, key=stage.get)
    ms_p = None if args.no_per_patch else timed_events(torch, stream, per_patch, max(3, args.steps_2d // 4), 1)[0]
    # each flux kernel covers one axis: Q read for the node fluxes, F of its axis written, 8 (4 + 4) B per cell
    kernel_bytes = 8 * (neq + neq) * cells
    kgbs = kernel_bytes / (stage[dominant] * 1e-3) / 1e9
    traffic = None
    if (args.N, args.patch, args.scheme) == (2048, 128, "WCNS5_JS_HLLC_HLL"):
        traffic = FLUX_TRAFFIC_2048_WCNS5.get(dominant)
    return {"workload": "2D single-species Euler four-quadrant Riemann, 2048^2 in 256 patches of 128^2, "
                        "4-cell halos, WCNS5_JS_HLLC_HLL, batched RHS" if (N, P) == (2048, 128) else
                        f"2D single-species four-quadrant Riemann, {N}^2 in patches of {P}^2, {args.scheme}",
            "cells": cells, "patches": len(pds), "ms_per_rhs_batched": ms, "value_batched": cells / (ms * 1e-3),
            "stage_ms": stage, "schedule": sched, "ms_per_rhs_per_patch": ms_p,
            "batched_speedup": (ms_p / ms) if ms_p else None,
            "flux_kernel": {"kernel": STAGE_KERNEL_SINGLE.get(dominant, dominant), "kernel_ms": stage[dominant],
                            "bytes_per_cell": 8 * (neq + neq), "achieved_gbs": kgbs, "hbm_frac": kgbs / HBM_PEAK_GBS,
                            "traffic": traffic, "bound": "valu",
                            "counts_source": "profiles/flux_kernels_2048sq_counters.md (73-77 % of VALU issue)"},
            "rhs_hbm_frac": 8 * (neq + 2 * neq) * cells / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}


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


def bench_two_level(args, torch, np):
    """Two-level advance (refinement ratio 2) of the shock / perturbed-interface
    field: a coarse level of 32^3 patches over the whole periodic domain and a
    fine level of uneven patches over a slab around the interface (full in y
    and z, so its coarse-fine boundary is two x-planes).  Times coarse steps of
    the hierarchy (a coarse SSP-RK3 step, two fine steps of dt / 2, coarse-fine
    fill, refluxing and coarsening) against the two levels' own steps."""
    from hamers_amd import ConvectiveFluxReconstructor, Hierarchy, Level

    rng = np.random.default_rng(77)
    Nc = (224, 192, 192)
    Nf = tuple(2 * v for v in Nc)
    dxc = tuple(1.0 / v for v in Nc)
    dxf = tuple(1.0 / v for v in Nf)
    modes = [(rng.uniform(0.0, 1.0), rng.uniform(0, 2 * np.pi), rng.uniform(0, 2 * np.pi)) for _ in range(7)]
    tot = sum(m[0] for m in modes)
    modes = [(m[0] / tot, m[1], m[2]) for m in modes]
    rec = ConvectiveFluxReconstructor("two-level", 3, "FIVE_EQN_ALLAIRE", "WCNS6_LD_HLLC_HLL", num_species=2,
                                      species_gamma=(1.4, 1.67), species_R=(1.0, 1.0))
    cboxes = [((i, j, k), (32, 32, 32)) for k in range(0, Nc[2], 32) for j in range(0, Nc[1], 32)
              for i in range(0, Nc[0], 32)]

    def even_segments(total):
        out = []
        while sum(out) < total:
            out.append(int(min(rng.choice((16, 32, 48, 64)), total - sum(out))))
        return out

    xs = [224, 64, 64]  # fine x: [224, 352), two patches of 64 (the interface at x = 0.6 is fine cell 269)
    ys, zs = even_segments(Nf[1]), even_segments(Nf[2])
    fboxes = []
    zo = 0
    for nz in zs:
        yo = 0
        for ny in ys:
            xo = xs[0]
            for nx in xs[1:]:
                fboxes.append(((xo, yo, zo), (nx, ny, nz)))
                xo += nx
            yo += ny
        zo += nz
    C = Level(rec, cboxes, ((0, 0, 0), Nc), (1, 1, 1), dxc)
    F = Level(rec, fboxes, ((0, 0, 0), Nf), (1, 1, 1), dxf)
    for lvl, boxes, dx in ((C, cboxes, dxc), (F, fboxes, dxf)):
        for i, (lo, n) in enumerate(boxes):
            Q = shock_interface_patch_torch(torch, lo, n, 0, dx, modes)
            lvl.write_patch(i, Q.cpu().numpy())
    torch.cuda.synchronize()
    H = Hierarchy(C, F)
    dt = 4.0e-5  # 0.4 / 224: the CFL of the headline level's steps on the coarse grid
    stream = torch.cuda.current_stream()
    ms_h, _ = timed_events(torch, stream, lambda: H.advance(dt, 1), 2, 1)
    ms_c, _ = timed_events(torch, stream, lambda: C.advance(dt, 1), 2, 1)
    ms_f, _ = timed_events(torch, stream, lambda: F.advance(0.5 * dt, 2), 2, 1)
    finite = all(bool(np.isfinite(F.read_patch(i)).all()) for i in range(0, len(fboxes), 17))
    del H
    cc = int(np.prod(Nc))
    fc = int(sum(n[0] * n[1] * n[2] for _, n in fboxes))
    return {
        "workload": "3D five-equation Allaire, Mach-1.5 shock and perturbed interface, WCNS6_LD_HLLC_HLL; coarse "
                    f"{Nc[0]}x{Nc[1]}x{Nc[2]} in 32^3 patches, fine level (ratio 2) of uneven patches over the slab "
                    "x in [0.5, 0.786) around the interface, periodic",
        "coarse_patches": len(cboxes), "coarse_cells": cc, "fine_patches": len(fboxes), "fine_cells": fc,
        "dt_coarse": dt,
        "ms_per_coarse_step": ms_h,
        "ms_coarse_level_step": ms_c, "ms_two_fine_level_steps": ms_f,
        "ms_coarse_fine_overhead": ms_h - ms_c - ms_f,
        "cell_updates_per_s": (cc + 2 * fc) / (ms_h * 1e-3),
        "finite": finite,
    }


def bench_two_level_ns(args, torch, np):
    """Two-level advance (refinement ratio 2) in the Navier-Stokes form: the
    four-equation two-species model with constant transport on a base level of
    n^3 cells in 32^3 patches and a refined level of uneven patches (extents
    across the interface drawn from UNEVEN_CHOICES) over a slab around the
    multimode interface, both with 6 ghost cells.  Times coarse steps of the
    hierarchy (coarse step, two fine steps of dt / 2, coarse-fine fill,
    convective and diffusive refluxing, coarsening) against the levels' own steps."""
    from hamers_amd import ConvectiveFluxReconstructor, DiffusiveFluxReconstructor, Hierarchy, Level

    rng = np.random.default_rng(256)
    N = args.two_level_ns_n
    Nc = (N, N, N)
    Nf = tuple(2 * v for v in Nc)
    dxc = tuple(1.0 / v for v in Nc)
    dxf = tuple(1.0 / v for v in Nf)
    ks = np.arange(4, 33)
    amps = ks ** -2.0
    amps = 0.02 * amps / amps.sum()
    modes = [(float(a), rng.uniform(0, 2 * np.pi), rng.uniform(0, 2 * np.pi)) for a in amps]
    gamma, R = (1.4, 1.67), (1.0, 0.25)
    rec = ConvectiveFluxReconstructor("two-level-ns", 3, "FOUR_EQN_CONSERVATIVE", "WCNS6_LD_HLLC_HLL", num_species=2,
                                      species_gamma=gamma, species_R=R)
    diff = DiffusiveFluxReconstructor(rec, species_mu=(1e-3, 2e-3), species_mu_v=(0.0, 0.0),
                                      species_kappa=(1.5e-3, 2.5e-3), species_D=(1e-3, 1e-3), species_M=(1.0, 4.0))
    cboxes = [((i, j, k), (32, 32, 32)) for k in range(0, N, 32) for j in range(0, N, 32) for i in range(0, N, 32)]

    def segments(total):
        out = []
        while sum(out) < total:
            out.append(int(min(rng.choice(UNEVEN_CHOICES), total - sum(out))))
        return out

    # fine slab x in [N - 32, N + 32) of the fine grid (the interface at x = 0.5), two patches of 32 cells
    xs = [N - 32, 32, 32]
    ys, zs = segments(Nf[1]), segments(Nf[2])
    fboxes = []
    zo = 0
    for nz in zs:
        yo = 0
        for ny in ys:
            xo = xs[0]
            for nx in xs[1:]:
                fboxes.append(((xo, yo, zo), (nx, ny, nz)))
                xo += nx
            yo += ny
        zo += nz
    C = Level(rec, cboxes, ((0, 0, 0), Nc), (1, 1, 1), dxc)
    F = Level(rec, fboxes, ((0, 0, 0), Nf), (1, 1, 1), dxf)
    for lvl, boxes, dx in ((C, cboxes, dxc), (F, fboxes, dxf)):
        for i, (lo, n) in enumerate(boxes):
            Q = rtm_interface_patch_torch(torch, lo, n, 0, dx, modes, gamma, R)
            lvl.write_patch(i, Q.cpu().numpy())
    torch.cuda.synchronize()
    H = Hierarchy(C, F)
    dt = 5.0e-4  # CFL ~0.2 of the light gas's sound speed on the base grid; well inside the viscous limit
    stream = torch.cuda.current_stream()
    ms_h, _ = timed_events(torch, stream, lambda: H.advance(dt, 1), 2, 1)
    ms_c, _ = timed_events(torch, stream, lambda: C.advance(dt, 1), 2, 1)
    ms_f, _ = timed_events(torch, stream, lambda: F.advance(0.5 * dt, 2), 2, 1)
    finite = all(bool(np.isfinite(F.read_patch(i)).all()) for i in range(0, len(fboxes), 17))
    del H
    del diff
    cc = int(np.prod(Nc))
    fc = int(sum(n[0] * n[1] * n[2] for _, n in fboxes))
    return {
        "workload": "3D four-equation two-species Navier-Stokes (constant transport), multimode interface with a "
                    f"k^-2 spectrum over k in [4, 32], WCNS6_LD_HLLC_HLL + sixth-order node diffusive flux; base "
                    f"{N}^3 in 32^3 patches, refined level (ratio 2) of uneven patches (extents across the interface "
                    "from {8,16,24,32,48,64}) over the slab around the interface, 6-cell halos, periodic",
        "coarse_patches": len(cboxes), "coarse_cells": cc, "fine_patches": len(fboxes), "fine_cells": fc,
        "dt_coarse": dt,
        "ms_per_coarse_step": ms_h,
        "ms_coarse_level_step": ms_c, "ms_two_fine_level_steps": ms_f,
        "ms_coarse_fine_overhead": ms_h - ms_c - ms_f,
        "cell_updates_per_s": (cc + 2 * fc) / (ms_h * 1e-3),
        "finite": finite,
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


def progress(msg):
    """One line per finished leg on stderr (the JSON line stays alone on stdout)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


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

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    if args.leg is not None:
        if args.leg == "2d":
            res = bench_2d(args, torch, np)
        elif args.leg == "ns":
            res = bench_ns_3d(args, torch, np, time)
        elif args.leg == "two_level":
            res = bench_two_level(args, torch, np)
        elif args.leg == "two_level_ns":
            res = bench_two_level_ns(args, torch, np)
        else:
            res = bench_adjoint(args, torch, np, time)
        print(json.dumps({"leg": args.leg, "result": res}), flush=True)
        return
    L = UnevenLevel(args, torch, np)
    progress("uneven 3D level built")
    ms, wall, stage, ms_p, sched = bench_headline(args, torch, np, L)
    progress(f"headline RHS {ms:.3f} ms (per-patch {ms_p} ms)")
    roof = headline_roofline(L, stage, ms)
    cpu = None if args.no_cpu else cpu_baseline_uneven(args, np, L)
    progress("CPU baseline done")
    config = L.describe(np)
    rk = bench_level_rk3(args, torch, np, L) if args.rk_steps > 0 else None
    progress("level RK3 step done")
    del L
    torch.cuda.empty_cache()

    rhs2d = ns3d = adjoint = two_level = two_level_ns = None
    if not args.headline_only:
        rhs2d = bench_2d(args, torch, np)
        progress("2D leg done")
        torch.cuda.empty_cache()
        if args.ns_seg > 0:
            ns3d = bench_ns_3d(args, torch, np, time)
            progress("Navier-Stokes leg done")
            torch.cuda.empty_cache()
        if args.adjoint_steps > 0:
            adjoint = bench_adjoint(args, torch, np, time)
            progress("adjoint leg done")
            torch.cuda.empty_cache()
        two_level = bench_two_level(args, torch, np)
        progress("two-level leg done")
        torch.cuda.empty_cache()
        if args.two_level_ns_n > 0:
            two_level_ns = bench_two_level_ns(args, torch, np)
            progress("Navier-Stokes two-level leg done")

    value = config["cells"] / (ms * 1e-3)
    out = {
        "metric": "FP64 cell-RHS-evals/s (batched convective RHS of an uneven-patch 3D five-equation level, "
                  "WCNS6-LD + hybrid HLLC/HLLC-HLL)",
        "value": value,
        "unit": "cell-RHS-evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": config,
        "roofline": roof,
        "stage_ms": stage,
        "schedule": sched,
        "per_patch": {"ms_per_step": ms_p, "value": (config["cells"] / (ms_p * 1e-3)) if ms_p else None,
                      "batched_speedup": (ms_p / ms) if ms_p else None},
        "host_wall_ms_per_step": wall,
        "rk3_level_step": rk,
        "cpu_baseline": cpu,
        "rhs_2d": rhs2d,
        "navier_stokes_3d": ns3d,
        "adjoint": adjoint,
        "two_level": two_level,
        "two_level_navier_stokes": two_level_ns,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
