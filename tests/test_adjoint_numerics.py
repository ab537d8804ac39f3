"""CPU checks of the adjoint: the hand-written reverse sweep of one face
(hamers_amd/csrc/hm_adjoint_face.hpp) against central differences of the
forward face code it reverses, compiled for the host with g++
(adjoint_numerics_check.cpp), and the harness's chain-rule pieces."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from hamers_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_face_adjoint_matches_differences(tmp_path):
    exe = tmp_path / "adjoint_check"
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-Wno-unknown-pragmas", "-I",
                           os.path.join(ROOT, "hamers_amd", "csrc"),
                           os.path.join(ROOT, "tests", "adjoint_numerics_check.cpp"), "-o", str(exe)])
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    print(out.stdout)
    assert out.returncode == 0, out.stdout
    assert "ALL OK" in out.stdout


def test_kh_mode_derivatives_match_differences():
    N = 32
    Q0, _, _, a0, _ = W.kelvin_helmholtz(N, seed=7)
    dQ = W.kelvin_helmholtz_mode_derivatives(N, Q0, seed=7)
    for k in range(8):
        h = 1e-3 * a0[k]  # Q0 is quadratic in a: central differences are exact up to rounding
        ap, am = a0.copy(), a0.copy()
        ap[k] += h
        am[k] -= h
        fd = (W.kelvin_helmholtz(N, amplitudes=ap, seed=7)[0] - W.kelvin_helmholtz(N, amplitudes=am, seed=7)[0]) / (2 * h)
        assert np.allclose(dQ[k], fd, rtol=1e-6, atol=1e-9 * np.abs(fd).max())


def test_kh_phases_do_not_depend_on_given_amplitudes():
    Q_default, _, _, a0, ph0 = W.kelvin_helmholtz(16, seed=7)
    Q_given, _, _, _, ph1 = W.kelvin_helmholtz(16, amplitudes=a0, seed=7)
    assert np.array_equal(ph0, ph1) and np.array_equal(Q_default, Q_given)


def test_objective_gradient_formula():
    from hamers_amd.adjoint import kinetic_energy_transverse, kinetic_energy_transverse_grad

    rng = np.random.default_rng(0)
    Q = np.stack([rng.uniform(0.5, 1.5, (8, 8)), rng.normal(size=(8, 8)), rng.normal(size=(8, 8)),
                  rng.uniform(2, 3, (8, 8))])
    dx = (0.1, 0.2)
    g = kinetic_energy_transverse_grad(Q, dx)
    d = rng.normal(size=Q.shape)
    h = 1e-6
    fd = (kinetic_energy_transverse(Q + h * d, dx) - kinetic_energy_transverse(Q - h * d, dx)) / (2 * h)
    assert abs(fd - np.sum(g * d)) <= 1e-8 * abs(fd)


def test_smooth_mode_derivatives_match_differences():
    N = 32
    Q0, _, _, a0, _ = W.smooth_modes(N)
    dQ = W.smooth_modes_derivatives(N, Q0)
    for k in range(8):
        h = 1e-3 * a0[k]  # Q0 is quadratic in a: central differences are exact up to rounding
        ap, am = a0.copy(), a0.copy()
        ap[k] += h
        am[k] -= h
        fd = (W.smooth_modes(N, amplitudes=ap)[0] - W.smooth_modes(N, amplitudes=am)[0]) / (2 * h)
        assert np.allclose(dQ[k], fd, rtol=1e-6, atol=1e-9 * np.abs(fd).max())


def _shock_patch_rhs(prm, Q, n, dx):
    from oracle import oracle as O

    F, _ = O.rhs_patch(prm, Q, n, 4, dx, 1e-3)
    return np.concatenate([f.ravel() for f in F])


def test_branch_tape_replays_decisions():
    """The oracle's branch tape: recording and replaying at the same state is
    bit-identical to the untaped scheme and consumes the whole tape."""
    from oracle import oracle as O

    Qi, n, dx = W.four_quadrant_riemann(24, seed=3, noise=1e-2)
    Q = W.fill_periodic(Qi, 4)
    for scheme in (0, 1):
        prm = O.Params(dim=2, scheme=scheme)
        plain = _shock_patch_rhs(prm, Q, n, dx)
        with O.frozen_branches() as fb:
            rec = _shock_patch_rhs(prm, Q, n, dx)
            assert fb.replay() > 0
            rep = _shock_patch_rhs(prm, Q, n, dx)
            assert fb.consumed()
        assert np.array_equal(plain, rec) and np.array_equal(plain, rep)
        assert np.array_equal(plain, _shock_patch_rhs(prm, Q, n, dx))  # tape off again


def test_branch_tape_differences_converge_across_switches():
    """On four-quadrant shock data free WCNS6-LD difference quotients of the
    RHS jump between steps (decisions flip), while replayed ones converge at
    second order: the piece the adjoint differentiates."""
    from oracle import oracle as O

    Qi, n, dx = W.four_quadrant_riemann(32, seed=3, noise=1e-2)
    Q = W.fill_periodic(Qi, 4)
    dQ = np.random.default_rng(5).normal(size=Q.shape) * (0.1 + np.abs(Q))
    prm = O.Params(dim=2, scheme=1)

    def quotients(frozen):
        out = []
        for h in (1e-7, 5e-8):
            with O.frozen_branches() as fb:
                if frozen:
                    _shock_patch_rhs(prm, Q, n, dx)
                    fb.replay()
                else:
                    O.branch_tape(0)
                fp = _shock_patch_rhs(prm, Q + h * dQ, n, dx)
                if frozen:
                    assert fb.consumed()
                    fb.replay()
                fm = _shock_patch_rhs(prm, Q - h * dQ, n, dx)
            out.append((fp - fm) / (2 * h))
        return np.max(np.abs(out[0] - out[1])) / np.max(np.abs(out[1]))

    assert quotients(True) <= 1e-5
    assert quotients(False) >= 1e-2
