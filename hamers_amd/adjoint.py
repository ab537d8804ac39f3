"""Gradient of a level objective with respect to initial-state mode
amplitudes (Kelvin-Helmholtz shear layer, or a smooth multi-mode flow)
through the discrete adjoint.

``KHObjective`` is a thin ``torch.autograd.Function`` used as a test and
benchmark harness: forward runs the level's SSP-RK3 steps on the GPU with
checkpoints (hm_level_advance_recorded) and returns
J = sum over cells of 1/2 rho v^2 dV at the final time (kinetic energy of the
transverse velocity); backward seeds dJ/dQ_T, runs hm_level_adjoint and
contracts dJ/dQ_0 with dQ_0/da for the mode amplitudes a.  Every RHS, VJP
and stage update runs in libhamers_amd.so; torch only carries the gradient.
"""
from __future__ import annotations

import numpy as np

from . import workloads as W
from .integrator import Level
from .reconstructor import ConvectiveFluxReconstructor


def kinetic_energy_transverse(Q: np.ndarray, dx) -> float:
    """J = sum 1/2 (rho v)^2 / rho dV."""
    dV = float(np.prod(dx))
    return float(0.5 * np.sum(Q[2] * Q[2] / Q[0]) * dV)


def kinetic_energy_transverse_grad(Q: np.ndarray, dx) -> np.ndarray:
    """dJ/dQ of kinetic_energy_transverse: d/d(rho) = -v^2/2, d/d(rho v) = v (times dV)."""
    dV = float(np.prod(dx))
    v = Q[2] / Q[0]
    g = np.zeros_like(Q)
    g[0] = -0.5 * v * v * dV
    g[2] = v * dV
    return g


_FLOWS = {
    # name: (initial state(N, amplitudes, seed), dQ0/da(N, Q0, seed), default seed)
    "kelvin_helmholtz": (W.kelvin_helmholtz, W.kelvin_helmholtz_mode_derivatives, 7),
    "smooth_modes": (W.smooth_modes, W.smooth_modes_derivatives, 11),
}


class ModalLevelProblem:
    """A periodic N^2 level split into patches of P^2 whose initial state has
    8 transverse-velocity mode amplitudes as parameters; J = transverse kinetic
    energy after ``nsteps`` SSP-RK3 steps of fixed dt."""

    def __init__(self, flow="kelvin_helmholtz", N=512, P=128, nsteps=200, checkpoint_every=20, cfl=0.4,
                 scheme="WCNS6_LD_HLLC_HLL", seed=None, use_graph=True):
        self._state, self._dstate, default_seed = _FLOWS[flow]
        self.flow = flow
        self.N, self.P, self.nsteps, self.every = N, P, nsteps, checkpoint_every
        self.seed = default_seed if seed is None else seed
        self.use_graph = use_graph
        self.rec = ConvectiveFluxReconstructor(flow, 2, "SINGLE_SPECIES", scheme, species_gamma=(1.4,))
        boxes = [((i, j), (P, P)) for j in range(0, N, P) for i in range(0, N, P)]
        self.dx = (1.0 / N, 1.0 / N)
        self.level = Level(self.rec, boxes, ((0, 0), (N, N)), (1, 1), self.dx)
        Q0, _, _, a0, _ = self._state(N, seed=self.seed)
        self.a0 = a0
        # fixed dt from the initial state's largest wave speed
        rho, u, v = Q0[0], Q0[1] / Q0[0], Q0[2] / Q0[0]
        p = (Q0[3] - 0.5 * rho * (u * u + v * v)) * 0.4
        smax = float(np.max(np.maximum(np.abs(u), np.abs(v)) + np.sqrt(1.4 * p / rho)))
        self.dt = cfl * self.dx[0] / smax

    def initial_state(self, a):
        Q0, _, _, _, _ = self._state(self.N, amplitudes=np.asarray(a, dtype=np.float64), seed=self.seed)
        return Q0

    def objective(self, a, record=False):
        """J(a) with a plain (or recorded) forward run."""
        Q0 = self.initial_state(a)
        self.level.write_global(Q0)
        if record:
            self.level.advance_recorded(self.dt, self.nsteps, self.every, use_graph=self.use_graph)
        else:
            self.level.advance(self.dt, self.nsteps, use_graph=self.use_graph)
        QT = self.level.read_global(Q0.shape)
        return kinetic_energy_transverse(QT, self.dx), Q0, QT

    def gradient(self, a):
        """(J, dJ/da) by one recorded forward run and one adjoint run."""
        J, Q0, QT = self.objective(a, record=True)
        self.level.write_adjoint_global(kinetic_energy_transverse_grad(QT, self.dx))
        self.level.adjoint()
        lam0 = self.level.read_adjoint_global(Q0.shape)
        dQ = self._dstate(self.N, Q0, seed=self.seed)
        return J, np.einsum("kcij,cij->k", dQ, lam0)


class KHProblem(ModalLevelProblem):
    """The Kelvin-Helmholtz double shear layer as a ModalLevelProblem."""

    def __init__(self, N=512, P=128, nsteps=200, checkpoint_every=20, cfl=0.4, scheme="WCNS6_LD_HLLC_HLL",
                 seed=7, use_graph=True):
        super().__init__("kelvin_helmholtz", N=N, P=P, nsteps=nsteps, checkpoint_every=checkpoint_every, cfl=cfl,
                         scheme=scheme, seed=seed, use_graph=use_graph)


def kh_objective_function():
    """The torch.autograd.Function of the harness (imports torch lazily)."""
    import torch

    class KHObjective(torch.autograd.Function):
        @staticmethod
        def forward(ctx, amplitudes, problem):
            a = amplitudes.detach().cpu().numpy().astype(np.float64)
            J, grad = problem.gradient(a)
            ctx.save_for_backward(torch.from_numpy(grad))
            return amplitudes.new_tensor(J)

        @staticmethod
        def backward(ctx, gJ):
            (grad,) = ctx.saved_tensors
            return gJ * grad.to(gJ.device, gJ.dtype), None

    return KHObjective
