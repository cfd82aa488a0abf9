"""Differentiable gray 2-D IMC step (torch.autograd).

``AbsorptionStep`` maps a per-cell opacity scale and a per-cell heat-capacity
scale to the energy absorbed in each cell during one IMC time step taken
from a given state (``IMC2D`` state: census and material temperature held
fixed).  Its backward is the device replay adjoint ``imc2d_adjoint``: every
history of the step's transport input is replayed from its Philox counter,
and pathwise (continuous absorption) plus score-function (scatter sampling)
terms give dL/dsigma_a and dL/dsigma_s per cell; the chain rule through the
Fleck factor gives the gradients of the scales.  PyTorch is only the autograd
plumbing: the forward and backward run in HIP.

Opacity sigma = scale * sigma_a_func(T); Fleck factor
f = 1/(1 + theta beta sigma c dt), beta = 4 a T^3 / (cv_scale c_v(T)),
sigma_a = f sigma, sigma_s = (1 - f) sigma, so
  dsigma_a/dsigma = f^2,  dsigma_s/dsigma = 1 - f^2,
  dsigma_a/dcv_scale = -dsigma_s/dcv_scale = sigma f (1 - f) / cv_scale.

Each forward restores the census it started from, so the step can be
evaluated repeatedly (finite differences with common random numbers, with
``fixed_emission_counts``).  A backward must follow its own forward: the
engine keeps one transport input pool.
"""
from __future__ import annotations

import numpy as np
import torch

from ._lib import A_RAD, C_LIGHT, SRC_EMISSION

__all__ = ["AbsorptionStep"]


def _cell_volumes(e1, e2, geometry):
    if geometry == "xy":
        return np.diff(e1)[:, None] * np.diff(e2)[None, :]
    return np.pi * (e1[1:] * e1[1:] - e1[:-1] * e1[:-1])[:, None] * np.diff(e2)[None, :]


def _face_energy(geometry, side, Tb, dt, e1, e2):
    if geometry == "xy":
        area = (e2[-1] - e2[0]) if side < 2 else (e1[-1] - e1[0])
    elif side < 2:
        area = 2.0 * np.pi * (e1[0] if side == 0 else e1[-1]) * (e2[-1] - e2[0])
    else:
        area = np.pi * (e1[-1] * e1[-1] - e1[0] * e1[0])
    return A_RAD * C_LIGHT * (Tb * Tb * Tb * Tb) / 4.0 * area * dt


class AbsorptionStep:
    """One differentiable IMC step from ``state`` (an IMC2D state).

    ``sigma_a_func(T)`` and ``cv(T)`` are the unscaled material callables;
    the other arguments mean what they mean in ``IMC2D.step``."""

    def __init__(self, state, Ntarget, Nboundary, T_boundary, dt, sigma_a_func, cv, reflect=(False,) * 4,
                 theta=1.0, rz_linear_source=True, max_events_per_particle=1_000_000):
        self.state = state
        self.eng = state._engine
        self.Ntarget, self.Nboundary = int(Ntarget), int(Nboundary)
        self.T_boundary = tuple(T_boundary)
        self.dt = float(dt)
        self.sigma_a_func, self.cv = sigma_a_func, cv
        self.reflect = tuple(bool(r) for r in reflect)
        self.theta = float(theta)
        self.rz_linear = bool(rz_linear_source)
        self.max_events = int(max_events_per_particle)
        e1, e2 = self.eng.e1, self.eng.e2
        self.vol = _cell_volumes(e1, e2, self.eng.geometry)
        self.fixed_emission_counts = None
        self.last_counts = None
        self._tag = 0
        self._saved = None
        self.last_forward_ms = 0.0
        self.last_backward_ms = 0.0

    # -- raw numpy interface ----------------------------------------------------
    def forward_np(self, scale, cv_scale):
        """Absorbed energy per cell (nx, ny) for the given scales (numpy)."""
        import time

        eng = self.eng
        T = self.state.temperature
        scale = np.broadcast_to(np.asarray(scale, dtype=np.float64), eng.shape)
        cv_scale = np.broadcast_to(np.asarray(cv_scale, dtype=np.float64), eng.shape)
        sigma0 = np.asarray(self.sigma_a_func(T), dtype=np.float64) * np.ones(eng.shape)
        sigma = scale * sigma0
        cvv = np.asarray(self.cv(T), dtype=np.float64) * np.ones(eng.shape) * cv_scale
        t0 = time.perf_counter()
        eng.snapshot(True)
        eng.enable_adjoint(True)
        eng.set_step(self.state.count)
        sa, ss = eng.set_material(T, sigma, cvv, self.theta, self.dt, clip=True, want_opacity=True)
        if self.Nboundary > 0:
            for side in range(4):
                Tb = self.T_boundary[side]
                Tb = Tb(self.state.time) if callable(Tb) else Tb
                if Tb > 0.0:
                    eng.sample_boundary(side, self.Nboundary,
                                        _face_energy(eng.geometry, side, float(Tb), self.dt, eng.e1, eng.e2), self.dt)
        if self.fixed_emission_counts is not None:
            eng.emission_counts(fix=self.fixed_emission_counts)
        eng.sample_volume(SRC_EMISSION, self.Ntarget, None, self.dt, self.rz_linear)
        if self.fixed_emission_counts is not None:
            eng.emission_counts(fix=False)
        self.last_counts = eng.emission_counts()
        dep, si, bl, st = eng.transport(self.dt, self.reflect, self.max_events, -1.0)
        self.last_forward_ms = eng.last_transport_ms()
        self.last_stats = st
        eng.snapshot(False)
        self._tag += 1
        f = np.where(sigma > 0.0, sa / np.where(sigma > 0.0, sigma, 1.0), 1.0)
        self._saved = (self._tag, sigma0, sigma, f, cv_scale)
        self.last_step_s = time.perf_counter() - t0
        return dep * self.vol

    def backward_np(self, lam, tag=None):
        """(dL/dscale, dL/dcv_scale) per cell for lam = dL/dE_abs per cell."""
        import time

        if self._saved is None or (tag is not None and tag != self._saved[0]):
            raise RuntimeError("AbsorptionStep: backward must follow its own forward")
        _, sigma0, sigma, f, cv_scale = self._saved
        t0 = time.perf_counter()
        gA, gS = self.eng.adjoint(self.dt, self.reflect, self.max_events, lam)
        self.last_backward_s = time.perf_counter() - t0
        g_scale = gA * f * f * sigma0 + gS * (1.0 - f * f) * sigma0
        g_cv = (gA - gS) * sigma * f * (1.0 - f) / cv_scale
        return g_scale, g_cv, gA, gS

    # -- torch interface --------------------------------------------------------
    def __call__(self, scale, cv_scale=None):
        if cv_scale is None:
            cv_scale = torch.ones_like(scale)
        return _StepFn.apply(scale, cv_scale, self)


class _StepFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, scale, cv_scale, step):
        E = step.forward_np(scale.detach().cpu().double().numpy(), cv_scale.detach().cpu().double().numpy())
        ctx.step = step
        ctx.tag = step._saved[0]
        ctx.meta = (scale.dtype, scale.device, cv_scale.dtype, cv_scale.device)
        return torch.from_numpy(E).to(device=scale.device, dtype=torch.float64)

    @staticmethod
    def backward(ctx, grad_E):
        lam = grad_E.detach().cpu().double().numpy()
        g_scale, g_cv, _, _ = ctx.step.backward_np(lam, ctx.tag)
        sd, sdev, cd, cdev = ctx.meta
        return (torch.from_numpy(g_scale).to(device=sdev, dtype=sd), torch.from_numpy(g_cv).to(device=cdev, dtype=cd),
                None)
