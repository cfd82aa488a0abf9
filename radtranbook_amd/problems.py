"""Deterministic inputs for the benchmark workloads (synthetic, seeded).

* ``marshak_2d_heterogeneous`` -- gray 2-D xy, 200x200 cells on [0,1]^2, per-cell
  sigma0 = exp(2.3 + 1.5 g) with g a unit-variance Gaussian random field
  (white noise FFT-convolved with a periodic Gaussian kernel of 8-cell
  width), sigma_a = sigma0 max(T,1e-3)^-3, c_v = 0.3, T0 = 1e-3 keV, left face
  at 0.5 keV, dt = 0.005 ns x 20 steps, 1e7 new particles per step split
  evenly between emission and the hot face, census combed to 1e7.
* ``marshak_slab`` -- gray 1-D slab, 100 cells on [0,0.2] cm, sigma = 300 T^-3.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


def gaussian_field(n1, n2, corr_cells, seed):
    """Zero-mean, unit-variance periodic Gaussian random field."""
    g = np.random.default_rng(seed).standard_normal((n1, n2))
    k1 = np.fft.fftfreq(n1)[:, None]
    k2 = np.fft.fftfreq(n2)[None, :]
    # FFT of a periodic Gaussian kernel of standard deviation corr_cells
    kern = np.exp(-2.0 * (np.pi * corr_cells) ** 2 * (k1 * k1 + k2 * k2))
    f = np.real(np.fft.ifft2(np.fft.fft2(g) * kern))
    f -= f.mean()
    f /= f.std()
    return f


@dataclass
class PowerLawOpacity:
    """sigma(T) = sigma0 * max(T, T_floor)^-power, per cell."""

    sigma0: np.ndarray
    power: float = 3.0
    T_floor: float = 1e-3

    def __call__(self, T):
        return self.sigma0 * np.maximum(T, self.T_floor) ** (-self.power)


@dataclass
class LinearEOS:
    cv: float

    def eos(self, T):
        return self.cv * T

    def inv_eos(self, u):
        return u / self.cv

    def cv_func(self, T):
        return self.cv + 0.0 * T


@dataclass
class Problem2D:
    name: str
    edges1: np.ndarray
    edges2: np.ndarray
    geometry: str
    opacity: PowerLawOpacity
    eos: LinearEOS
    T0: float
    T_boundary: tuple
    reflect: tuple
    dt: float
    n_steps: int
    n_particles: int
    max_events: int = 1_000_000

    @property
    def shape(self):
        return (self.edges1.size - 1, self.edges2.size - 1)

    @property
    def Ntarget(self):
        return self.n_particles // 2

    @property
    def Nboundary(self):
        n_hot = sum(1 for t in self.T_boundary if t > 0)
        return (self.n_particles - self.Ntarget) // max(n_hot, 1)

    @property
    def Nmax(self):
        return self.n_particles


def marshak_2d_heterogeneous(n=200, n_particles=10_000_000, n_steps=20, seed=1):
    e = np.linspace(0.0, 1.0, n + 1)
    g = gaussian_field(n, n, 8.0, seed)
    sigma0 = np.exp(2.3 + 1.5 * g)
    return Problem2D(
        name=f"marshak_2d_heterogeneous_{n}x{n}",
        edges1=e, edges2=e.copy(), geometry="xy",
        opacity=PowerLawOpacity(sigma0, 3.0, 1e-3), eos=LinearEOS(0.3), T0=1e-3,
        T_boundary=(0.5, 0.0, 0.0, 0.0), reflect=(False, False, False, False),
        dt=0.005, n_steps=n_steps, n_particles=int(n_particles),
    )
