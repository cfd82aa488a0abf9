"""Deterministic inputs for the benchmark workloads (synthetic, seeded).

* ``marshak_2d_heterogeneous`` -- gray 2-D xy, 200x200 cells on [0,1]^2, per-cell
  sigma0 = exp(2.3 + 1.5 g) with g a unit-variance Gaussian random field
  (white noise FFT-convolved with a periodic Gaussian kernel of 8-cell
  width), sigma_a = sigma0 max(T,1e-3)^-3, c_v = 0.3, T0 = 1e-3 keV, left face
  at 0.5 keV, dt = 0.005 ns x 20 steps, 1e7 new particles per step split
  evenly between emission and the hot face, census combed to 1e7.
* ``marshak_slab`` -- gray 1-D slab, 100 uniform cells on [0,0.2] cm,
  sigma_a = 300 max(T,1e-3)^-3, c_v = 0.3, T0 = 1e-3 keV, left face 1 keV,
  dt = 0.01 ns x 10 steps, 1e5 new particles per step, census combed to 1e5.
* ``batch_slabs`` -- 256 independent gray slabs of length 0.2 cm: cells
  floor(exp U(ln 50, ln 2000)), particles per step floor(exp U(ln 1e4, ln 1e6))
  (half emission, half from the hot face, census combed to that count),
  sigma0 ~ U(10, 1000) with sigma_a = sigma0 max(T,1e-3)^-3, c_v ~ U(0.05, 0.5),
  T_b ~ U(0.1, 1) keV on the left face, T0 = 1e-3 keV, dt = 0.01 ns x 10.
* ``marshak_wave_compare`` -- the reference's Fleck-Cummings vs Carter-Forest
  comparison case (IMC/marshak_wave_profile_compare.py, defaults): gray slab,
  50 cells on [0, 0.2] cm, sigma_a = 300 T^-3 (no floor), c_v = 0.3,
  T0 = Tr0 = 1e-4 keV, left face 1 keV, right face reflecting,
  dt = 0.025 ns to 1 ns, 1e4 emission and 1e4 face particles per step,
  census combed to 4e4.
* ``crooked_pipe_mg`` -- multigroup 2-D xy, 240x240 cells on [0,2]^2, 10 groups
  with edges logspace(-4, 1, 11) keV and nu_g = sqrt(lo hi).  The thin
  channel is every cell within 0.05 cm of a polyline from (0, y0) to (2, y5)
  through x = 0.4, 0.8, 1.2, 1.6 with all six y ~ U(0.3, 1.7) (seed 1):
  sigma_g = 0.2 cm^-1, c_v = 0.01.  Elsewhere sigma_g = 200 (nu_g/T)^-3 T^-0.5,
  c_v = 0.3.  Inlet: the left-face cells inside the channel at 0.5 keV;
  T0 = 0.05 keV; dt = 0.002 ns x 20; 2e7 particles per step split between the
  inlet and the material in proportion to their energies (fmin 0.1).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .materials import LinearEOS, PowerLawOpacity  # noqa: F401  (the problems' material models)


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


@dataclass
class SlabProblem:
    name: str
    mesh: np.ndarray  # (I, 2) cell edges
    opacity: PowerLawOpacity
    eos: LinearEOS
    T0: float
    T_boundary: tuple
    reflect: tuple
    dt: float
    n_steps: int
    n_particles: int

    @property
    def I(self):
        return self.mesh.shape[0]

    @property
    def Ntarget(self):
        return self.n_particles // 2

    @property
    def Nboundary(self):
        return self.n_particles - self.n_particles // 2

    @property
    def Nmax(self):
        return self.n_particles


def _slab_mesh(I, L):
    e = np.linspace(0.0, L, I + 1)
    return np.column_stack([e[:-1], e[1:]])


def marshak_slab(I=100, n_particles=100_000, n_steps=10):
    """The reference's own CPU-runnable slab case (1e5 new particles per step)."""
    return SlabProblem(
        name=f"marshak_slab_{I}", mesh=_slab_mesh(I, 0.2),
        opacity=PowerLawOpacity(np.full(I, 300.0), 3.0, 1e-3), eos=LinearEOS(0.3), T0=1e-3,
        T_boundary=(1.0, 0.0), reflect=(False, False), dt=0.01, n_steps=n_steps, n_particles=int(n_particles))


@dataclass
class WaveCompareCase:
    mesh: np.ndarray
    T0: float = 1e-4
    T_boundary: tuple = (1.0, 0.0)
    reflect: tuple = (False, True)
    dt: float = 0.025
    t_final: float = 1.0
    n_target: int = 10_000
    n_boundary: int = 10_000
    n_max: int = 40_000
    cv: float = 0.3
    geometry: str = "slab"

    @property
    def I(self):
        return self.mesh.shape[0]

    @staticmethod
    def sigma_a(T):
        return 300.0 * T ** -3

    def eos(self, T):
        return self.cv * T

    def inv_eos(self, u):
        return u / self.cv

    def cv_func(self, T):
        return self.cv + 0.0 * T


def marshak_wave_compare(cells=50, length=0.2, geometry="slab"):
    """IMC/marshak_wave_profile_compare.py:65-98 with its default arguments.
    ``geometry="spherical"`` is the driver's ``--geometry spherical`` case
    (lines 85-91): reflecting centre, a 1 keV bath at the outer radius."""
    if geometry == "slab":
        return WaveCompareCase(mesh=_slab_mesh(cells, length))
    if geometry == "spherical":
        return WaveCompareCase(mesh=_slab_mesh(cells, length), T_boundary=(0.0, 1.0), reflect=(True, False),
                               geometry="spherical")
    raise ValueError(f"geometry must be 'slab' or 'spherical', got {geometry!r}")


def batch_slabs(P=256, seed=1, n_steps=10, scale=1.0):
    """P independent slabs with seeded, log-uniform cell and particle counts.
    ``scale`` multiplies every particle count (for bounded samples)."""
    rng = np.random.default_rng(seed)
    cells = np.floor(np.exp(rng.uniform(np.log(50), np.log(2000), P))).astype(int)
    parts = np.floor(np.exp(rng.uniform(np.log(1e4), np.log(1e6), P))).astype(np.int64)
    sig0 = rng.uniform(10.0, 1000.0, P)
    cv = rng.uniform(0.05, 0.5, P)
    Tb = rng.uniform(0.1, 1.0, P)
    parts = np.maximum((parts * scale).astype(np.int64), 2)
    return [SlabProblem(
        name=f"batch_slab_{p}", mesh=_slab_mesh(int(cells[p]), 0.2),
        opacity=PowerLawOpacity(np.full(int(cells[p]), sig0[p]), 3.0, 1e-3), eos=LinearEOS(float(cv[p])), T0=1e-3,
        T_boundary=(float(Tb[p]), 0.0), reflect=(False, False), dt=0.01, n_steps=n_steps,
        n_particles=int(parts[p])) for p in range(P)]


@dataclass
class MGProblem:
    name: str
    edges1: np.ndarray
    edges2: np.ndarray
    energy_edges: np.ndarray
    geometry: str
    thin: np.ndarray          # (nx, ny) bool
    cv: np.ndarray            # (nx, ny)
    T0: float
    dt: float
    n_steps: int
    n_particles: int
    inlet: tuple              # (y_lo, y_hi) of the left-face inlet
    T_inlet: float

    @property
    def shape(self):
        return (self.edges1.size - 1, self.edges2.size - 1)

    @property
    def G(self):
        return self.energy_edges.size - 1

    def sigma_funcs(self):
        """One callable per group.  200 (nu/T)^-3 T^-0.5 = 200 nu^-3 T^2.5; the
        T^2.5 factor is shared by the groups of one call sequence."""
        nu = np.sqrt(self.energy_edges[:-1] * self.energy_edges[1:])
        thin = self.thin
        cache = {}

        def t25(T):
            if cache.get("T") is not T:
                Tc = np.maximum(T, 1e-3)
                cache["T"] = T
                cache["v"] = Tc * Tc * np.sqrt(Tc)
            return cache["v"]

        def make(g):
            c = 200.0 * nu[g] ** -3

            def f(T):
                return np.where(thin, 0.2, c * t25(T))
            return f
        return [make(g) for g in range(self.G)]

    def eos(self, T):
        return self.cv * T

    def inv_eos(self, u):
        return u / self.cv

    def cv_func(self, T):
        return self.cv + 0.0 * T

    def boundary_source_func(self, x, y, side):
        lo, hi = self.inlet
        return self.T_inlet if (side == "left" and lo <= y <= hi) else 0.0


def _dist_to_polyline(px, py, vx, vy):
    d = np.full(px.shape, np.inf)
    for k in range(len(vx) - 1):
        ax, ay, bx, by = vx[k], vy[k], vx[k + 1], vy[k + 1]
        ex, ey = bx - ax, by - ay
        t = np.clip(((px - ax) * ex + (py - ay) * ey) / (ex * ex + ey * ey), 0.0, 1.0)
        d = np.minimum(d, np.hypot(px - (ax + t * ex), py - (ay + t * ey)))
    return d


def crooked_pipe_mg(n=240, n_particles=20_000_000, n_steps=20, seed=1):
    e = np.linspace(0.0, 2.0, n + 1)
    c = 0.5 * (e[:-1] + e[1:])
    rng = np.random.default_rng(seed)
    vy = rng.uniform(0.3, 1.7, 6)
    vx = np.array([0.0, 0.4, 0.8, 1.2, 1.6, 2.0])
    X, Y = np.meshgrid(c, c, indexing="ij")
    thin = _dist_to_polyline(X, Y, vx, vy) <= 0.05
    cv = np.where(thin, 0.01, 0.3)
    return MGProblem(name=f"crooked_pipe_mg_{n}x{n}", edges1=e, edges2=e.copy(),
                     energy_edges=np.logspace(-4, 1, 11), geometry="xy", thin=thin, cv=cv, T0=0.05, dt=0.002,
                     n_steps=n_steps, n_particles=int(n_particles), inlet=(vy[0] - 0.05, vy[0] + 0.05),
                     T_inlet=0.5)


@dataclass
class MGPowerLawMarshak:
    """The reference's 10-group power-law Marshak wave
    (MG_IMC/problems/test_marshak_wave_multigroup_powerlaw.py:149-320, 726-836,
    ``--mode publication``): 140 uniform cells on [0, 7] cm x one cell of
    height 1 (xy, top and bottom reflecting), rho = 0.01 g/cm^3,
    sigma_a(T, E) = min(10 rho max(T, 1e-2)^-1/2 E^-3, 1e14) with each group
    opacity the geometric mean of its edge values, c_v = 0.05 rho, T0 = Tr0 =
    0.005 keV, left face ramped from 0.05 to 0.25 keV over 5 ns, right face
    at T0, dt = 0.01 ns; 5e5 emission and 5e5 face particles per step, census
    combed to Nmax = 2e6 growing by 1e5 per step to 2e7; T_emit_floor 0.02 keV
    (the driver's command-line default); max_events_per_particle 1e5."""

    nx: int = 140
    n_groups: int = 10
    rho: float = 0.01
    dt: float = 0.01
    T0: float = 0.005
    Ntarget: int = 500_000
    Nboundary: int = 500_000
    Nmax: int = 2_000_000
    Nmax_growth: int = 100_000
    Nmax_final: int = 20_000_000
    T_emit_floor: float = 0.02
    max_events: int = 100_000

    @property
    def edges1(self):
        return np.linspace(0.0, 7.0, self.nx + 1)

    @property
    def edges2(self):
        return np.array([0.0, 1.0])

    @property
    def energy_edges(self):
        return np.logspace(np.log10(1e-4), np.log10(10.0), self.n_groups + 1)

    @property
    def shape(self):
        return (self.nx, 1)

    @property
    def cv(self):
        return 0.05 * self.rho

    def _sigma_at(self, T, E):
        return np.minimum(10.0 * self.rho * np.maximum(T, 1e-2) ** -0.5 * E ** -3.0, 1e14)

    @property
    def sigma_a_funcs(self):
        e = self.energy_edges

        def make(lo, hi):
            return lambda T: np.sqrt(self._sigma_at(T, lo) * self._sigma_at(T, hi))
        return [make(e[g], e[g + 1]) for g in range(self.n_groups)]

    def T_left(self, t):
        return 0.05 + (0.25 - 0.05) * (t / 5.0) if t < 5.0 else 0.25

    @property
    def T_boundary(self):
        return (self.T_left, self.T0, 0.0, 0.0)

    @property
    def reflect(self):
        return (False, False, True, True)

    def eos(self, T):
        return self.cv * T

    def inv_eos(self, u):
        return u / self.cv

    def cv_func(self, T):
        return self.cv * np.ones_like(T)


def mg_powerlaw_marshak(**kw):
    return MGPowerLawMarshak(**kw)


@dataclass
class CrookedPipeMGRZ:
    """The reference's multigroup crooked-pipe driver
    (MG_IMC/problems/crooked_pipe_multigroup_imc.py, default arguments,
    ``--mode quick``): rz, r in [0, 2] cm (40 cells), z in [0, 7] cm (140
    cells); thin where r < 0.5 and (z < 3 or z > 4), r < 1.5 and 2.5 < z < 3,
    r < 1.5 and 4 < z < 4.5, or 1 <= r < 1.5 and 2.5 < z < 4.5 (density 0.1),
    thick elsewhere (density 2; is_optically_thick and material_density,
    172-218); sigma_a(T, E) = min(10 rho max(T, 1e-2)^-1/2 E^-3, 1e14) with
    each group the geometric mean of its edge values (237-263), 10 groups
    with edges logspace(-2, 1, 11) keV (809); c_v = 0.05 rho (72); T0 = 0.05
    keV (69); a 0.5 keV surface source on z = 0 for r < 0.5, given by
    boundary_source_func (919-931), every other face open except the axis
    (reflect = (True, False, False, False), 975); dt from 1e-4 ns growing by
    1.1 per step to 0.01 ns; T_emit_floor 0.075 keV; 2e4 emission and 2e4
    face particles per step, census combed to 1e5 growing by 1e4 per step to
    5e5 (759-780)."""

    nr: int = 40
    nz: int = 140
    n_groups: int = 10
    T0: float = 0.05
    T_bc: float = 0.5
    dt_initial: float = 1e-4
    dt_growth: float = 1.1
    dt_max: float = 0.01
    T_emit_floor: float = 0.075
    Ntarget: int = 20_000
    Nboundary: int = 20_000
    Nmax: int = 100_000
    Nmax_growth: int = 10_000
    Nmax_final: int = 500_000
    rho_thick: float = 2.0
    rho_thin: float = 0.1
    cv_mass: float = 0.05

    @property
    def edges1(self):
        return np.linspace(0.0, 2.0, self.nr + 1)

    @property
    def edges2(self):
        return np.linspace(0.0, 7.0, self.nz + 1)

    @property
    def energy_edges(self):
        return np.logspace(-2.0, 1.0, self.n_groups + 1)

    @property
    def shape(self):
        return (self.nr, self.nz)

    @property
    def thick(self):
        r = 0.5 * (self.edges1[:-1] + self.edges1[1:])[:, None]
        z = 0.5 * (self.edges2[:-1] + self.edges2[1:])[None, :]
        thin = ((r < 0.5) & ((z < 3.0) | (z > 4.0))) | ((r < 1.5) & (z > 2.5) & (z < 3.0)) | \
               ((r < 1.5) & (z > 4.0) & (z < 4.5)) | ((r >= 1.0) & (r < 1.5) & (z > 2.5) & (z < 4.5))
        return ~thin

    @property
    def rho(self):
        return np.where(self.thick, self.rho_thick, self.rho_thin)

    def sigma_at(self, T, E):
        return np.minimum(10.0 * self.rho * np.maximum(T, 1e-2) ** -0.5 * E ** -3.0, 1e14)

    @property
    def sigma_a_funcs(self):
        e = self.energy_edges

        def make(lo, hi):
            return lambda T: np.sqrt(self.sigma_at(T, lo) * self.sigma_at(T, hi))
        return [make(e[g], e[g + 1]) for g in range(self.n_groups)]

    @property
    def cv(self):
        return self.cv_mass * self.rho

    def eos(self, T):
        return self.cv * T

    def inv_eos(self, u):
        return u / self.cv

    def cv_func(self, T):
        return self.cv + 0.0 * T

    def boundary_source_func(self, r, z, side):
        return self.T_bc if (side == "zmin" and r < 0.5) else 0.0

    T_boundary = (0.0, 0.0, 0.0, 0.0)
    reflect = (True, False, False, False)

    def dt_schedule(self, t_final):
        """The driver's steps: dt_initial growing by dt_growth to dt_max, the
        last step cut to land on t_final."""
        t, dt, out = 0.0, self.dt_initial, []
        while t < t_final - 1e-15:
            d = min(dt, t_final - t)
            out.append(d)
            t += d
            dt = min(dt * self.dt_growth, self.dt_max)
        return out


def crooked_pipe_mg_rz(**kw):
    return CrookedPipeMGRZ(**kw)
