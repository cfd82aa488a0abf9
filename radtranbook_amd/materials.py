"""Material models the engine can evaluate on the device.

The drop-in ``step`` functions take the reference's host callables
(``sigma_a_func(T)``, ``cv(T)``, ``inv_eos(u)``; IMC/IMC2D.py:1646-1651,
1800-1810).  When those callables are a ``PowerLawOpacity`` and the bound
``cv_func`` / ``inv_eos`` of one ``LinearEOS`` -- the form every benchmark
configuration uses -- the gray 2-D step keeps T and u on the GPU and evaluates
them there (``imc2d_set_material_powerlaw`` / ``imc2d_fleck_powerlaw`` /
``imc2d_update_material``).  Any other callable runs on the host, as in the
reference.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class PowerLawOpacity:
    """sigma(T) = sigma0 * max(T, T_floor)^-power, per cell."""

    sigma0: np.ndarray
    power: float = 3.0
    T_floor: float = 1e-3

    def __call__(self, T):
        m = np.maximum(T, self.T_floor)
        if self.power == 3.0:
            return self.sigma0 / (m * m * m)  # the generic power is several times slower on the host
        return self.sigma0 * m ** (-self.power)


@dataclass
class LinearEOS:
    """u = c_v T with c_v a constant or a per-cell array."""

    cv: float

    def eos(self, T):
        return self.cv * T

    def inv_eos(self, u):
        return u / self.cv

    def cv_func(self, T):
        return self.cv + 0.0 * T


def device_material(sigma_a_func, cv, inv_eos, shape):
    """(sigma0, power, T_floor, c_v) per cell when the step's callables have
    the device form, else None."""
    if not isinstance(sigma_a_func, PowerLawOpacity):
        return None
    eos = getattr(cv, "__self__", None)
    if not isinstance(eos, LinearEOS) or getattr(inv_eos, "__self__", None) is not eos:
        return None
    if getattr(cv, "__func__", None) is not LinearEOS.cv_func or getattr(inv_eos, "__func__", None) is not LinearEOS.inv_eos:
        return None
    sigma0 = np.broadcast_to(np.asarray(sigma_a_func.sigma0, dtype=np.float64), shape)
    cvv = np.broadcast_to(np.asarray(eos.cv, dtype=np.float64), shape)
    if not (np.all(np.isfinite(sigma0)) and np.all(cvv > 0.0)):
        return None
    return sigma0, float(sigma_a_func.power), float(sigma_a_func.T_floor), cvv
