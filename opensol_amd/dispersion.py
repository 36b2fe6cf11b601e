"""Dispersive poles of the E_z stepper: Drude and Lorentz media with their auxiliary-equation coefficients, stability
terms and the device records of their cell boxes.

The coefficient formulas are the reference's (physics/em/materials.py:132-190): the update runs in velocity form,
so a float32 coefficient near 1 (or 2) never carries the small damping or restoring term. Pole parameters are stored
as float32 and the coefficients rounded to float32, as there.
"""
import math

import numpy as np

from .constants import DTYPE, EPS0


class DrudePole:
    """Drude medium dJ/dt + gamma J = eps0 omega_p^2 E over the background eps_inf."""

    def __init__(self, omega_p, gamma, eps_inf=1.0):
        self.omega_p = DTYPE(omega_p)
        self.gamma = DTYPE(gamma)
        self.eps_inf = DTYPE(eps_inf)

    def ade_coefficients(self, dt):
        """(damping, drive) of J += damping J + drive E^n (physics/em/materials.py:139-149)."""
        dt, gamma, wp = float(dt), float(self.gamma), float(self.omega_p)
        denom = 1.0 + gamma * dt / 2.0
        return DTYPE(-gamma * dt / denom), DTYPE((EPS0 * wp ** 2 * dt) / denom)

    def susceptibility(self, omega):
        """chi(omega) = -omega_p^2 / (omega^2 - j gamma omega) in exp(+j omega t)."""
        omega = np.asarray(omega, dtype=np.float64)
        chi = -float(self.omega_p) ** 2 / (omega ** 2 - 1j * float(self.gamma) * omega)
        return complex(chi) if np.ndim(chi) == 0 else chi


class LorentzPole:
    """Lorentz medium P'' + gamma P' + omega_0^2 P = eps0 delta_eps omega_0^2 E."""

    def __init__(self, delta_eps, omega_0, gamma):
        self.delta_eps = DTYPE(delta_eps)
        self.omega_0 = DTYPE(omega_0)
        self.gamma = DTYPE(gamma)

    def ade_coefficients(self, dt):
        """(damping, restoring, drive) of dP += damping dP - restoring P + drive E^n, P += dP
        (physics/em/materials.py:167-180)."""
        dt, w0, g, de = float(dt), float(self.omega_0), float(self.gamma), float(self.delta_eps)
        denom = 1.0 + g * dt / 2.0
        return (DTYPE(-g * dt / denom), DTYPE(w0 ** 2 * dt ** 2 / denom),
                DTYPE(EPS0 * de * w0 ** 2 * dt ** 2 / denom))

    def susceptibility(self, omega):
        """chi(omega) = delta_eps omega_0^2 / (omega_0^2 - omega^2 + j gamma omega)."""
        omega = np.asarray(omega, dtype=np.float64)
        w0 = float(self.omega_0)
        chi = float(self.delta_eps) * w0 ** 2 / (w0 ** 2 - omega ** 2 + 1j * float(self.gamma) * omega)
        return complex(chi) if np.ndim(chi) == 0 else chi


def pole_from_dict(pole, number=float):
    """DrudePole / LorentzPole of a dispersion entry {'model': 'drude', 'omega_p', 'gamma'[, 'eps_inf']} or
    {'model': 'lorentz', 'delta_eps', 'omega_0', 'gamma'} (physics/em/definition_common.py:1079-1099); `number`
    evaluates each value. Unknown or missing keys, negative values and eps_inf <= 0 are refused."""
    model = pole.get('model') if isinstance(pole, dict) else None
    keys = {'drude': (('omega_p', 'gamma'), ('eps_inf',)),
            'lorentz': (('delta_eps', 'omega_0', 'gamma'), ())}.get(model)
    if keys is None:
        raise ValueError(f"dispersion pole model must be 'drude' or 'lorentz', got {model!r}")
    required, optional = keys
    if set(pole) - {'model', *required, *optional} or any(k not in pole for k in required):
        raise ValueError(f'{model} pole needs {required} (optional {optional}), got {sorted(pole)}')
    numbers = {k: number(pole[k]) for k in (*required, *optional) if k in pole}
    if any(v < 0 for v in numbers.values()) or numbers.get('eps_inf', 1.0) <= 0:
        raise ValueError(f'{model} pole values must be >= 0 (eps_inf > 0), got {numbers}')
    return DrudePole(**numbers) if model == 'drude' else LorentzPole(**numbers)


def background_permittivity(poles):
    """eps_r of cells whose material gives poles but no eps_r: the first Drude pole's eps_inf, else 1
    (physics/em/definition_common.py:1102-1111)."""
    return next((float(p.eps_inf) for p in poles if isinstance(p, DrudePole)), 1.0)


def pole_step_term(pole, dt):
    """Pole term of the stability bound (physics/em/materials.py:193-205): Drude (omega_p dt)^2 / 4, Lorentz
    delta_eps (omega_0 dt)^2 / (4 - (omega_0 dt)^2); ValueError when omega_0 dt >= 2."""
    if isinstance(pole, DrudePole):
        return (float(pole.omega_p) * dt) ** 2 / 4.0
    step = (float(pole.omega_0) * dt) ** 2
    if step >= 4.0:
        raise ValueError(f'Lorentz pole: omega_0 dt = {math.sqrt(step):.4g} >= 2, so the update is unstable; lower '
                         f'courant_factor or refine the mesh')
    return float(pole.delta_eps) * step / (4.0 - step)


def bounding_box(mask):
    """((i0, i1), (j0, j1)) of the True cells of a 2D mask, None when it has none."""
    rows, cols = np.nonzero(np.any(mask, axis=1))[0], np.nonzero(np.any(mask, axis=0))[0]
    if rows.size == 0:
        return None
    return (int(rows[0]), int(rows[-1]) + 1), (int(cols[0]), int(cols[-1]) + 1)


__all__ = ['DrudePole', 'LorentzPole', 'bounding_box', 'pole_step_term']
