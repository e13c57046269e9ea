"""Isotropic moduli helpers used to prepare solver inputs on the host.

Same formulas as mgkmc/elasticity.py:3-10 (compute_lame_2d) and
mgkmc/elasticity.py:55-59 (compute_lame_3d).  These run once per simulation
(moduli are fixed for a run); the per-iteration stress evaluation happens on
the GPU inside the LS kernels.
"""
import numpy as np


def compute_lame_2d(E, nu, plane_mode="plane_strain"):
    """Lame parameters for 2D isotropic elasticity (plane strain or stress)."""
    mu = E / (2 * (1 + nu))
    if plane_mode == "plane_stress":
        lam_2d = E * nu / (1 - nu ** 2)
    else:
        lam_2d = E * nu / ((1 + nu) * (1 - 2 * nu))
    return lam_2d, mu


def compute_lame_3d(E, nu):
    """Lame parameters for 3D isotropic elasticity."""
    mu = E / (2 * (1 + nu))
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    return lam, mu


def reference_medium(E, nu, d, plane_mode="plane_strain"):
    """(lam0, mu0) of the LS reference medium: Lame pair of (mean E, mean nu),
    as mgkmc/linear_elastic_simulator.py:22-27 (2D) and :284-289 (3D)."""
    E_avg = np.asarray(E).mean()
    nu_avg = np.asarray(nu).mean()
    if d == 3:
        return compute_lame_3d(E_avg, nu_avg)
    return compute_lame_2d(E_avg, nu_avg, plane_mode)
