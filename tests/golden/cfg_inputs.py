"""Seeded inputs of two headline workloads, shared by the fixture generator
(make_fixtures.py, which runs the oracle on them) and the GPU tests that
replay them on the device.

  cfg0: 2D 64 x 64 plane strain, FP64: mu ~ U[0.9, 1.1] mu_unit, nu = 0.35,
        eigenstrain xx, yy, xy ~ N(0, 1e-3) i.i.d. then mean-subtracted per
        component (xy mirrored), seed 0; athermal (T = 0) AQS tension along
        xx to 2 % in 200 increments of 1e-4 with sigma_yy = 0, LS tol 1e-8.
  cfg1: 3D 128^3, the same mu / nu / eigenstrain distributions (6 symmetric
        components) from seed 1; kT/E0 = 0.05 (E0 = 2.2 eV, 1276.5 K); 500
        KMC events, the numpy stream and the selection stream seeded 1.

Material and STZ constants are the reference's config.yaml defaults: M = 20
modes, gamma0 = 0.14, pixel 0.7 nm, jp = 20, jt = 0, softening cap 0.916,
gaussian barriers 2.2 +- 0.45 eV clipped at 0, nu0 = 1e13 /s, strain rate
1e7 /s, mixed-BC tolerance 1 MPa; mu_unit = 70 GPa / 2.6 (E = 70 GPa, nu =
0.3).
"""
import numpy as np

MU_UNIT = 70e9 / 2.6
KB = 8.617e-5
STZ = dict(softening_cap=0.916, jp=20.0, jt=0.0)
BARRIERS = dict(mean=2.2, std=0.45, min_cutoff=0.0)

CFG0 = dict(n=64, M=20, gamma0=0.14, pixel=0.7, temperature=0.0, strain_rate=1e7, step_size=1e-4, steps=200,
            ls_tol=1e-8, mixed_tol=1e6, seed=0, sel_seed=0)
CFG1 = dict(n=128, M=20, gamma0=0.14, pixel=0.7, temperature=0.05 * 2.2 / KB, strain_rate=1e7, step_size=1e-4,
            events=500, ls_tol=1e-6, mixed_tol=1e6, seed=1, sel_seed=1)


def cfg0_inputs(n=64):
    """(E, nu, eps_plastic (n, n, 2, 2)) of cfg0."""
    rng = np.random.default_rng(0)
    mu = MU_UNIT * rng.uniform(0.9, 1.1, (n, n))
    nu = np.full((n, n), 0.35)
    E = 2.0 * mu * (1.0 + nu)
    g = rng.standard_normal((3, n, n)) * 1e-3
    g -= g.mean(axis=(1, 2), keepdims=True)
    ep = np.zeros((n, n, 2, 2))
    ep[..., 0, 0], ep[..., 1, 1], ep[..., 0, 1] = g[0], g[1], g[2]
    ep[..., 1, 0] = g[2]
    return E, nu, ep


def cfg1_inputs(n=128):
    """(E, nu, eps_plastic (n, n, n, 3, 3)) of cfg1."""
    rng = np.random.default_rng(1)
    mu = MU_UNIT * rng.uniform(0.9, 1.1, (n, n, n))
    nu = np.full((n, n, n), 0.35)
    E = 2.0 * mu * (1.0 + nu)
    g = rng.standard_normal((6, n, n, n)) * 1e-3
    g -= g.mean(axis=(1, 2, 3), keepdims=True)
    ep = np.zeros((n, n, n, 3, 3))
    for c, (i, j) in enumerate(((0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1))):
        ep[..., i, j] = g[c]
        ep[..., j, i] = g[c]
    return E, nu, ep
