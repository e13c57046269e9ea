"""Oracle: fast patching, the predictor step of the reference's KMC loops.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Restates mgkmc/aqs.py:196-256 (3D set-up: five trace-free basis eigenstrains
at the grid centre, each solved with update_stress_fft_full, the stress
rolled so the centre sits at the origin, cropped to a (2R+1)^3 box (a flat
axis keeps one plane), and the "missing mean" full_mean - crop_sum / NV),
mgkmc/aqs.py:1021-1054 (per KMC event: the macro-unit stress scaled by
strain_rate * t_wait, the patch c0 K0 + ... + c4 K4 with
(c0..c4) = (C00, -C22, C01, C02, C12) added over the periodic box around the
flipped voxel, then the combined missing mean added everywhere; a full
re-solve every sync_interval flips), and the 2D counterparts
mgkmc/kmc_simulator.py:229-261 (two bases, pure and simple shear, solved with
spectral_solver_2d on the modulus field), :764-792 (cascade batches: patches
from the catalog entry after its redraw, no re-solve per batch, one re-solve
after a cascade that flipped anything) and :1044-1098 (KMC events, and the
macro-unit stress over the remaining time of an increment once it exists).

Arithmetic order is numpy's: patch = ((((c0 K0 + c1 K1) + c2 K2) + c3 K3) + c4 K4)
elementwise, the stamp adds it voxel by voxel in (dx, dy, dz) order, so a
voxel that the box covers twice (2R + 1 > n) receives both values in that
order.

Pinned in tests/test_oracle_kmc.py: on an odd grid with a homogeneous
modulus and 2R + 1 = n the box covers every voxel once and the predictor is
the exact superposition of point solutions, so the patched field after a flip
equals a full re-solve to the LS tolerance and the missing mean vanishes; a
patch stamped at a translated voxel equals the translated point solution;
each basis reproduces the mode it stands for (a flipped pure-shear and
simple-shear mode give the full solve's field).
"""
import numpy as np

from . import ls as ol

BASES_3D = [
    np.array([[1.0, 0, 0], [0, -1.0, 0], [0, 0, 0]]),
    np.array([[0, 0, 0], [0, 1.0, 0], [0, 0, -1.0]]),
    np.array([[0, 1.0, 0], [1.0, 0, 0], [0, 0, 0]]),
    np.array([[0, 0, 1.0], [0, 0, 0], [1.0, 0, 0]]),
    np.array([[0, 0, 0], [0, 0, 1.0], [0, 1.0, 0]]),
]
BASES_2D = [np.array([[1.0, 0], [0, -1.0]]), np.array([[0, 1.0], [1.0, 0]])]


def patch_kernels_3d(E, nu, R, pixel=1.0, tol=1e-6, max_iter=200):
    """(kernels [5 x (2R+1, 2R+1, W, 3, 3)], missing means [5 x (3, 3)])
    (mgkmc/aqs.py:220-256)."""
    nx, ny, nz = E.shape
    kernels, missing = [], []
    for P in BASES_3D:
        ep = np.zeros((nx, ny, nz, 3, 3))
        ep[nx // 2, ny // 2, nz // 2] = P
        _, s, _, _ = ol.update_stress_fft_full(ep, np.zeros((3, 3)), E, nu, pixel=pixel, tol=tol, max_iter=max_iter)
        rolled = np.roll(s, shift=(-(nx // 2), -(ny // 2), -(nz // 2)), axis=(0, 1, 2))
        crop = np.zeros((2 * R + 1, 2 * R + 1, 1 if nz == 1 else 2 * R + 1, 3, 3))
        for dx in range(-R, R + 1):
            for dy in range(-R, R + 1):
                for dz in range(-R if nz > 1 else 0, R + 1 if nz > 1 else 1):
                    crop[dx + R, dy + R, dz + R if nz > 1 else 0] = rolled[dx % nx, dy % ny, dz % nz]
        kernels.append(crop)
        missing.append(np.mean(s, axis=(0, 1, 2)) - np.sum(crop, axis=(0, 1, 2)) / (nx * ny * max(1, nz)))
    return kernels, missing


def patch_kernels_2d(E, nu, R, pixel=1.0, plane_mode="plane_strain", tol=1e-6, max_iter=200):
    """(kernels [2 x (2R+1, 2R+1, 2, 2)], missing means [2 x (2, 2)])
    (mgkmc/kmc_simulator.py:229-261)."""
    nx, ny = E.shape
    kernels, missing = [], []
    for P in BASES_2D:
        ep = np.zeros((nx, ny, 2, 2))
        ep[nx // 2, ny // 2] = P
        _, s, _, _ = ol.spectral_solver_2d(E, nu, np.zeros((2, 2)), eps_plastic=ep, pixel=pixel,
                                           plane_mode=plane_mode, tol=tol, max_iter=max_iter)
        rolled = np.roll(s, shift=(-(nx // 2), -(ny // 2)), axis=(0, 1))
        crop = np.zeros((2 * R + 1, 2 * R + 1, 2, 2))
        for dx in range(-R, R + 1):
            for dy in range(-R, R + 1):
                crop[dx + R, dy + R] = rolled[dx % nx, dy % ny]
        kernels.append(crop)
        missing.append(np.mean(s, axis=(0, 1)) - np.sum(crop, axis=(0, 1)) / (nx * ny))
    return kernels, missing


def coefficients_3d(C):
    """Basis weights of a trace-free mode (mgkmc/aqs.py:1030)."""
    return C[0, 0], -C[2, 2], C[0, 1], C[0, 2], C[1, 2]


def coefficients_2d(C):
    """Basis weights of a 2D mode (mgkmc/kmc_simulator.py:765, :1055)."""
    return C[0, 0], C[0, 1]


def stamp(sig, center, coefs, kernels, missing, R):
    """Adds the combined patch over the periodic box around `center`, then
    the combined missing mean everywhere (mgkmc/aqs.py:1031-1047,
    kmc_simulator.py:766-774).  sig: (nx, ny[, nz], d, d), in place."""
    patch = coefs[0] * kernels[0]
    for c, k in zip(coefs[1:], kernels[1:]):
        patch = patch + c * k
    shift = coefs[0] * missing[0]
    for c, mm in zip(coefs[1:], missing[1:]):
        shift = shift + c * mm
    if sig.ndim == 5:
        nx, ny, nz = sig.shape[:3]
        x, y, z = center
        for dx in range(-R, R + 1):
            for dy in range(-R, R + 1):
                for dz in range(-R if nz > 1 else 0, R + 1 if nz > 1 else 1):
                    sig[(x + dx) % nx, (y + dy) % ny, (z + dz) % nz] += patch[dx + R, dy + R, dz + R if nz > 1 else 0]
    else:
        nx, ny = sig.shape[:2]
        x, y = center
        for dx in range(-R, R + 1):
            for dy in range(-R, R + 1):
                sig[(x + dx) % nx, (y + dy) % ny] += patch[dx + R, dy + R]
    sig += shift
