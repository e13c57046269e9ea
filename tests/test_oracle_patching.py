"""Pin the fast-patching oracle (oracle/patching.py) against what linearity
fixes: in a homogeneous medium on an odd grid with 2R + 1 = n the box covers
every voxel once, so the predicted field equals a full LS re-solve to the LS
tolerance, the missing mean vanishes, and the KMC loops with and without the
predictor pick the same events."""
import numpy as np
import pytest

from oracle import ls as ol
from oracle import patching as opt
from oracle import aqs_loop as oal
from oracle import kmc2d_loop as ok2


def _mode3(rng):
    g = rng.standard_normal((3, 3))
    g = 0.5 * (g + g.T)
    return 0.01 * (g - np.trace(g) / 3 * np.eye(3))


def test_patch_3d_full_box_equals_resolve():
    rng = np.random.default_rng(3)
    n, R = 5, 2
    E = np.full((n, n, n), 70e9)
    nu = np.full((n, n, n), 0.3)
    K, mm = opt.patch_kernels_3d(E, nu, R, tol=1e-13, max_iter=500)
    for m in mm:
        assert np.abs(m).max() < 1e-9 * np.abs(K[0]).max()
    ep = np.zeros((n, n, n, 3, 3))
    ep[1, 2, 3] = _mode3(rng)
    _, s0, _, _ = ol.update_stress_fft_full(ep, np.zeros((3, 3)), E, nu, tol=1e-13, max_iter=500)
    for v in [(0, 0, 0), (4, 1, 2), (2, 4, 4)]:
        C = _mode3(rng)
        s = s0.copy()
        opt.stamp(s, v, opt.coefficients_3d(C), K, mm, R)
        ep1 = ep.copy()
        ep1[v] += C
        _, s1, _, _ = ol.update_stress_fft_full(ep1, np.zeros((3, 3)), E, nu, tol=1e-13, max_iter=500)
        assert np.abs(s - s1).max() < 1e-8 * np.abs(s1).max()
        # a wrong basis weight (the sign of C22) breaks the superposition
        bad = list(opt.coefficients_3d(C))
        bad[1] = -bad[1]
        sb = s0.copy()
        opt.stamp(sb, v, bad, K, mm, R)
        if abs(C[2, 2]) > 1e-4:
            assert np.abs(sb - s1).max() > 1e-3 * np.abs(s1).max()


def test_patch_3d_translated_point_solution():
    # a patch stamped at voxel v is the point solution translated to v
    n, R = 7, 3
    E = np.full((n, n, n), 50e9)
    nu = np.full((n, n, n), 0.25)
    K, mm = opt.patch_kernels_3d(E, nu, R, tol=1e-13, max_iter=500)
    for b, P in enumerate(opt.BASES_3D):
        ep = np.zeros((n, n, n, 3, 3))
        ep[6, 0, 2] = P
        _, s1, _, _ = ol.update_stress_fft_full(ep, np.zeros((3, 3)), E, nu, tol=1e-13, max_iter=500)
        coefs = [0.0] * 5
        coefs[b] = 1.0
        s = np.zeros_like(s1)
        opt.stamp(s, (6, 0, 2), coefs, K, mm, R)
        assert np.abs(s - s1).max() < 1e-8 * np.abs(s1).max()


def test_patch_3d_truncated_box_keeps_mean():
    # R < n/2: the missing mean restores the field mean of the point solution
    rng = np.random.default_rng(5)
    n, R = 8, 2
    E = 70e9 * rng.uniform(0.8, 1.2, (n, n, n))
    nu = np.full((n, n, n), 0.3)
    K, mm = opt.patch_kernels_3d(E, nu, R, tol=1e-12, max_iter=500)
    for b, P in enumerate(opt.BASES_3D):
        ep = np.zeros((n, n, n, 3, 3))
        ep[n // 2, n // 2, n // 2] = P
        _, s1, _, _ = ol.update_stress_fft_full(ep, np.zeros((3, 3)), E, nu, tol=1e-12, max_iter=500)
        coefs = [0.0] * 5
        coefs[b] = 1.0
        s = np.zeros_like(s1)
        opt.stamp(s, (n // 2, n // 2, n // 2), coefs, K, mm, R)
        assert np.allclose(s.mean(axis=(0, 1, 2)), s1.mean(axis=(0, 1, 2)), rtol=0, atol=1e-9 * np.abs(s1).max())
        # inside the box the patch is the solution itself
        c = n // 2
        box = s1[c - R:c + R + 1, c - R:c + R + 1, c - R:c + R + 1]
        got = s[c - R:c + R + 1, c - R:c + R + 1, c - R:c + R + 1]
        assert np.abs(got - box - mm[b]).max() < 1e-9 * np.abs(s1).max()


def test_patch_wrapped_box_adds_twice():
    # 2R + 1 > n: a voxel covered twice receives both patch entries
    n, R = 3, 2
    K = [np.arange((2 * R + 1) ** 2 * 4, dtype=float).reshape(2 * R + 1, 2 * R + 1, 2, 2), np.zeros((5, 5, 2, 2))]
    mm = [np.zeros((2, 2)), np.zeros((2, 2))]
    s = np.zeros((n, n, 2, 2))
    opt.stamp(s, (0, 0), (1.0, 0.0), K, mm, R)
    ref = np.zeros_like(s)
    for dx in range(-R, R + 1):
        for dy in range(-R, R + 1):
            ref[dx % n, dy % n] += K[0][dx + R, dy + R]
    assert np.array_equal(s, ref)
    assert np.isclose(s.sum(), K[0].sum())


def test_patch_2d_full_box_equals_resolve():
    rng = np.random.default_rng(7)
    n, R = 9, 4
    E = np.full((n, n), 60e9)
    nu = np.full((n, n), 0.35)
    for pm in ("plane_strain", "plane_stress"):
        K, mm = opt.patch_kernels_2d(E, nu, R, plane_mode=pm, tol=1e-13, max_iter=500)
        ep = np.zeros((n, n, 2, 2))
        ep[3, 5] = [[0.01, 0.004], [0.004, -0.01]]
        _, s0, _, _ = ol.spectral_solver_2d(E, nu, np.zeros((2, 2)), eps_plastic=ep, plane_mode=pm, tol=1e-13,
                                            max_iter=500)
        gxx, gxy = rng.standard_normal(2) * 0.01
        C = np.array([[gxx, gxy], [gxy, -gxx]])
        s = s0.copy()
        opt.stamp(s, (8, 1), opt.coefficients_2d(C), K, mm, R)
        ep[8, 1] += C
        _, s1, _, _ = ol.spectral_solver_2d(E, nu, np.zeros((2, 2)), eps_plastic=ep, plane_mode=pm, tol=1e-13,
                                            max_iter=500)
        assert np.abs(s - s1).max() < 1e-8 * np.abs(s1).max()


def test_loop_3d_full_box_predictor_picks_the_same_events():
    n = 5
    kw = dict(nx=n, ny=n, nz=n, M=4, gamma0=0.05, E=np.full((n, n, n), 70e9), nu=np.full((n, n, n), 0.3),
              pixel=1.0, temperature=300.0, strain_rate=1e7, step_size=2e-4, component=(0, 0), n_events=12,
              seed=11, sel_seed=12, ls_tol=1e-10, stress_targets={}, mixed_max_iter=0)
    ev0, st0 = oal.run_events(**kw)
    ev1, st1 = oal.run_events(**kw, fast_patching=dict(enabled=True, patch_radius=2, sync_interval=1000))
    assert [e[:4] for e in ev0] == [e[:4] for e in ev1]
    assert np.allclose([e[4] for e in ev0], [e[4] for e in ev1], rtol=1e-9, atol=0)
    h0, h1 = np.array(st0["stress_history"]), np.array(st1["stress_history"])
    assert np.abs(h0 - h1).max() < 1e-6 * np.abs(h0).max()


def test_loop_2d_full_box_predictor_picks_the_same_events():
    n = 7
    kw = dict(nx=n, ny=n, M=20, gamma0=0.14, E=np.full((n, n), 70e9), nu=np.full((n, n), 0.35), pixel=0.7,
              temperature=1276.5, strain_rate=1e7, step_size=1e-2, component=(0, 0), n_events=15, seed=21,
              sel_seed=22, ls_tol=1e-10, mixed_max_iter=0)
    ev0, st0 = ok2.run_events_2d(**kw)
    ev1, st1 = ok2.run_events_2d(**kw, fast_patching=dict(enabled=True, patch_radius=3, sync_interval=1000))
    assert [e[:3] for e in ev0] == [e[:3] for e in ev1]
    assert np.allclose([e[3] for e in ev0], [e[3] for e in ev1], rtol=1e-9, atol=0)
