"""Pins of the 2D finite-strain oracle (oracle/finite_strain2d.py): the 3D
oracle on the same plane problem embedded as an (nx, 1, ny) grid, closed
forms, finite differences of P(F), equilibrium."""
import numpy as np
import pytest

from oracle import finite_strain as fs
from oracle import finite_strain2d as f2

LANDAU = dict(v1=-236.5e9, v2=-27.4e9, v3=11.8e9, g1=-5640.7e9, g2=2207.1e9, g3=-332.5e9, g4=-305.2e9)
M3 = [0, 2]  # in-plane axes of the embedding (the singleton is y)


def embed(A2, fill=1.0):
    """(2, 2, nx, ny) -> (3, 3, nx, 1, ny); the out-of-plane diagonal = fill."""
    nx, ny = A2.shape[2:]
    A3 = np.zeros((3, 3, nx, 1, ny))
    for a in range(2):
        for b in range(2):
            A3[M3[a], M3[b], :, 0, :] = A2[a, b]
    A3[1, 1] = fill
    return A3


def extract(A3):
    return np.stack([np.stack([A3[M3[a], M3[b], :, 0, :] for b in range(2)]) for a in range(2)])


def _problem(sh, seed, plane_mode="plane_strain", spread=0.15):
    rng = np.random.default_rng(seed)
    E = 70e9 * rng.uniform(1 - spread, 1 + spread, sh)
    nu = np.full(sh, 0.3)
    I2, I4, I4rt, I4s, II = f2.identity_tensors_2d(*sh)
    C4 = f2.build_C4_2d(E, nu, I4s, II, plane_mode)
    lam, mu = C4[0, 0, 1, 1], C4[0, 1, 0, 1]
    sh3 = (sh[0], 1, sh[1])
    I2_3, I4_3, I4rt_3, I4s_3, II_3 = fs.identity_tensors(*sh3)
    C4_3 = lam.reshape(sh3)[None, None, None, None] * II_3 + 2.0 * mu.reshape(sh3)[None, None, None, None] * I4s_3
    return E, nu, (I2, I4, I4rt, C4), (I2_3, I4_3, I4rt_3, C4_3), sh3


@pytest.mark.parametrize("model,kw,fp,mode", [
    ("svk", {}, False, "plane_strain"), ("svk", {}, True, "plane_strain"), ("svk", {}, False, "plane_stress"),
    ("landau", LANDAU, False, "plane_strain"), ("landau", LANDAU, True, "plane_strain"),
    ("landau", dict(LANDAU, strain_capping_enabled=True, strain_capping_type="smooth"), False, "plane_strain"),
    ("murnaghan", dict(A_m=-100e9, B_m=-100e9, C_m=-100e9), False, "plane_strain")])
def test_2d_step_equals_the_3d_oracle_on_the_embedded_grid(model, kw, fp, mode):
    sh = (8, 6)
    E, nu, t2, t3, sh3 = _problem(sh, 4, mode)
    Fbar = np.eye(2)
    Fbar[0, 1] = 0.02
    Fbar[1, 1] = 0.996
    Fp = None
    if fp:
        rng = np.random.default_rng(5)
        Fp = t2[0] + 0.01 * rng.standard_normal((2, 2) + sh)
    s2, s3 = {}, {}
    F2, P2, S2, _, Fb2 = f2.solver_step_2d(t2[0].copy(), Fbar, f2.build_ghat4_2d(*sh, *sh), t2[3], *t2[:3], Fp=Fp,
                                           model_type=model, plane_mode=mode, tol_NW=1e-8, tol_CG=1e-8, stats=s2,
                                           **kw)
    F3, P3, S3, _, Fb3 = fs.solver_step(t3[0].copy(), embed(Fbar[:, :, None, None])[:, :, 0, 0, 0],
                                        fs.build_ghat4(*sh3, *sh3), t3[3], *t3[:3],
                                        Fp=None if Fp is None else embed(Fp), model_type=model, tol_NW=1e-8,
                                        tol_CG=1e-8, stats=s3, **kw)
    assert [n[:2] for n in s2["newton"]] == [n[:2] for n in s3["newton"]]
    assert np.abs(F2 - extract(F3)).max() < 1e-13
    assert np.abs(P2 - extract(P3)).max() < 1e-12 * np.abs(P2).max()
    if mode == "plane_strain":
        assert np.abs(S2 - extract(S3)).max() < 1e-12 * np.abs(S2).max()
    # the out-of-plane block of the embedded solution stays untouched
    assert np.all(F3[1, 1] == 1.0) and not np.any(F3[1, [0, 2]]) and not np.any(F3[[0, 2], 1])


def test_2d_plane_stress_svk_closed_form():
    # homogeneous plane stress: E33 = -lam* tr E / (2 mu) makes sigma_33 = 0 of
    # the 3D SVK law with E_33 = E33 (lam* = 2 lam mu / (lam + 2 mu))
    sh = (4, 4)
    E, nu = np.full(sh, 70e9), np.full(sh, 0.3)
    I2, I4, I4rt, I4s, II = f2.identity_tensors_2d(*sh)
    C4 = f2.build_C4_2d(E, nu, I4s, II, "plane_stress")
    F = I2.copy()
    F[0, 0] += 0.01
    F[0, 1] += 0.004
    P, K4, F33 = f2.constitutive_2d(F, C4, I2, I4, I4rt, plane_mode="plane_stress")
    lam3, mu = 70e9 * 0.3 / (1.3 * 0.4), 70e9 / 2.6
    Eg = 0.5 * (F[:, :, 0, 0].T @ F[:, :, 0, 0] - np.eye(2))
    E33 = 0.5 * (F33[0, 0] ** 2 - 1.0)
    assert abs(lam3 * (np.trace(Eg) + E33) + 2 * mu * E33) < 1e-9 * lam3 * abs(np.trace(Eg))
    S = lam3 * (np.trace(Eg) + E33) * np.eye(2) + 2 * mu * Eg
    assert np.allclose(P[:, :, 1, 2], F[:, :, 0, 0] @ S, rtol=1e-12)
    sig = f2.cauchy_2d(P, F, F33)[:, :, 0, 0]
    J = np.linalg.det(F[:, :, 0, 0]) * F33[0, 0]
    assert np.allclose(sig, F[:, :, 0, 0] @ S @ F[:, :, 0, 0].T / J, rtol=1e-12)


MURN = dict(A_m=-100e9, B_m=-100e9, C_m=-100e9)


@pytest.mark.parametrize("model,kw,mode", [("landau", LANDAU, "plane_strain"), ("murnaghan", MURN, "plane_strain"),
                                           ("neo_hookean", {}, "plane_strain"), ("landau", LANDAU, "plane_stress"),
                                           ("murnaghan", MURN, "plane_stress"), ("neo_hookean", {}, "plane_stress")])
def test_2d_tangent_is_the_derivative_of_P(model, kw, mode):
    # in plane stress P depends on F also through E33(F): the condensed
    # tangent is that total derivative
    sh = (3, 2)
    E, nu, t2, _, _ = _problem(sh, 7, plane_mode=mode)
    rng = np.random.default_rng(2)
    F = t2[0] + 0.02 * rng.standard_normal((2, 2) + sh)
    dF = rng.standard_normal((2, 2) + sh)
    P, K4, _ = f2.constitutive_2d(F, t2[3], *t2[:3], model_type=model, plane_mode=mode, **kw)
    h = 1e-6
    Pp = f2.constitutive_2d(F + h * dF, t2[3], *t2[:3], model_type=model, plane_mode=mode, need_tangent=False,
                            **kw)[0]
    Pm = f2.constitutive_2d(F - h * dF, t2[3], *t2[:3], model_type=model, plane_mode=mode, need_tangent=False,
                            **kw)[0]
    fd = (Pp - Pm) / (2 * h)
    lin = f2._trans2(f2._ddot42(K4, f2._trans2(dF)))
    assert np.abs(fd - lin).max() < 1e-6 * np.abs(lin).max()


@pytest.mark.parametrize("model,kw", [("landau", LANDAU), ("murnaghan", MURN)])
def test_2d_plane_stress_laws_are_the_3d_law_at_zero_s33(model, kw):
    # plane stress: with F_33 = sqrt(1 + 2 E33) the 3D oracle's law (its 3D
    # lam = E nu / ((1 + nu)(1 - 2 nu)) equals the one recovered from lam*)
    # gives P_33 = 0 and the same in-plane P; Cauchy's J carries F33
    sh = (4, 3)
    rng = np.random.default_rng(5)
    E, nu = 70e9 * rng.uniform(0.8, 1.2, sh), np.full(sh, 0.3)
    I2, I4, I4rt, I4s, II = f2.identity_tensors_2d(*sh)
    C4 = f2.build_C4_2d(E, nu, I4s, II, "plane_stress")
    F = I2 + 0.03 * rng.standard_normal((2, 2) + sh)
    P, _, F33 = f2.constitutive_2d(F, C4, I2, I4, I4rt, model_type=model, plane_mode="plane_stress",
                                   need_tangent=False, **kw)
    sh3 = sh + (1,)
    F3 = np.zeros((3, 3) + sh3)
    F3[:2, :2, ..., 0] = F
    F3[2, 2, ..., 0] = F33
    i2, i4, i4rt, i4s, ii = fs.identity_tensors(*sh3)
    C43 = fs.build_C4(E.reshape(sh3), nu.reshape(sh3), i4s, ii)
    P3 = fs.constitutive(F3, C43, i2, i4, i4rt, model_type=model, need_tangent=False, **kw)[0]
    scale = np.abs(P).max()
    assert np.abs(P3[2, 2]).max() < 1e-9 * scale
    assert np.abs(P3[:2, :2, ..., 0] - P).max() < 1e-9 * scale
    sig = f2.cauchy_2d(P, F, F33)
    sig3 = fs.cauchy_from_P(P3, F3)[:2, :2, ..., 0]
    assert np.abs(sig - sig3).max() < 1e-9 * np.abs(sig3).max()


def test_2d_plane_stress_neo_hookean_is_the_3d_law_at_zero_s33():
    # an independently written 3D neo-Hookean law, S = mu I + (-mu + lam ln J) C^-1
    # with J = det F3, at F_33 = sqrt(x): S_33 = 0 and the same in-plane P
    sh = (4, 3)
    rng = np.random.default_rng(6)
    E, nu = 70e9 * rng.uniform(0.8, 1.2, sh), np.full(sh, 0.3)
    I2, I4, I4rt, I4s, II = f2.identity_tensors_2d(*sh)
    C4 = f2.build_C4_2d(E, nu, I4s, II, "plane_stress")
    for amp in (0.03, 0.2):
        F = I2 + amp * rng.standard_normal((2, 2) + sh)
        P, _, F33 = f2.constitutive_2d(F, C4, I2, I4, I4rt, model_type="neo_hookean", plane_mode="plane_stress",
                                       need_tangent=False)
        mu = E / (2 * (1 + nu))
        lam3 = E * nu / ((1 + nu) * (1 - 2 * nu))
        for x in range(sh[0]):
            for y in range(sh[1]):
                F3 = np.zeros((3, 3))
                F3[:2, :2] = F[:, :, x, y]
                F3[2, 2] = F33[x, y]
                C = F3.T @ F3
                S3 = mu[x, y] * np.eye(3) + (-mu[x, y] + lam3[x, y] * np.log(np.linalg.det(F3))) * np.linalg.inv(C)
                P3 = F3 @ S3
                scale = np.abs(P[:, :, x, y]).max()
                assert abs(S3[2, 2]) < 1e-9 * mu[x, y]
                assert np.abs(P3[:2, :2] - P[:, :, x, y]).max() < 1e-9 * scale


@pytest.mark.parametrize("kind", ["piecewise", "smooth"])
def test_2d_plane_stress_capped_landau_zero_s33_and_uncapped_limit(kind):
    # capped plane stress: S_33 of the capped 3D law at F_33 = sqrt(1 + 2 E33)
    # vanishes to what the 1e-9 step tolerance of the linearly converging
    # iteration leaves; with the cap out of reach the uncapped root is found
    sh = (5, 4)
    rng = np.random.default_rng(8)
    E, nu = 70e9 * rng.uniform(0.8, 1.2, sh), np.full(sh, 0.3)
    I2, I4, I4rt, I4s, II = f2.identity_tensors_2d(*sh)
    C4 = f2.build_C4_2d(E, nu, I4s, II, "plane_stress")
    F = I2 + 0.03 * rng.standard_normal((2, 2) + sh)
    cap = dict(LANDAU, strain_capping_enabled=True, strain_capping_limit=0.012, strain_capping_type=kind,
               strain_capping_smooth_power=2.0)
    P, _, F33 = f2.constitutive_2d(F, C4, I2, I4, I4rt, model_type="landau", plane_mode="plane_stress",
                                   need_tangent=False, **cap)
    sh3 = sh + (1,)
    F3 = np.zeros((3, 3) + sh3)
    F3[:2, :2, ..., 0] = F
    F3[2, 2, ..., 0] = F33
    i2, i4, i4rt, i4s, ii = fs.identity_tensors(*sh3)
    C43 = fs.build_C4(E.reshape(sh3), nu.reshape(sh3), i4s, ii)
    P3 = fs.constitutive(F3, C43, i2, i4, i4rt, model_type="landau", need_tangent=False, **cap)[0]
    scale = np.abs(P).max()
    assert np.abs(P3[2, 2]).max() < 1e-6 * scale
    assert np.abs(P3[:2, :2, ..., 0] - P).max() < 1e-6 * scale
    # the cap was active somewhere (the uncapped stress differs)
    Pu = f2.constitutive_2d(F, C4, I2, I4, I4rt, model_type="landau", plane_mode="plane_stress",
                            need_tangent=False, **LANDAU)[0]
    assert np.abs(Pu - P).max() > 1e-3 * scale
    far = dict(cap, strain_capping_limit=10.0)
    Pf = f2.constitutive_2d(F, C4, I2, I4, I4rt, model_type="landau", plane_mode="plane_stress",
                            need_tangent=False, **far)[0]
    if kind == "piecewise":
        assert np.abs(Pf - Pu).max() < 1e-9 * scale


def test_2d_step_is_equilibrated_and_keeps_the_mean():
    sh = (10, 8)
    E, nu, t2, _, _ = _problem(sh, 9, spread=0.3)
    G = f2.build_ghat4_2d(*sh, *sh)
    Fbar = np.eye(2)
    Fbar[0, 0] = 1.01
    Pt, Pm = np.zeros((2, 2)), np.zeros((2, 2), bool)
    Pm[1, 1] = True
    F, P, S, _, Fb = f2.solver_step_2d(t2[0].copy(), Fbar, G, t2[3], *t2[:3], P_target=Pt, P_mask=Pm,
                                       E_avg=E.mean(), nu_avg=0.3, tol_NW=1e-7, tol_CG=1e-7, tol_macro=1e3)
    assert np.allclose(F.mean(axis=(2, 3)), Fb, atol=1e-14)
    assert np.abs(f2.project2d(P, G)).max() < 1e-6 * np.abs(P).max()
    assert abs(S.mean(axis=(2, 3))[1, 1]) < 1e3 and Fb[1, 1] < 1.0


def test_2d_neo_hookean_limits():
    # S = 0 at F = I; at small strain the law is linear plane-strain
    # elasticity (S = lam tr E I + 2 mu E to O(E^2)); under a pure rotation
    # P = R S(I) = 0
    sh = (3, 2)
    E, nu = np.full(sh, 70e9), np.full(sh, 0.3)
    I2, I4, I4rt, I4s, II = f2.identity_tensors_2d(*sh)
    C4 = f2.build_C4_2d(E, nu, I4s, II)
    lam, mu = 70e9 * 0.3 / (1.3 * 0.4), 70e9 / 2.6
    P0 = f2.constitutive_2d(I2.copy(), C4, I2, I4, I4rt, model_type="neo_hookean", need_tangent=False)[0]
    assert np.abs(P0).max() < 1e-3
    F = I2.copy()
    F[0, 0] += 1e-5
    F[0, 1] += 4e-6
    P = f2.constitutive_2d(F, C4, I2, I4, I4rt, model_type="neo_hookean", need_tangent=False)[0][:, :, 0, 0]
    Eg = 0.5 * (F[:, :, 0, 0].T @ F[:, :, 0, 0] - np.eye(2))
    lin = F[:, :, 0, 0] @ (lam * np.trace(Eg) * np.eye(2) + 2 * mu * Eg)
    assert np.abs(P - lin).max() < 1e-4 * np.abs(lin).max()
    th = 0.3
    R = np.array([[np.cos(th), -np.sin(th)], [np.sin(th), np.cos(th)]])
    Fr = np.einsum("ij,xy->ijxy", R, np.ones(sh))
    Pr = f2.constitutive_2d(Fr, C4, I2, I4, I4rt, model_type="neo_hookean", need_tangent=False)[0]
    assert np.abs(Pr).max() < 1e-2


def test_2d_dbfft_vs_newton_cg_reference_case():
    # tests/test_finite_strain_kmc.py:314-365 of the reference: 16^2, an 8^2
    # square at 68 GPa in 70 GPa, F_bar = diag(1.0005, 0.9995): relative F
    # difference < 1e-5, P < 1e-2; and the homogeneous closed form in two
    # DBFFT iterations
    sh = (16, 16)
    E = np.full(sh, 70.0e9)
    E[4:12, 4:12] = 68.0e9
    nu = np.full(sh, 0.3)
    I2, I4, I4rt, I4s, II = f2.identity_tensors_2d(*sh)
    C4 = f2.build_C4_2d(E, nu, I4s, II, "plane_strain")
    G = f2.build_ghat4_2d(*sh, *sh)
    Fbar = np.diag([1.0005, 0.9995])
    kw = dict(E_avg=E.mean(), nu_avg=0.3, tol_NW=1e-8, tol_CG=1e-9, max_NW=30)
    Fn, Pn, _, _, _ = f2.solver_step_2d(I2.copy(), Fbar, G, C4, I2, I4, I4rt, solver="newton_cg", **kw)
    Fd, Pd, _, _, _ = f2.solver_step_2d(I2.copy(), Fbar, G, C4, I2, I4, I4rt, solver="dbfft", **kw)
    assert np.linalg.norm(Fd - Fn) / np.linalg.norm(Fn) < 1e-5
    assert np.linalg.norm(Pd - Pn) / np.linalg.norm(Pn) < 1e-2
    C4h = f2.build_C4_2d(np.full(sh, 70e9), nu, I4s, II, "plane_strain")
    st = {}
    Fh, _, _, _, _ = f2.solver_step_2d(I2.copy(), Fbar, G, C4h, I2, I4, I4rt, solver="dbfft", stats=st)
    assert st["dbfft"] == [2] and np.abs(Fh - Fbar[:, :, None, None]).max() < 1e-15


@pytest.mark.parametrize("mode", ["plane_strain", "plane_stress"])
def test_2d_dbfft_reaches_the_newton_equilibrium(mode):
    # odd grid: the displacement and projection discretisations coincide, so
    # both 2D solvers end at one equilibrium
    sh = (7, 9)
    E, nu, t2, _, _ = _problem(sh, 6, mode)
    G = f2.build_ghat4_2d(*sh, *sh)
    Fbar = np.eye(2)
    Fbar[0, 1] = 0.02
    kw = dict(plane_mode=mode, tol_NW=1e-9, tol_CG=1e-9)
    Fn = f2.solver_step_2d(t2[0].copy(), Fbar, G, t2[3], *t2[:3], solver="newton_cg", **kw)[0]
    Fd = f2.solver_step_2d(t2[0].copy(), Fbar, G, t2[3], *t2[:3], solver="dbfft", **kw)[0]
    assert np.abs(Fd - Fn).max() < 1e-11


def test_2d_al_closed_form_and_heterogeneous_failure():
    # the AL update adds G of a stress to F: homogeneous O(1) moduli give the
    # closed form after two iterations, any heterogeneity diverges (ValueError)
    sh = (6, 4)
    nu = np.full(sh, 0.3)
    I2, I4, I4rt, I4s, II = f2.identity_tensors_2d(*sh)
    G = f2.build_ghat4_2d(*sh, *sh)
    Fbar = np.eye(2)
    Fbar[0, 1] = 0.02
    C4h = f2.build_C4_2d(np.full(sh, 1.0), nu, I4s, II)
    st = {}
    F, P, _, _, _ = f2.solver_step_2d(I2.copy(), Fbar, G, C4h, I2, I4, I4rt, solver="al", tol_NW=1e-12, stats=st)
    assert st["al"] == [2] and np.allclose(F, Fbar[:, :, None, None], atol=1e-14)
    C4 = f2.build_C4_2d(70e9 * np.random.default_rng(9).uniform(0.85, 1.15, sh), nu, I4s, II)
    with pytest.raises(ValueError, match="AL solver did not converge or exploded"):
        f2.solver_step_2d(I2.copy(), Fbar, G, C4, I2, I4, I4rt, solver="al")


def test_2d_simulation_driver_meets_the_reference_tests_assertions():
    # the reference's plane-strain Landau AL simulation
    # (tests/test_landau_finite_strain.py:4-34): shapes and the driven F_xx;
    # plane strain keeps F_yy = 1 (no mixed target), so sig_yy > 0 from the
    # Poisson coupling
    L = dict(v1=-100.0e9, v2=-50.0e9, v3=-10.0e9, g1=-1000.0e9, g2=500.0e9, g3=-100.0e9, g4=-200.0e9)
    E = np.full((8, 8), 70.0e9)
    nu = np.full((8, 8), 0.3)
    Fm, Sm, Pm = f2.simulation_2d(E, nu, (0, 0), 0.01, 5, plane_mode="plane_strain", model_type="landau",
                                  solver="al", **L)
    assert Fm.shape == Sm.shape == Pm.shape == (6, 2, 2)
    np.testing.assert_allclose(Fm[-1, 0, 0], 1.01, atol=1e-5)
    np.testing.assert_allclose(Fm[:, 1, 1], 1.0, atol=1e-12)
    assert Sm[-1, 1, 1] > 0 and Sm[-1, 0, 0] > Sm[-1, 1, 1]


def test_2d_plane_stress_landau_simulation_is_the_3d_uniaxial_one():
    # the reference's plane-stress Landau AL case (tests/test_landau_finite_strain.py:36-67):
    # its assertions, and on a homogeneous grid uniaxial plane stress is the
    # 3D uniaxial state (sigma_yy = sigma_zz = 0), so sigma_xx matches the 3D
    # driver's to the macro tolerance
    L = dict(v1=-100.0e9, v2=-50.0e9, v3=-10.0e9, g1=-1000.0e9, g2=500.0e9, g3=-100.0e9, g4=-200.0e9)
    Fm, Sm, _ = f2.simulation_2d(np.full((8, 8), 70e9), np.full((8, 8), 0.3), (0, 0), 0.01, 5,
                                 mixed_targets={(1, 1): 0.0}, plane_mode="plane_stress", model_type="landau",
                                 solver="al", tol_macro=1e5, **L)
    np.testing.assert_allclose(Fm[-1, 0, 0], 1.01, atol=1e-5)
    np.testing.assert_allclose(Sm[-1, 1, 1], 0.0, atol=2e5)
    F3, S3, _ = fs.simulation(np.full((4, 4, 4), 70e9), np.full((4, 4, 4), 0.3), (0, 0), 0.01, 5,
                              mixed_targets={(1, 1): 0.0, (2, 2): 0.0}, model_type="landau", solver="al",
                              tol_macro=1e5, **L)
    assert abs(Sm[-1, 0, 0] - S3[-1, 0, 0]) < 1e-4 * S3[-1, 0, 0]
    assert abs(Fm[-1, 1, 1] - F3[-1, 1, 1]) < 1e-5


# the reference's capped plane-stress case (tests/test_finite_strain_capped_plane_stress.py:19-116):
# its Landau constants, moduli and production capping; S33 of the capped law at the
# returned E33 must stay below 1e3 Pa (1e-9 on E33 times O(100 GPa))
REF_CAP_LANDAU = dict(v1=-97.4e9, v2=-67.0e9, v3=-19.1e9, g1=5472.5e9, g2=-1209.0e9, g3=348.7e9, g4=-238.3e9)
REF_CAP_LAM, REF_CAP_MU = 80.9e9, 23.75e9
REF_CAPPING = dict(strain_capping_enabled=True, strain_capping_limit=0.10, strain_capping_tangent_ratio=0.1,
                   strain_capping_smooth_power=1.0)


def capped_s33(E2d, E33, cap, kind):
    """S33 of the capped landau law at diag-block(E_2d, E33), written from the law:
    deviator scaled toward E_cap (smooth: tanh map), invariants of the capped strain,
    plus the residual-stiffness term 2 G_t (E33 - c33) w."""
    lam, mu, c = REF_CAP_LAM, REF_CAP_MU, REF_CAP_LANDAU
    E = np.zeros(E33.shape + (3, 3))
    E[..., 0, 0], E[..., 1, 1], E[..., 0, 1], E[..., 1, 0] = E2d[0, 0], E2d[1, 1], E2d[0, 1], E2d[1, 0]
    E[..., 2, 2] = E33
    tr = np.trace(E, axis1=-2, axis2=-1)
    dev = E - tr[..., None, None] / 3.0 * np.eye(3)
    eq = np.sqrt(np.maximum(0.0, 2.0 / 3.0 * np.einsum("...ij,...ij->...", dev, dev)))
    Ecap = cap["strain_capping_limit"]
    if kind == "smooth":
        p = cap["strain_capping_smooth_power"]
        r = np.where(eq > 1e-12, eq, 1.0) / Ecap
        f = np.where(eq > 1e-12, np.tanh(np.minimum(100.0, r ** p)) ** (1.0 / p) / r, 1.0)
        w = np.tanh(np.minimum(100.0, (eq / Ecap) ** p)) ** 2
    else:
        f = np.where(eq > Ecap, Ecap / np.maximum(eq, 1e-300), 1.0)
        w = (eq > Ecap).astype(float)
    Ec = dev * f[..., None, None] + tr[..., None, None] / 3.0 * np.eye(3)
    I1 = np.trace(Ec, axis1=-2, axis2=-1)
    E2 = Ec @ Ec
    I2 = np.trace(E2, axis1=-2, axis2=-1)
    I3 = np.trace(E2 @ Ec, axis1=-2, axis2=-1)
    A = (lam * I1 + 0.5 * c["v1"] * I1 ** 2 + c["v2"] * I2 + c["g1"] * I1 ** 3 / 6.0 + c["g2"] * I1 * I2
         + 4.0 / 3.0 * c["g3"] * I3)
    B = 2.0 * (mu + c["v2"] * I1 + 0.5 * c["g2"] * I1 ** 2 + c["g4"] * I2)
    C = 4.0 * (c["v3"] + c["g3"] * I1)
    c33 = Ec[..., 2, 2]
    return A + B * c33 + C * c33 ** 2 + 2.0 * cap["strain_capping_tangent_ratio"] * mu * (E33 - c33) * w


def reference_capped_case(kind, amp, n=8):
    rng = np.random.default_rng(3)
    E = REF_CAP_MU * (3 * REF_CAP_LAM + 2 * REF_CAP_MU) / (REF_CAP_LAM + REF_CAP_MU)
    nu = REF_CAP_LAM / (2 * (REF_CAP_LAM + REF_CAP_MU))
    I2, I4, I4rt, I4s, II = f2.identity_tensors_2d(n, n)
    C4 = f2.build_C4_2d(np.full((n, n), E), np.full((n, n), nu), I4s, II, "plane_stress")
    F = I2.copy()
    F[0, 0] += amp * rng.random((n, n))
    F[1, 1] -= 0.3 * amp * rng.random((n, n))
    F[0, 1] += 0.3 * amp * (2 * rng.random((n, n)) - 1)
    return F, C4, (I2, I4, I4rt), dict(REF_CAPPING, strain_capping_type=kind)


@pytest.mark.parametrize("kind", ["smooth", "piecewise"])
@pytest.mark.parametrize("amp", [0.05, 0.15])
def test_2d_capped_plane_stress_reference_case(kind, amp):
    F, C4, t, cap = reference_capped_case(kind, amp)
    P, K4, F33 = f2.constitutive_2d(F, C4, *t, model_type="landau", plane_mode="plane_stress", **REF_CAP_LANDAU,
                                    **cap)
    E33 = 0.5 * (F33 ** 2 - 1.0)
    E2d = 0.5 * (np.einsum("jixy,jkxy->ikxy", F, F) - t[0])
    assert np.abs(capped_s33(E2d, E33, cap, kind)).max() < 1e3
