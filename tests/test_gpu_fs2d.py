"""GPU parity of the 2D finite-strain step (finite_strain_solver_step_2d,
newton_cg) with the 2D oracle: svk in both plane modes, landau (plain,
capped, with Fp) and murnaghan in plane strain, mixed BCs."""
import numpy as np
import pytest

from oracle import finite_strain2d as f2

pytestmark = pytest.mark.gpu

LANDAU = dict(v1=-236.5e9, v2=-27.4e9, v3=11.8e9, g1=-5640.7e9, g2=2207.1e9, g3=-332.5e9, g4=-305.2e9)
# capped landau with a limit the 2% shear reaches (piecewise and smooth maps)
CAPP = dict(LANDAU, strain_capping_enabled=True, strain_capping_limit=0.012)
CAPS = dict(LANDAU, strain_capping_enabled=True, strain_capping_limit=0.012, strain_capping_type="smooth",
            strain_capping_smooth_power=2.0)


def _rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.mark.parametrize("sh,model,kw,fp,mode", [
    ((8, 6), "svk", {}, False, "plane_strain"), ((8, 6), "svk", {}, True, "plane_strain"),
    ((8, 6), "svk", {}, False, "plane_stress"), ((8, 6), "landau", LANDAU, False, "plane_strain"),
    ((8, 6), "landau", LANDAU, True, "plane_strain"),
    ((12, 10), "landau", dict(LANDAU, strain_capping_enabled=True, strain_capping_type="smooth"), False,
     "plane_strain"),
    ((8, 6), "murnaghan", dict(A_m=-100e9, B_m=-100e9, C_m=-100e9), False, "plane_strain"),
    ((64, 48), "svk", {}, False, "plane_stress"), ((8, 6), "neo_hookean", {}, False, "plane_strain"),
    ((10, 12), "neo_hookean", {}, True, "plane_strain"), ((8, 6), "landau", LANDAU, False, "plane_stress"),
    ((10, 12), "landau", LANDAU, True, "plane_stress"),
    ((8, 6), "murnaghan", dict(A_m=-100e9, B_m=-100e9, C_m=-100e9), False, "plane_stress"),
    ((64, 48), "landau", LANDAU, False, "plane_stress"),
    ((8, 6), "neo_hookean", {}, False, "plane_stress"), ((10, 12), "neo_hookean", {}, True, "plane_stress"),
    ((12, 10), "landau", CAPP, False, "plane_stress"), ((10, 12), "landau", CAPS, True, "plane_stress"),
    ((64, 48), "landau", CAPP, False, "plane_stress")])
def test_2d_step_parity(gpu, sh, model, kw, fp, mode):
    from mgkmc_amd.finite_strain_simulator import finite_strain_solver_step_2d
    rng = np.random.default_rng(4)
    E = 70e9 * rng.uniform(0.85, 1.15, sh)
    nu = np.full(sh, 0.3)
    I2, I4, I4rt, I4s, II = f2.identity_tensors_2d(*sh)
    C4 = f2.build_C4_2d(E, nu, I4s, II, mode)
    G = f2.build_ghat4_2d(*sh, *sh)
    Fbar = np.eye(2)
    Fbar[0, 1] = 0.02
    Fbar[1, 1] = 0.996
    Fp = None
    if fp:
        Fp = I2 + 0.01 * np.random.default_rng(5).standard_normal((2, 2) + sh)
    print(f"[fs2d] {sh} {model} {mode} Fp={fp}", flush=True)
    st = {}
    args = dict(Fp=Fp, model_type=model, plane_mode=mode, tol_NW=1e-8, tol_CG=1e-8, **kw)
    try:
        ref = f2.solver_step_2d(I2.copy(), Fbar, G, C4, I2, I4, I4rt, stats=st, **args)
    except ValueError as e:
        # the reference's own failure (the capped law's approximate tangent on the
        # 64 x 48 grid converges too slowly for its 50 Newton steps): the device
        # step fails the same way
        print(f"[fs2d] oracle: {e}", flush=True)
        with pytest.raises(ValueError, match=str(e)[:30]):
            finite_strain_solver_step_2d(I2.copy(), Fbar, G, C4, I2, I4, I4rt, solver="newton_cg", **args)
        return
    got = finite_strain_solver_step_2d(I2.copy(), Fbar, G, C4, I2, I4, I4rt, solver="newton_cg", **args)
    eF, eP, eS = _rel(got[0], ref[0]), _rel(got[1], ref[1]), _rel(got[2], ref[2])
    print(f"[fs2d] newton={st['newton'][-1]} relF={eF:.2e} relP={eP:.2e} relSig={eS:.2e}", flush=True)
    assert eF < 1e-12 and eP < 1e-10 and eS < 1e-10


def test_2d_mixed_bc_parity(gpu):
    from mgkmc_amd.finite_strain_simulator import finite_strain_solver_step_2d
    sh = (10, 8)
    rng = np.random.default_rng(9)
    E = 70e9 * rng.uniform(0.7, 1.3, sh)
    nu = np.full(sh, 0.3)
    I2, I4, I4rt, I4s, II = f2.identity_tensors_2d(*sh)
    for mode in ("plane_strain", "plane_stress"):
        C4 = f2.build_C4_2d(E, nu, I4s, II, mode)
        G = f2.build_ghat4_2d(*sh, *sh)
        Fbar = np.eye(2)
        Fbar[0, 0] = 1.01
        Pt, Pm = np.zeros((2, 2)), np.zeros((2, 2), bool)
        Pm[1, 1] = True
        # tolerances the reference's own loop completes at (at 1e-7 its plane-stress
        # macro loop ends in a line-search failure, which the oracle reproduces)
        kw = dict(P_target=Pt, P_mask=Pm, E_avg=E.mean(), nu_avg=0.3, tol_NW=1e-6, tol_CG=1e-6, tol_macro=1e3,
                  plane_mode=mode)
        ref = f2.solver_step_2d(I2.copy(), Fbar, G, C4, I2, I4, I4rt, **kw)
        got = finite_strain_solver_step_2d(I2.copy(), Fbar, G, C4, I2, I4, I4rt, solver="newton_cg", **kw)
        print(f"[fs2d mixed] {mode} Fbar ref={np.diag(ref[4])} got={np.diag(got[4])}", flush=True)
        assert np.abs(got[4] - ref[4]).max() < 1e-12
        assert _rel(got[0], ref[0]) < 1e-11 and _rel(got[2], ref[2]) < 1e-10


def test_2d_unsupported_solver_raises(gpu):
    from mgkmc_amd.finite_strain_simulator import finite_strain_solver_step_2d
    sh = (4, 4)
    I2, I4, I4rt, I4s, II = f2.identity_tensors_2d(*sh)
    C4 = f2.build_C4_2d(np.full(sh, 70e9), np.full(sh, 0.3), I4s, II)
    with pytest.raises(NotImplementedError):
        finite_strain_solver_step_2d(I2.copy(), np.eye(2), None, C4, I2, I4, I4rt, solver="bogus")


def test_2d_capped_plane_stress_iteration_count_is_the_grids(gpu):
    # the capped E33 Newton takes the grid's iteration count in every voxel:
    # the device law through mgk_fs_constitutive equals the oracle's law (P and
    # the tangent action) on a grid where capped and uncapped voxels converge
    # at different rates
    from mgkmc_amd.finite_strain_simulator import constitutive_hyperelastic_2d
    sh = (9, 7)
    rng = np.random.default_rng(12)
    E = 70e9 * rng.uniform(0.8, 1.2, sh)
    nu = np.full(sh, 0.3)
    I2, I4, I4rt, I4s, II = f2.identity_tensors_2d(*sh)
    C4 = f2.build_C4_2d(E, nu, I4s, II, "plane_stress")
    F = I2 + 0.03 * rng.standard_normal((2, 2) + sh)
    for kw in (CAPP, CAPS):
        P, K4, F33 = f2.constitutive_2d(F, C4, I2, I4, I4rt, model_type="landau", plane_mode="plane_stress", **kw)
        Pg, K4g, F33g = constitutive_hyperelastic_2d(F, C4, I2, I4, I4rt, model_type="landau",
                                                     plane_mode="plane_stress", **kw)
        assert _rel(Pg, P) < 1e-12 and _rel(F33g, F33) < 1e-13
        assert _rel(K4g, K4) < 1e-10


@pytest.mark.parametrize("model,mode,kw,drv,targets", [
    ("svk", "plane_stress", {}, (0, 0), {(1, 1): 0.0}),
    ("landau", "plane_strain", dict(LANDAU, strain_capping_enabled=True), (0, 1), {(0, 0): 0.0, (1, 1): 0.0}),
    ("landau", "plane_stress", LANDAU, (0, 1), {(0, 0): 0.0, (1, 1): 0.0}),
    ("murnaghan", "plane_stress", dict(A_m=-100e9, B_m=-100e9, C_m=-100e9), (0, 1), {(0, 0): 0.0, (1, 1): 0.0}),
    ("neo_hookean", "plane_stress", {}, (0, 0), {(1, 1): 0.0}),
    ("landau", "plane_stress", CAPP, (0, 1), {(0, 0): 0.0, (1, 1): 0.0})])
def test_2d_simulation_parity(gpu, model, mode, kw, drv, targets):
    # finite_strain_simulation_2d: the macro F, Cauchy and P lists of every
    # load step equal the oracle driver's
    from mgkmc_amd.finite_strain_simulator import finite_strain_simulation_2d
    sh = (16, 12)
    rng = np.random.default_rng(2)
    E = 70e9 * rng.uniform(0.8, 1.2, sh)
    nu = np.full(sh, 0.3)
    args = dict(mixed_targets=targets, plane_mode=mode, tol_NW=1e-6, tol_CG=1e-6, tol_macro=1e4, model_type=model,
                **kw)
    ref = f2.simulation_2d(E, nu, drv, 0.01, 3, **args)
    got = finite_strain_simulation_2d(E, nu, drv, 0.01, 3, solver="newton_cg", store=False, **args)
    print(f"[fs2d sim] {model} {mode} F_macro ref={ref[0][-1].ravel()} got={got[0][-1].ravel()}", flush=True)
    for a, b in zip(got[:3], ref):
        assert np.abs(a - b).max() <= 1e-10 * max(1.0, np.abs(b).max())


@pytest.mark.parametrize("sh,model,mode,fp", [((8, 6), "svk", "plane_strain", False),
                                              ((8, 6), "svk", "plane_stress", True),
                                              ((5, 9), "landau", "plane_strain", False),
                                              ((10, 12), "neo_hookean", "plane_strain", True),
                                              ((8, 6), "landau", "plane_stress", False),
                                              ((10, 12), "landau", "plane_stress", True),
                                              ((8, 6), "neo_hookean", "plane_stress", False),
                                              ((10, 12), "neo_hookean", "plane_stress", True)])
def test_2d_dbfft_parity(gpu, sh, model, mode, fp):
    # solver="dbfft" in 2D (the displacement step on the (nx, 1, ny) grid, its
    # stop test over the in-plane |P|): F, P, sigma of the oracle
    from mgkmc_amd.finite_strain_simulator import finite_strain_solver_step_2d
    rng = np.random.default_rng(4)
    E = 70e9 * rng.uniform(0.85, 1.15, sh)
    nu = np.full(sh, 0.3)
    I2, I4, I4rt, I4s, II = f2.identity_tensors_2d(*sh)
    C4 = f2.build_C4_2d(E, nu, I4s, II, mode)
    G = f2.build_ghat4_2d(*sh, *sh)
    Fbar = np.eye(2)
    Fbar[0, 1] = 0.02
    Fbar[1, 1] = 0.996
    Fp = I2 + 0.01 * np.random.default_rng(5).standard_normal((2, 2) + sh) if fp else None
    kw = dict(LANDAU) if model == "landau" else {}
    print(f"[fs2d dbfft] {sh} {model} {mode} Fp={fp}", flush=True)
    st = {}
    ref = f2.solver_step_2d(I2.copy(), Fbar, G, C4, I2, I4, I4rt, Fp=Fp, model_type=model, plane_mode=mode,
                            tol_NW=1e-8, tol_CG=1e-8, solver="dbfft", stats=st, **kw)
    got = finite_strain_solver_step_2d(I2.copy(), Fbar, G, C4, I2, I4, I4rt, Fp=Fp, model_type=model,
                                       plane_mode=mode, tol_NW=1e-8, tol_CG=1e-8, solver="dbfft", **kw)
    eF, eP, eS = _rel(got[0], ref[0]), _rel(got[1], ref[1]), _rel(got[2], ref[2])
    print(f"[fs2d dbfft] oracle newton={st['dbfft']} krylov={st['dbfft_krylov']} relF={eF:.2e} relP={eP:.2e} "
          f"relSig={eS:.2e}", flush=True)
    assert eF < 1e-12 and eP < 1e-10 and eS < 1e-10


def test_2d_dbfft_vs_newton_cg_reference_case(gpu):
    # tests/test_finite_strain_kmc.py:314-365 of the reference on the device
    from mgkmc_amd.finite_strain_simulator import finite_strain_solver_step_2d
    sh = (16, 16)
    E = np.full(sh, 70.0e9)
    E[4:12, 4:12] = 68.0e9
    nu = np.full(sh, 0.3)
    I2, I4, I4rt, I4s, II = f2.identity_tensors_2d(*sh)
    C4 = f2.build_C4_2d(E, nu, I4s, II, "plane_strain")
    Fbar = np.diag([1.0005, 0.9995])
    kw = dict(E_avg=E.mean(), nu_avg=0.3, tol_NW=1e-8, tol_CG=1e-9, max_NW=30)
    Fn, Pn = finite_strain_solver_step_2d(I2.copy(), Fbar, None, C4, I2, I4, I4rt, solver="newton_cg", **kw)[:2]
    Fd, Pd = finite_strain_solver_step_2d(I2.copy(), Fbar, None, C4, I2, I4, I4rt, solver="dbfft", **kw)[:2]
    dF = np.linalg.norm(Fd - Fn) / np.linalg.norm(Fn)
    dP = np.linalg.norm(Pd - Pn) / np.linalg.norm(Pn)
    print(f"[fs2d dbfft vs ncg] F {dF:.3e} P {dP:.3e}", flush=True)
    assert dF < 1e-5 and dP < 1e-2


@pytest.mark.parametrize("mode", ["plane_strain", "plane_stress"])
def test_2d_al_matches_oracle(gpu, mode):
    # solver="al" in 2D: the homogeneous closed form as the oracle (in-plane
    # residual), and the reference's ValueError on a heterogeneous grid
    from mgkmc_amd.finite_strain_simulator import finite_strain_solver_step_2d
    sh = (6, 4)
    nu = np.full(sh, 0.3)
    I2, I4, I4rt, I4s, II = f2.identity_tensors_2d(*sh)
    G = f2.build_ghat4_2d(*sh, *sh)
    Fbar = np.eye(2)
    Fbar[0, 1] = 0.02
    C4h = f2.build_C4_2d(np.full(sh, 1.0), nu, I4s, II, mode)
    ref = f2.solver_step_2d(I2.copy(), Fbar, G, C4h, I2, I4, I4rt, solver="al", plane_mode=mode, tol_NW=1e-12)
    got = finite_strain_solver_step_2d(I2.copy(), Fbar, G, C4h, I2, I4, I4rt, solver="al", plane_mode=mode,
                                       tol_NW=1e-12)
    assert _rel(got[0], ref[0]) < 1e-14 and _rel(got[1], ref[1]) < 1e-12
    C4 = f2.build_C4_2d(70e9 * np.random.default_rng(9).uniform(0.85, 1.15, sh), nu, I4s, II, mode)
    for fn in (f2.solver_step_2d, finite_strain_solver_step_2d):
        with pytest.raises(ValueError, match="AL solver did not converge or exploded"):
            fn(I2.copy(), Fbar, G, C4, I2, I4, I4rt, solver="al", plane_mode=mode)


def test_2d_reference_landau_al_case(gpu):
    # the reference's plane-strain Landau AL simulation
    # (tests/test_landau_finite_strain.py:4-34): parity with the oracle driver
    # and the reference test's own assertions
    from mgkmc_amd.finite_strain_simulator import finite_strain_simulation_2d
    L = dict(v1=-100.0e9, v2=-50.0e9, v3=-10.0e9, g1=-1000.0e9, g2=500.0e9, g3=-100.0e9, g4=-200.0e9)
    E = np.full((8, 8), 70.0e9)
    nu = np.full((8, 8), 0.3)
    ref = f2.simulation_2d(E, nu, (0, 0), 0.01, 5, plane_mode="plane_strain", model_type="landau", solver="al", **L)
    got = finite_strain_simulation_2d(E, nu, (0, 0), 0.01, 5, plane_mode="plane_strain", model_type="landau",
                                      solver="al", enable_console=False, **L)
    for a, b in zip(got[:3], ref):
        assert a.shape == (6, 2, 2)
        assert np.abs(a - b).max() <= 1e-10 * max(1.0, np.abs(b).max())
    np.testing.assert_allclose(got[0][-1, 0, 0], 1.01, atol=1e-5)


def test_2d_reference_landau_plane_stress_al_case(gpu):
    # the reference's plane-stress Landau AL simulation with sigma_yy = 0
    # (tests/test_landau_finite_strain.py:36-67): parity with the oracle driver
    # and the reference test's own assertions
    from mgkmc_amd.finite_strain_simulator import finite_strain_simulation_2d
    L = dict(v1=-100.0e9, v2=-50.0e9, v3=-10.0e9, g1=-1000.0e9, g2=500.0e9, g3=-100.0e9, g4=-200.0e9)
    E = np.full((8, 8), 70.0e9)
    nu = np.full((8, 8), 0.3)
    args = dict(mixed_targets={(1, 1): 0.0}, plane_mode="plane_stress", model_type="landau", solver="al",
                tol_macro=1e5, **L)
    ref = f2.simulation_2d(E, nu, (0, 0), 0.01, 5, **args)
    got = finite_strain_simulation_2d(E, nu, (0, 0), 0.01, 5, enable_console=False, **args)
    for a, b in zip(got[:3], ref):
        assert a.shape == (6, 2, 2)
        assert np.abs(a - b).max() <= 1e-10 * max(1.0, np.abs(b).max())
    np.testing.assert_allclose(got[0][-1, 0, 0], 1.01, atol=1e-5)
    np.testing.assert_allclose(got[1][-1, 1, 1], 0.0, atol=2e5)


def test_2d_per_voxel_landau_constants_parity(gpu):
    # the 2D law with landau constants as grid arrays: pointwise (P, K4, F33) and
    # the plane-stress Newton-CG step, uncapped and capped
    from mgkmc_amd.finite_strain_simulator import finite_strain_solver_step_2d, constitutive_hyperelastic_2d
    sh = (12, 10)
    rng = np.random.default_rng(23)
    E = 70e9 * rng.uniform(0.85, 1.15, sh)
    nu = np.full(sh, 0.3)
    per = {k: v * rng.uniform(0.9, 1.1, sh) for k, v in LANDAU.items()}
    I2, I4, I4rt, I4s, II = f2.identity_tensors_2d(*sh)
    G = f2.build_ghat4_2d(*sh, *sh)
    Fbar = np.eye(2)
    Fbar[0, 1] = 0.02
    Fbar[1, 1] = 0.996
    for mode, cap in (("plane_strain", {}), ("plane_stress", {}),
                      ("plane_stress", dict(strain_capping_enabled=True, strain_capping_limit=0.012))):
        C4 = f2.build_C4_2d(E, nu, I4s, II, mode)
        F = I2 + 0.03 * rng.standard_normal((2, 2) + sh)
        P, K4, F33 = f2.constitutive_2d(F, C4, I2, I4, I4rt, model_type="landau", plane_mode=mode, **per, **cap)
        Pg, K4g, F33g = constitutive_hyperelastic_2d(F, C4, I2, I4, I4rt, model_type="landau", plane_mode=mode,
                                                     **per, **cap)
        assert _rel(Pg, P) < 1e-12 and _rel(K4g, K4) < 1e-10 and _rel(F33g, F33) < 1e-13
        args = dict(model_type="landau", plane_mode=mode, tol_NW=1e-8, tol_CG=1e-8, **per, **cap)
        ref = f2.solver_step_2d(I2.copy(), Fbar, G, C4, I2, I4, I4rt, **args)
        got = finite_strain_solver_step_2d(I2.copy(), Fbar, G, C4, I2, I4, I4rt, solver="newton_cg", **args)
        eF, eP, eS = _rel(got[0], ref[0]), _rel(got[1], ref[1]), _rel(got[2], ref[2])
        print(f"[fs2d per-voxel] {mode} cap={bool(cap)} relF={eF:.2e} relP={eP:.2e} relSig={eS:.2e}", flush=True)
        assert eF < 1e-12 and eP < 1e-10 and eS < 1e-10


@pytest.mark.parametrize("kind", ["smooth", "piecewise"])
@pytest.mark.parametrize("amp", [0.05, 0.15])
def test_2d_capped_plane_stress_reference_case_on_device(gpu, kind, amp):
    # the reference's own capped plane-stress check (tests/test_finite_strain_capped_plane_stress.py:
    # 63-91) on the device law: S33 of the capped law at the returned E33 below 1e3 Pa, and
    # P, K4, F33 equal to the oracle's
    from mgkmc_amd.finite_strain_simulator import constitutive_hyperelastic_2d
    from test_oracle_fs2d import REF_CAP_LANDAU, capped_s33, reference_capped_case
    F, C4, t, cap = reference_capped_case(kind, amp)
    P, K4, F33 = constitutive_hyperelastic_2d(F, C4, *t, model_type="landau", plane_mode="plane_stress",
                                              **REF_CAP_LANDAU, **cap)
    E33 = 0.5 * (F33 ** 2 - 1.0)
    E2d = 0.5 * (np.einsum("jixy,jkxy->ikxy", F, F) - t[0])
    s33 = np.abs(capped_s33(E2d, E33, cap, kind)).max()
    print(f"[fs2d ref capped] {kind} amp={amp}: max|S33| = {s33:.3e} Pa", flush=True)
    assert s33 < 1e3
    Pr, K4r, F33r = f2.constitutive_2d(F, C4, *t, model_type="landau", plane_mode="plane_stress",
                                       **REF_CAP_LANDAU, **cap)
    assert _rel(P, Pr) < 1e-12 and _rel(F33, F33r) < 1e-13 and _rel(K4, K4r) < 1e-10
