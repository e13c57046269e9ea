"""Pin the LS oracle (oracle/ls.py) against facts that do not come from its
own formulas: the reference's KATs, an exact laminate solution, and the
equilibrium / compatibility conditions that fix the converged field."""
import numpy as np
import pytest

from oracle import ls as ol


def _rng(seed=0):
    return np.random.default_rng(seed)


def test_homogeneous_kat_sigma_xx():
    # reference tests/test_barrier_physics.py:49-98: 16x16x1, E=70 GPa, nu=0.3,
    # eps_xx = 1% => sigma_xx = (lam+2mu)*0.01 = 0.94231 GPa
    shape = (16, 16, 1)
    E = np.full(shape, 70e9)
    nu = np.full(shape, 0.3)
    eb = np.zeros((3, 3))
    eb[0, 0] = 0.01
    info = {}
    eps, sig, em, sm = ol.spectral_solver_3d(E, nu, eb, max_iter=50, tol=1e-6, info=info)
    lam = 70e9 * 0.3 / (1.3 * 0.4)
    mu = 70e9 / 2.6
    assert abs(sm[0, 0] - (lam + 2 * mu) * 0.01) < 1e-6 * abs(sm[0, 0])
    assert abs(sm[0, 0] / 1e9 - 0.9423077) < 1e-6
    assert np.allclose(eps, eb)
    assert info["iters"] == 1


def test_eshelby_sign_kat():
    # reference tests/test_plasticity.py:10-61: positive plastic shear at one
    # voxel => negative residual shear stress there
    shape = (16, 16, 1)
    E = np.full(shape, 70e9)
    nu = np.full(shape, 0.3)
    ep = np.zeros(shape + (3, 3))
    ep[8, 8, 0, 0, 1] = ep[8, 8, 0, 1, 0] = 0.05
    eps, sig, _, _ = ol.spectral_solver_3d(E, nu, np.zeros((3, 3)), eps_plastic=ep)
    assert sig[8, 8, 0, 0, 1] < 0
    assert 0 < eps[8, 8, 0, 0, 1] < 0.05


@pytest.mark.parametrize("d", [2, 3])
def test_laminate_exact(d):
    # layers normal to x: sigma_xx uniform, eps_xx(x) = S/(lam+2mu)(x),
    # transverse strains uniform, S fixed by <eps_xx> = e
    rng = _rng(1)
    n = 16
    shape = (n,) * d
    Ex = 70e9 * rng.uniform(0.5, 1.5, n)
    E = np.broadcast_to(Ex.reshape((n,) + (1,) * (d - 1)), shape).copy()
    nu = np.full(shape, 0.3)
    e = 0.01
    eb = np.zeros((d, d))
    eb[0, 0] = e
    if d == 3:
        eps, sig, em, sm = ol.spectral_solver_3d(E, nu, eb, tol=1e-14, max_iter=2000)
        lam, mu = ol.compute_lame_3d(Ex, 0.3)
    else:
        eps, sig, em, sm = ol.spectral_solver_2d(E, nu, eb, tol=1e-14, max_iter=2000)
        lam, mu = ol.compute_lame_2d(Ex, 0.3)
    M = lam + 2 * mu
    S = e / np.mean(1.0 / M)
    exx = S / M
    got = eps[(slice(None),) + (0,) * (d - 1) + (0, 0)]
    assert np.allclose(got, exx, rtol=1e-9, atol=0)
    assert np.allclose(sig[..., 0, 0], S, rtol=1e-9)


@pytest.mark.parametrize("shape", [(8, 8), (6, 5, 4), (4, 4, 4)])
def test_converged_field_is_equilibrated_and_compatible(shape):
    # the LS solution is the unique field with mean eps_bar, compatible eps and
    # divergence-free sigma; check both conditions in Fourier space
    d = len(shape)
    rng = _rng(2)
    E = 70e9 * rng.uniform(0.8, 1.2, shape)
    nu = np.full(shape, 0.35)
    g = rng.standard_normal(shape + (d, d)) * 1e-3
    ep = 0.5 * (g + np.swapaxes(g, -1, -2))
    eb = np.eye(d) * 0.002
    eb[0, 1] = eb[1, 0] = 0.001
    solve = ol.spectral_solver_3d if d == 3 else ol.spectral_solver_2d
    eps, sig, em, sm = solve(E, nu, eb, eps_plastic=ep, tol=1e-14, max_iter=5000)
    assert np.allclose(em, eb, atol=1e-15)
    # equilibrium: residual relative to stress magnitude
    assert ol.equilibrium_residual(sig) < 1e-10
    # compatibility: for k != 0 (non-Nyquist), eps_hat = sym(k (x) u) for some u,
    # i.e. the projection onto compatible fields leaves it unchanged
    q = ol.wave_vectors(shape)
    eh = np.stack([np.stack([np.fft.fftn(eps[..., i, j]) for j in range(d)], -1) for i in range(d)], -2)
    k2 = sum(qi * qi for qi in q)
    for idx in np.ndindex(shape):
        if all(i == 0 for i in idx):
            continue
        if any(n % 2 == 0 and i == n // 2 for i, n in zip(idx, shape)):
            continue
        kv = np.array([qi[idx] for qi in q])
        E_ = eh[idx]
        # u = 2 E k / |k|^2 - (k.E.k) k / |k|^4 (unique u with sym(k u) = E when compatible)
        u = (2 * E_ @ kv) / k2[idx] - (kv @ E_ @ kv) * kv / k2[idx] ** 2
        rec = 0.5 * (np.outer(kv, u) + np.outer(u, kv))
        assert np.abs(rec - E_).max() <= 1e-9 * np.abs(eh).max()


@pytest.mark.parametrize("d", [2, 3])
def test_gamma_projector(d):
    # Gamma : C0 : Gamma = Gamma and Gamma(0) = 0
    shape = (6,) * d
    q = ol.wave_vectors(shape)
    lam0, mu0 = 40e9, 27e9
    G = ol.green_operator(q, lam0, mu0)
    I = np.eye(d)
    C0 = (lam0 * np.einsum("ij,kl->ijkl", I, I)
          + mu0 * (np.einsum("ik,jl->ijkl", I, I) + np.einsum("il,jk->ijkl", I, I)))
    assert np.all(G[(0,) * d] == 0)
    for idx in [(1,) + (2,) * (d - 1), (3,) + (0,) * (d - 1), (2,) * d]:
        g = G[idx]
        gcg = np.einsum("khij,ijab,abmn->khmn", g, C0, g)
        assert np.allclose(gcg, g, rtol=1e-12, atol=1e-12 * np.abs(g).max())


def test_iteration_trajectory_stops_like_reference():
    # the stop rule is diff < tol after the update, silent at max_iter
    shape = (8, 8)
    rng = _rng(3)
    E = 70e9 * rng.uniform(0.9, 1.1, shape)
    nu = np.full(shape, 0.35)
    eb = np.diag([0.01, 0.0])
    info = {}
    ol.spectral_solver_2d(E, nu, eb, tol=1e-8, info=info)
    diffs = info["diffs"]
    assert diffs[-1] < 1e-8 and all(x >= 1e-8 for x in diffs[:-1])
    info2 = {}
    ol.spectral_solver_2d(E, nu, eb, tol=1e-30, max_iter=3, info=info2)
    assert info2["iters"] == 3


@pytest.mark.parametrize("d,plane_mode", [(2, "plane_strain"), (2, "plane_stress"), (3, "plane_strain")])
def test_mixed_bc_driver_homogeneous_closed_form(d, plane_mode):
    # homogeneous medium, eps_xx driven, other normal stresses zero: the macro
    # loop must reach uniaxial stress: sig_xx = E' eps_xx, eps_yy = -nu' eps_xx
    E0, nu0 = 70e9, 0.3
    shape = (8,) * d
    E = np.full(shape, E0)
    nu = np.full(shape, nu0)
    mask = np.zeros((d, d), bool)
    mask[0, 0] = True
    for i in range(d):
        for j in range(d):
            if i != j:
                mask[i, j] = True  # shear strains held at 0
    tv = np.zeros((d, d))
    tv[0, 0] = 0.01
    em, sm, el, sl, its = ol.linear_elastic_simulation(d, E, nu, mask, tv, n_steps=2, tol_ls=1e-10, tol_macro=1.0,
                                                      plane_mode=plane_mode)
    if d == 3 or plane_mode == "plane_stress":
        Ep, nup = E0, nu0
    else:
        Ep, nup = E0 / (1 - nu0 ** 2), nu0 / (1 - nu0)
    exx = em[-1][0, 0]
    assert abs(exx - 0.01) < 1e-15
    assert abs(sm[-1][0, 0] - Ep * exx) < 1e-6 * Ep * exx
    for k in range(1, d):
        assert abs(em[-1][k, k] + nup * exx) < 1e-6 * exx
        assert abs(sm[-1][k, k]) < 1.0
    assert np.allclose(em[1], 0.5 * em[2], rtol=1e-6, atol=1e-14)
    assert len(el) == 3 and el[-1].shape == shape + (d, d)


LANDAU_PA = dict(v1=-236.5e9, v2=-27.4e9, v3=11.8e9, g1=-5640.7e9, g2=2207.1e9, g3=-332.5e9, g4=-305.2e9)


def test_landau_stress_is_energy_derivative():
    # sigma_ij = dW/d eps_ij (central differences of the independently written energy)
    rng = np.random.default_rng(4)
    g = rng.standard_normal((3, 3)) * 0.01
    eps = 0.5 * (g + g.T)
    lam, mu = 60e9, 27e9
    sig = ol.stress_landau_3d(eps, lam, mu, **LANDAU_PA)
    h = 1e-7
    for i in range(3):
        for j in range(3):
            d = np.zeros((3, 3))
            d[i, j] += 0.5 * h
            d[j, i] += 0.5 * h
            fd = (ol.landau_energy_3d(eps + d, lam, mu, **LANDAU_PA) -
                  ol.landau_energy_3d(eps - d, lam, mu, **LANDAU_PA)) / (2 * h)
            assert abs(fd - sig[i, j]) <= 1e-6 * np.abs(sig).max(), (i, j, fd, sig[i, j])


def test_landau_ls_linear_limit_and_capping():
    # v_i = g_i = 0: the Landau LS converges to the linear LS solution (the
    # reference media differ, the fixed point does not); capping leaves small
    # strains alone and holds the deviatoric strain at E_cap for large ones
    shape = (6, 5, 4)
    rng = np.random.default_rng(2)
    E = 70e9 * rng.uniform(0.9, 1.1, shape)
    nu = np.full(shape, 0.3)
    lam, mu = ol.compute_lame_3d(E, nu)
    g = rng.standard_normal(shape + (3, 3)) * 1e-3
    ep = 0.5 * (g + np.swapaxes(g, -1, -2))
    eb = np.diag([0.01, -0.002, 0.0])
    zero = {k: 0.0 for k in LANDAU_PA}
    a = ol.spectral_solver_landau_3d(lam, mu, eps_bar=eb, eps_plastic=ep, tol=1e-12, **zero)
    b = ol.spectral_solver_3d(E, nu, eb, eps_plastic=ep, tol=1e-12)
    assert np.abs(a[0] - b[0]).max() < 1e-10 * np.abs(b[0]).max()
    assert np.abs(a[1] - b[1]).max() < 1e-9 * np.abs(b[1]).max()
    small = np.diag([1e-4, -5e-5, -5e-5])
    e, w = ol.cap_strain_3d(small, 27e9, -305e9, True, 0.1, 0.1, "piecewise", 1.0)
    assert np.array_equal(e, small) and w == 0.0
    big = np.array([[0.0, 0.3, 0.0], [0.3, 0.0, 0.0], [0.0, 0.0, 0.0]]) + 0.01 * np.eye(3)
    e, w = ol.cap_strain_3d(big, 27e9, -305e9, True, 0.1, 0.1, "piecewise", 1.0)
    dev = e - np.trace(e) / 3 * np.eye(3)
    assert abs(np.sqrt(2.0 / 3.0 * np.sum(dev * dev)) - 0.1) < 1e-14 and w == 1.0
    assert abs(np.trace(e) - np.trace(big)) < 1e-15


@pytest.mark.parametrize("cap", [{}, dict(strain_capping_enabled=True, strain_capping_limit=0.03),
                                 dict(strain_capping_enabled=True, strain_capping_type="smooth",
                                      strain_capping_smooth_power=2.0)])
def test_landau_2d_stress_against_3d_law(cap):
    # plane strain: the 2D block-diagonal form equals the 3D tensor law at e33 = 0;
    # plane stress: the Newton e33 makes the 3D law's sigma33 vanish, and the
    # in-plane stress equals the 3D law at that e33 (to the Newton step tolerance)
    rng = np.random.default_rng(6)
    shape = (5, 4)
    g = rng.standard_normal(shape + (2, 2)) * 0.01
    e2 = 0.5 * (g + np.swapaxes(g, -1, -2))
    lam = 60e9 * rng.uniform(0.9, 1.1, shape)
    mu = 27e9 * rng.uniform(0.9, 1.1, shape)
    s2 = ol.stress_landau_2d(e2, lam, mu, **LANDAU_PA, **cap)
    e3 = np.zeros(shape + (3, 3))
    e3[..., :2, :2] = e2
    s3 = ol.stress_landau_3d(e3, lam, mu, **LANDAU_PA, **cap)
    assert np.abs(s2 - s3[..., :2, :2]).max() < 1e-9 * np.abs(s3).max()
    st = {}
    s2p = ol.stress_landau_2d(e2, lam, mu, **LANDAU_PA, plane_mode="plane_stress", e33_state=st, **cap)
    e3[..., 2, 2] = st["e33"]
    s3p = ol.stress_landau_3d(e3, lam, mu, **LANDAU_PA, **cap)
    assert np.abs(s3p[..., 2, 2]).max() < 1e-6 * np.abs(s3p).max()
    assert np.abs(s2p - s3p[..., :2, :2]).max() < 1e-6 * np.abs(s3p).max()


def test_landau_2d_anderson_reaches_the_picard_fixed_point():
    shape = (12, 10)
    rng = np.random.default_rng(8)
    E = 70e9 * rng.uniform(0.8, 1.2, shape)
    lam, mu = ol.compute_lame_2d(E, np.full(shape, 0.3), "plane_strain")
    g = rng.standard_normal(shape + (2, 2)) * 2e-3
    ep = 0.5 * (g + np.swapaxes(g, -1, -2))
    eb = np.array([[0.01, 0.03], [0.03, -0.004]])
    kw = dict(eps_plastic=ep, tol=1e-10, plane_mode="plane_stress", strain_capping_enabled=True,
              strain_capping_limit=0.03)
    ia, ip = {}, {}
    a = ol.spectral_solver_landau_2d(lam, mu, **LANDAU_PA, eps_bar=eb, anderson_m=5, info=ia, e33_state={}, **kw)
    b = ol.spectral_solver_landau_2d(lam, mu, **LANDAU_PA, eps_bar=eb, anderson_m=0, info=ip, e33_state={}, **kw)
    assert ia["iters"] < ip["iters"]
    assert np.abs(a[0] - b[0]).max() < 1e-8 * np.abs(b[0]).max()
    assert np.abs(a[1] - b[1]).max() < 1e-7 * np.abs(b[1]).max()


# The Newton-CG branch differentiates the stress by central differences with a
# fixed h = 1e-7 along G(p); with moduli in Pa, G(p) of a Krylov vector is
# ~1e-11 of p, so h G(p) falls below the rounding unit of the strain and the
# differenced stress is zero (the reference's BiCGSTAB then stalls at maxiter).
# Its tests therefore run in reduced units (moduli and Landau constants / 70 GPa)
# and on odd grids: on an even grid -Gamma0 at the Nyquist frequencies is not a
# projector (G(C0 G(t)) != -G(t) there), so the Newton-CG and the fixed point
# settle on different fields and the Krylov solves stall.
LANDAU_RED = {k: v / 70e9 for k, v in LANDAU_PA.items()}


def _landau_red_case(seed, shape):
    rng = np.random.default_rng(seed)
    E = rng.uniform(0.8, 1.2, shape)
    lam, mu = ol.compute_lame_2d(E, np.full(shape, 0.3), "plane_strain")
    g = rng.standard_normal(shape + (2, 2)) * 2e-3
    return lam, mu, 0.5 * (g + np.swapaxes(g, -1, -2))


@pytest.mark.parametrize("plane_mode,cap", [("plane_strain", {}),
                                            ("plane_stress", dict(strain_capping_enabled=True,
                                                                  strain_capping_limit=0.03))])
def test_landau_2d_newton_cg_reaches_the_fixed_point(plane_mode, cap):
    # the Newton-CG branch and the plain fixed point solve one equilibrium
    # problem: at tight tolerances they agree to the fixed point's accuracy,
    # with far fewer Newton steps than Picard iterations; the Newton-CG field
    # is equilibrated (-Gamma0 : sigma = 0) and keeps the prescribed mean
    shape = (11, 9)
    lam, mu, ep = _landau_red_case(8, shape)
    eb = np.array([[0.01, 0.02], [0.02, -0.004]])
    kw = dict(eps_plastic=ep, plane_mode=plane_mode, **cap)
    inc, ip = {}, {}
    a = ol.spectral_solver_landau_2d(lam, mu, **LANDAU_RED, eps_bar=eb, tol=1e-6, solver="newton_cg", info=inc, **kw)
    b = ol.spectral_solver_landau_2d(lam, mu, **LANDAU_RED, eps_bar=eb, tol=1e-13, anderson_m=0, info=ip,
                                     e33_state={}, **kw)
    assert inc["iters"] < ip["iters"] and 0 < inc["krylov"] < 100
    assert np.abs(a[0] - b[0]).max() < 1e-6 * np.abs(b[0]).max()
    assert np.abs(a[1] - b[1]).max() < 1e-6 * np.abs(b[1]).max()
    assert np.abs(a[2] - eb).max() < 1e-14
    Gam = ol.green_operator(ol.wave_vectors(shape, 1.0), float(lam.mean()), float(mu.mean()))
    th = np.stack([np.stack([ol.fft_field(a[1][..., i, j]) for j in range(2)], -1) for i in range(2)], -2)
    t0 = np.stack([np.stack([ol.fft_field(ol.stress_landau_2d(np.broadcast_to(eb, shape + (2, 2)) - ep, lam, mu,
                                                              **LANDAU_RED, plane_mode=plane_mode, **cap)[..., i, j])
                             for j in range(2)], -1) for i in range(2)], -2)
    r = np.einsum("xykhij,xyij->xykh", Gam, th)
    r0 = np.einsum("xykhij,xyij->xykh", Gam, t0)
    assert np.linalg.norm(r) < 1e-6 * np.linalg.norm(r0)


def test_landau_2d_newton_cg_linear_limit():
    # zero Landau coefficients: the law is linear, C_h is exact up to the
    # differencing round-off, and one Newton step after the (empty) first
    # reaches the linear LS solution
    shape = (9, 13)
    lam, mu, ep = _landau_red_case(4, shape)
    eb = np.array([[0.01, 0.0], [0.0, 0.0]])
    zero = {k: 0.0 for k in LANDAU_PA}
    info = {}
    a = ol.spectral_solver_landau_2d(lam, mu, **zero, eps_bar=eb, eps_plastic=ep, tol=1e-6, solver="newton_cg",
                                     info=info)
    b = ol.spectral_solver_landau_2d(lam, mu, **zero, eps_bar=eb, eps_plastic=ep, tol=1e-13, anderson_m=0)
    assert info["iters"] <= 3
    assert np.abs(a[0] - b[0]).max() < 1e-7 * np.abs(b[0]).max()


def test_landau_2d_newton_cg_fd_underflow_in_pa_units():
    # the reference's fixed-step difference along G(p) with moduli in Pa: the
    # step h G(p) is below the strain's rounding unit and C_h returns zeros
    shape = (8, 6)
    rng = np.random.default_rng(2)
    lam, mu = ol.compute_lame_2d(70e9 * rng.uniform(0.8, 1.2, shape), np.full(shape, 0.3), "plane_strain")
    Gam = ol.green_operator(ol.wave_vectors(shape, 1.0), float(lam.mean()), float(mu.mean()))
    e = np.broadcast_to(np.array([[0.01, 0.0], [0.0, 0.0]]), shape + (2, 2)).copy()
    p = rng.standard_normal(shape + (2, 2)) * 1e-2
    p = 0.5 * (p + np.swapaxes(p, -1, -2))
    th = np.stack([np.stack([ol.fft_field(p[..., i, j]) for j in range(2)], -1) for i in range(2)], -2)
    gh = -np.einsum("xykhij,xyij->xykh", Gam, th)
    d = np.stack([np.stack([ol.ifft_field(gh[..., i, j]) for j in range(2)], -1) for i in range(2)], -2)
    assert np.abs(1e-7 * d).max() < 0.5 * np.spacing(0.01)
    sp = ol.stress_landau_2d(e + 1e-7 * d, lam, mu, **LANDAU_PA)
    sm = ol.stress_landau_2d(e - 1e-7 * d, lam, mu, **LANDAU_PA)
    assert np.all((sp - sm)[..., 0, 0] == 0.0)


# ---------------------------------------------------------------------------
# secant-degradation and Rose laws (oracle/ls.py stress_secant, stress_rose_2d)
# ---------------------------------------------------------------------------
def _vm_3x3(eps):
    """Von Mises strain from the full 3x3 deviator (a 2D strain padded with
    e33 = e13 = e23 = 0), written independently of the oracle's component form."""
    d = eps.shape[-1]
    E3 = np.zeros(eps.shape[:-2] + (3, 3))
    E3[..., :d, :d] = eps
    dev = E3 - np.trace(E3, axis1=-2, axis2=-1)[..., None, None] / 3.0 * np.eye(3)
    return np.sqrt(2.0 / 3.0 * np.einsum("...ij,...ij->...", dev, dev))


def _sym(rng, shape, d, s=0.01):
    a = rng.normal(0, s, shape + (d, d))
    return 0.5 * (a + np.swapaxes(a, -1, -2))


@pytest.mark.parametrize("d", [2, 3])
def test_secant_law_closed_forms(d):
    rng = _rng(11)
    shape = (5, 4) if d == 2 else (3, 4, 2)
    lam = rng.uniform(30e9, 60e9, shape)
    mu = rng.uniform(20e9, 40e9, shape)
    e = _sym(rng, shape, d)
    assert np.allclose(ol.von_mises_strain(e), _vm_3x3(e), rtol=1e-13, atol=0)
    lin = lam[..., None, None] * np.trace(e, axis1=-2, axis2=-1)[..., None, None] * np.eye(d) \
        + 2 * mu[..., None, None] * e
    assert np.allclose(ol.stress_secant(e, lam, mu, 0.0, 50.0), lin, rtol=1e-14, atol=0)
    # pure shear gamma: e_eq = gamma/sqrt3; k e_eq = ln 2 => mu_sec = mu (1 - d/2)
    g, dd = 0.02, 0.6
    es = np.zeros((d, d))
    es[0, 1] = es[1, 0] = g / 2
    k = np.log(2.0) * np.sqrt(3.0) / g
    s = ol.stress_secant(es, 40e9, 30e9, dd, k)
    assert abs(s[0, 1] - 2 * 30e9 * (1 - dd / 2) * g / 2) < 1e-12 * 30e9 * g
    assert np.all(s[np.eye(d) > 0] == 0)
    # equibiaxial / hydrostatic a I: 3D deviator zero (mu_sec = mu); 2D keeps
    # e33 = 0 in the deviator, e_eq = 2a/3
    a = 0.01
    s = ol.stress_secant(a * np.eye(d), 40e9, 30e9, dd, 20.0)
    if d == 3:
        assert np.allclose(np.diag(s), (3 * 40e9 + 2 * 30e9) * a, rtol=1e-14)
    else:
        ms = 30e9 * (1 - dd * (1 - np.exp(-20.0 * 2 * a / 3)))
        assert np.allclose(np.diag(s), 2 * 40e9 * a + 2 * ms * a, rtol=1e-14)
    # full degradation at large k: sigma -> lam tr(e) I
    s = ol.stress_secant(e, lam, mu, 1.0, 1e9)
    assert np.allclose(s, lam[..., None, None] * np.trace(e, axis1=-2, axis2=-1)[..., None, None] * np.eye(d),
                       rtol=0, atol=1e-6 * np.abs(s).max())


def test_rose_law_closed_forms():
    rng = _rng(12)
    shape = (6, 5)
    lam = rng.uniform(30e9, 60e9, shape)
    mu = rng.uniform(20e9, 40e9, shape)
    e = _sym(rng, shape, 2)
    tr = np.trace(e, axis1=-2, axis2=-1)[..., None, None]
    L, M = lam[..., None, None], mu[..., None, None]
    # eta = 0: linear plane strain, and exact linear plane stress with
    # lam* = 2 lam mu / (lam + 2 mu)
    assert np.allclose(ol.stress_rose_2d(e, lam, mu, 0.0), L * tr * np.eye(2) + 2 * M * e, rtol=1e-13, atol=0)
    ls = 2 * L * M / (L + 2 * M)
    assert np.allclose(ol.stress_rose_2d(e, lam, mu, 0.0, plane_mode="plane_stress"),
                       ls * tr * np.eye(2) + 2 * M * e, rtol=1e-12, atol=1e-9 * np.abs(L * e).max())
    # pure shear: eta e_eq = ln 2 => G = mu/2; a huge eta hits the floor
    g = 0.02
    es = np.zeros((2, 2))
    es[0, 1] = es[1, 0] = g / 2
    eta = np.log(2.0) * np.sqrt(3.0) / g
    for mode in ("plane_strain", "plane_stress"):
        s = ol.stress_rose_2d(es, 40e9, 30e9, eta, 0.1, plane_mode=mode)
        assert abs(s[0, 1] - 30e9 * g / 2) < 1e-12 * 30e9 * g
        s = ol.stress_rose_2d(es, 40e9, 30e9, 1e6, 0.1, plane_mode=mode)
        assert abs(s[0, 1] - 2 * 0.1 * 30e9 * g / 2) < 1e-12 * 30e9 * g
    # constant bulk modulus: the in-plane trace of the plane-strain stress is
    # 2 (K - 2/3 G) tr + 2 G tr with G from the 3x3 deviator
    eta = 40.0
    s = ol.stress_rose_2d(e, lam, mu, eta)
    G = np.maximum(mu * np.exp(-eta * _vm_3x3(e)), 0.1 * mu)
    K = lam + 2 * mu / 3
    t = np.trace(e, axis1=-2, axis2=-1)
    assert np.allclose(s[..., 0, 0] + s[..., 1, 1], 2 * (K - 2 * G / 3) * t + 2 * G * t, rtol=1e-12)
    assert np.allclose(s[..., 0, 1], 2 * G * e[..., 0, 1], rtol=1e-12)
    # plane stress: five updates approach the root of sigma_33(e33) = 0 of the 3D law
    from scipy.optimize import brentq
    e33 = np.zeros(shape)
    ol.stress_rose_2d(e, lam, mu, eta, plane_mode="plane_stress", e33_out=e33)
    for i in np.ndindex(shape):
        def s33(z, i=i):
            E3 = np.zeros((3, 3))
            E3[:2, :2] = e[i]
            E3[2, 2] = z
            dev = E3 - np.trace(E3) / 3 * np.eye(3)
            Gz = max(mu[i] * np.exp(-eta * np.sqrt(2 / 3 * np.sum(dev * dev))), 0.1 * mu[i])
            return (K[i] - 2 * Gz / 3) * np.trace(E3) + 2 * Gz * z
        root = brentq(s33, -0.1, 0.1, xtol=1e-16)
        assert abs(e33[i] - root) <= 1e-3 * (abs(root) + 1e-4)
        # exactly five updates from the linear guess (the 3x3 form above,
        # not the oracle's components): the truncation error is visible
        z = -lam[i] / (lam[i] + 2 * mu[i]) * (e[i][0, 0] + e[i][1, 1])
        for _ in range(5):
            E3 = np.zeros((3, 3))
            E3[:2, :2] = e[i]
            E3[2, 2] = z
            dev = E3 - np.trace(E3) / 3 * np.eye(3)
            Gz = max(mu[i] * np.exp(-eta * np.sqrt(2 / 3 * np.sum(dev * dev))), 0.1 * mu[i])
            z = -(3 * K[i] - 2 * Gz) / (3 * K[i] + 4 * Gz) * (e[i][0, 0] + e[i][1, 1])
        assert abs(e33[i] - z) <= 1e-14 * abs(z) + 1e-20


def _shear_laminate(law_shear, mus, gbar):
    """Two-layer laminate (layers normal to x, equal fractions) under mean
    shear strain gbar/2: the shear stress is uniform, the layer shear strains
    average to gbar/2, normal strains vanish.  Returns the layer e_xy."""
    from scipy.optimize import brentq
    ea_of = lambda eb: gbar - eb  # e_a + e_b = 2 * (gbar/2)

    def f(eb):
        return law_shear(ea_of(eb), mus[0]) - law_shear(eb, mus[1])
    eb = brentq(f, 1e-9 * gbar, gbar - 1e-9 * gbar, xtol=1e-18, rtol=1e-15)
    return ea_of(eb), eb


@pytest.mark.parametrize("law", ["secant3", "secant2", "rose_strain", "rose_stress"])
def test_secant_ls_laminate_exact(law):
    mus = (20e9, 36e9)
    lam_v = 45e9
    gbar = 0.02
    d = 3 if law == "secant3" else 2
    n = 8
    shape = (n, 2, 2) if d == 3 else (n, 2)
    mu = np.empty(shape)
    mu[: n // 2] = mus[0]
    mu[n // 2:] = mus[1]
    lam = np.full(shape, lam_v)
    eb = np.zeros((d, d))
    eb[0, 1] = eb[1, 0] = gbar / 2
    info = {}
    if law.startswith("secant"):
        dd, k = 0.7, 60.0
        shear = lambda e, m: 2 * m * (1 - dd * (1 - np.exp(-k * 2 * abs(e) / np.sqrt(3)))) * e
        fn = ol.spectral_solver_secant_3d if d == 3 else ol.spectral_solver_secant_2d
        eps, sig, em, sm = fn(lam, mu, dd, k, eb, tol=1e-13, max_iter=3000, info=info)
    else:
        eta, fl = 30.0, 0.1
        shear = lambda e, m: 2 * max(m * np.exp(-eta * 2 * abs(e) / np.sqrt(3)), fl * m) * e
        mode = "plane_stress" if law == "rose_stress" else "plane_strain"
        eps, sig, em, sm = ol.spectral_solver_rose_2d(lam, mu, eta, fl, eps_bar=eb, tol=1e-13, max_iter=3000,
                                                      plane_mode=mode, info=info)
    ea, ebb = _shear_laminate(shear, mus, gbar)
    assert info["iters"] < 3000
    assert np.allclose(eps[: n // 2, ..., 0, 1], ea, rtol=1e-9, atol=0)
    assert np.allclose(eps[n // 2:, ..., 0, 1], ebb, rtol=1e-9, atol=0)
    assert np.abs(eps[..., 0, 0]).max() < 1e-12
    assert np.allclose(sig[..., 0, 1], shear(ea, mus[0]), rtol=1e-8)
    assert np.allclose(em, eb, atol=1e-15)


def test_secant_ls_linear_limit_and_equilibrium():
    rng = _rng(13)
    shape = (8, 6, 4)
    E = 70e9 * rng.uniform(0.7, 1.3, shape)
    nu = np.full(shape, 0.3)
    lam, mu = ol.compute_lame_3d(E, nu)
    eb = np.zeros((3, 3))
    eb[0, 0], eb[0, 1], eb[1, 0] = 0.01, 0.004, 0.004
    ep = _sym(rng, shape, 3, 2e-3)
    a = ol.spectral_solver_secant_3d(lam, mu, 0.0, 30.0, eb, eps_plastic=ep, tol=1e-12, max_iter=2000)
    b = ol.spectral_solver_3d(E, nu, eb, eps_plastic=ep, tol=1e-12, max_iter=2000)
    assert np.allclose(a[0], b[0], rtol=0, atol=1e-9 * np.abs(b[0]).max())
    c = ol.spectral_solver_secant_3d(lam, mu, 0.5, 30.0, eb, eps_plastic=ep, tol=1e-11, max_iter=2000)
    assert ol.equilibrium_residual(c[1]) < 1e-8
    assert np.allclose(c[2], eb, atol=1e-15)
    e2 = _sym(rng, (8, 8), 2, 2e-3)
    lam2, mu2 = lam[:, :, 0].repeat(2, 1)[:, :8], mu[:, :, 0].repeat(2, 1)[:, :8]
    r = ol.spectral_solver_rose_2d(lam2, mu2, 30.0, 0.1, eps_bar=eb[:2, :2], eps_plastic=e2, tol=1e-11,
                                   max_iter=3000, plane_mode="plane_stress")
    assert ol.equilibrium_residual(r[1]) < 1e-8


@pytest.mark.parametrize("law,d,plane_mode", [("secant", 3, "plane_strain"), ("secant", 2, "plane_strain"),
                                              ("rose", 2, "plane_strain"), ("rose", 2, "plane_stress")])
def test_secant_driver_uniaxial_closed_form(law, d, plane_mode):
    # homogeneous nonlinear medium, eps_xx driven, transverse normal stresses
    # zero, shear strains held: the macro loop reaches the uniaxial state whose
    # transverse strain t solves sigma_yy(eps_xx, t) = 0 of the law (brentq on
    # the 3x3 deviator form), at each of the ramped steps
    from scipy.optimize import brentq
    lam0, mu0 = 45e9, 27e9
    shape = (6,) * d
    lam = np.full(shape, lam0)
    mu = np.full(shape, mu0)
    mask = ~np.eye(d, dtype=bool)
    mask[0, 0] = True
    tv = np.zeros((d, d))
    tv[0, 0] = 0.02
    p1, p2 = (0.6, 40.0) if law == "secant" else (25.0, 0.1)
    em, sm, el, sl, its = ol.secant_simulation(d, law, lam, mu, p1, p2, mask, tv, n_steps=2, tol_ls=1e-10,
                                               tol_macro=1.0, plane_mode=plane_mode)

    def stress(ex, t):
        E3 = np.diag([ex] + [t] * (d - 1) + [0.0] * (3 - d))
        if law == "rose" and plane_mode == "plane_stress":
            # sigma_33 = 0 fixes e33 (exact root here; the law's five updates
            # are within 1e-3 of it, far below the tolerance below)
            def s33(z):
                E3[2, 2] = z
                dev_ = E3 - np.trace(E3) / 3 * np.eye(3)
                G = max(mu0 * np.exp(-p1 * np.sqrt(2 / 3 * np.sum(dev_ * dev_))), p2 * mu0)
                return (lam0 + 2 * mu0 / 3 - 2 * G / 3) * np.trace(E3) + 2 * G * z
            E3[2, 2] = brentq(s33, -0.1, 0.1, xtol=1e-18)
        dev_ = E3 - np.trace(E3) / 3 * np.eye(3)
        eq = np.sqrt(2 / 3 * np.sum(dev_ * dev_))
        if law == "secant":
            ms = mu0 * (1 - p1 * (1 - np.exp(-p2 * eq)))
            return lam0 * np.trace(E3[:d, :d]) + 2 * ms * E3[1, 1], lam0 * np.trace(E3[:d, :d]) + 2 * ms * ex
        G = max(mu0 * np.exp(-p1 * eq), p2 * mu0)
        le = lam0 + 2 * mu0 / 3 - 2 * G / 3
        return le * np.trace(E3) + 2 * G * E3[1, 1], le * np.trace(E3) + 2 * G * ex

    rt = 1e-4 if plane_mode == "plane_stress" else 1e-8
    for s, ex in ((1, 0.01), (2, 0.02)):
        t = brentq(lambda t: stress(ex, t)[0], -0.05, 0.05, xtol=1e-18)
        assert abs(em[s][0, 0] - ex) < 1e-15
        for k in range(1, d):
            assert abs(em[s][k, k] - t) < rt * abs(t)
            assert abs(sm[s][k, k]) < 1.0
        assert abs(sm[s][0, 0] - stress(ex, t)[1]) < rt * abs(sm[s][0, 0])
    assert len(el) == 3 and el[-1].shape == shape + (d, d)


@pytest.mark.parametrize("d,plane_mode", [(3, "plane_strain"), (2, "plane_strain"), (2, "plane_stress")])
def test_landau_driver_uniaxial_closed_form(d, plane_mode):
    # homogeneous Landau medium: the secant-slope macro loop reaches the
    # uniaxial state whose transverse strain t zeroes sigma_yy of the law
    # (brentq on the single-voxel law)
    from scipy.optimize import brentq
    lam0, mu0 = 45e9, 27e9
    shape = (6,) * d
    lam = np.full(shape, lam0)
    mu = np.full(shape, mu0)
    mask = ~np.eye(d, dtype=bool)
    mask[0, 0] = True
    tv = np.zeros((d, d))
    tv[0, 0] = 0.02
    em, sm, el, sl, its = ol.landau_simulation(d, lam, mu, **LANDAU_PA, target_strain_mask=mask, target_values=tv,
                                               n_steps=2, tol_ls=1e-10, tol_macro=1.0, plane_mode=plane_mode)

    def stress(ex, t):
        e = np.diag([ex] + [t] * (d - 1)).reshape((1,) * d + (d, d))
        one = np.ones((1,) * d)
        if d == 3:
            return ol.stress_landau_3d(e, lam0 * one, mu0 * one, **LANDAU_PA).reshape(d, d)
        return ol.stress_landau_2d(e, lam0 * one, mu0 * one, **LANDAU_PA, plane_mode=plane_mode).reshape(d, d)

    for s, ex in ((1, 0.01), (2, 0.02)):
        t = brentq(lambda t: stress(ex, t)[1, 1], -0.05, 0.05, xtol=1e-18)
        assert abs(em[s][0, 0] - ex) < 1e-15
        for k in range(1, d):
            assert abs(em[s][k, k] - t) < 1e-7 * abs(t)
            assert abs(sm[s][k, k]) < 1.0
        assert abs(sm[s][0, 0] - stress(ex, t)[0, 0]) < 1e-7 * abs(sm[s][0, 0])


def test_matrix_free_ls_equals_materialised_ls():
    # spectral_solver_3d_mf: the reference's iteration with Gamma0 applied on
    # r2c spectra (Nyquist: the Hermitian average) = the materialised
    # full-spectrum solver to round-off, with the same iteration count; odd,
    # even and mixed grids (two even axes exercise the Nyquist sign pairing)
    import numpy as np
    from oracle import ls as ol
    for shape in [(8, 6, 10), (7, 9, 5), (6, 7, 4), (5, 8, 8)]:
        rng = np.random.default_rng(3)
        E = 70e9 * rng.uniform(0.9, 1.1, shape)
        nu = np.full(shape, 0.3)
        g = rng.standard_normal(shape + (3, 3)) * 2e-3
        ep = 0.5 * (g + np.swapaxes(g, -1, -2))
        eb = np.array([[0.01, 0.002, 0.0], [0.002, -0.003, 0.001], [0.0, 0.001, 0.0]])
        i0, i1 = {}, {}
        r0 = ol.spectral_solver_3d(E, nu, eb, eps_plastic=ep, tol=1e-10, info=i0)
        r1 = ol.spectral_solver_3d_mf(E, nu, eb, eps_plastic=ep, tol=1e-10, info=i1)
        assert i0["iters"] == i1["iters"], shape
        for a, b in zip(r1, r0):
            assert np.abs(a - b).max() <= 1e-13 * np.abs(b).max(), shape
