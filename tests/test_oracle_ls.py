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
