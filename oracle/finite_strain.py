"""Oracle: de Geus finite-strain Newton-Krylov solve (CPU restatement).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Follows mgkmc/finite_strain_simulator.py:
  :39-97        tensor einsum helpers (_trans2_3d, _ddot42_3d, ...)
  :104-119      _make_identity_tensors_3d
  :279-300      _invert_Fp_3d
  :409-483      _landau_C4eff_mandel_3d
  :1121-1167    build_ghat4_3d (zero at DC and on even-axis Nyquist planes)
  :1174-1188    _fft3_tensor / _ifft3_tensor / _project_3d
  :1628-1640    build_C4_3d
  :1647-1885    constitutive_hyperelastic_3d (svk / landau / murnaghan,
                capping, Fp, Mandel and einsum tangents)
  :1892-1904    cauchy_from_P_3d
  :2852-2899    build_finite_strain_bc_3d
  :2145-2242    _al_step_3d (augmented Lagrangian)
  :1212-1379,1508-1620  _get_frequencies_3d, DBFFT helpers, solve_dbfft_linear_system_3d,
                _dbfft_step_3d (displacement-based Newton with the (xi.<K4>.xi)^-1 preconditioner)
  :2967-3168    finite_strain_solver_step_3d, solver="newton_cg" / "al" / "dbfft"
  :3175-3413    finite_strain_simulation_3d (load stepping, store=False)
The Krylov solver is SciPy's own scipy.sparse.linalg.bicgstab (the reference's
dependency, SciPy 1.15.3 here), called exactly as the reference calls it.

Pinned by (tests/test_oracle_fs.py): central finite differences of P(F)
against the assembled tangent (no Fp), the Mandel assembly against the einsum
("reference") assembly (tests/test_landau_tangent_mandel.py of the reference),
the SVK limit of Landau, projector identities of G (G:G = G, G annihilates
compatible... fields with zero mean), the converged Newton solution
(G:P = 0 to tolerance, <F> = F_bar), and a homogeneous closed form.  AL: the
homogeneous closed form after two iterations and the reference's failure on
heterogeneous grids.  DBFFT: the homogeneous closed form, the same equilibrium
as the Newton-CG solution, an exactly compatible F - F_bar, and its operators
in the real-field form the device uses (grad / div with K4 transposed).
"""
import numpy as np
import scipy.fft as _sfft
import scipy.sparse.linalg as sp

FFT_WORKERS = 1


def _trans2(A):
    return np.einsum("ijxyz->jixyz", A)


def _ddot42(A4, B2):
    return np.einsum("ijklxyz,lkxyz->ijxyz", A4, B2)


def _dot22(A, B):
    return np.einsum("ijxyz,jkxyz->ikxyz", A, B)


def identity_tensors(nx, ny, nz):
    i = np.eye(3)
    ones = np.ones((nx, ny, nz))
    I2 = np.einsum("ij,xyz->ijxyz", i, ones)
    I4 = np.einsum("ijkl,xyz->ijklxyz", np.einsum("il,jk", i, i), ones)
    I4rt = np.einsum("ijkl,xyz->ijklxyz", np.einsum("ik,jl", i, i), ones)
    I4s = 0.5 * (I4 + I4rt)
    II = np.einsum("ijxyz,klxyz->ijklxyz", I2, I2)
    return I2, I4, I4rt, I4s, II


def invert_Fp(Fp):
    det = (Fp[0, 0] * (Fp[1, 1] * Fp[2, 2] - Fp[1, 2] * Fp[2, 1])
           - Fp[0, 1] * (Fp[1, 0] * Fp[2, 2] - Fp[1, 2] * Fp[2, 0])
           + Fp[0, 2] * (Fp[1, 0] * Fp[2, 1] - Fp[1, 1] * Fp[2, 0]))
    ds = np.where(np.abs(det) < 1e-14, 1e-14, det)
    inv = np.zeros_like(Fp)
    inv[0, 0] = (Fp[1, 1] * Fp[2, 2] - Fp[1, 2] * Fp[2, 1]) / ds
    inv[0, 1] = -(Fp[0, 1] * Fp[2, 2] - Fp[0, 2] * Fp[2, 1]) / ds
    inv[0, 2] = (Fp[0, 1] * Fp[1, 2] - Fp[0, 2] * Fp[1, 1]) / ds
    inv[1, 0] = -(Fp[1, 0] * Fp[2, 2] - Fp[1, 2] * Fp[2, 0]) / ds
    inv[1, 1] = (Fp[0, 0] * Fp[2, 2] - Fp[0, 2] * Fp[2, 0]) / ds
    inv[1, 2] = -(Fp[0, 0] * Fp[1, 2] - Fp[0, 2] * Fp[1, 0]) / ds
    inv[2, 0] = (Fp[1, 0] * Fp[2, 1] - Fp[1, 1] * Fp[2, 0]) / ds
    inv[2, 1] = -(Fp[0, 0] * Fp[2, 1] - Fp[0, 1] * Fp[2, 0]) / ds
    inv[2, 2] = (Fp[0, 0] * Fp[1, 1] - Fp[0, 1] * Fp[1, 0]) / ds
    return inv


def build_ghat4(nx, ny, nz, Lx, Ly, Lz):
    xs = [np.fft.fftfreq(n, d=L / n) for n, L in ((nx, Lx), (ny, Ly), (nz, Lz))]
    X = np.meshgrid(*xs, indexing="ij")
    xi = np.stack(X, 0)
    xi2 = X[0] ** 2 + X[1] ** 2 + X[2] ** 2
    nyq = np.zeros((nx, ny, nz), bool)
    for a, (n, L) in enumerate(((nx, Lx), (ny, Ly), (nz, Lz))):
        if n % 2 == 0:
            nyq |= X[a] == -n / (2.0 * L)
    s = xi2.copy()
    s[xi2 == 0] = 1.0
    G = np.zeros((3, 3, 3, 3, nx, ny, nz))
    d = np.eye(3)
    for i in range(3):
        for j in range(3):
            for l in range(3):
                for m in range(3):
                    v = d[i, m] * xi[j] * xi[l] / s
                    v[xi2 == 0] = 0.0
                    v[nyq] = 0.0
                    G[i, j, l, m] = v
    return G


def project(A2, G):
    Ah = _sfft.fftn(A2, axes=(-3, -2, -1), workers=FFT_WORKERS)
    rh = np.einsum("ijklxyz,lkxyz->ijxyz", G, Ah)
    return _sfft.ifftn(rh, axes=(-3, -2, -1), workers=FFT_WORKERS).real


def build_C4(E, nu, I4s, II):
    mu = E / (2.0 * (1.0 + nu))
    lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
    return lam[None, None, None, None] * II + 2.0 * mu[None, None, None, None] * I4s


_S2 = np.sqrt(2.0)


def _landau_mandel(E_GL, E2, trE, trE2, lam, v1, v2, g1, g2, g3, g4, B, C, capping, Gt, w):
    sp_ = trE.shape
    alpha = lam + v1 * trE + 0.5 * g1 * trE ** 2 + g2 * trE2
    beta = 2.0 * (v2 + g2 * trE)
    gamma = 4.0 * g3

    def mand(T):
        v = np.empty((6,) + sp_)
        v[0], v[1], v[2] = T[0, 0], T[1, 1], T[2, 2]
        v[3], v[4], v[5] = _S2 * T[1, 2], _S2 * T[0, 2], _S2 * T[0, 1]
        return v

    e, f = mand(E_GL), mand(E2)
    iv = np.zeros((6,) + (1,) * 3)
    iv[0] = iv[1] = iv[2] = 1.0
    M = (alpha * (iv[:, None] * iv[None, :]) + beta * (iv[:, None] * e[None, :] + e[:, None] * iv[None, :])
         + gamma * (iv[:, None] * f[None, :] + f[:, None] * iv[None, :]) + (4.0 * g4) * (e[:, None] * e[None, :]))
    for p in range(6):
        M[p, p] += B
    E11, E22, E33 = E_GL[0, 0], E_GL[1, 1], E_GL[2, 2]
    E23, E13, E12 = E_GL[1, 2], E_GL[0, 2], E_GL[0, 1]
    KE = np.zeros_like(M)
    KE[0, 0], KE[1, 1], KE[2, 2] = 2 * E11, 2 * E22, 2 * E33
    KE[3, 3], KE[4, 4], KE[5, 5] = E22 + E33, E11 + E33, E11 + E22
    KE[0, 4] = KE[4, 0] = _S2 * E13
    KE[0, 5] = KE[5, 0] = _S2 * E12
    KE[1, 3] = KE[3, 1] = _S2 * E23
    KE[1, 5] = KE[5, 1] = _S2 * E12
    KE[2, 3] = KE[3, 2] = _S2 * E23
    KE[2, 4] = KE[4, 2] = _S2 * E13
    KE[3, 4] = KE[4, 3] = E12
    KE[3, 5] = KE[5, 3] = E13
    KE[4, 5] = KE[5, 4] = E23
    M += C * KE
    if capping:
        cw = 2.0 * Gt * w
        for p in range(6):
            M[p, p] += cw
        for a in range(3):
            for b in range(3):
                M[a, b] -= cw / 3.0
    C4 = np.empty((3, 3, 3, 3) + sp_)
    pairs = [((0, 0), 1.0), ((1, 1), 1.0), ((2, 2), 1.0), ((1, 2), 1 / _S2), ((0, 2), 1 / _S2), ((0, 1), 1 / _S2)]
    for p, ((i, j), wp) in enumerate(pairs):
        for q, ((k, l), wq) in enumerate(pairs):
            val = M[p, q] * (wp * wq)
            C4[i, j, k, l] = C4[j, i, k, l] = C4[i, j, l, k] = C4[j, i, l, k] = val
    return C4


def landau_constants(v):
    """Autoconversion of Landau constants given in GPa (|x| < 1e6, x != 0)
    (finite_strain_simulator.py:1687-1690)."""
    a = np.array(v, dtype=float, copy=True)
    mask = (np.abs(a) < 1e6) & (a != 0.0)
    a[mask] *= 1e9
    return a


def constitutive(F, C4, I2, I4, I4rt, Fp=None, model_type="svk", A_m=0.0, B_m=0.0, C_m=0.0, v1=0.0, v2=0.0,
                 v3=0.0, g1=0.0, g2=0.0, g3=0.0, g4=0.0, strain_capping_enabled=False, strain_capping_limit=None,
                 strain_capping_tangent_ratio=0.1, strain_capping_type="piecewise", strain_capping_smooth_power=1.0,
                 need_tangent=True, tangent_method="mandel"):
    """P and K4 (finite_strain_simulator.py:1647-1885)."""
    if Fp is not None:
        Fpi = invert_Fp(Fp)
        Fe = np.einsum("ijxyz,jkxyz->ikxyz", F, Fpi)
    else:
        Fe = F
    sh = Fe.shape[2:]
    if model_type == "landau":
        cs = [np.full(sh, float(landau_constants(x))) if np.isscalar(x) else landau_constants(x)
              for x in (v1, v2, v3, g1, g2, g3, g4)]
        v1a, v2a, v3a, g1a, g2a, g3a, g4a = cs
        lam = C4[0, 0, 1, 1]
        mu = C4[0, 1, 0, 1]
        E_GL = 0.5 * (np.einsum("jixyz,jkxyz->ikxyz", Fe, Fe) - I2)
        w = np.zeros(sh)
        if strain_capping_enabled:
            g4v = g4a.mean()
            muv = mu.mean()
            if strain_capping_limit is None or strain_capping_limit <= 0:
                Ecap = np.sqrt(0.2 * muv / np.abs(g4v)) if g4v < 0 else 999.0
            else:
                Ecap = strain_capping_limit
            tr0 = E_GL[0, 0] + E_GL[1, 1] + E_GL[2, 2]
            Ed = E_GL.copy()
            t3 = tr0 / 3.0
            for i in range(3):
                Ed[i, i] -= t3
            dd = np.zeros_like(tr0)
            for i in range(3):
                for j in range(3):
                    dd += Ed[i, j] * Ed[j, i]
            Eeq = np.sqrt(np.maximum(0.0, (2.0 / 3.0) * dd))
            if strain_capping_type == "smooth":
                f = np.ones_like(Eeq)
                nz = Eeq > 1e-12
                p = strain_capping_smooth_power
                ratio = Eeq[nz] / Ecap
                rp = np.minimum(100.0, ratio ** p)
                f[nz] = (Ecap / Eeq[nz]) * (np.tanh(rp)) ** (1.0 / p)
                w = np.tanh(np.minimum(100.0, (Eeq / Ecap) ** p)) ** 2
            else:
                capped = Eeq > Ecap
                f = np.ones_like(Eeq)
                np.divide(Ecap, Eeq, out=f, where=capped)
                w = capped.astype(float)
            Ec = Ed * f[None, None]
            for i in range(3):
                Ec[i, i] += t3
            E_orig = E_GL
            E_GL = Ec
        trE = E_GL[0, 0] + E_GL[1, 1] + E_GL[2, 2]
        trE2 = np.einsum("ijxyz,jixyz->xyz", E_GL, E_GL)
        trE3 = np.einsum("ijxyz,jkxyz,kixyz->xyz", E_GL, E_GL, E_GL)
        E2 = np.einsum("ijxyz,jkxyz->ikxyz", E_GL, E_GL)
        A = (lam * trE + 0.5 * v1a * trE ** 2 + v2a * trE2 + (1.0 / 6.0) * g1a * trE ** 3 + g2a * trE * trE2
             + (4.0 / 3.0) * g3a * trE3)
        B = 2.0 * (mu + v2a * trE + 0.5 * g2a * trE ** 2 + g4a * trE2)
        Cc = 4.0 * (v3a + g3a * trE)
        S = A[None, None] * I2 + B[None, None] * E_GL + Cc[None, None] * E2
        Gt = None
        if strain_capping_enabled:
            Gt = strain_capping_tangent_ratio * mu
            S += 2.0 * Gt[None, None] * (E_orig - E_GL) * w[None, None]
        C4e = None
        if need_tangent and tangent_method == "mandel":
            C4e = _landau_mandel(E_GL, E2, trE, trE2, lam, v1a, v2a, g1a, g2a, g3a, g4a, B, Cc,
                                 strain_capping_enabled, Gt, w)
        elif need_tangent:
            dA = ((lam + v1a * trE + 0.5 * g1a * trE ** 2 + g2a * trE2)[None, None] * I2
                  + 2.0 * (v2a + g2a * trE)[None, None] * E_GL + 4.0 * g3a[None, None] * E2)
            dB = 2.0 * (v2a + g2a * trE)[None, None] * I2 + 4.0 * g4a[None, None] * E_GL
            dC = 4.0 * g3a[None, None] * I2
            I4s = 0.5 * (I4 + I4rt)
            II = np.einsum("ijxyz,klxyz->ijklxyz", I2, I2)
            t1 = np.einsum("ijxyz,klxyz->ijklxyz", I2, dA)
            t2 = np.einsum("ijxyz,klxyz->ijklxyz", E_GL, dB)
            t3_ = np.einsum("ijxyz,klxyz->ijklxyz", E2, dC)
            t4 = B[None, None, None, None] * I4s
            p1 = np.einsum("ikxyz,jlxyz->ijklxyz", I2, E_GL) + np.einsum("ikxyz,jlxyz->ijklxyz", E_GL, I2)
            p2 = np.einsum("ilxyz,jkxyz->ijklxyz", I2, E_GL) + np.einsum("ilxyz,jkxyz->ijklxyz", E_GL, I2)
            C4e = t1 + t2 + t3_ + t4 + Cc[None, None, None, None] * 0.5 * (p1 + p2)
            if strain_capping_enabled:
                Kd = I4s - (1.0 / 3.0) * II
                C4e = C4e + (2.0 * Gt)[None, None, None, None] * Kd * w[None, None, None, None]
    elif model_type == "murnaghan":
        Aa = np.full(sh, float(A_m))
        Ba = np.full(sh, float(B_m))
        Ca = np.full(sh, float(C_m))
        lam = C4[0, 0, 1, 1]
        mu = C4[0, 1, 0, 1]
        E_GL = 0.5 * (np.einsum("jixyz,jkxyz->ikxyz", Fe, Fe) - I2)
        trE = E_GL[0, 0] + E_GL[1, 1] + E_GL[2, 2]
        trE2 = np.einsum("ijxyz,jixyz->xyz", E_GL, E_GL)
        E2 = np.einsum("ijxyz,jkxyz->ikxyz", E_GL, E_GL)
        S = ((lam * trE + Aa * trE ** 2 + Ba * trE2)[None, None] * I2 + 2.0 * (mu + Ba * trE)[None, None] * E_GL
             + Ca[None, None] * E2)
        C4e = None
        if need_tangent:
            I4s = 0.5 * (I4 + I4rt)
            II = np.einsum("ijxyz,klxyz->ijklxyz", I2, I2)
            t1 = (lam + 2.0 * Aa * trE)[None, None, None, None] * II
            t2 = (2.0 * Ba)[None, None, None, None] * (np.einsum("ijxyz,klxyz->ijklxyz", I2, E_GL)
                                                      + np.einsum("ijxyz,klxyz->ijklxyz", E_GL, I2))
            t3_ = (2.0 * (mu + Ba * trE))[None, None, None, None] * I4s
            p1 = np.einsum("ikxyz,jlxyz->ijklxyz", I2, E_GL) + np.einsum("ikxyz,jlxyz->ijklxyz", E_GL, I2)
            p2 = np.einsum("ilxyz,jkxyz->ijklxyz", I2, E_GL) + np.einsum("ilxyz,jkxyz->ijklxyz", E_GL, I2)
            C4e = t1 + t2 + t3_ + Ca[None, None, None, None] * 0.5 * (p1 + p2)
    else:
        E_GL = 0.5 * (np.einsum("jixyz,jkxyz->ikxyz", Fe, Fe) - I2)
        S = np.einsum("ijklxyz,lkxyz->ijxyz", C4, E_GL)
        C4e = C4
    if Fp is not None:
        P = np.einsum("ikxyz,klxyz,jlxyz->ijxyz", Fe, S, Fpi)
        if need_tangent:
            Sref = np.einsum("mkxyz,klxyz,jlxyz->mjxyz", Fpi, S, Fpi)
            A = np.einsum("klmnxyz,jlxyz->kjmnxyz", C4e, Fpi)
            B_ = np.einsum("kjmnxyz,bmxyz->kjbnxyz", A, Fpi)
            Cx = np.einsum("kjbnxyz,ikxyz->ijbnxyz", B_, Fe)
            t2 = np.einsum("ijbnxyz,anxyz->ijabxyz", Cx, Fe)
            t1 = np.einsum("bjxyz,ia->ijabxyz", Sref, np.eye(3))
            K4 = t1 + t2
        else:
            K4 = None
    else:
        P = np.einsum("ijxyz,jkxyz->ikxyz", Fe, S)
        if need_tangent:
            FC4 = np.einsum("ijxyz,jkmnxyz->ikmnxyz", Fe, C4e)
            Ft = np.einsum("ijxyz->jixyz", Fe)
            FC4Ft = np.einsum("ijklxyz,lmxyz->ijkmxyz", FC4, Ft)
            K4 = np.ascontiguousarray(FC4Ft.transpose(1, 0, 3, 2, 4, 5, 6))
            for k in range(3):
                K4[:, k, k] += S
        else:
            K4 = None
    return P, K4


def cauchy_from_P(P, F):
    J = (F[0, 0] * (F[1, 1] * F[2, 2] - F[1, 2] * F[2, 1]) - F[0, 1] * (F[1, 0] * F[2, 2] - F[1, 2] * F[2, 0])
         + F[0, 2] * (F[1, 0] * F[2, 1] - F[1, 1] * F[2, 0]))
    Js = np.where(np.abs(J) < 1e-14, 1e-14, J)
    return _dot22(P, _trans2(F)) / Js[None, None]


def build_bc(driving_component, eps_target_step, mixed_targets, F_bar_initial=None):
    """build_finite_strain_bc_3d (finite_strain_simulator.py:2852-2899)."""
    F_bar = np.eye(3) if F_bar_initial is None else F_bar_initial.copy()
    F_mask = np.zeros((3, 3), bool)
    P_target = np.zeros((3, 3))
    P_mask = np.zeros((3, 3), bool)
    i, j = driving_component
    F_bar[i, j] = 1.0 + eps_target_step if i == j else eps_target_step
    F_mask[i, j] = True
    for a in range(3):
        for b in range(3):
            if a == b or (a, b) == (i, j):
                continue
            F_bar[a, b] = 0.0
            F_mask[a, b] = True
    for k in range(3):
        if k == i and i == j:
            continue
        if (k, k) in mixed_targets:
            P_target[k, k] = mixed_targets[(k, k)]
            P_mask[k, k] = True
        else:
            F_bar[k, k] = 1.0
            F_mask[k, k] = True
    return F_bar, F_mask, P_target, P_mask


def al_step(F, F_bar, Ghat4, C4, I2, I4, I4rt, Fp=None, tol=1e-5, max_iter=200, model_type="svk", **mat):
    """_al_step_3d (finite_strain_simulator.py:2145-2242): reference medium L0 =
    voxel mean of C4; F + (F_bar - <F>); lam = P; per iteration dF = -G(lam -
    L0:F), the first beta in 1, 1/2, .., 1/16 whose F + beta dF (mean reset to
    F_bar) has det F_e > 1e-4 everywhere and a finite P; lam = P; stop when
    |G(P)| / (sqrt(size) (<|P|> + 1e-20)) < tol after the first iteration.
    Returns (F, P, iterations); ValueError otherwise."""
    nx, ny, nz = F.shape[2:]
    L0 = C4.mean(axis=(-3, -2, -1))
    L0f = np.einsum("ijkl,xyz->ijklxyz", L0, np.ones((nx, ny, nz)))
    Fpi = invert_Fp(Fp) if Fp is not None else None
    F = F + np.einsum("ij,xyz->ijxyz", F_bar - F.mean(axis=(2, 3, 4)), np.ones((nx, ny, nz)))
    P, _ = constitutive(F, C4, I2, I4, I4rt, Fp=Fp, model_type=model_type, need_tangent=False, **mat)
    lam = P.copy()
    for i in range(max_iter):
        dF = -project(lam - _ddot42(L0f, F), Ghat4)
        ok = False
        for beta in (1.0, 0.5, 0.25, 0.125, 0.0625):
            Ft = F + beta * dF
            Ft = Ft + np.einsum("ij,xyz->ijxyz", F_bar - Ft.mean(axis=(2, 3, 4)), np.ones((nx, ny, nz)))
            Fe = np.einsum("ijxyz,jkxyz->ikxyz", Ft, Fpi) if Fpi is not None else Ft
            Je = (Fe[0, 0] * (Fe[1, 1] * Fe[2, 2] - Fe[1, 2] * Fe[2, 1])
                  - Fe[0, 1] * (Fe[1, 0] * Fe[2, 2] - Fe[1, 2] * Fe[2, 0])
                  + Fe[0, 2] * (Fe[1, 0] * Fe[2, 1] - Fe[1, 1] * Fe[2, 0]))
            Pn, _ = constitutive(Ft, C4, I2, I4, I4rt, Fp=Fp, model_type=model_type, need_tangent=False, **mat)
            if not np.any(np.isnan(Pn)) and np.all(Je > 1e-4):
                F, P, ok = Ft, Pn, True
                break
        if not ok:
            break
        lam = P.copy()
        GP = project(P, Ghat4)
        res = np.linalg.norm(GP) / (np.sqrt(GP.size) * (np.mean(np.abs(P)) + 1e-20))
        if res < tol and i > 0:
            return F, P, i + 1
    GP = project(P, Ghat4)
    res = np.linalg.norm(GP) / (np.sqrt(GP.size) * (np.mean(np.abs(P)) + 1e-20))
    raise ValueError(f"AL solver did not converge or exploded. Final relative residual: {res:.2e}")


def frequencies(nx, ny, nz, pixel=1.0):
    """_get_frequencies_3d (finite_strain_simulator.py:1212-1229) with L = n pixel:
    2 pi fftfreq(n, L / n) per axis, entries equal to -pi n / L set to zero."""
    ns = (nx, ny, nz)
    Ls = tuple(n * pixel for n in ns)
    X = list(np.meshgrid(*[2.0 * np.pi * np.fft.fftfreq(n, d=L / n) for n, L in zip(ns, Ls)], indexing="ij"))
    for a in range(3):
        if ns[a] % 2 == 0:
            X[a][X[a] == -np.pi * ns[a] / Ls[a]] = 0.0
    return np.stack(X, axis=0)


def _fftn(A):
    return _sfft.fftn(A, axes=(-3, -2, -1), workers=FFT_WORKERS)


def _ifftn_real(A):
    return _sfft.ifftn(A, axes=(-3, -2, -1), workers=FFT_WORKERS).real


def inv3_field(A):
    """_invert_matrix_field_3d (:1250-1276): adjugate / det, det -> 1 where
    |det| < 1e-14, zero at k = 0."""
    det = (A[0, 0] * (A[1, 1] * A[2, 2] - A[1, 2] * A[2, 1]) - A[0, 1] * (A[1, 0] * A[2, 2] - A[1, 2] * A[2, 0])
           + A[0, 2] * (A[1, 0] * A[2, 1] - A[1, 1] * A[2, 0]))
    d = np.where(np.abs(det) < 1e-14, 1.0, det)
    M = np.empty_like(A)
    M[0, 0] = (A[1, 1] * A[2, 2] - A[1, 2] * A[2, 1]) / d
    M[0, 1] = -(A[0, 1] * A[2, 2] - A[0, 2] * A[2, 1]) / d
    M[0, 2] = (A[0, 1] * A[1, 2] - A[0, 2] * A[1, 1]) / d
    M[1, 0] = -(A[1, 0] * A[2, 2] - A[1, 2] * A[2, 0]) / d
    M[1, 1] = (A[0, 0] * A[2, 2] - A[0, 2] * A[2, 0]) / d
    M[1, 2] = -(A[0, 0] * A[1, 2] - A[0, 2] * A[1, 0]) / d
    M[2, 0] = (A[1, 0] * A[2, 1] - A[1, 1] * A[2, 0]) / d
    M[2, 1] = -(A[0, 0] * A[2, 1] - A[0, 1] * A[2, 0]) / d
    M[2, 2] = (A[0, 0] * A[1, 1] - A[0, 1] * A[1, 0]) / d
    M[:, :, 0, 0, 0] = 0.0
    return M


def u_from_F(F, F_bar, Xi):
    """_reconstruct_u_from_F_3d (:1294-1306)."""
    gh = _fftn(F - F_bar[:, :, None, None, None])
    xi2 = np.sum(Xi ** 2, axis=0)
    xi2 = np.where(xi2 < 1e-14, 1.0, xi2)
    uh = -1j * np.einsum("jxyz,ijxyz->ixyz", Xi, gh) / xi2
    uh[:, 0, 0, 0] = 0.0
    return _ifftn_real(uh)


def grad_u(u, Xi):
    """get_grad_u_3d (:1315-1318): grad_ij = ifft(i xi_j u_hat_i).real."""
    return _ifftn_real(1j * Xi[None] * _fftn(u)[:, None])


def dbfft_operators(Xi, K4, Kavg):
    """(A, M) of solve_dbfft_linear_system_3d (:1351-1379) on flattened full
    complex spectra: A u_hat = i xi . fft(K4 : ifft(i xi u_hat).real), zero at
    k = 0; M = (xi . Kavg . xi)^-1."""
    sh = Xi.shape[1:]
    Minv = inv3_field(np.einsum("jxyz,ijlk,lxyz->ikxyz", Xi, Kavg, Xi))

    def A(uf):
        uh = uf.reshape((3,) + sh)
        g = _ifftn_real(1j * Xi[None] * uh[:, None])
        dh = _fftn(_ddot42(K4, g))
        out = 1j * np.einsum("jxyz,ijxyz->ixyz", Xi, dh)
        out[:, 0, 0, 0] = 0.0
        return out.reshape(-1)

    def M(rf):
        return np.einsum("ijxyz,jxyz->ixyz", Minv, rf.reshape((3,) + sh)).reshape(-1)

    return A, M


def dbfft_step(F, F_bar, Xi, C4, I2, I4, I4rt, Fp=None, tol=1e-5, tol_CG=1e-6, max_iter=20, model_type="svk",
               stats=None, **mat):
    """_dbfft_step_3d (:1508-1620): u from F, then Newton on the displacement:
    F = F_bar + grad u; b = -i xi . fft(P); stop when |b| / (|P| + 1e-20) < tol
    after the first iteration; BiCGSTAB (SciPy, M = (xi.<K4>.xi)^-1, maxiter
    150) for du; line search alpha = 1 .. 2^-15 on det F_e > 1e-4 and a NaN-free
    P, whose P is reused by the next iteration.  Returns (F, P, iterations)."""
    n = F.shape[2:]
    Fbg = np.einsum("ij,xyz->ijxyz", F_bar, np.ones(n))
    u = u_from_F(F, F_bar, Xi)
    Fpi = invert_Fp(Fp) if Fp is not None else None
    carry = None
    rel = np.nan
    for i in range(max_iter):
        Fc = Fbg + grad_u(u, Xi)
        if carry is not None:
            P, K4 = carry, None
            carry = None
        else:
            P, K4 = constitutive(Fc, C4, I2, I4, I4rt, Fp=Fp, model_type=model_type, **mat)
        bh = -1j * np.einsum("jxyz,ijxyz->ixyz", Xi, _fftn(P))
        rel = np.linalg.norm(bh) / (np.linalg.norm(P) + 1e-20)
        if rel < tol and i > 0:
            return Fc, P, i + 1
        if K4 is None:
            P, K4 = constitutive(Fc, C4, I2, I4, I4rt, Fp=Fp, model_type=model_type, **mat)
        A, M = dbfft_operators(Xi, K4, K4.mean(axis=(-3, -2, -1)))
        nk = {"n": 0}

        def cb(xk):
            nk["n"] += 1

        N = bh.size
        duh, _ = sp.bicgstab(sp.LinearOperator((N, N), matvec=A, dtype="complex128"), bh.reshape(-1),
                             M=sp.LinearOperator((N, N), matvec=M, dtype="complex128"), rtol=tol_CG, maxiter=150,
                             callback=cb)
        if stats is not None:
            stats.setdefault("dbfft_krylov", []).append(nk["n"])
        if np.any(np.isnan(duh)) or np.any(np.isinf(duh)):
            raise ValueError("Linear solver returned NaN or Inf (BiCGSTAB breakdown).")
        du = _ifftn_real(duh.reshape((3,) + n))
        alpha = 1.0
        for _ in range(16):
            ut = u + alpha * du
            Ft = Fbg + grad_u(ut, Xi)
            Fe = np.einsum("ijxyz,jkxyz->ikxyz", Ft, Fpi) if Fpi is not None else Ft
            Je = (Fe[0, 0] * (Fe[1, 1] * Fe[2, 2] - Fe[1, 2] * Fe[2, 1])
                  - Fe[0, 1] * (Fe[1, 0] * Fe[2, 2] - Fe[1, 2] * Fe[2, 0])
                  + Fe[0, 2] * (Fe[1, 0] * Fe[2, 1] - Fe[1, 1] * Fe[2, 0]))
            if np.any(Je <= 1e-4) or np.any(np.isnan(Je)):
                alpha *= 0.5
                continue
            Pt, _ = constitutive(Ft, C4, I2, I4, I4rt, Fp=Fp, model_type=model_type, need_tangent=False, **mat)
            if np.any(np.isnan(Pt)):
                alpha *= 0.5
                continue
            u = ut
            carry = Pt
            break
        else:
            raise ValueError("Constitutive solver failed to converge in line search (potential negative Jacobian/NaN).")
    raise ValueError(f"DBFFT solver did not converge within {max_iter} iterations. Final relative residual: {rel:.2e}")


def solver_step(F, F_bar, Ghat4, C4, I2, I4, I4rt, Fp=None, driving_component=(0, 0), P_target=None, P_mask=None,
                E_avg=100e9, nu_avg=0.3, tol_NW=1e-5, tol_CG=1e-6, max_NW=50, tol_macro=1e6, max_iter_macro=20,
                model_type="svk", stats=None, solver="newton_cg", pixel=1.0, **mat):
    """finite_strain_solver_step_3d, newton_cg branch (:2967-3168), or its "al"
    / "dbfft" branch (al_step / dbfft_step in the same macro loop).  ``stats``
    (oracle extension) collects Newton and BiCGSTAB iteration counts (AL: "al"
    iterations; DBFFT: "dbfft" iterations and "dbfft_krylov" counts)."""
    nx, ny, nz = F.shape[2:]
    P_target = np.zeros((3, 3)) if P_target is None else P_target
    P_mask = np.zeros((3, 3), bool) if P_mask is None else P_mask
    F_start = F.copy()
    Fbc = F_bar.copy()
    i_drv, j_drv = driving_component
    upd = np.zeros((3, 3), bool)
    for a in range(3):
        if P_mask[a, a] and (a, a) != (i_drv, j_drv):
            upd[a, a] = True
    prev_F = prev_s = None
    max_err = 0.0
    res = None
    for it_mac in range(max_iter_macro):
        if solver == "al":
            F_curr, P_al, nit = al_step(F_start.copy(), Fbc, Ghat4, C4, I2, I4, I4rt, Fp=Fp, tol=tol_NW,
                                        max_iter=max_NW, model_type=model_type, **mat)
            if stats is not None:
                stats.setdefault("al", []).append(nit)
            cur = {"P": P_al, "K4": None}
        elif solver == "dbfft":
            F_curr, P_db, nit = dbfft_step(F_start.copy(), Fbc, frequencies(nx, ny, nz, pixel), C4, I2, I4, I4rt,
                                           Fp=Fp, tol=tol_NW, tol_CG=tol_CG, max_iter=max_NW, model_type=model_type,
                                           stats=stats, **mat)
            if stats is not None:
                stats.setdefault("dbfft", []).append(nit)
            cur = {"P": P_db, "K4": None}
        else:
            F_curr, cur = _newton(F_start, Fbc, Ghat4, C4, I2, I4, I4rt, Fp, tol_NW, tol_CG, max_NW, model_type,
                                  stats, mat)
        Sig = cauchy_from_P(cur["P"], F_curr)
        Sm = Sig.mean(axis=(2, 3, 4))
        res = (F_curr, cur["P"], Sig, cur["K4"])
        if not np.any(P_mask):
            break
        err = np.zeros((3, 3))
        err[P_mask] = P_target[P_mask] - Sm[P_mask]
        max_err = np.max(np.abs(err[P_mask]))
        if max_err < tol_macro:
            break
        dFm_ = np.clip((err - nu_avg * np.trace(err) * np.eye(3)) / E_avg, -0.01, 0.01)
        dsc = dFm_[upd]
        Fsc = Fbc[upd].copy()
        ssc = Sm[upd].copy()
        if prev_F is not None:
            dh = Fsc - prev_F
            dsh = ssc - prev_s
            with np.errstate(divide="ignore", invalid="ignore"):
                slope = dsh / dh
            esc = err[upd]
            ok = np.isfinite(slope) & (np.abs(dh) > 1e-14) & (slope > 0.01 * E_avg)
            dsc[ok] = np.clip(esc[ok] / slope[ok], -0.01, 0.01)
        prev_F, prev_s = Fsc, ssc
        Fbc[upd] += dsc
        F_start = F_curr
    else:
        if np.any(P_mask):
            raise ValueError(f"Outer BC loop did not converge after {max_iter_macro} iterations.")
    return res[0], res[1], res[2], res[3], Fbc


def _newton(F_start, Fbc, Ghat4, C4, I2, I4, I4rt, Fp, tol_NW, tol_CG, max_NW, model_type, stats, mat):
    """The newton_cg step of one macro iteration (:3040-3119)."""
    nx, ny, nz = F_start.shape[2:]
    DbarF = Fbc - F_start.mean(axis=(2, 3, 4))
    DbarF_grid = np.einsum("ij,xyz->ijxyz", DbarF, np.ones((nx, ny, nz)))
    cur = {}
    cur["P"], cur["K4"] = constitutive(F_start, C4, I2, I4, I4rt, Fp=Fp, model_type=model_type, **mat)

    def G_op(A2):
        return project(A2, Ghat4)

    def K_dF(x):
        return _trans2(_ddot42(cur["K4"], _trans2(x.reshape(3, 3, nx, ny, nz))))

    ncg = {"n": 0}

    def cb(xk):
        ncg["n"] += 1

    A_op = sp.LinearOperator(shape=(F_start.size,) * 2, matvec=lambda x: G_op(K_dF(x)).reshape(-1),
                             dtype="float64")
    F_curr = F_start.copy()
    Fpi = invert_Fp(Fp) if Fp is not None else None
    for i_NW in range(max_NW):
        rhs = (-G_op(K_dF(DbarF_grid.reshape(-1))).reshape(-1) if i_NW == 0 else -G_op(cur["P"]).reshape(-1))
        dFm, _ = sp.bicgstab(A_op, rhs, rtol=tol_CG, maxiter=150, callback=cb)
        dF = dFm.reshape(3, 3, nx, ny, nz)
        dX = (DbarF_grid + dF) if i_NW == 0 else dF
        alpha = 1.0
        for _ in range(16):
            Ft = F_curr + alpha * dX
            Fe = np.einsum("ijxyz,jkxyz->ikxyz", Ft, Fpi) if Fpi is not None else Ft
            Je = (Fe[0, 0] * (Fe[1, 1] * Fe[2, 2] - Fe[1, 2] * Fe[2, 1])
                  - Fe[0, 1] * (Fe[1, 0] * Fe[2, 2] - Fe[1, 2] * Fe[2, 0])
                  + Fe[0, 2] * (Fe[1, 0] * Fe[2, 1] - Fe[1, 1] * Fe[2, 0]))
            if np.any(Je <= 1e-4) or np.any(np.isnan(Je)):
                alpha *= 0.5
                continue
            Pt, Kt = constitutive(Ft, C4, I2, I4, I4rt, Fp=Fp, model_type=model_type, **mat)
            if np.any(np.isnan(Pt)) or np.any(np.isnan(Kt)):
                alpha *= 0.5
                continue
            F_curr, cur["P"], cur["K4"] = Ft, Pt, Kt
            break
        else:
            raise ValueError("Newton-CG line search failed to find a valid step.")
        if stats is not None:
            stats.setdefault("newton", []).append((i_NW, ncg["n"], alpha))
        if np.linalg.norm(dFm) / (np.linalg.norm(F_curr) + 1e-20) < tol_NW and i_NW > 0:
            break
    else:
        raise ValueError(f"Newton-CG inner solver did not converge within {max_NW} iterations.")
    return F_curr, cur



def simulation(E, nu, driving_component, eps_target, n_steps, mixed_targets=None, pixel=1.0, tol_NW=1e-5,
               tol_CG=1e-6, max_NW=20, tol_macro=1e6, max_iter_macro=20, model_type="svk", solver="al", **mat):
    """finite_strain_simulation_3d (finite_strain_simulator.py:3175-3413),
    store=False: load steps over linspace(0, eps_target, n_steps + 1)
    (:3270), build_bc from the previous step's F_bar (:3276-3279), one
    solver_step per load step (:3281-3298) -> (F_macro, Sig_macro, P_macro)
    as voxel means (:3301-3303)."""
    mixed_targets = {} if mixed_targets is None else mixed_targets
    nx, ny, nz = E.shape
    I2, I4, I4rt, I4s, II = identity_tensors(nx, ny, nz)
    C4 = build_C4(E, nu, I4s, II)
    G = build_ghat4(nx, ny, nz, nx * pixel, ny * pixel, nz * pixel)
    F = I2.copy()
    Fb_init = None
    Fm, Sm, Pm = [], [], []
    for eps_s in np.linspace(0.0, eps_target, n_steps + 1):
        Fb, _, Pt, Pmask = build_bc(driving_component, eps_s, mixed_targets, Fb_init)
        F, P, Sig, _, Fb = solver_step(F, Fb, G, C4, I2, I4, I4rt, driving_component=driving_component,
                                       P_target=Pt, P_mask=Pmask, E_avg=E.mean(), nu_avg=nu.mean(), tol_NW=tol_NW,
                                       tol_CG=tol_CG, max_NW=max_NW, tol_macro=tol_macro,
                                       max_iter_macro=max_iter_macro, model_type=model_type, solver=solver,
                                       pixel=pixel, **mat)
        Fb_init = Fb
        Fm.append(F.mean(axis=(2, 3, 4)))
        Sm.append(Sig.mean(axis=(2, 3, 4)))
        Pm.append(P.mean(axis=(2, 3, 4)))
    return np.array(Fm), np.array(Sm), np.array(Pm)
