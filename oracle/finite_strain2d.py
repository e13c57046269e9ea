"""Oracle: 2D finite-strain Newton-Krylov step (CPU restatement).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Follows mgkmc/finite_strain_simulator.py:
  :125-140      _make_identity_tensors_2d
  :147-199      build_ghat4_2d (zero at DC and on even-axis Nyquist lines)
  :206-225      _fft2_tensor / _ifft2_tensor / _project
  :233-261      build_C4_2d (plane strain: 3D lambda; plane stress: E nu / (1 - nu^2))
  :591-1091     constitutive_hyperelastic_2d: svk (both plane modes, F33 of the
                plane-stress E33 = -lam* tr E / (2 mu)), landau and murnaghan
                (plane strain: E33 = 0; plane stress: the 3D lam from lam*,
                E33 of S33 = 0 by landau's Newton (:762-811, uncapped) or
                murnaghan's quadratic root (:931-954), the tangent condensed
                on the 33 slot (:1026-1031)), the einsum tangent of the 3D law
                cut to the in-plane block, neo-Hookean in plane strain
                (:686-712: S = mu I + (-mu + lam ln J) C^-1 and its tangent)
                and in plane stress (:631-685: x = F33^2 of
                mu x + lam ln(x) / 2 = mu - lam ln J by Newton from x = 1,
                at most 10 steps, x >= 1e-10, stop when the grid's max |dx| <
                1e-12; S = mu I - mu x C^-1, tangent lam* C^-1 C^-1 +
                mu x (C^-1 C^-1 sym) with lam* = 2 lam mu x / max(1e-14,
                lam + 2 mu x), lam the 3D lam from lam*), capped landau in
                plane stress (:486-585 _solve_E33_capped_landau_2d: Newton on
                S33 of the capped law with the cap factor held fixed in the
                derivative, at most 20 steps, stop when the grid's max |dE33| <
                1e-9, FloatingPointError on a non-finite update), Fp, the
                push-forward to K4
  :1094-1114    cauchy_from_P (J = det F * F33)
  :2356-2572    finite_strain_solver_step_2d, solver "newton_cg", "dbfft" or "al"
  :1195-1210,1232-1248,1279-1313,1321-1349,1381-1507  the 2D DBFFT step (frequencies,
                matrix-field inverse, u from F, grad u, the preconditioned system)
  :2007-2140    _al_step_2d
  :2580-2850    finite_strain_simulation_2d (load stepping, store=False)
The Krylov solver is SciPy's own bicgstab, called as the reference calls it.

Pinned by (tests/test_oracle_fs2d.py): the homogeneous closed form, the 3D
oracle on the same problem embedded as an (nx, 1, ny) grid (plane strain:
identical fields and Newton / Krylov counts), the plane-stress SVK law against
its closed form (sigma_33 = 0 of the linearised law), the plane-stress landau
and murnaghan laws against the 3D oracle's law at F_33 = sqrt(1 + 2 E33)
(P_33 = 0 and the same in-plane P), central differences of P(F) against K4
(in plane stress: the derivative through E33(F), which pins the condensation),
equilibrium of the converged step, the reference's DBFFT-vs-NCG and Landau AL
simulation cases; plane-stress neo-Hookean against an independently written
3D neo-Hookean law at F_33 = sqrt(x) (S_33 = 0, same in-plane S) and central
differences of P(F) (its tangent is the exact derivative through x); capped
plane-stress landau: S_33 of the capped 3D law at the returned E33 within
what the 1e-9 step tolerance leaves, the uncapped root when the cap is out of
reach.
"""
import numpy as np
import scipy.fft as _sfft
import scipy.sparse.linalg as sp

from .finite_strain import landau_constants


def identity_tensors_2d(nx, ny):
    i = np.eye(2)
    ones = np.ones((nx, ny))
    I2 = np.einsum("ij,xy->ijxy", i, ones)
    I4 = np.einsum("ijkl,xy->ijklxy", np.einsum("il,jk", i, i), ones)
    I4rt = np.einsum("ijkl,xy->ijklxy", np.einsum("ik,jl", i, i), ones)
    I4s = 0.5 * (I4 + I4rt)
    II = np.einsum("ijxy,klxy->ijklxy", I2, I2)
    return I2, I4, I4rt, I4s, II


def build_ghat4_2d(nx, ny, Lx, Ly):
    X = np.meshgrid(np.fft.fftfreq(nx, d=Lx / nx), np.fft.fftfreq(ny, d=Ly / ny), indexing="ij")
    xi = np.stack(X, 0)
    xi2 = X[0] ** 2 + X[1] ** 2
    nyq = np.zeros((nx, ny), bool)
    if nx % 2 == 0:
        nyq |= X[0] == -nx / (2.0 * Lx)
    if ny % 2 == 0:
        nyq |= X[1] == -ny / (2.0 * Ly)
    s = np.where(xi2 == 0, 1.0, xi2)
    G = np.zeros((2, 2, 2, 2, nx, ny))
    d = np.eye(2)
    for i in range(2):
        for j in range(2):
            for l in range(2):
                for m in range(2):
                    v = d[i, m] * xi[j] * xi[l] / s
                    v[xi2 == 0] = 0.0
                    v[nyq] = 0.0
                    G[i, j, l, m] = v
    return G


def project2d(A2, G):
    Ah = _sfft.fftn(A2, axes=(-2, -1))
    return _sfft.ifftn(np.einsum("ijklxy,lkxy->ijxy", G, Ah), axes=(-2, -1)).real


def build_C4_2d(E, nu, I4s, II, plane_mode="plane_strain"):
    mu = E / (2.0 * (1.0 + nu))
    lam = E * nu / (1.0 - nu ** 2) if plane_mode == "plane_stress" else E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
    return lam[None, None, None, None] * II + 2.0 * mu[None, None, None, None] * I4s


def _trans2(A):
    return np.einsum("ijxy->jixy", A)


def _ddot42(A4, B2):
    return np.einsum("ijklxy,lkxy->ijxy", A4, B2)


def _inv2(A):
    det = A[0, 0] * A[1, 1] - A[0, 1] * A[1, 0]
    ds = np.where(np.abs(det) < 1e-14, 1e-14, det)
    out = np.empty_like(A)
    out[0, 0], out[1, 1] = A[1, 1] / ds, A[0, 0] / ds
    out[0, 1], out[1, 0] = -A[0, 1] / ds, -A[1, 0] / ds
    return out


def _i4s3(sh):
    i3 = np.eye(3)
    t = 0.5 * (np.einsum("ik,jl->ijkl", i3, i3) + np.einsum("il,jk->ijkl", i3, i3))
    return np.broadcast_to(t[..., None, None], (3, 3, 3, 3) + sh)


def _landau_arrays(mat, sh):
    """The seven landau constants as grid arrays, scalars broadcast, each element
    converted from GPa (finite_strain_simulator.py:723-733)."""
    out = []
    for k in ("v1", "v2", "v3", "g1", "g2", "g3", "g4"):
        x = mat.get(k, 0.0)
        a = np.full(sh, float(x)) if np.isscalar(x) else np.array(x, dtype=float).reshape(sh)
        out.append(landau_constants(a))
    return out


def _law3d(E_GL, lam, mu, model_type, mat, need_tangent):
    """S (3D) and dS/dE (3D einsum form) of the landau / murnaghan laws at a
    plane-strain E_GL (finite_strain_simulator.py:700-964)."""
    sh = lam.shape
    I3 = np.broadcast_to(np.eye(3)[..., None, None], (3, 3) + sh)
    II3 = np.einsum("ijxy,klxy->ijklxy", I3, I3)
    trE = E_GL[0, 0] + E_GL[1, 1] + E_GL[2, 2]
    trE2 = np.einsum("ijxy,jixy->xy", E_GL, E_GL)
    E2 = np.einsum("ijxy,jkxy->ikxy", E_GL, E_GL)
    KE = 0.5 * (np.einsum("ikxy,jlxy->ijklxy", I3, E_GL) + np.einsum("ikxy,jlxy->ijklxy", E_GL, I3)
                + np.einsum("ilxy,jkxy->ijklxy", I3, E_GL) + np.einsum("ilxy,jkxy->ijklxy", E_GL, I3))
    if model_type == "murnaghan":
        A, B, C = (np.full(sh, float(mat.get(k, 0.0))) for k in ("A_m", "B_m", "C_m"))
        S = ((lam * trE + A * trE ** 2 + B * trE2)[None, None] * I3 + 2.0 * (mu + B * trE)[None, None] * E_GL
             + C[None, None] * E2)
        if not need_tangent:
            return S, None, None
        C4 = ((lam + 2.0 * A * trE)[None, None, None, None] * II3
              + (2.0 * B)[None, None, None, None] * (np.einsum("ijxy,klxy->ijklxy", I3, E_GL)
                                                     + np.einsum("ijxy,klxy->ijklxy", E_GL, I3))
              + (2.0 * (mu + B * trE))[None, None, None, None] * _i4s3(sh) + C[None, None, None, None] * KE)
        return S, C4, None
    v1, v2, v3, g1, g2, g3, g4 = _landau_arrays(mat, sh)
    cap = mat.get("strain_capping_enabled", False)
    if cap:
        lim = mat.get("strain_capping_limit")
        g4m, mum = g4.mean(), mu.mean()
        Ecap = lim if (lim is not None and lim > 0) else (np.sqrt(0.2 * mum / abs(g4m)) if g4m < 0 else 999.0)
        tr3 = trE / 3.0
        Ed = E_GL.copy()
        for i in range(3):
            Ed[i, i] -= tr3
        Eeq = np.sqrt(np.maximum(0.0, (2.0 / 3.0) * np.einsum("ijxy,jixy->xy", Ed, Ed)))
        if mat.get("strain_capping_type", "piecewise") == "smooth":
            p = mat.get("strain_capping_smooth_power", 1.0)
            f = np.ones_like(Eeq)
            nz = Eeq > 1e-12
            f[nz] = (Ecap / Eeq[nz]) * np.tanh(np.minimum(100.0, (Eeq[nz] / Ecap) ** p)) ** (1.0 / p)
            w = np.tanh(np.minimum(100.0, (Eeq / Ecap) ** p)) ** 2
        else:
            capd = Eeq > Ecap
            f = np.ones_like(Eeq)
            np.divide(Ecap, Eeq, out=f, where=capd)
            w = capd.astype(float)
        Eorig = E_GL
        E_GL = Ed * f[None, None]
        for i in range(3):
            E_GL[i, i] += tr3
        trE = E_GL[0, 0] + E_GL[1, 1] + E_GL[2, 2]
        trE2 = np.einsum("ijxy,jixy->xy", E_GL, E_GL)
        E2 = np.einsum("ijxy,jkxy->ikxy", E_GL, E_GL)
        KE = 0.5 * (np.einsum("ikxy,jlxy->ijklxy", I3, E_GL) + np.einsum("ikxy,jlxy->ijklxy", E_GL, I3)
                    + np.einsum("ilxy,jkxy->ijklxy", I3, E_GL) + np.einsum("ilxy,jkxy->ijklxy", E_GL, I3))
    trE3 = np.einsum("ijxy,jkxy,kixy->xy", E_GL, E_GL, E_GL)
    A = lam * trE + 0.5 * v1 * trE ** 2 + v2 * trE2 + g1 * trE ** 3 / 6.0 + g2 * trE * trE2 + (4.0 / 3.0) * g3 * trE3
    B = 2.0 * (mu + v2 * trE + 0.5 * g2 * trE ** 2 + g4 * trE2)
    C = 4.0 * (v3 + g3 * trE)
    S = A[None, None] * I3 + B[None, None] * E_GL + C[None, None] * E2
    Gt = mat.get("strain_capping_tangent_ratio", 0.1) * mu if cap else None
    if cap:
        S = S + 2.0 * Gt[None, None] * (Eorig - E_GL) * w[None, None]
    if not need_tangent:
        return S, None, None
    dA = ((lam + v1 * trE + 0.5 * g1 * trE ** 2 + g2 * trE2)[None, None] * I3
          + 2.0 * (v2 + g2 * trE)[None, None] * E_GL + 4.0 * g3[None, None] * E2)
    dB = 2.0 * (v2 + g2 * trE)[None, None] * I3 + 4.0 * g4[None, None] * E_GL
    dC = 4.0 * g3[None, None] * I3
    C4 = (np.einsum("ijxy,klxy->ijklxy", I3, dA) + np.einsum("ijxy,klxy->ijklxy", E_GL, dB)
          + np.einsum("ijxy,klxy->ijklxy", E2, dC) + B[None, None, None, None] * _i4s3(sh)
          + C[None, None, None, None] * KE)
    if cap:
        C4 = C4 + (2.0 * Gt * w)[None, None, None, None] * (_i4s3(sh) - II3 / 3.0)
    return S, C4, None


def _plane_stress_e33(E2d, lam, mu, model_type, mat):
    """Out-of-plane E33 of a plane-stress state (lam the 3D lam).  landau,
    uncapped (finite_strain_simulator.py:762-811): Newton on S33(E33) = 0 from
    -(lam / (lam + 2 mu)) tr E_2d, C3333 -> 1e-12 where |C3333| < 1e-12, at
    most 20 iterations, stop when the grid's max |dE33| < 1e-12.  murnaghan
    (:931-954): a = A + 3B + C, b = lam + 2 mu + 2 (A + B) tr E_2d,
    c = lam tr E_2d + A tr^2 + B tr(E_2d^2); E33 = -c / max(1e-14, b) where
    |a| < 1e-12, else (-b + sign(b) sqrt(max(0, b^2 - 4ac))) / (2a), sign 0 -> 1."""
    t = E2d[0, 0] + E2d[1, 1]
    q = E2d[0, 0] ** 2 + E2d[1, 1] ** 2 + 2.0 * E2d[0, 1] * E2d[1, 0]
    sh = lam.shape
    if model_type == "murnaghan":
        A, B, C = (np.full(sh, float(mat.get(k, 0.0))) for k in ("A_m", "B_m", "C_m"))
        a = A + 3.0 * B + C
        b = lam + 2.0 * mu + 2.0 * (A + B) * t
        c = lam * t + A * t ** 2 + B * q
        e = np.zeros(sh)
        sm = np.abs(a) < 1e-12
        e[sm] = -c[sm] / np.maximum(1e-14, b[sm])
        disc = np.maximum(0.0, b[~sm] ** 2 - 4.0 * a[~sm] * c[~sm])
        sg = np.sign(b[~sm])
        sg[sg == 0] = 1.0
        e[~sm] = (-b[~sm] + sg * np.sqrt(disc)) / (2.0 * a[~sm])
        return e
    v1, v2, v3, g1, g2, g3, g4 = _landau_arrays(mat, sh)
    r = np.einsum("ijxy,jkxy,kixy->xy", E2d, E2d, E2d)
    e = -(lam / (lam + 2.0 * mu)) * t
    for _ in range(20):
        I1, I2, I3 = t + e, q + e ** 2, r + e ** 3
        A = lam * I1 + 0.5 * v1 * I1 ** 2 + v2 * I2 + g1 * I1 ** 3 / 6.0 + g2 * I1 * I2 + (4.0 / 3.0) * g3 * I3
        B = 2.0 * (mu + v2 * I1 + 0.5 * g2 * I1 ** 2 + g4 * I2)
        C = 4.0 * (v3 + g3 * I1)
        dA = lam + v1 * I1 + 0.5 * g1 * I1 ** 2 + g2 * I2 + 2.0 * (v2 + g2 * I1) * e + 4.0 * g3 * e ** 2
        dB = 2.0 * (v2 + g2 * I1) + 4.0 * g4 * e
        K = dA + dB * e + 4.0 * g3 * e ** 2 + B + 2.0 * C * e
        K = np.where(np.abs(K) < 1e-12, 1e-12, K)
        de = -(A + B * e + C * e ** 2) / K
        e = e + de
        if np.max(np.abs(de)) < 1e-12:
            break
    return e


def _plane_stress_e33_capped(E2d, lam, mu, mat):
    """E33 with S33 = 0 of the capped landau law (finite_strain_simulator.py:486-585):
    from the linear guess -(lam / (lam + 2 mu)) tr E_2d, the cap map of the
    deviator inside the residual (S33 of the landau law at the capped strain
    plus 2 G_t (E33 - c33) w), its derivative without the cap factor's chain
    rule, E33 += -S33 / (dS33 + 1e-12), at most 20 steps, stop when the grid's
    max |update| < 1e-9."""
    E11, E22, E12 = E2d[0, 0], E2d[1, 1], E2d[0, 1]
    sh = lam.shape
    v1, v2, v3, g1, g2, g3, g4 = _landau_arrays(mat, sh)
    lim = mat.get("strain_capping_limit")
    g4m, mum = g4.mean(), mu.mean()
    Ecap = lim if (lim is not None and lim > 0) else (np.sqrt(0.2 * mum / abs(g4m)) if g4m < 0 else 999.0)
    Gt = mat.get("strain_capping_tangent_ratio", 0.1) * mu
    smooth = mat.get("strain_capping_type", "piecewise") == "smooth"
    p = mat.get("strain_capping_smooth_power", 1.0)
    E33 = -(lam / (lam + 2.0 * mu)) * (E11 + E22)
    for _ in range(20):
        tr = E11 + E22 + E33
        t3 = tr / 3.0
        d11, d22, d33 = E11 - t3, E22 - t3, E33 - t3
        Eeq = np.sqrt(np.maximum(0.0, (2.0 / 3.0) * (d11 ** 2 + d22 ** 2 + d33 ** 2 + 2.0 * E12 ** 2)))
        if smooth:
            f = np.ones_like(Eeq)
            nz = Eeq > 1e-12
            f[nz] = (Ecap / Eeq[nz]) * np.tanh(np.minimum(100.0, (Eeq[nz] / Ecap) ** p)) ** (1.0 / p)
            w = np.tanh(np.minimum(100.0, (Eeq / Ecap) ** p)) ** 2
        else:
            capd = Eeq > Ecap
            f = np.ones_like(Eeq)
            np.divide(Ecap, Eeq, out=f, where=capd)
            w = capd.astype(float)
        c11, c22, c33, c12 = f * d11 + t3, f * d22 + t3, f * d33 + t3, f * E12
        I1 = tr
        I2 = c11 ** 2 + c22 ** 2 + c33 ** 2 + 2.0 * c12 ** 2
        t2 = c11 + c22
        det2 = c11 * c22 - c12 ** 2
        I3 = (t2 ** 3 - 3.0 * t2 * det2) + c33 ** 3
        A = lam * I1 + 0.5 * v1 * I1 ** 2 + v2 * I2 + g1 * I1 ** 3 / 6.0 + g2 * I1 * I2 + (4.0 / 3.0) * g3 * I3
        B = 2.0 * (mu + v2 * I1 + 0.5 * g2 * I1 ** 2 + g4 * I2)
        C = 4.0 * (v3 + g3 * I1)
        S33 = A + B * c33 + C * c33 ** 2 + 2.0 * Gt * (E33 - c33) * w
        dA = lam + v1 * I1 + 0.5 * g1 * I1 ** 2 + g2 * I2 + 2.0 * (v2 + g2 * I1) * c33 + 4.0 * g3 * c33 ** 2
        dB = 2.0 * (v2 + g2 * I1) + 4.0 * g4 * c33
        dS33 = dA + dB * c33 + B + 4.0 * g3 * c33 ** 2 + 2.0 * C * c33 + 2.0 * Gt * w
        delta = -S33 / (dS33 + 1e-12)
        if not np.all(np.isfinite(delta)):
            raise FloatingPointError("constitutive_hyperelastic_2d: non-finite capped E33 Newton update")
        E33 = E33 + delta
        if np.max(np.abs(delta)) < 1e-9:
            break
    return E33


def _nh_plane_stress_x(Je, lam3, mu):
    """x = F33^2 of the plane-stress neo-Hookean law (finite_strain_simulator.py:652-663)."""
    x = np.ones_like(Je)
    target = mu - lam3 * np.log(np.maximum(1e-14, Je))
    for _ in range(10):
        f = mu * x + 0.5 * lam3 * np.log(x) - target
        df = mu + 0.5 * lam3 / x
        dx = -f / df
        x = np.maximum(1e-10, x + dx)
        if np.max(np.abs(dx)) < 1e-12:
            break
    return x


def constitutive_2d(F, C4, I2, I4, I4rt, Fp=None, model_type="svk", plane_mode="plane_strain", need_tangent=True,
                    **mat):
    """(P, K4, F33) of constitutive_hyperelastic_2d (:591-1091) for the laws
    listed in the module docstring."""
    if model_type not in ("svk", "landau", "murnaghan", "neo_hookean"):
        raise NotImplementedError(f"{model_type} / {plane_mode}")
    Fpi = _inv2(Fp) if Fp is not None else None
    Fe = np.einsum("ijxy,jkxy->ikxy", F, Fpi) if Fpi is not None else F
    lam, mu = C4[0, 0, 1, 1], C4[0, 1, 0, 1]
    E2d = 0.5 * (np.einsum("jixy,jkxy->ikxy", Fe, Fe) - I2)
    if model_type == "neo_hookean" and plane_mode == "plane_stress":
        lam3 = (2.0 * mu * lam) / np.maximum(1e-14, 2.0 * mu - lam)
        Ce = np.einsum("jixy,jkxy->ikxy", Fe, Fe)
        x = _nh_plane_stress_x(Fe[0, 0] * Fe[1, 1] - Fe[0, 1] * Fe[1, 0], lam3, mu)
        F33 = np.sqrt(np.maximum(1e-12, x))
        Ci = _inv2(Ce)
        S = mu[None, None] * I2 + (-mu * x)[None, None] * Ci
        C4e = None
        if need_tangent:
            ct = mu * x
            ls = (2.0 * lam3 * ct) / np.maximum(1e-14, lam3 + 2.0 * ct)
            C4e = (ls[None, None, None, None] * np.einsum("ijxy,klxy->ijklxy", Ci, Ci)
                   + ct[None, None, None, None] * (np.einsum("ikxy,jlxy->ijklxy", Ci, Ci)
                                                   + np.einsum("ilxy,jkxy->ijklxy", Ci, Ci)))
    elif model_type == "neo_hookean":
        Ce = np.einsum("jixy,jkxy->ikxy", Fe, Fe)
        lnJ = np.log(np.maximum(1e-14, Fe[0, 0] * Fe[1, 1] - Fe[0, 1] * Fe[1, 0]))
        Ci = _inv2(Ce)
        S = mu[None, None] * I2 + (-mu + lam * lnJ)[None, None] * Ci
        C4e = None
        if need_tangent:
            C4e = (lam[None, None, None, None] * np.einsum("ijxy,klxy->ijklxy", Ci, Ci)
                   + (mu - lam * lnJ)[None, None, None, None] * (np.einsum("ikxy,jlxy->ijklxy", Ci, Ci)
                                                                 + np.einsum("ilxy,jkxy->ijklxy", Ci, Ci)))
        F33 = np.ones_like(lam)
    elif model_type == "svk":
        S = np.einsum("ijklxy,lkxy->ijxy", C4, E2d)
        C4e = C4
        if plane_mode == "plane_stress":
            F33 = np.sqrt(np.maximum(1e-12, 1.0 + 2.0 * (-(lam / (2.0 * mu)) * (E2d[0, 0] + E2d[1, 1]))))
        else:
            F33 = np.ones_like(lam)
    else:
        EG = np.zeros((3, 3) + lam.shape)
        EG[:2, :2] = E2d
        F33 = np.ones_like(lam)
        if plane_mode == "plane_stress":
            lam = (2.0 * mu * lam) / np.maximum(1e-14, 2.0 * mu - lam)
            if model_type == "landau" and mat.get("strain_capping_enabled"):
                EG[2, 2] = _plane_stress_e33_capped(E2d, lam, mu, mat)
            else:
                EG[2, 2] = _plane_stress_e33(E2d, lam, mu, model_type, mat)
            F33 = np.sqrt(np.maximum(1e-12, 1.0 + 2.0 * EG[2, 2]))
        S3, C43, _ = _law3d(EG, lam, mu, model_type, mat, need_tangent)
        S = S3[:2, :2]
        C4e = None
        if need_tangent and plane_mode == "plane_stress":
            C2222 = np.maximum(1e-14, C43[2, 2, 2, 2])
            C22 = C43[:, :, 2, 2]
            C4e = C43[:2, :2, :2, :2] - C22[:2, :2, None, None] * C22[None, None, :2, :2] / C2222
        elif need_tangent:
            C4e = C43[:2, :2, :2, :2]
    if Fpi is not None:
        P = np.einsum("ikxy,klxy,jlxy->ijxy", Fe, S, Fpi)
        if not need_tangent:
            return P, None, F33
        Sref = np.einsum("mkxy,klxy,jlxy->mjxy", Fpi, S, Fpi)
        A = np.einsum("klmnxy,jlxy->kjmnxy", C4e, Fpi)
        B = np.einsum("kjmnxy,bmxy->kjbnxy", A, Fpi)
        Cc = np.einsum("kjbnxy,ikxy->ijbnxy", B, Fe)
        K4 = np.einsum("bjxy,ia->ijabxy", Sref, np.eye(2)) + np.einsum("ijbnxy,anxy->ijabxy", Cc, Fe)
        return P, K4, F33
    P = np.einsum("ijxy,jkxy->ikxy", Fe, S)
    if not need_tangent:
        return P, None, F33
    FC4 = np.einsum("ijxy,jkmnxy->ikmnxy", Fe, C4e)
    FC4Ft = np.einsum("ijklxy,lmxy->ijkmxy", FC4, _trans2(Fe))
    K4 = np.ascontiguousarray(FC4Ft.transpose(1, 0, 3, 2, 4, 5))
    for k in range(2):
        K4[:, k, k] += S
    return P, K4, F33


def frequencies_2d(nx, ny, pixel=1.0):
    """_get_frequencies_2d (:1195-1210): 2 pi fftfreq(n, L / n), L = n pixel,
    entries equal to -pi n / L set to zero."""
    ns, Ls = (nx, ny), (nx * pixel, ny * pixel)
    X = list(np.meshgrid(*[2.0 * np.pi * np.fft.fftfreq(n, d=L / n) for n, L in zip(ns, Ls)], indexing="ij"))
    for a in range(2):
        if ns[a] % 2 == 0:
            X[a][X[a] == -np.pi * ns[a] / Ls[a]] = 0.0
    return np.stack(X, 0)


def dbfft_step_2d(F, F_bar, Xi, C4, I2, I4, I4rt, Fp=None, tol=1e-5, tol_CG=1e-6, max_iter=20, model_type="svk",
                  plane_mode="plane_strain", stats=None, **mat):
    """_dbfft_step_2d (:1381-1507) with solve_dbfft_linear_system_2d
    (:1321-1349): u from F; Newton on u with b = -i xi . fft(P), stop on
    |b| / (|P| + 1e-20) < tol after the first iteration, SciPy BiCGSTAB with
    M = (xi . <K4> . xi)^-1 (det -> 1 below 1e-14, zero at k = 0), line search
    1 .. 2^-15 on det F_e > 1e-4 and a NaN-free P (reused with its F33).
    Returns (F, P, F33, iterations)."""
    n = F.shape[2:]
    fft = lambda a: _sfft.fftn(a, axes=(-2, -1))
    ifft = lambda a: _sfft.ifftn(a, axes=(-2, -1)).real
    Fbg = np.einsum("ij,xy->ijxy", F_bar, np.ones(n))
    gh = fft(F - Fbg)
    xi2 = np.sum(Xi ** 2, axis=0)
    uh = -1j * np.einsum("jxy,ijxy->ixy", Xi, gh) / np.where(xi2 < 1e-14, 1.0, xi2)
    uh[:, 0, 0] = 0.0
    u = ifft(uh)
    grad = lambda w: ifft(1j * Xi[None] * fft(w)[:, None])
    Fpi = _inv2(Fp) if Fp is not None else None
    carry = None
    rel = np.nan
    for i in range(max_iter):
        Fc = Fbg + grad(u)
        if carry is not None:
            (P, F33), K4 = carry, None
            carry = None
        else:
            P, K4, F33 = constitutive_2d(Fc, C4, I2, I4, I4rt, Fp=Fp, model_type=model_type, plane_mode=plane_mode,
                                         **mat)
        bh = -1j * np.einsum("jxy,ijxy->ixy", Xi, fft(P))
        rel = np.linalg.norm(bh) / (np.linalg.norm(P) + 1e-20)
        if rel < tol and i > 0:
            return Fc, P, F33, i + 1
        if K4 is None:
            P, K4, F33 = constitutive_2d(Fc, C4, I2, I4, I4rt, Fp=Fp, model_type=model_type, plane_mode=plane_mode,
                                         **mat)
        A = np.einsum("jxy,ijlk,lxy->ikxy", Xi, K4.mean(axis=(-2, -1)), Xi)
        det = A[0, 0] * A[1, 1] - A[0, 1] * A[1, 0]
        ds = np.where(np.abs(det) < 1e-14, 1.0, det)
        Mi = np.empty_like(A)
        Mi[0, 0], Mi[1, 1] = A[1, 1] / ds, A[0, 0] / ds
        Mi[0, 1], Mi[1, 0] = -A[0, 1] / ds, -A[1, 0] / ds
        Mi[:, :, 0, 0] = 0.0

        def A_op(uf):
            g = ifft(1j * Xi[None] * uf.reshape((2,) + n)[:, None])
            out = 1j * np.einsum("jxy,ijxy->ixy", Xi, fft(_ddot42(K4, g)))
            out[:, 0, 0] = 0.0
            return out.reshape(-1)

        N = bh.size
        nk = {"n": 0}

        def cb(xk):
            nk["n"] += 1

        duh, _ = sp.bicgstab(sp.LinearOperator((N, N), matvec=A_op, dtype="complex128"), bh.reshape(-1),
                             M=sp.LinearOperator((N, N), dtype="complex128",
                                                 matvec=lambda r: np.einsum("ijxy,jxy->ixy", Mi,
                                                                            r.reshape((2,) + n)).reshape(-1)),
                             rtol=tol_CG, maxiter=150, callback=cb)
        if stats is not None:
            stats.setdefault("dbfft_krylov", []).append(nk["n"])
        if np.any(np.isnan(duh)) or np.any(np.isinf(duh)):
            raise ValueError("Linear solver returned NaN or Inf (BiCGSTAB breakdown).")
        du = ifft(duh.reshape((2,) + n))
        alpha = 1.0
        for _ in range(16):
            ut = u + alpha * du
            Ft = Fbg + grad(ut)
            Fe = np.einsum("ijxy,jkxy->ikxy", Ft, Fpi) if Fpi is not None else Ft
            Je = Fe[0, 0] * Fe[1, 1] - Fe[0, 1] * Fe[1, 0]
            if np.any(Je <= 1e-4) or np.any(np.isnan(Je)):
                alpha *= 0.5
                continue
            Pt, _, F33t = constitutive_2d(Ft, C4, I2, I4, I4rt, Fp=Fp, model_type=model_type, plane_mode=plane_mode,
                                          need_tangent=False, **mat)
            if np.any(np.isnan(Pt)):
                alpha *= 0.5
                continue
            u = ut
            carry = (Pt, F33t)
            break
        else:
            raise ValueError("Constitutive solver failed to converge in line search (potential negative Jacobian/NaN).")
    raise ValueError(f"DBFFT solver did not converge within {max_iter} iterations. Final relative residual: {rel:.2e}")


def al_step_2d(F, F_bar, Ghat4, C4, I2, I4, I4rt, Fp=None, tol=1e-5, max_iter=200, model_type="svk",
               plane_mode="plane_strain", **mat):
    """_al_step_2d (:2007-2140): L0 = grid mean of C4; F + (F_bar - <F>); lam
    = P; per iteration dF = -G(lam - L0:F), the first beta in 1 .. 1/16 whose
    trial (mean reset to F_bar) has det F_e > 1e-4 and a NaN-free P; lam = P;
    stop when |G(P)| / (sqrt(size) (<|P|> + 1e-20)) < tol after the first
    iteration.  Returns (F, P, F33, iterations); ValueError otherwise."""
    n = F.shape[2:]
    L0 = np.einsum("ijkl,xy->ijklxy", C4.mean(axis=(-2, -1)), np.ones(n))
    Fpi = _inv2(Fp) if Fp is not None else None
    F = F + np.einsum("ij,xy->ijxy", F_bar - F.mean(axis=(2, 3)), np.ones(n))
    P, _, F33 = constitutive_2d(F, C4, I2, I4, I4rt, Fp=Fp, model_type=model_type, plane_mode=plane_mode,
                                need_tangent=False, **mat)
    lam = P.copy()
    for i in range(max_iter):
        dF = -project2d(lam - _ddot42(L0, F), Ghat4)
        ok = False
        for beta in (1.0, 0.5, 0.25, 0.125, 0.0625):
            Ft = F + beta * dF
            Ft = Ft + np.einsum("ij,xy->ijxy", F_bar - Ft.mean(axis=(2, 3)), np.ones(n))
            Fe = np.einsum("ijxy,jkxy->ikxy", Ft, Fpi) if Fpi is not None else Ft
            Je = Fe[0, 0] * Fe[1, 1] - Fe[0, 1] * Fe[1, 0]
            Pn, _, F33n = constitutive_2d(Ft, C4, I2, I4, I4rt, Fp=Fp, model_type=model_type, plane_mode=plane_mode,
                                          need_tangent=False, **mat)
            if not np.any(np.isnan(Pn)) and np.all(Je > 1e-4):
                F, P, F33, ok = Ft, Pn, F33n, True
                break
        if not ok:
            break
        lam = P.copy()
        GP = project2d(P, Ghat4)
        res = np.linalg.norm(GP) / (np.sqrt(GP.size) * (np.mean(np.abs(P)) + 1e-20))
        if res < tol and i > 0:
            return F, P, F33, i + 1
    GP = project2d(P, Ghat4)
    res = np.linalg.norm(GP) / (np.sqrt(GP.size) * (np.mean(np.abs(P)) + 1e-20))
    raise ValueError(f"AL solver did not converge or exploded. Final relative residual: {res:.2e}")


def cauchy_2d(P, F, F33=None):
    """cauchy_from_P (:1094-1114): sigma = P F^T / (det F * F33)."""
    J = F[0, 0] * F[1, 1] - F[0, 1] * F[1, 0]
    if F33 is not None:
        J = J * F33
    J = np.where(np.abs(J) < 1e-14, 1e-14, J)
    return np.einsum("ijxy,kjxy->ikxy", P, F) / J[None, None]


def solver_step_2d(F, F_bar, Ghat4, C4, I2, I4, I4rt, Fp=None, driving_component=(0, 0), P_target=None, P_mask=None,
                   E_avg=100e9, nu_avg=0.3, tol_NW=1e-5, tol_CG=1e-6, max_NW=50, tol_macro=1e6, max_iter_macro=20,
                   model_type="svk", plane_mode="plane_strain", stats=None, solver="newton_cg", pixel=1.0, **mat):
    """finite_strain_solver_step_2d, newton_cg or dbfft branch (:2356-2572) ->
    (F, P, Sig, K4, F_bar).  ``stats`` collects (Newton iteration, Krylov
    iterations so far in this macro iteration, alpha) per Newton step under
    "newton", as the 3D oracle does."""
    nx, ny = F.shape[2:]
    P_target = np.zeros((2, 2)) if P_target is None else P_target
    P_mask = np.zeros((2, 2), bool) if P_mask is None else P_mask
    F_start = F.copy()
    Fbc = np.array(F_bar, dtype=float)
    i_drv, j_drv = driving_component
    upd = np.zeros((2, 2), bool)
    for a in range(2):
        if P_mask[a, a] and (a, a) != (i_drv, j_drv):
            upd[a, a] = True
    prev_F = prev_s = None
    Fpi = _inv2(Fp) if Fp is not None else None
    res = None
    max_err = 0.0
    for _ in range(max_iter_macro):
        if solver == "al":
            F_curr, P_al, F33_al, nit = al_step_2d(F_start.copy(), Fbc, Ghat4, C4, I2, I4, I4rt, Fp=Fp, tol=tol_NW,
                                                   max_iter=max_NW, model_type=model_type, plane_mode=plane_mode,
                                                   **mat)
            if stats is not None:
                stats.setdefault("al", []).append(nit)
            cur = {"P": P_al, "K4": None, "F33": F33_al}
        elif solver == "dbfft":
            F_curr, P_db, F33_db, nit = dbfft_step_2d(F_start.copy(), Fbc, frequencies_2d(nx, ny, pixel), C4, I2, I4,
                                                      I4rt, Fp=Fp, tol=tol_NW, tol_CG=tol_CG, max_iter=max_NW,
                                                      model_type=model_type, plane_mode=plane_mode, stats=stats, **mat)
            if stats is not None:
                stats.setdefault("dbfft", []).append(nit)
            cur = {"P": P_db, "K4": None, "F33": F33_db}
        else:
            DbarF = Fbc - F_start.mean(axis=(2, 3))
            Dg = np.einsum("ij,xy->ijxy", DbarF, np.ones((nx, ny)))
            cur = {}
            cur["P"], cur["K4"], cur["F33"] = constitutive_2d(F_start, C4, I2, I4, I4rt, Fp=Fp, model_type=model_type,
                                                              plane_mode=plane_mode, **mat)
            G_op = lambda A2: project2d(A2, Ghat4)
            K_dF = lambda x: _trans2(_ddot42(cur["K4"], _trans2(x.reshape(2, 2, nx, ny))))
            nk = {"n": 0}

            def cb(xk):
                nk["n"] += 1

            A_op = sp.LinearOperator((F_start.size,) * 2, matvec=lambda x: G_op(K_dF(x)).reshape(-1), dtype="float64")
            F_curr = F_start.copy()
            for i_NW in range(max_NW):
                rhs = -G_op(K_dF(Dg.reshape(-1))).reshape(-1) if i_NW == 0 else -G_op(cur["P"]).reshape(-1)
                dFm, _ = sp.bicgstab(A_op, rhs, rtol=tol_CG, maxiter=150, callback=cb)
                dF = dFm.reshape(2, 2, nx, ny)
                dX = Dg + dF if i_NW == 0 else dF
                alpha = 1.0
                for _ in range(16):
                    Ft = F_curr + alpha * dX
                    Fe = np.einsum("ijxy,jkxy->ikxy", Ft, Fpi) if Fpi is not None else Ft
                    Je = Fe[0, 0] * Fe[1, 1] - Fe[0, 1] * Fe[1, 0]
                    if np.any(Je <= 1e-4) or np.any(np.isnan(Je)):
                        alpha *= 0.5
                        continue
                    Pt, Kt, F33t = constitutive_2d(Ft, C4, I2, I4, I4rt, Fp=Fp, model_type=model_type,
                                                   plane_mode=plane_mode, **mat)
                    if np.any(np.isnan(Pt)) or np.any(np.isnan(Kt)):
                        alpha *= 0.5
                        continue
                    F_curr, cur["P"], cur["K4"], cur["F33"] = Ft, Pt, Kt, F33t
                    break
                else:
                    raise ValueError("Newton-CG line search failed to find a valid step.")
                if stats is not None:
                    stats.setdefault("newton", []).append((i_NW, nk["n"], alpha))
                if np.linalg.norm(dFm) / (np.linalg.norm(F_curr) + 1e-20) < tol_NW and i_NW > 0:
                    break
            else:
                raise ValueError(f"Newton-CG inner solver did not converge within {max_NW} iterations.")
        Sig = cauchy_2d(cur["P"], F_curr, F33=cur["F33"])
        Sm = Sig.mean(axis=(2, 3))
        res = (F_curr, cur["P"], Sig, cur["K4"])
        if not np.any(P_mask):
            break
        err = np.zeros((2, 2))
        err[P_mask] = P_target[P_mask] - Sm[P_mask]
        max_err = np.max(np.abs(err[P_mask]))
        if max_err < tol_macro:
            break
        dFm_ = np.clip((err - nu_avg * np.trace(err) * np.eye(2)) / E_avg, -0.01, 0.01)
        dsc = dFm_[upd]
        Fsc, ssc = Fbc[upd].copy(), Sm[upd].copy()
        if prev_F is not None:
            dh, dsh = Fsc - prev_F, ssc - prev_s
            with np.errstate(divide="ignore", invalid="ignore"):
                slope = dsh / dh
            ok = np.isfinite(slope) & (np.abs(dh) > 1e-14) & (slope > 0.01 * E_avg)
            dsc[ok] = np.clip(err[upd][ok] / slope[ok], -0.01, 0.01)
        prev_F, prev_s = Fsc, ssc
        Fbc[upd] += dsc
        F_start = F_curr
    else:
        if np.any(P_mask):
            raise ValueError(f"Outer BC loop did not converge after {max_iter_macro} iterations.")
    return res[0], res[1], res[2], res[3], Fbc


def build_bc_2d(driving_component, eps, mixed_targets, F_bar_initial=None):
    """build_finite_strain_bc (:1912-1974) -> (F_bar, P_target, P_mask): the
    driven component (1 + eps on the diagonal, eps off it), other shears 0,
    undriven diagonals stress-controlled when targeted, else 1."""
    Fb = np.eye(2) if F_bar_initial is None else np.array(F_bar_initial, dtype=float)
    Pt, Pm = np.zeros((2, 2)), np.zeros((2, 2), bool)
    i, j = driving_component
    Fb[i, j] = 1.0 + eps if i == j else eps
    for a in range(2):
        for b in range(2):
            if a != b and (a, b) != (i, j):
                Fb[a, b] = 0.0
    for k in range(2):
        if k == i and i == j:
            continue
        if (k, k) in mixed_targets:
            Pt[k, k] = mixed_targets[(k, k)]
            Pm[k, k] = True
        else:
            Fb[k, k] = 1.0
    return Fb, Pt, Pm


def simulation_2d(E, nu, driving_component, eps_target, n_steps, mixed_targets=None, plane_mode="plane_strain",
                  pixel=1.0, tol_NW=1e-5, tol_CG=1e-6, max_NW=20, tol_macro=1e6, max_iter_macro=20,
                  model_type="svk", **mat):
    """finite_strain_simulation_2d (:2580-2845), store=False: load steps over
    linspace(0, eps_target, n_steps + 1) with build_finite_strain_bc (:1912-1974)
    from the previous F_bar -> (F_macro, Sig_macro, P_macro)."""
    mixed_targets = {} if mixed_targets is None else mixed_targets
    nx, ny = E.shape
    I2, I4, I4rt, I4s, II = identity_tensors_2d(nx, ny)
    C4 = build_C4_2d(E, nu, I4s, II, plane_mode)
    G = build_ghat4_2d(nx, ny, nx * pixel, ny * pixel)
    F = I2.copy()
    Fb_init = None
    Fm, Sm, Pm = [], [], []
    for eps_s in np.linspace(0.0, eps_target, n_steps + 1):
        Fb, Pt, Pmask = build_bc_2d(driving_component, eps_s, mixed_targets, Fb_init)
        F, P, Sig, _, Fb = solver_step_2d(F, Fb, G, C4, I2, I4, I4rt, driving_component=driving_component,
                                          P_target=Pt, P_mask=Pmask, E_avg=E.mean(), nu_avg=nu.mean(),
                                          tol_NW=tol_NW, tol_CG=tol_CG, max_NW=max_NW, tol_macro=tol_macro,
                                          max_iter_macro=max_iter_macro, model_type=model_type,
                                          plane_mode=plane_mode, **mat)
        Fb_init = Fb
        Fm.append(F.mean(axis=(2, 3)))
        Sm.append(Sig.mean(axis=(2, 3)))
        Pm.append(P.mean(axis=(2, 3)))
    return np.array(Fm), np.array(Sm), np.array(Pm)
