"""Oracle: small-strain Lippmann-Schwinger fixed point (CPU restatement).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Follows, step for step:
  mgkmc/elasticity.py:3-10      compute_lame_2d
  mgkmc/elasticity.py:12-18     stress_from_strain_2d
  mgkmc/elasticity.py:20-53     green_operator_2d
  mgkmc/elasticity.py:55-59     compute_lame_3d
  mgkmc/elasticity.py:62-67     stress_from_strain_3d
  mgkmc/elasticity.py:70-105    green_operator_3d
  mgkmc/fft.py:9-31             wave vectors, fft_field / ifft_field (ortho)
  mgkmc/linear_elastic_simulator.py:12-88    spectral_solver_2d
  mgkmc/linear_elastic_simulator.py:274-376  spectral_solver_3d
  mgkmc/stz/update_fft.py:4-22  update_stress_fft_full
  mgkmc/elasticity.py:641-810,868-943 + linear_elastic_simulator.py:614-950
                                secant-degradation and Rose laws and LS loops

The reference uses pyfftw (FFTW3) complex fftn with norm="ortho"; this file
uses scipy.fft.fftn/ifftn (same DFT definition and normalisation; FFTW is not
installable here).  ``FFT_WORKERS`` sets scipy's thread count, the stand-in
for pyfftw.config.NUM_THREADS.

Pinned by (tests/test_oracle_ls.py): homogeneous closed form
sigma_xx = (lam+2mu)*eps (tests/test_barrier_physics.py:49-98 of the
reference), the Eshelby sign KAT (tests/test_plasticity.py:10-61), the exact
laminate solution (uniform sigma_xx, strain ratio = inverse stiffness ratio),
Gamma projector identities (Gamma:C0:Gamma = Gamma, Gamma(0) = 0), and the
two conditions that fix the converged field independently of the iteration:
divergence-free stress and compatible strain with mean eps_bar (checked in
Fourier space away from Nyquist planes).
"""
import numpy as np
import scipy.fft as _sfft

FFT_WORKERS = 1


def compute_lame_2d(E, nu, plane_mode="plane_strain"):
    """Lame pair for 2D (elasticity.py:3-10)."""
    mu = E / (2 * (1 + nu))
    if plane_mode == "plane_stress":
        lam = E * nu / (1 - nu ** 2)
    else:
        lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    return lam, mu


def compute_lame_3d(E, nu):
    """Lame pair for 3D (elasticity.py:55-59)."""
    mu = E / (2 * (1 + nu))
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    return lam, mu


def stress_from_strain_2d(eps, E, nu, plane_mode="plane_strain"):
    """sigma = 2 mu eps + lam tr(eps) I on (nx,ny,2,2) (elasticity.py:12-18)."""
    lam, mu = compute_lame_2d(E, nu, plane_mode)
    tr = np.trace(eps, axis1=2, axis2=3)[..., None, None]
    return 2 * mu[..., None, None] * eps + lam[..., None, None] * tr * np.eye(2)[None, None]


def stress_from_strain_3d(eps, E, nu):
    """sigma = 2 mu eps + lam tr(eps) I on (nx,ny,nz,3,3) (elasticity.py:62-67)."""
    lam, mu = compute_lame_3d(E, nu)
    tr = np.trace(eps, axis1=3, axis2=4)[..., None, None]
    return 2 * mu[..., None, None] * eps + lam[..., None, None] * tr * np.eye(3)[None, None, None]


def wave_vectors(shape, pixel=1.0):
    """k_i = 2 pi fftfreq(n_i, d=L_i/n_i), meshgrid 'ij' (fft.py:9-20)."""
    ks = [2 * np.pi * np.fft.fftfreq(n, d=(n * pixel) / n) for n in shape]
    return np.meshgrid(*ks, indexing="ij")


def green_operator(q, lam0, mu0):
    """Materialised isotropic strain Green operator Gamma0_khij(k), spatial
    dims first, Gamma(0) = 0, Nyquist kept (elasticity.py:20-53, 70-105)."""
    d = len(q)
    shape = q[0].shape
    k2 = sum(qi * qi for qi in q)
    k2s = k2.copy()
    k2s[(0,) * d] = 1.0
    A = 1.0 / (4 * mu0)
    B = (lam0 + mu0) / (mu0 * (lam0 + 2 * mu0))
    G = np.zeros(shape + (d, d, d, d))
    for k in range(d):
        for h in range(d):
            for i in range(d):
                for j in range(d):
                    t1 = 0.0
                    if k == i:
                        t1 = t1 + q[h] * q[j]
                    if h == i:
                        t1 = t1 + q[k] * q[j]
                    if k == j:
                        t1 = t1 + q[h] * q[i]
                    if h == j:
                        t1 = t1 + q[k] * q[i]
                    t1 = A * t1 / k2s
                    t2 = B * (q[k] * q[h] * q[i] * q[j]) / (k2s * k2s)
                    G[..., k, h, i, j] = t1 - t2
    G[(0,) * d] = 0
    return G


def fft_field(f):
    """Complex fftn, norm='ortho' (fft.py:22-25)."""
    return _sfft.fftn(f, norm="ortho", workers=FFT_WORKERS)


def ifft_field(f_hat):
    """Complex ifftn, norm='ortho', real part (fft.py:28-31)."""
    return _sfft.ifftn(f_hat, norm="ortho", workers=FFT_WORKERS).real


def _ls_solve(d, E, nu, eps_bar, eps_plastic, max_iter, tol, pixel, lame, stress,
              Gamma, eps_init, record):
    shape = E.shape
    E_avg = E.mean()
    nu_avg = nu.mean()
    if Gamma is None:
        lam0, mu0 = lame(E_avg, nu_avg)
        Gamma = green_operator(wave_vectors(shape, pixel), lam0, mu0)
    if eps_init is not None:
        eps = eps_init.copy()
        for i in range(d):
            for j in range(d):
                eps[..., i, j] += eps_bar[i, j] - eps_init[..., i, j].mean()
    else:
        eps = np.zeros(shape + (d, d))
        for i in range(d):
            for j in range(d):
                eps[..., i, j] = eps_bar[i, j]
    letters = "xyz"[:d]
    ein = f"{letters}khij,{letters}ij->{letters}kh"
    it_done = 0
    diff = np.inf
    for it in range(max_iter):
        if eps_plastic is not None:
            sig = stress(eps - eps_plastic, E, nu)
        else:
            sig = stress(eps, E, nu)
        sig0 = stress(eps, E * 0 + E_avg, nu * 0 + nu_avg)
        tau = sig - sig0
        tau_hat = np.zeros(shape + (d, d), dtype=complex)
        for i in range(d):
            for j in range(d):
                tau_hat[..., i, j] = fft_field(tau[..., i, j])
        et_hat = -np.einsum(ein, Gamma, tau_hat)
        et = np.zeros(shape + (d, d))
        for i in range(d):
            for j in range(d):
                et[..., i, j] = ifft_field(et_hat[..., i, j])
        for i in range(d):
            for j in range(d):
                et[..., i, j] -= et[..., i, j].mean()
        eps_new = np.zeros(shape + (d, d))
        for i in range(d):
            for j in range(d):
                eps_new[..., i, j] = eps_bar[i, j] + et[..., i, j]
        diff = np.linalg.norm(eps_new - eps) / (np.linalg.norm(eps) + 1e-20)
        eps = eps_new
        it_done = it + 1
        if record is not None:
            record.append(diff)
        if diff < tol:
            break
    if eps_plastic is not None:
        sig_out = stress(eps - eps_plastic, E, nu)
    else:
        sig_out = stress(eps, E, nu)
    axes = tuple(range(d))
    return eps, sig_out, eps.mean(axis=axes), sig_out.mean(axis=axes), it_done, diff


def spectral_solver_2d(E, nu, eps_bar, eps_plastic=None, max_iter=200, tol=1e-6,
                       verbose=False, pixel=1.0, plane_mode="plane_strain",
                       Gamma=None, eps_init=None, info=None):
    """2D LS fixed point (linear_elastic_simulator.py:12-88).  ``info`` (a dict,
    oracle-only extension) receives the iteration count and per-iteration diffs."""
    rec = [] if info is not None else None
    out = _ls_solve(2, np.asarray(E, float), np.asarray(nu, float), np.asarray(eps_bar, float),
                    eps_plastic, max_iter, tol, pixel,
                    lambda E_, nu_: compute_lame_2d(E_, nu_, plane_mode),
                    lambda e, E_, nu_: stress_from_strain_2d(e, E_, nu_, plane_mode),
                    Gamma, eps_init, rec)
    if info is not None:
        info.update(iters=out[4], diff=out[5], diffs=rec)
    return out[:4]


def spectral_solver_3d(E, nu, eps_bar, eps_plastic=None, max_iter=200, tol=1e-6,
                       verbose=False, pixel=1.0, Gamma=None, eps_init=None, info=None):
    """3D LS fixed point (linear_elastic_simulator.py:274-376)."""
    rec = [] if info is not None else None
    out = _ls_solve(3, np.asarray(E, float), np.asarray(nu, float), np.asarray(eps_bar, float),
                    eps_plastic, max_iter, tol, pixel, compute_lame_3d, stress_from_strain_3d,
                    Gamma, eps_init, rec)
    if info is not None:
        info.update(iters=out[4], diff=out[5], diffs=rec)
    return out[:4]


def update_stress_fft_full(eps_plastic_field, eps_macro, E, nu, pixel=1.0,
                           max_iter=200, tol=1e-6, verbose=False):
    """stz/update_fft.py:4-22: cold-start 3D LS, Gamma rebuilt."""
    return spectral_solver_3d(E, nu, eps_macro, eps_plastic=eps_plastic_field,
                              max_iter=max_iter, tol=tol, verbose=verbose, pixel=pixel)


_SYM3 = ((0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1))


def _gamma_sym3(q, t, A2, Bc):
    """-Gamma0(q):t for symmetric t given as its 6 components (xx, yy, zz, yz,
    xz, xy) on a spectrum: (q_h (t q)_k + q_k (t q)_h) / (2 mu0 |q|^2) -
    B q_k q_h (q.t.q) / |q|^4 (elasticity.py:70-105 contracted with a
    symmetric tau), zero at q = 0."""
    k2 = q[0] * q[0] + q[1] * q[1] + q[2] * q[2]
    zero = k2 == 0
    k2 = np.where(zero, 1.0, k2)
    a = A2 / k2
    bb = Bc / (k2 * k2)
    T = {}
    for c, (i, j) in enumerate(_SYM3):
        T[(i, j)] = T[(j, i)] = t[c]
    u = [sum(T[(k, j)] * q[j] for j in range(3)) for k in range(3)]
    qtq = (q[0] * u[0] + q[1] * u[1] + q[2] * u[2]) * bb
    out = []
    for k, h in _SYM3:
        r = -(a * (q[h] * u[k] + q[k] * u[h]) - q[k] * q[h] * qtq)
        out.append(np.where(zero, 0.0, r))
    return out


def spectral_solver_3d_mf(E, nu, eps_bar, eps_plastic=None, max_iter=200, tol=1e-6, pixel=1.0, info=None):
    """Matrix-free restatement of spectral_solver_3d (linear_elastic_simulator.py:
    274-376) for large grids: the same iteration (sigma = C:(eps - eps_p),
    sigma0 = C0:eps, tau = sigma - sigma0, eps~ = -F^-1[Gamma0:F[tau]] minus its
    mean, eps_new = eps_bar + eps~, diff = |eps_new - eps| / (|eps| + 1e-20),
    cold start) on the 6 symmetric components, with Gamma0 applied from q on
    real-to-complex spectra.  The reference keeps .real of a full-spectrum
    inverse, which for a real tau is the inverse of the Hermitian part
    1/2 (Gamma0(q) + Gamma0(R q)) tau_hat (R negates the Nyquist components of
    q); that average is applied here.  Equal to spectral_solver_3d to
    round-off with the same iteration count (tests/test_oracle_ls.py);
    materialises no Gamma0 and no (d, d) spectra.  Inputs and outputs in the
    reference's (nx, ny, nz, 3, 3) layout."""
    E = np.asarray(E, float)
    nu = np.asarray(nu, float)
    shape = E.shape
    lam, mu = compute_lame_3d(E, nu)
    lam0, mu0 = compute_lame_3d(E.mean(), nu.mean())
    A2 = 1.0 / (2.0 * mu0)
    Bc = (lam0 + mu0) / (mu0 * (lam0 + 2.0 * mu0))
    # the half axis keeps numpy's fftfreq sign convention (-1/2 at an even
    # length's Nyquist bin), which the Nyquist average pairs with R q
    qs = [np.fft.fftfreq(n) for n in shape[:2]] + [np.fft.fftfreq(shape[2])[: shape[2] // 2 + 1]]
    q = np.meshgrid(*qs, indexing="ij", sparse=True)
    nyq = []
    for a, n in enumerate(shape):
        v = (np.arange(len(qs[a])) == n // 2) & (n % 2 == 0)
        sh = [1, 1, 1]
        sh[a] = len(qs[a])
        nyq.append(v.reshape(sh))
    qR = [np.where(nyq[a], -q[a], q[a]) for a in range(3)]
    eb = [float(eps_bar[i][j]) if i == j else 0.5 * (eps_bar[i][j] + eps_bar[j][i]) for i, j in _SYM3]
    anti = np.asarray(eps_bar, float) - np.asarray(eps_bar, float).T
    e = [np.full(shape, v) for v in eb]
    ep = [np.zeros(shape)] * 6 if eps_plastic is None else \
        [0.5 * (eps_plastic[..., i, j] + eps_plastic[..., j, i]) for i, j in _SYM3]
    w = [1.0, 1.0, 1.0, 2.0, 2.0, 2.0]
    ax = (0, 1, 2)
    it_done, diff = 0, np.inf
    for it in range(max_iter):
        d = [e[c] - ep[c] for c in range(6)]
        tr = d[0] + d[1] + d[2]
        tr0 = e[0] + e[1] + e[2]
        tau = [(2 * mu * d[c] + (lam * tr if c < 3 else 0.0)) - (2 * mu0 * e[c] + (lam0 * tr0 if c < 3 else 0.0))
               for c in range(6)]
        th = [_sfft.rfftn(t, axes=ax, workers=FFT_WORKERS) for t in tau]
        o1 = _gamma_sym3(q, th, A2, Bc)
        o2 = _gamma_sym3(qR, th, A2, Bc)
        new = []
        for c in range(6):
            et = _sfft.irfftn(0.5 * (o1[c] + o2[c]), s=shape, axes=ax, workers=FFT_WORKERS)
            et -= et.mean()
            new.append(eb[c] + et)
        num = sum(w[c] * np.sum((new[c] - e[c]) ** 2) for c in range(6))
        den = sum(w[c] * np.sum(e[c] ** 2) for c in range(6)) + 0.25 * np.sum(anti ** 2) * E.size
        diff = np.sqrt(num) / (np.sqrt(den) + 1e-20)
        e = new
        it_done = it + 1
        if diff < tol:
            break
    d = [e[c] - ep[c] for c in range(6)]
    tr = d[0] + d[1] + d[2]
    sig = [2 * mu * d[c] + (lam * tr if c < 3 else 0.0) for c in range(6)]
    eps = np.zeros(shape + (3, 3))
    sg = np.zeros(shape + (3, 3))
    for c, (i, j) in enumerate(_SYM3):
        eps[..., i, j] = e[c]
        sg[..., i, j] = sig[c]
        if i != j:
            eps[..., j, i] = e[c]
            sg[..., j, i] = sig[c]
    eps = eps + 0.5 * anti
    if info is not None:
        info.update(iters=it_done, diff=diff)
    return eps, sg, eps.mean(axis=ax), sg.mean(axis=ax)


def equilibrium_residual(sig, pixel=1.0):
    """FP64 equilibrium residual of a stress field (nx,..,d,d):
    sqrt(sum_k |q_j sig_hat_ij|^2 / |q|^2) / sqrt(sum_k |sig_hat|^2), k != 0 and
    k off the Nyquist planes.
    A build-defined measure (the reference has none); used for the FP32 +
    refinement form's acceptance test."""
    d = sig.shape[-1]
    shape = sig.shape[:-2]
    q = wave_vectors(shape, pixel)
    k2 = sum(qi * qi for qi in q)
    k2[(0,) * d] = 1.0
    sh = np.stack([np.stack([np.fft.fftn(sig[..., i, j]) for j in range(d)], -1)
                   for i in range(d)], -2)
    sh[(0,) * d] = 0
    # Nyquist planes of even axes are excluded: there the reference's
    # full-spectrum Re(ifft) applies the average of Gamma over +-q_nyq, whose
    # kernel is not {q.sigma = 0}.
    for a, n in enumerate(shape):
        if n % 2 == 0:
            sl = [slice(None)] * d
            sl[a] = n // 2
            sh[tuple(sl)] = 0
    div = np.zeros(shape + (d,), complex)
    for i in range(d):
        for j in range(d):
            div[..., i] += q[j] * sh[..., i, j]
    num = np.sum(np.abs(div) ** 2 / k2[..., None])
    den = np.sum(np.abs(sh) ** 2)
    return float(np.sqrt(num) / (np.sqrt(den) + 1e-300))


def linear_elastic_simulation(d, E, nu, target_strain_mask, target_values, n_steps=20, pixel=1.0, tol_ls=1e-6,
                              max_iter_ls=200, tol_macro=1e-4, max_iter_macro=20, store=True,
                              plane_mode="plane_strain"):
    """Mixed stress/strain load stepping (linear_elastic_simulator.py:90-272 in
    2D, :378-566 in 3D): targets ramp linearly over n_steps + 1 steps; per step
    a macro loop sets the strain-controlled entries, solves warm-started from
    the previous step's field (eps_init), and corrects the stress-controlled
    entries by (err - nu_avg tr(err) I) / E_avg until max |err| < tol_macro.
    Returns (eps_macro (n+1,d,d), sig_macro (n+1,d,d), eps list, sig list, LS
    iteration counts per solve)."""
    E = np.asarray(E, float)
    nu = np.asarray(nu, float)
    shape = E.shape
    if d == 3:
        lam0, mu0 = compute_lame_3d(E.mean(), nu.mean())
    else:
        lam0, mu0 = compute_lame_2d(E.mean(), nu.mean(), plane_mode)
    G = green_operator(wave_vectors(shape, pixel), lam0, mu0)
    E_avg, nu_avg = E.mean(), nu.mean()
    mask = np.asarray(target_strain_mask, bool)
    smask = ~mask
    tv = np.asarray(target_values, float)
    em_l, sm_l, e_l, s_l, its = [], [], [], [], []
    eb = np.zeros((d, d))
    warm = None
    solve = spectral_solver_3d if d == 3 else (
        lambda *a, **k: spectral_solver_2d(*a, plane_mode=plane_mode, **k))
    for s in range(n_steps + 1):
        tgt = tv * (s / n_steps)
        for _ in range(max_iter_macro):
            eb[mask] = tgt[mask]
            info = {}
            eps, sig, epsM, sigM = solve(E, nu, eb, max_iter=max_iter_ls, tol=tol_ls, pixel=pixel, Gamma=G,
                                         eps_init=warm, info=info)
            its.append(info["iters"])
            err = np.zeros((d, d))
            err[smask] = tgt[smask] - sigM[smask]
            mx = np.max(np.abs(err[smask])) if np.any(smask) else 0.0
            if mx < tol_macro:
                break
            de = (err - nu_avg * np.trace(err) * np.eye(d)) / E_avg
            eb[smask] += de[smask]
        warm = eps
        em_l.append(epsM)
        sm_l.append(sigM)
        if store:
            e_l.append(eps)
            s_l.append(sig)
    return np.array(em_l), np.array(sm_l), e_l, s_l, its


# ---------------------------------------------------------------------------
# Landau small-strain model, 3D (elasticity.py:297-335, :811-861) and its LS
# fixed point (linear_elastic_simulator.py:1231-1350, the default "dbfft" path)
# ---------------------------------------------------------------------------
def cap_limit(mu, g4, strain_capping_limit):
    """Resolved capping strain E_cap (elasticity.py:312-319)."""
    mu_val = mu.mean() if isinstance(mu, np.ndarray) else mu
    if strain_capping_limit is None or strain_capping_limit <= 0:
        return np.sqrt(0.2 * mu_val / np.abs(g4)) if g4 < 0 else 999.0
    return strain_capping_limit


def cap_strain_3d(eps, mu, g4, enabled, limit, ratio, kind, power):
    """(capped strain, capping weight) (elasticity.py:297-335)."""
    if not enabled:
        return eps, np.zeros(eps.shape[:-2])
    I1 = np.trace(eps, axis1=-2, axis2=-1)
    I = np.broadcast_to(np.eye(3), eps.shape)
    dev = eps - I1[..., None, None] * I / 3.0
    dd = np.einsum("...ij,...ji->...", dev, dev)
    Eeq = np.sqrt(np.maximum(0.0, (2.0 / 3.0) * dd))
    Ecap = cap_limit(mu, g4, limit)
    f = np.ones_like(Eeq)
    nz = Eeq > 1e-12
    if kind == "smooth":
        rp = np.minimum(100.0, (Eeq[nz] / Ecap) ** power)
        f[nz] = (Ecap / Eeq[nz]) * np.tanh(rp) ** (1.0 / power)
        w = np.tanh(np.minimum(100.0, (Eeq / Ecap) ** power)) ** 2
    else:
        capped = Eeq > Ecap
        np.divide(Ecap, Eeq, out=f, where=capped)
        w = capped.astype(float)
    return dev * f[..., None, None] + I1[..., None, None] * I / 3.0, w


def stress_landau_3d(eps, lam, mu, v1, v2, v3, g1, g2, g3, g4, strain_capping_enabled=False,
                     strain_capping_limit=None, strain_capping_tangent_ratio=0.1, strain_capping_type="piecewise",
                     strain_capping_smooth_power=1.0):
    """sigma = A I + B eps + C eps^2 with the Landau invariant coefficients,
    plus the capping tangent term (elasticity.py:811-861)."""
    e0 = eps
    e, w = cap_strain_3d(eps, mu, g4, strain_capping_enabled, strain_capping_limit, strain_capping_tangent_ratio,
                         strain_capping_type, strain_capping_smooth_power)
    I1 = np.trace(e, axis1=-2, axis2=-1)
    e2 = np.einsum("...ij,...jk->...ik", e, e)
    I2 = np.trace(e2, axis1=-2, axis2=-1)
    I3 = np.trace(np.einsum("...ij,...jk->...ik", e2, e), axis1=-2, axis2=-1)
    lam = np.asarray(lam)
    mu = np.asarray(mu)
    A = lam * I1 + 0.5 * v1 * I1 ** 2 + v2 * I2 + (1.0 / 6.0) * g1 * I1 ** 3 + g2 * I1 * I2 + (4.0 / 3.0) * g3 * I3
    B = 2.0 * (mu + v2 * I1 + 0.5 * g2 * I1 ** 2 + g4 * I2)
    C = 4.0 * (v3 + g3 * I1)
    sig = A[..., None, None] * np.eye(3) + B[..., None, None] * e + C[..., None, None] * e2
    if strain_capping_enabled:
        sig = sig + 2.0 * (strain_capping_tangent_ratio * mu)[..., None, None] * (e0 - e) * w[..., None, None]
    return sig


def landau_energy_3d(eps, lam, mu, v1, v2, v3, g1, g2, g3, g4):
    """Stored energy whose strain derivative is the Landau stress (no capping):
    W = lam/2 I1^2 + mu I2 + v1/6 I1^3 + v2 I1 I2 + 4/3 v3 I3 + g1/24 I1^4
        + g2/2 I1^2 I2 + 4/3 g3 I1 I3 + g4/2 I2^2  (used to pin the stress by finite differences)."""
    I1 = np.trace(eps, axis1=-2, axis2=-1)
    e2 = np.einsum("...ij,...jk->...ik", eps, eps)
    I2 = np.trace(e2, axis1=-2, axis2=-1)
    I3 = np.trace(np.einsum("...ij,...jk->...ik", e2, eps), axis1=-2, axis2=-1)
    return (0.5 * lam * I1 ** 2 + mu * I2 + v1 / 6.0 * I1 ** 3 + v2 * I1 * I2 + 4.0 / 3.0 * v3 * I3
            + g1 / 24.0 * I1 ** 4 + 0.5 * g2 * I1 ** 2 * I2 + 4.0 / 3.0 * g3 * I1 * I3 + 0.5 * g4 * I2 ** 2)


def spectral_solver_landau_3d(lam, mu, v1, v2, v3, g1, g2, g3, g4, eps_bar, eps_plastic=None, max_iter=400, tol=1e-6,
                              verbose=False, pixel=1.0, Gamma=None, eps_init=None, strain_capping_enabled=False,
                              strain_capping_limit=None, strain_capping_tangent_ratio=0.1,
                              strain_capping_type="piecewise", strain_capping_smooth_power=1.0, solver="dbfft",
                              info=None):
    """Nonlinear LS fixed point with the Landau stress; reference medium from
    (mean lam, mean mu); FloatingPointError on a non-finite update
    (linear_elastic_simulator.py:1231-1350)."""
    if solver == "newton_cg":
        raise NotImplementedError("newton_cg is not available for 3D small strain")
    lam = np.asarray(lam, float)
    mu = np.asarray(mu, float)
    shape = lam.shape
    la, ma = float(lam.mean()), float(mu.mean())
    if Gamma is None:
        Gamma = green_operator(wave_vectors(shape, pixel), la, ma)
    if eps_init is not None:
        eps = eps_init.copy()
        for i in range(3):
            for j in range(3):
                eps[..., i, j] += eps_bar[i, j] - eps_init[..., i, j].mean()
    else:
        eps = np.zeros(shape + (3, 3))
        for i in range(3):
            for j in range(3):
                eps[..., i, j] = eps_bar[i, j]
    cap = dict(strain_capping_enabled=strain_capping_enabled, strain_capping_limit=strain_capping_limit,
               strain_capping_tangent_ratio=strain_capping_tangent_ratio, strain_capping_type=strain_capping_type,
               strain_capping_smooth_power=strain_capping_smooth_power)
    it_done = 0
    for it in range(max_iter):
        e_el = eps - eps_plastic if eps_plastic is not None else eps
        sig = stress_landau_3d(e_el, lam, mu, v1, v2, v3, g1, g2, g3, g4, **cap)
        tr = np.trace(eps, axis1=3, axis2=4)[..., None, None]
        tau = sig - (la * tr * np.eye(3) + 2.0 * ma * eps)
        tau_hat = np.zeros(shape + (3, 3), dtype=complex)
        for i in range(3):
            for j in range(3):
                tau_hat[..., i, j] = fft_field(tau[..., i, j])
        et_hat = -np.einsum("xyzkhij,xyzij->xyzkh", Gamma, tau_hat)
        et = np.zeros(shape + (3, 3))
        for i in range(3):
            for j in range(3):
                et[..., i, j] = ifft_field(et_hat[..., i, j])
                et[..., i, j] -= et[..., i, j].mean()
        eps_new = eps_bar[None, None, None] + et
        diff = np.linalg.norm(eps_new - eps) / (np.linalg.norm(eps) + 1e-20)
        eps = eps_new
        it_done = it + 1
        if not np.isfinite(diff):
            raise FloatingPointError("spectral_solver_landau_3d: non-finite strain update")
        if diff < tol:
            break
    e_el = eps - eps_plastic if eps_plastic is not None else eps
    sig_out = stress_landau_3d(e_el, lam, mu, v1, v2, v3, g1, g2, g3, g4, **cap)
    if info is not None:
        info.update(iters=it_done, diff=diff)
    return eps, sig_out, eps.mean(axis=(0, 1, 2)), sig_out.mean(axis=(0, 1, 2))


# ---------------------------------------------------------------------------
# Landau small-strain model, 2D (elasticity.py:487-638, the block-diagonal
# scalar form) and its LS fixed point with Anderson mixing
# (linear_elastic_simulator.py:952-1228, default "dbfft" path)
# ---------------------------------------------------------------------------
def stress_landau_2d(eps, lam, mu, v1, v2, v3, g1, g2, g3, g4, plane_mode="plane_strain",
                     strain_capping_enabled=False, strain_capping_limit=None, strain_capping_tangent_ratio=0.1,
                     strain_capping_type="piecewise", strain_capping_smooth_power=1.0, e33_state=None, info=None):
    """In-plane Cauchy stress of the 3D Landau law for the block-diagonal strain
    [[a11, a12, 0], [a21, a22, 0], [0, 0, e33]]: e33 = 0 (plane strain) or the
    root of sigma33 = 0 by at most 20 Newton steps that stop when every voxel's
    step is below 1e-9, warm-started from e33_state["e33"] (plane stress).  The
    stress comes from the last evaluated state (before the final e33 update)."""
    a11, a22, a12, a21 = eps[..., 0, 0], eps[..., 1, 1], eps[..., 0, 1], eps[..., 1, 0]
    lam = np.asarray(lam, float)
    mu = np.asarray(mu, float)
    cap = strain_capping_enabled
    if cap:
        Ecap = cap_limit(mu, g4, strain_capping_limit)
        Gt = strain_capping_tangent_ratio * mu

    def scal(e33):
        I1r = a11 + a22 + e33
        m3 = I1r / 3.0
        d11, d22, d33 = a11 - m3, a22 - m3, e33 - m3
        if cap:
            Eeq = np.sqrt(np.maximum(0.0, (2.0 / 3.0) * (d11 * d11 + d22 * d22 + d33 * d33 + 2.0 * a12 * a21)))
            f = np.ones_like(Eeq)
            nz = Eeq > 1e-12
            if strain_capping_type == "smooth":
                p = strain_capping_smooth_power
                rp = np.minimum(100.0, (Eeq[nz] / Ecap) ** p)
                f[nz] = (Ecap / Eeq[nz]) * np.tanh(rp) ** (1.0 / p)
                w = np.tanh(np.minimum(100.0, (Eeq / Ecap) ** p)) ** 2
            else:
                capped = Eeq > Ecap
                np.divide(Ecap, Eeq, out=f, where=capped)
                w = capped.astype(float)
            c11, c22, c12, c21, c33 = f * d11 + m3, f * d22 + m3, f * a12, f * a21, f * d33 + m3
        else:
            w = None
            c11, c22, c12, c21, c33 = a11, a22, a12, a21, e33
        I1 = c11 + c22 + c33
        cross = c12 * c21
        tr2 = c11 * c11 + c22 * c22 + 2.0 * cross
        I2 = tr2 + c33 * c33
        I3 = (c11 + c22) * (tr2 - (c11 * c22 - cross)) + c33 ** 3
        A = lam * I1 + 0.5 * v1 * I1 ** 2 + v2 * I2 + (1.0 / 6.0) * g1 * I1 ** 3 + g2 * I1 * I2 + (4.0 / 3.0) * g3 * I3
        B = 2.0 * (mu + v2 * I1 + 0.5 * g2 * I1 ** 2 + g4 * I2)
        C = 4.0 * (v3 + g3 * I1)
        return I1, I2, c11, c22, c12, c21, c33, w, A, B, C

    n_newton = 0
    if plane_mode == "plane_stress":
        warm = e33_state.get("e33") if e33_state is not None else None
        e33 = warm.copy() if warm is not None and warm.shape == a11.shape else -lam / (lam + 2.0 * mu) * (a11 + a22)
        for _ in range(20):
            I1, I2, c11, c22, c12, c21, c33, w, A, B, C = scal(e33)
            s33 = A + B * c33 + C * c33 ** 2
            dA = lam + v1 * I1 + 0.5 * g1 * I1 ** 2 + g2 * I2 + 2.0 * (v2 + g2 * I1) * c33 + 4.0 * g3 * c33 ** 2
            dB = 2.0 * (v2 + g2 * I1) + 4.0 * g4 * c33
            ds33 = dA + dB * c33 + B + 4.0 * g3 * c33 ** 2 + 2.0 * C * c33
            if cap:
                s33 = s33 + 2.0 * Gt * (e33 - c33) * w
                ds33 = ds33 + 2.0 * Gt * w
            delta = -s33 / (ds33 + 1e-12)
            if not np.all(np.isfinite(delta)):
                raise FloatingPointError("stress_from_strain_landau_2d: non-finite e33 Newton update")
            e33 = e33 + delta
            n_newton += 1
            if np.all(np.abs(delta) < 1e-9):
                break
        if e33_state is not None:
            e33_state["e33"] = e33
    else:
        e33 = np.zeros_like(a11)
        I1, I2, c11, c22, c12, c21, c33, w, A, B, C = scal(e33)
    if info is not None:
        info["newton"] = n_newton
    cross = c12 * c21
    t = c11 + c22
    sig = np.empty(eps.shape)
    sig[..., 0, 0] = A + B * c11 + C * (c11 * c11 + cross)
    sig[..., 1, 1] = A + B * c22 + C * (c22 * c22 + cross)
    sig[..., 0, 1] = B * c12 + C * (c12 * t)
    sig[..., 1, 0] = B * c21 + C * (c21 * t)
    if cap:
        r = 2.0 * Gt * w
        sig[..., 0, 0] += r * (a11 - c11)
        sig[..., 1, 1] += r * (a22 - c22)
        sig[..., 0, 1] += r * (a12 - c12)
        sig[..., 1, 0] += r * (a21 - c21)
    return sig


def spectral_solver_landau_2d(lam, mu, v1, v2, v3, g1, g2, g3, g4, eps_bar, eps_plastic=None, max_iter=400, tol=1e-6,
                              verbose=False, pixel=1.0, plane_mode="plane_strain", Gamma=None, eps_init=None,
                              strain_capping_enabled=False, strain_capping_limit=None, strain_capping_tangent_ratio=0.1,
                              strain_capping_type="piecewise", strain_capping_smooth_power=1.0, solver="dbfft",
                              e33_state=None, anderson_m=5, info=None):
    """2D nonlinear LS with type-II Anderson mixing of the last m iterates
    (numpy lstsq on the residual differences; plain step and history restart
    when the mixed iterate is not finite); reference medium (mean lam, mean mu)."""
    lam = np.asarray(lam, float)
    mu = np.asarray(mu, float)
    shape = lam.shape
    la, ma = float(lam.mean()), float(mu.mean())
    if Gamma is None:
        Gamma = green_operator(wave_vectors(shape, pixel), la, ma)
    if eps_init is not None:
        eps = eps_init.copy()
        for i in range(2):
            for j in range(2):
                eps[..., i, j] += eps_bar[i, j] - eps_init[..., i, j].mean()
    else:
        eps = np.zeros(shape + (2, 2))
        for i in range(2):
            for j in range(2):
                eps[..., i, j] = eps_bar[i, j]
    cap = dict(plane_mode=plane_mode, strain_capping_enabled=strain_capping_enabled,
               strain_capping_limit=strain_capping_limit, strain_capping_tangent_ratio=strain_capping_tangent_ratio,
               strain_capping_type=strain_capping_type, strain_capping_smooth_power=strain_capping_smooth_power,
               e33_state=e33_state)
    if solver == "newton_cg":
        law = dict(cap, e33_state=None)
        return _landau_2d_newton_cg(lam, mu, (v1, v2, v3, g1, g2, g3, g4), eps_bar, eps_plastic, eps, Gamma,
                                    max_iter, tol, law, info)
    aa_g, aa_f = [], []
    it_done = 0
    diff = np.inf
    for it in range(max_iter):
        e_el = eps - eps_plastic if eps_plastic is not None else eps
        sig = stress_landau_2d(e_el, lam, mu, v1, v2, v3, g1, g2, g3, g4, **cap)
        tr = np.trace(eps, axis1=2, axis2=3)[..., None, None]
        tau = sig - (la * tr * np.eye(2) + 2.0 * ma * eps)
        tau_hat = np.zeros(shape + (2, 2), dtype=complex)
        for i in range(2):
            for j in range(2):
                tau_hat[..., i, j] = fft_field(tau[..., i, j])
        et_hat = -np.einsum("xykhij,xyij->xykh", Gamma, tau_hat)
        et = np.zeros(shape + (2, 2))
        for i in range(2):
            for j in range(2):
                et[..., i, j] = ifft_field(et_hat[..., i, j])
                et[..., i, j] -= et[..., i, j].mean()
        eps_new = eps_bar[None, None] + et
        diff = np.linalg.norm(eps_new - eps) / (np.linalg.norm(eps) + 1e-20)
        it_done = it + 1
        if not np.isfinite(diff):
            raise FloatingPointError("spectral_solver_landau_2d: non-finite strain update")
        if diff < tol:
            eps = eps_new
            break
        if anderson_m > 0:
            aa_g.append(eps_new.reshape(-1))
            aa_f.append((eps_new - eps).reshape(-1))
            if len(aa_f) > anderson_m + 1:
                aa_g.pop(0)
                aa_f.pop(0)
        if anderson_m > 0 and len(aa_f) >= 2:
            dF = np.stack([aa_f[i + 1] - aa_f[i] for i in range(len(aa_f) - 1)], axis=1)
            dG = np.stack([aa_g[i + 1] - aa_g[i] for i in range(len(aa_g) - 1)], axis=1)
            try:
                gam, *_ = np.linalg.lstsq(dF, aa_f[-1], rcond=None)
                eps_aa = aa_g[-1] - dG @ gam
            except np.linalg.LinAlgError:
                eps_aa = None
            if eps_aa is not None and np.all(np.isfinite(eps_aa)):
                eps = eps_aa.reshape(shape + (2, 2))
            else:
                eps = eps_new
                aa_g, aa_f = aa_g[-1:], aa_f[-1:]
        else:
            eps = eps_new
    e_el = eps - eps_plastic if eps_plastic is not None else eps
    sig_out = stress_landau_2d(e_el, lam, mu, v1, v2, v3, g1, g2, g3, g4, **cap)
    if info is not None:
        info.update(iters=it_done, diff=diff)
    return eps, sig_out, eps.mean(axis=(0, 1)), sig_out.mean(axis=(0, 1))


def _landau_2d_newton_cg(lam, mu, coef, eps_bar, eps_plastic, eps, Gamma, max_iter, tol, law, info):
    """Newton-CG branch of spectral_solver_landau_2d (linear_elastic_simulator.py:
    1011-1117): Newton steps with SciPy's bicgstab (rtol = tol, maxiter 150) on
    A = G C_h G, G the -Gamma0 convolution with the mean subtracted, C_h the
    central difference (h = 1e-7) of the Landau stress with a cold plane-stress
    e33 Newton per evaluation; b = -G C_h (eps_bar - <eps>) first, then -G sigma;
    stop when |dx| / (|eps| + 1e-20) < tol after the first step.  info gets the
    Newton steps, the BiCGSTAB iterations in total and the last ratio."""
    import scipy.sparse.linalg as spla
    shape = lam.shape
    nx, ny = shape

    def sigma(e_el):
        return stress_landau_2d(e_el, lam, mu, *coef, **law)

    def G(tau):
        th = np.zeros(shape + (2, 2), dtype=complex)
        for i in range(2):
            for j in range(2):
                th[..., i, j] = fft_field(tau[..., i, j])
        et_hat = -np.einsum("xykhij,xyij->xykh", Gamma, th)
        et = np.zeros(shape + (2, 2))
        for i in range(2):
            for j in range(2):
                et[..., i, j] = ifft_field(et_hat[..., i, j])
                et[..., i, j] -= et[..., i, j].mean()
        return et

    h = 1e-7
    state = {"eps": eps}

    def C_h(d):
        e_el = state["eps"] - eps_plastic if eps_plastic is not None else state["eps"]
        return (sigma(e_el + h * d) - sigma(e_el - h * d)) / (2.0 * h)

    def matvec(x):
        dc = G(x.reshape(nx, ny, 2, 2))
        return G(C_h(dc)).reshape(-1)

    n = eps.size
    A = spla.LinearOperator(shape=(n, n), matvec=matvec, dtype="float64")
    dbar = eps_bar - eps.mean(axis=(0, 1))
    dbar_grid = np.einsum("ij,xy->xyij", dbar, np.ones((nx, ny)))
    e_el = eps - eps_plastic if eps_plastic is not None else eps
    sig = sigma(e_el)
    krylov = 0
    newton = 0
    res = np.inf
    for it in range(max_iter):
        rhs = -G(C_h(dbar_grid)).reshape(-1) if it == 0 else -G(sig).reshape(-1)
        count = [0]

        def cb(_xk):
            count[0] += 1
        dx, _ = spla.bicgstab(A, rhs, rtol=tol, maxiter=150, callback=cb)
        krylov += count[0]
        d = dx.reshape(nx, ny, 2, 2)
        eps = eps + dbar_grid + G(d) if it == 0 else eps + G(d)
        state["eps"] = eps
        e_el = eps - eps_plastic if eps_plastic is not None else eps
        sig = sigma(e_el)
        res = np.linalg.norm(dx) / (np.linalg.norm(eps) + 1e-20)
        newton = it + 1
        if res < tol and it > 0:
            break
    if info is not None:
        info.update(iters=newton, krylov=krylov, diff=res)
    return eps, sig, eps.mean(axis=(0, 1)), sig.mean(axis=(0, 1))


# ---------------------------------------------------------------------------
# Secant-modulus small-strain models and their LS fixed points:
#   exponential degradation, 2D/3D (elasticity.py:641-810;
#   linear_elastic_simulator.py:614-850),
#   Rose law, 2D plane strain / plane stress (elasticity.py:868-943;
#   linear_elastic_simulator.py:852-950).
# ---------------------------------------------------------------------------
def von_mises_strain(eps):
    """sqrt(2/3 e':e'); a 2D tensor is taken with e33 = 0, whose deviator
    still has e'33 = -(e11 + e22)/3 (elasticity.py:641-700)."""
    d = eps.shape[-1]
    if d == 2:
        e11, e22, e12, e21 = eps[..., 0, 0], eps[..., 1, 1], eps[..., 0, 1], eps[..., 1, 0]
        tr3 = (e11 + e22) / 3.0
        ep11, ep22, ep33 = e11 - tr3, e22 - tr3, -(e11 + e22) / 3.0
        return np.sqrt((2.0 / 3.0) * (ep11 ** 2 + ep22 ** 2 + ep33 ** 2 + 0.5 * (e12 + e21) ** 2))
    e = eps
    tr3 = (e[..., 0, 0] + e[..., 1, 1] + e[..., 2, 2]) / 3.0
    ep11, ep22, ep33 = e[..., 0, 0] - tr3, e[..., 1, 1] - tr3, e[..., 2, 2] - tr3
    return np.sqrt((2.0 / 3.0) * (ep11 ** 2 + ep22 ** 2 + ep33 ** 2
                                  + 0.5 * ((e[..., 0, 1] + e[..., 1, 0]) ** 2 + (e[..., 0, 2] + e[..., 2, 0]) ** 2
                                           + (e[..., 1, 2] + e[..., 2, 1]) ** 2)))


def stress_secant(eps, lam, mu, d, k):
    """sigma = lam tr(e) I + 2 mu_sec e with mu_sec = mu (1 - d (1 - exp(-k e_eq)))
    (elasticity.py:703-810; the 2D form takes lam as given for either plane mode)."""
    dim = eps.shape[-1]
    mu_sec = mu * (1.0 - d * (1.0 - np.exp(-k * von_mises_strain(eps))))
    tr = np.trace(eps, axis1=-2, axis2=-1)[..., None, None]
    return np.asarray(lam)[..., None, None] * tr * np.eye(dim) + 2.0 * mu_sec[..., None, None] * eps


def rose_shear(eps_eq, mu, eta, mu_floor_fraction=0.1):
    """G_sec = max(mu exp(-eta e_eq), floor mu) (elasticity.py:868-876)."""
    return np.maximum(mu * np.exp(-eta * eps_eq), mu_floor_fraction * mu)


def stress_rose_2d(eps, lam, mu, eta, mu_floor_fraction=0.1, plane_mode="plane_strain", e33_out=None):
    """In-plane stress of the Rose law at constant bulk modulus K = lam + 2/3 mu:
    sigma = (K - 2/3 G) tr I + 2 G e; plane stress runs five fixed-point updates
    e33 = -(3K - 2G)/(3K + 4G) (e11 + e22) from the linear guess and evaluates G
    once more (elasticity.py:879-943)."""
    exx, eyy, exy = eps[..., 0, 0], eps[..., 1, 1], eps[..., 0, 1]
    K = lam + (2.0 / 3.0) * mu

    def shear(ezz):
        tr = exx + eyy + ezz
        tr3 = tr / 3.0
        eq = np.sqrt((2.0 / 3.0) * ((exx - tr3) ** 2 + (eyy - tr3) ** 2 + (ezz - tr3) ** 2 + 2.0 * exy ** 2))
        return tr, rose_shear(eq, mu, eta, mu_floor_fraction)

    if plane_mode == "plane_stress":
        ezz = -lam / (lam + 2.0 * mu) * (exx + eyy)
        for _ in range(5):
            _, G = shear(ezz)
            ezz = -((3.0 * K - 2.0 * G) / (3.0 * K + 4.0 * G)) * (exx + eyy)
    else:
        ezz = np.zeros_like(exx)
    tr, G = shear(ezz)
    if e33_out is not None:
        e33_out[...] = ezz
    sig = np.zeros_like(eps)
    le = K - (2.0 / 3.0) * G
    sig[..., 0, 0] = le * tr + 2.0 * G * exx
    sig[..., 1, 1] = le * tr + 2.0 * G * eyy
    sig[..., 0, 1] = 2.0 * G * exy
    sig[..., 1, 0] = sig[..., 0, 1]
    return sig


def _secant_ls(stress, dim, shape, lam_avg, mu_ref, eps_bar, eps_plastic, max_iter, tol, pixel, Gamma, eps_init,
               info):
    """LS fixed point eps = eps_bar - Gamma0 * (sigma(eps - eps_p) - C0:eps) with
    C0 = (lam_avg, mu_ref) and the relative-change stop test; no non-finite
    check (linear_elastic_simulator.py:614-950)."""
    if Gamma is None:
        Gamma = green_operator(wave_vectors(shape, pixel), lam_avg, mu_ref)
    eps_bar = np.asarray(eps_bar, float)
    if eps_init is not None:
        eps = eps_init.copy()
        for i in range(dim):
            for j in range(dim):
                eps[..., i, j] += eps_bar[i, j] - eps_init[..., i, j].mean()
    else:
        eps = np.zeros(shape + (dim, dim))
        for i in range(dim):
            for j in range(dim):
                eps[..., i, j] = eps_bar[i, j]
    letters = "xyz"[:dim]
    ein = f"{letters}khij,{letters}ij->{letters}kh"
    it_done = 0
    diff = np.inf
    for it in range(max_iter):
        e_el = eps - eps_plastic if eps_plastic is not None else eps
        sig = stress(e_el)
        tr = np.trace(eps, axis1=-2, axis2=-1)[..., None, None]
        tau = sig - (lam_avg * tr * np.eye(dim) + 2.0 * mu_ref * eps)
        tau_hat = np.zeros(shape + (dim, dim), dtype=complex)
        for i in range(dim):
            for j in range(dim):
                tau_hat[..., i, j] = fft_field(tau[..., i, j])
        et_hat = -np.einsum(ein, Gamma, tau_hat)
        et = np.zeros(shape + (dim, dim))
        for i in range(dim):
            for j in range(dim):
                et[..., i, j] = ifft_field(et_hat[..., i, j])
                et[..., i, j] -= et[..., i, j].mean()
        eps_new = eps_bar + et
        diff = np.linalg.norm(eps_new - eps) / (np.linalg.norm(eps) + 1e-20)
        eps = eps_new
        it_done = it + 1
        if diff < tol:
            break
    e_el = eps - eps_plastic if eps_plastic is not None else eps
    sig_out = stress(e_el)
    if info is not None:
        info.update(iters=it_done, diff=diff)
    axes = tuple(range(dim))
    return eps, sig_out, eps.mean(axis=axes), sig_out.mean(axis=axes)


def _field_shape(lam, mu, dim):
    return lam.shape if lam.ndim == dim else mu.shape


def spectral_solver_secant_2d(lam, mu, d, k, eps_bar, eps_plastic=None, max_iter=400, tol=1e-6, verbose=False,
                              pixel=1.0, plane_mode="plane_strain", Gamma=None, eps_init=None, info=None):
    lam = np.asarray(lam, float)
    mu = np.asarray(mu, float)
    shape = _field_shape(lam, mu, 2)
    la = float(lam.mean()) if lam.ndim > 0 else float(lam)
    ma = float(mu.mean()) if mu.ndim > 0 else float(mu)
    return _secant_ls(lambda e: stress_secant(e, lam, mu, d, k), 2, shape, la, ma, eps_bar, eps_plastic, max_iter,
                      tol, pixel, Gamma, eps_init, info)


def spectral_solver_secant_3d(lam, mu, d, k, eps_bar, eps_plastic=None, max_iter=400, tol=1e-6, verbose=False,
                              pixel=1.0, Gamma=None, eps_init=None, info=None):
    lam = np.asarray(lam, float)
    mu = np.asarray(mu, float)
    shape = _field_shape(lam, mu, 3)
    return _secant_ls(lambda e: stress_secant(e, lam, mu, d, k), 3, shape, float(lam.mean()), float(mu.mean()),
                      eps_bar, eps_plastic, max_iter, tol, pixel, Gamma, eps_init, info)


def spectral_solver_rose_2d(lam, mu, eta, mu_floor_fraction=0.1, eps_bar=None, eps_plastic=None, max_iter=400,
                            tol=1e-6, verbose=False, pixel=1.0, plane_mode="plane_strain", Gamma=None, eps_init=None,
                            info=None):
    """Reference shear modulus mu_ref = (mu_avg + floor mu_avg)/2, lam_avg as given."""
    lam = np.asarray(lam, float)
    mu = np.asarray(mu, float)
    shape = _field_shape(lam, mu, 2)
    la = float(lam.mean()) if lam.ndim > 0 else float(lam)
    ma = float(mu.mean()) if mu.ndim > 0 else float(mu)
    mu_ref = 0.5 * (ma + mu_floor_fraction * ma)
    return _secant_ls(lambda e: stress_rose_2d(e, lam, mu, eta, mu_floor_fraction, plane_mode), 2, shape, la, mu_ref,
                      eps_bar, eps_plastic, max_iter, tol, pixel, Gamma, eps_init, info)


def _macro_stepping(d, solve, E_avg, nu_avg, target_strain_mask, target_values, n_steps, tol_macro, max_iter_macro,
                    store, secant_slope):
    """Load stepping shared by the nonlinear drivers: targets ramp over
    n_steps + 1 steps; each macro iteration sets the strain-controlled
    entries, solves from the previous step's field and corrects the
    stress-controlled entries by the reference-medium compliance
    (err - nu_avg tr(err) I) / E_avg; with secant_slope (the Landau drivers,
    linear_elastic_simulator.py:2155-2205, :2390-2438) an entry whose
    last two macro iterates give a finite slope above 0.01 E_avg (and a strain
    change above 1e-14) takes err / slope instead; the history restarts each
    step."""
    mask = np.asarray(target_strain_mask, bool)
    smask = ~mask
    tv = np.asarray(target_values, float)
    em_l, sm_l, e_l, s_l, its = [], [], [], [], []
    eb = np.zeros((d, d))
    warm = None
    for s in range(n_steps + 1):
        tgt = tv * (s / n_steps)
        prev_e = prev_s = None
        for _ in range(max_iter_macro):
            eb[mask] = tgt[mask]
            eps, sig, epsM, sigM, it = solve(eb, warm)
            its.append(it)
            err = np.zeros((d, d))
            err[smask] = tgt[smask] - sigM[smask]
            mx = np.max(np.abs(err[smask])) if np.any(smask) else 0.0
            if mx < tol_macro:
                break
            de = (err - nu_avg * np.trace(err) * np.eye(d)) / E_avg
            dsc = de[smask]
            if secant_slope:
                e_sc = eb[smask].copy()
                s_sc = sigM[smask].copy()
                if prev_e is not None:
                    dh = e_sc - prev_e
                    with np.errstate(divide="ignore", invalid="ignore"):
                        slope = (s_sc - prev_s) / dh
                    ok = np.isfinite(slope) & (np.abs(dh) > 1e-14) & (slope > 0.01 * E_avg)
                    dsc[ok] = err[smask][ok] / slope[ok]
                prev_e, prev_s = e_sc, s_sc
            eb[smask] += dsc
        warm = eps
        em_l.append(epsM)
        sm_l.append(sigM)
        if store:
            e_l.append(eps)
            s_l.append(sig)
    return np.array(em_l), np.array(sm_l), e_l, s_l, its


def _medium_compliance(la, ma):
    return ma * (3.0 * la + 2.0 * ma) / (la + ma), la / (2.0 * (la + ma))


def secant_simulation(d, law, lam, mu, p1, p2, target_strain_mask, target_values, n_steps=20, pixel=1.0, tol_ls=1e-6,
                      max_iter_ls=400, tol_macro=1e-4, max_iter_macro=20, store=True, plane_mode="plane_strain"):
    """Mixed stress/strain load stepping of the secant-modulus laws
    (linear_elastic_simulator.py:1353-1600 secant 2D, :1603-1855 Rose 2D,
    :1857-2052 secant 3D): law "secant" (p1, p2 = d, k) or "rose" (eta,
    floor); compliance of the reference medium (mean lam, mean mu), mu_ref for
    the Rose law.  Returns (eps_macro, sig_macro, eps list, sig list, LS
    iteration counts per solve)."""
    lam = np.asarray(lam, float)
    mu = np.asarray(mu, float)
    la, ma = float(lam.mean()), float(mu.mean())
    if law == "rose":
        ma = 0.5 * (ma + p2 * ma)
    E_avg, nu_avg = _medium_compliance(la, ma)

    def solve(eb, warm):
        info = {}
        kw = dict(max_iter=max_iter_ls, tol=tol_ls, pixel=pixel, eps_init=warm, info=info)
        if law == "rose":
            r = spectral_solver_rose_2d(lam, mu, p1, p2, eps_bar=eb, plane_mode=plane_mode, **kw)
        elif d == 3:
            r = spectral_solver_secant_3d(lam, mu, p1, p2, eb, **kw)
        else:
            r = spectral_solver_secant_2d(lam, mu, p1, p2, eb, plane_mode=plane_mode, **kw)
        return r + (info["iters"],)
    return _macro_stepping(d, solve, E_avg, nu_avg, target_strain_mask, target_values, n_steps, tol_macro,
                           max_iter_macro, store, False)


def landau_simulation(d, lam, mu, v1, v2, v3, g1, g2, g3, g4, target_strain_mask, target_values, n_steps=20,
                      pixel=1.0, tol_ls=1e-6, max_iter_ls=400, tol_macro=1e-4, max_iter_macro=20, store=True,
                      plane_mode="plane_strain", anderson_m=5, solver="dbfft", **cap):
    """Mixed stress/strain load stepping of the Landau law
    (linear_elastic_simulator.py:2055-2284 in 2D, :2287-2516 in 3D): secant
    slope correction, compliance of (mean lam, mean mu); in 2D plane stress
    one e33 state carries across every solve of the run."""
    lam = np.asarray(lam, float)
    mu = np.asarray(mu, float)
    E_avg, nu_avg = _medium_compliance(float(lam.mean()), float(mu.mean()))
    e33_state = {} if (d == 2 and plane_mode == "plane_stress") else None

    def solve(eb, warm):
        info = {}
        kw = dict(max_iter=max_iter_ls, tol=tol_ls, pixel=pixel, eps_init=warm, info=info, **cap)
        if d == 3:
            r = spectral_solver_landau_3d(lam, mu, v1, v2, v3, g1, g2, g3, g4, eb, **kw)
        else:
            r = spectral_solver_landau_2d(lam, mu, v1, v2, v3, g1, g2, g3, g4, eb, plane_mode=plane_mode,
                                          e33_state=e33_state, anderson_m=anderson_m, solver=solver, **kw)
        return r + (info["iters"],)
    return _macro_stepping(d, solve, E_avg, nu_avg, target_strain_mask, target_values, n_steps, tol_macro,
                           max_iter_macro, store, True)
