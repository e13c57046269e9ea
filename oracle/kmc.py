"""Oracle: STZ barrier, KMC rates, event selection, flips (CPU restatement).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Follows:
  mgkmc/stz/barriers.py:4-100             compute_barrier (3D)
  mgkmc/kmc_simulator_functions.py:100-160 compute_barrier_2d
  mgkmc/stz/kmc.py:4-73                   compute_rates / _compute_rates_jit
  mgkmc/kmc_simulator_functions.py:5-56   compute_rates_2d
  mgkmc/stz/kmc.py:76-92, :94-103         select_event, decode_index
  mgkmc/kmc_simulator_functions.py:58-76  select_event_2d, decode_index_2d
  mgkmc/stz/cascade.py:4-41               find_unstable
  mgkmc/kmc_simulator_functions.py:162-181 find_unstable_2d
  mgkmc/stz/cascade.py:43-113             apply_flip_soa
  mgkmc/kmc_simulator_functions.py:183-218 apply_flip_soa_2d
  mgkmc/stz/catalog.py:7-40               stz_mode_glass / stz_catalog_glass
  mgkmc/kmc_simulator_functions.py:78-98  stz_catalog_glass_2d
  mgkmc/stz/barrier_generators.py:4-30    gaussian barrier generator

The reference loops are numba-jitted scalar loops; this restatement is
vectorised over voxels/modes but keeps every per-element operation and every
sum in the reference's order (e.g. the 9-term work sum sigma_ij*gamma_ij in
(0,0),(0,1),...,(2,2) order, the rate total and the selection running sum as
strictly sequential FP64 sums via np.cumsum).

Selection stream: the reference draws r = np.random.uniform(0, total) inside
numba (numba's own MT19937, not seeded by np.random.seed).  Here the draw is
an explicit argument; the build's contract is r = total * u with u from
numpy.random.RandomState(seed).random_sample() (MT19937 init_genrand +
genrand_res53), i.e. what numba produces after a jitted np.random.seed(seed).

Pinned by (tests/test_oracle_kmc.py): the reference's barrier KAT
Q = 1 - 0.6241509 eV (tests/test_barrier_physics.py:10-47), directional
modifier limits (cos = 1 -> 1, cos = -1 -> 0, orthogonal -> 1/4), brute-force
scalar loops on tiny grids (a literal transcription, used only as a checker),
the Arrhenius closed form, the tau closed form of tests/verify_tau.py:4-29, and
the von Mises increment of a pure shear (hand computed).
"""
import numpy as np

KB = 8.617e-5
GPA_NM3_TO_EV = 6.241509


def _work_sum(sig, cat, d):
    """sum_ij sig_ij * cat_ij in the reference's (i, j) loop order.
    sig: (..., d, d) broadcast against cat (..., M, d, d)."""
    w = np.zeros(np.broadcast_shapes(sig[..., None, 0, 0].shape, cat[..., 0, 0].shape))
    for i in range(d):
        for j in range(d):
            w = w + sig[..., None, i, j] * cat[..., i, j]
    return w


def _dir_modifier(cat, prev, d):
    dot = np.zeros(cat.shape[:-2])
    nm = np.zeros(cat.shape[:-2])
    npv = np.zeros(cat.shape[:-2])
    for i in range(d):
        for j in range(d):
            vm = cat[..., i, j]
            vp = prev[..., None, i, j]
            dot = dot + vm * vp
            nm = nm + vm ** 2
            npv = npv + vp ** 2
    npv = np.broadcast_to(npv, dot.shape)
    n_prev = np.sqrt(npv)
    n_mode = np.sqrt(nm)
    small = (n_prev < 1e-12) | (n_mode < 1e-12)
    with np.errstate(divide="ignore", invalid="ignore"):
        cos = dot / (n_mode * n_prev)
        mod = (1.0 + cos) ** 2 / 4.0
    return np.where(small, 1.0, mod)


def _barrier(Q, Q0, sigma, catalog, volume, soft_prop, last_event_time, current_time,
             prev_strain_dir, softening_cap, scheme, tau, d):
    g_t = soft_prop[..., 1]
    t_last = last_event_time
    if tau == np.inf:
        g_tc = np.where(g_t > 0, g_t, 0.0)
    else:
        dt = current_time - t_last
        dt = np.where(dt < 0, 0.0, dt)
        with np.errstate(invalid="ignore", over="ignore"):
            dec = g_t * np.exp(-dt / tau)
        g_tc = np.where(g_t > 0, np.where(t_last == -np.inf, 0.0, dec), 0.0)
    g_base = soft_prop[..., 0] + g_tc
    if softening_cap > 0:
        g_base = np.where(g_base > softening_cap, softening_cap, g_base)
    if scheme == 1:
        mod = _dir_modifier(catalog, prev_strain_dir, d)
    else:
        mod = np.ones(Q.shape)
    g_eff = mod * g_base[..., None]
    w = _work_sum(sigma, catalog, d)
    w_val = 0.5 * volume * (w / 1e9) * GPA_NM3_TO_EV
    Q[...] = Q0 * np.exp(-g_eff) - w_val


def compute_barrier_planar(Q_field, Q0_field, sigma_field, gxx, gxy, volume, soft_prop, last_event_time,
                           current_time, softening_cap, tau=np.inf):
    """compute_barrier (stz/barriers.py:4-100) for the isotropic scheme and
    planar catalogs [[gxx, gxy, 0], [gxy, -gxx, 0], [0, 0, 0]] held as their
    two parameters (..., M): the work sum keeps the reference's (i, j) order
    over the nonzero entries (the zero entries add exact zeros), so Q is the
    same, bit for bit, as compute_barrier on the full catalog
    (tests/test_oracle_kmc.py), without the (..., M, 3, 3) array."""
    g_t = soft_prop[..., 1]
    t_last = last_event_time
    if tau == np.inf:
        g_tc = np.where(g_t > 0, g_t, 0.0)
    else:
        dt = current_time - t_last
        dt = np.where(dt < 0, 0.0, dt)
        with np.errstate(invalid="ignore", over="ignore"):
            dec = g_t * np.exp(-dt / tau)
        g_tc = np.where(g_t > 0, np.where(t_last == -np.inf, 0.0, dec), 0.0)
    g_base = soft_prop[..., 0] + g_tc
    if softening_cap > 0:
        g_base = np.where(g_base > softening_cap, softening_cap, g_base)
    s = sigma_field
    w = s[..., None, 0, 0] * gxx
    w += s[..., None, 0, 1] * gxy
    w += s[..., None, 1, 0] * gxy
    w += s[..., None, 1, 1] * (-gxx)
    w_val = 0.5 * volume * (w / 1e9) * GPA_NM3_TO_EV
    np.multiply(Q0_field, np.exp(-(1.0 * g_base[..., None])), out=Q_field)
    Q_field -= w_val


def compute_barrier(Q_field, Q0_field, sigma_field, catalog, volume, soft_prop, last_event_time,
                    current_time, prev_strain_dir, softening_cap, softening_scheme_idx=0, tau=np.inf):
    """3D, in place on Q_field (stz/barriers.py:4-100)."""
    _barrier(Q_field, Q0_field, sigma_field, catalog, volume, soft_prop, last_event_time, current_time,
             prev_strain_dir, softening_cap, softening_scheme_idx, tau, 3)


def compute_barrier_2d(Q_field, Q0_field, sigma_field, catalog, volume, soft_prop, last_event_time,
                       current_time, prev_strain_dir, softening_cap, softening_scheme_idx=0, tau=np.inf):
    """2D, in place on Q_field (kmc_simulator_functions.py:100-160)."""
    _barrier(Q_field, Q0_field, sigma_field, catalog, volume, soft_prop, last_event_time, current_time,
             prev_strain_dir, softening_cap, softening_scheme_idx, tau, 2)


def _rates(Q, volume, temperature, nu0, cascade_mode):
    sp = Q.shape[:-1]
    M = Q.shape[-1]
    T = np.full(sp, float(temperature)) if np.isscalar(temperature) else np.asarray(temperature, float)
    with np.errstate(divide="ignore"):
        beta = np.where(T > 0.0, 1.0 / (KB * np.where(T > 0, T, 1.0)), np.inf)
    q = Q.reshape(-1, M)
    b = np.broadcast_to(beta.reshape(-1, 1), q.shape)
    with np.errstate(invalid="ignore", over="ignore"):
        r = np.where(q <= 0, volume * nu0, volume * nu0 * np.exp(-q * b))
    flat = np.arange(q.size, dtype=np.int64)
    if cascade_mode:
        keep = (q > 0).ravel()
    else:
        keep = np.ones(q.size, bool)
    rates = r.ravel()[keep]
    indices = flat[keep]
    total = float(np.cumsum(rates)[-1]) if rates.size else 0.0
    return rates, indices, total


def compute_rates(Q_field, volume, temperature, nu0=1e13, cascade_mode=True):
    """(rates, indices, total_rate) in (x,y,z,m) order (stz/kmc.py:4-73)."""
    return _rates(Q_field, volume, temperature, nu0, cascade_mode)


def compute_rates_2d(Q_field, volume, temperature, nu0=1e13, cascade_mode=True):
    """2D twin (kmc_simulator_functions.py:5-56)."""
    return _rates(Q_field, volume, temperature, nu0, cascade_mode)


def select_event(rates, total_rate, r):
    """First i with sequential running sum >= r; -1 if total <= 0; last index on
    fall-through (stz/kmc.py:76-92).  r is the draw U(0, total)."""
    if total_rate <= 0:
        return -1
    cs = np.cumsum(rates)
    i = int(np.searchsorted(cs, r, side="left"))
    return i if i < len(rates) else len(rates) - 1


def decode_index(flat_idx, ny, nz, M):
    m = flat_idx % M
    t = flat_idx // M
    z = t % nz
    t = t // nz
    y = t % ny
    x = t // ny
    return x, y, z, m


def decode_index_2d(flat_idx, ny, M):
    m = flat_idx % M
    t = flat_idx // M
    return t // ny, t % ny, m


def find_unstable(Q_field, threshold=0.0):
    """(count, 4) int64 array of (x, y, z, m) with Q < threshold, C order."""
    return np.argwhere(Q_field < threshold).astype(np.int64)


def find_unstable_2d(Q_field, threshold=0.0):
    return np.argwhere(Q_field < threshold).astype(np.int64)


def vm_sq_3d(c):
    e11, e22, e33 = c[0, 0], c[1, 1], c[2, 2]
    e12, e13, e23 = c[0, 1], c[0, 2], c[1, 2]
    return (e12 ** 2 + e23 ** 2 + e13 ** 2) + ((e22 - e33) ** 2 + (e33 - e11) ** 2 + (e11 - e22) ** 2) / 6.0


def vm_sq_2d(c):
    e11, e22, e12 = c[0, 0], c[1, 1], c[0, 1]
    return (e12 ** 2) + (e22 ** 2 + e11 ** 2 + (e11 - e22) ** 2) / 6.0


def apply_flip_soa(eps_plastic_field, key_field, soft_prop_field, last_event_time_field, catalog,
                   x, y, z, m, current_time, jp, jt, g_max, jn_frac):
    """In place (stz/cascade.py:43-113)."""
    nx, ny, nz = eps_plastic_field.shape[:3]
    c = catalog[x, y, z, m]
    eps_plastic_field[x, y, z] += c
    s = vm_sq_3d(c)
    gp = soft_prop_field[x, y, z, 0] + jp * s
    if g_max > 0 and gp > g_max:
        gp = g_max
    soft_prop_field[x, y, z, 0] = gp
    soft_prop_field[x, y, z, 1] = jt * s
    last_event_time_field[x, y, z] = current_time
    if jn_frac > 0.0:
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    if dx == 0 and dy == 0 and dz == 0:
                        continue
                    a, b, cc = (x + dx + nx) % nx, (y + dy + ny) % ny, (z + dz + nz) % nz
                    g = soft_prop_field[a, b, cc, 0] + jn_frac * jp * s
                    if g_max > 0 and g > g_max:
                        g = g_max
                    soft_prop_field[a, b, cc, 0] = g
                    soft_prop_field[a, b, cc, 1] = soft_prop_field[a, b, cc, 1] + jn_frac * jt * s


def apply_flip_soa_2d(eps_plastic_field, soft_prop_field, last_event_time_field, catalog,
                      x, y, m, current_time, jp, jt, g_max, jn_frac):
    """In place (kmc_simulator_functions.py:183-218)."""
    nx, ny = eps_plastic_field.shape[:2]
    c = catalog[x, y, m]
    eps_plastic_field[x, y] += c
    s = vm_sq_2d(c)
    gp = soft_prop_field[x, y, 0] + jp * s
    if g_max > 0 and gp > g_max:
        gp = g_max
    soft_prop_field[x, y, 0] = gp
    soft_prop_field[x, y, 1] = jt * s
    last_event_time_field[x, y] = current_time
    if jn_frac > 0.0:
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                if dx == 0 and dy == 0:
                    continue
                a, b = (x + dx + nx) % nx, (y + dy + ny) % ny
                g = soft_prop_field[a, b, 0] + jn_frac * jp * s
                if g_max > 0 and g > g_max:
                    g = g_max
                soft_prop_field[a, b, 0] = g
                soft_prop_field[a, b, 1] += jn_frac * jt * s


def stz_catalog_glass(M, gamma0, rs):
    """M planar modes [[gxx,gxy,0],[gxy,-gxx,0],[0,0,0]], gxx then gxy drawn per
    mode from rs.normal() (stz/catalog.py:7-40)."""
    cat = np.zeros((M, 3, 3))
    for i in range(M):
        gxx = 0.5 * gamma0 * rs.normal()
        gxy = 0.5 * gamma0 * rs.normal()
        cat[i, 0, 0] = gxx
        cat[i, 1, 1] = -gxx
        cat[i, 0, 1] = cat[i, 1, 0] = gxy
    return cat


def stz_catalog_glass_3d(M, gamma0, rs):
    """M fully 3D modes (stz/catalog.py:47-82): per mode a 3x3 standard-normal
    matrix, its symmetric deviatoric part G scaled by 1 / (sqrt(2) |G|) (zero,
    with no further draw, when |G| < 1e-12), times gamma0 sqrt(-2 ln u) with
    u = rs.uniform(1e-15, 1)."""
    cat = np.zeros((M, 3, 3))
    for i in range(M):
        A = rs.normal(size=(3, 3))
        S = 0.5 * (A + A.T)
        G = S - (1.0 / 3.0) * np.trace(S) * np.eye(3)
        nrm = np.sqrt(np.sum(G ** 2))
        if nrm < 1e-12:
            continue
        G = G * (1.0 / (np.sqrt(2.0) * nrm))
        cat[i] = G * (gamma0 * np.sqrt(-2.0 * np.log(rs.uniform(1e-15, 1.0))))
    return cat


def stz_catalog_glass_2d(M, gamma0, rs, stz_mode="pure_shear"):
    """2x2 modes (kmc_simulator_functions.py:78-98): gxx, gxy from rs.normal()
    per mode; pure shear [[gxx, gxy], [gxy, -gxx]], or simple shear with the
    same symmetric part, [[gxx, gxy + s R], [gxy - s R, -gxx]], R = |(gxx, gxy)|,
    s = +1 if rs.rand() > 0.5 else -1 (drawn after the mode's two normals)."""
    cat = np.zeros((M, 2, 2))
    for i in range(M):
        gxx = 0.5 * gamma0 * rs.normal()
        gxy = 0.5 * gamma0 * rs.normal()
        cat[i, 0, 0] = gxx
        cat[i, 1, 1] = -gxx
        if stz_mode == "pure_shear":
            cat[i, 0, 1] = cat[i, 1, 0] = gxy
        else:
            R = np.sqrt(gxx ** 2 + gxy ** 2)
            sgn = 1.0 if rs.rand() > 0.5 else -1.0
            cat[i, 0, 1] = gxy + sgn * R
            cat[i, 1, 0] = gxy - sgn * R
    return cat


def gaussian_barriers(shape, rs, mean=2.0, std=0.6, min_cutoff=0.1, max_cutoff=None):
    """np.random.normal(mean, std, shape) then clip (barrier_generators.py:4-30)."""
    b = rs.normal(loc=mean, scale=std, size=shape)
    if min_cutoff is not None or max_cutoff is not None:
        b = np.clip(b, a_min=min_cutoff, a_max=max_cutoff)
    return b


def rayleigh_barriers(shape, rs, loc=1.0, scale=0.5, min_cutoff=0.1, max_cutoff=None):
    """scipy.stats.rayleigh(loc, scale).rvs then clip (barrier_generators.py:32-58)."""
    from scipy.stats import rayleigh
    b = rayleigh(loc=loc, scale=scale).rvs(size=shape, random_state=rs)
    if min_cutoff is not None or max_cutoff is not None:
        b = np.clip(b, a_min=min_cutoff, a_max=max_cutoff)
    return b


def _mod_rayleigh(N, mean, std, rs):
    samples = np.empty(N)
    k = 0
    lo, hi = max(0.0, mean - 6 * std), mean + 6 * std
    xm = (mean + np.sqrt(mean ** 2 + 4 * std ** 2)) / 2.0
    cm = xm * np.exp(-(xm - mean) ** 2 / (2 * std ** 2))
    while k < N:
        need = N - k
        m = max(int(need * 3), 100)
        x = rs.uniform(lo, hi, m)
        u = rs.uniform(0, cm, m)
        ok_ = x[u < x * np.exp(-(x - mean) ** 2 / (2 * std ** 2))]
        t = min(len(ok_), need)
        samples[k:k + t] = ok_[:t]
        k += t
    return samples


def modified_rayleigh_barriers(shape, rs, mean=2.0, std=0.6, min_cutoff=0.1, max_cutoff=None):
    """Rejection sampling of x exp(-(x - mean)^2 / (2 std^2)) (barrier_generators.py:60-107)."""
    N = shape if isinstance(shape, int) else int(np.prod(shape))
    b = _mod_rayleigh(N, mean, std, rs).reshape(shape)
    if min_cutoff is not None or max_cutoff is not None:
        b = np.clip(b, a_min=min_cutoff, a_max=max_cutoff)
    return b


def modified_rayleigh_exp_barriers(shape, rs, mean=2.0, std=0.6, epsilon=0.1, ratio=0.8, min_cutoff=0.1,
                                   max_cutoff=None):
    """Modified Rayleigh / exponential mixture (barrier_generators.py:111-170)."""
    N = shape if isinstance(shape, int) else int(np.prod(shape))
    b = np.empty(N)
    m_exp = rs.uniform(0.0, 1.0, N) < ratio
    n_ray = int(np.sum(~m_exp))
    if n_ray > 0:
        b[~m_exp] = _mod_rayleigh(n_ray, mean, std, rs)
    if N - n_ray > 0:
        b[m_exp] = rs.exponential(scale=epsilon, size=N - n_ray)
    b = b.reshape(shape)
    if min_cutoff is not None or max_cutoff is not None:
        b = np.clip(b, a_min=min_cutoff, a_max=max_cutoff)
    return b
