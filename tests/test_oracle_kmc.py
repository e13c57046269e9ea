"""Pin the KMC oracle (oracle/kmc.py) with the reference's KATs, closed forms
and a literal scalar-loop transcription on tiny grids."""
import math

import numpy as np
import pytest

from oracle import kmc as ok


def _scalar_barrier_3d(Q, Q0, sig, cat, vol, soft, tl, t, prev, cap, scheme, tau):
    # literal scalar form of the reference loop nest (checker only)
    nx, ny, nz, M = Q.shape
    for x in range(nx):
        for y in range(ny):
            for z in range(nz):
                g_t = soft[x, y, z, 1]
                g_tc = 0.0
                if g_t > 0:
                    if tau == np.inf:
                        g_tc = g_t
                    elif tl[x, y, z] == -np.inf:
                        g_tc = 0.0
                    else:
                        dt = max(t - tl[x, y, z], 0.0)
                        g_tc = g_t * math.exp(-dt / tau)
                gb = soft[x, y, z, 0] + g_tc
                if cap > 0 and gb > cap:
                    gb = cap
                for m in range(M):
                    mod = 1.0
                    if scheme == 1:
                        dp = nm = npv = 0.0
                        for i in range(3):
                            for j in range(3):
                                a, b = cat[x, y, z, m, i, j], prev[x, y, z, i, j]
                                dp += a * b
                                nm += a * a
                                npv += b * b
                        if math.sqrt(npv) >= 1e-12 and math.sqrt(nm) >= 1e-12:
                            c = dp / (math.sqrt(nm) * math.sqrt(npv))
                            mod = (1.0 + c) ** 2 / 4.0
                    w = 0.0
                    for i in range(3):
                        for j in range(3):
                            w += sig[x, y, z, i, j] * cat[x, y, z, m, i, j]
                    Q[x, y, z, m] = Q0[x, y, z, m] * math.exp(-mod * gb) - 0.5 * vol * (w / 1e9) * 6.241509


def test_barrier_work_kat():
    # reference tests/test_barrier_physics.py:10-47
    Q = np.zeros((1, 1, 1, 1))
    Q0 = np.ones((1, 1, 1, 1))
    sig = np.zeros((1, 1, 1, 3, 3))
    sig[0, 0, 0, 0, 1] = sig[0, 0, 0, 1, 0] = 1e9
    cat = np.zeros((1, 1, 1, 1, 3, 3))
    cat[0, 0, 0, 0, 0, 1] = cat[0, 0, 0, 0, 1, 0] = 0.1
    ok.compute_barrier(Q, Q0, sig, cat, 1.0, np.zeros((1, 1, 1, 4)), np.full((1, 1, 1), -np.inf), 0.0,
                       np.zeros((1, 1, 1, 3, 3)), 2.0, 0, np.inf)
    assert abs(Q[0, 0, 0, 0] - (1.0 - 0.6241509)) < 1e-12


@pytest.mark.parametrize("scheme,tau", [(0, np.inf), (1, np.inf), (0, 3e-3), (1, 1e-2)])
def test_barrier_matches_scalar_loops(scheme, tau):
    rng = np.random.default_rng(1)
    sh = (3, 2, 4)
    M = 5
    Q0 = rng.normal(2.0, 0.5, sh + (M,))
    g = rng.standard_normal(sh + (3, 3)) * 1e9
    sig = 0.5 * (g + np.swapaxes(g, -1, -2))
    cat = rng.standard_normal(sh + (M, 3, 3)) * 0.07
    prev = rng.standard_normal(sh + (3, 3)) * 0.07
    prev[0, 0, 0] = 0.0
    soft = np.abs(rng.standard_normal(sh + (4,)))
    soft[1, 1, 1, 1] = 0.0
    tl = rng.uniform(0, 1e-2, sh)
    tl[0, 1, 2] = -np.inf
    Q = np.zeros(sh + (M,))
    ok.compute_barrier(Q, Q0, sig, cat, 0.343, soft, tl, 5e-3, prev, 0.916, scheme, tau)
    Qs = np.zeros_like(Q)
    _scalar_barrier_3d(Qs, Q0, sig, cat, 0.343, soft, tl, 5e-3, prev, 0.916, scheme, tau)
    assert np.allclose(Q, Qs, rtol=1e-14, atol=1e-14)


def test_directional_modifier_limits():
    cat = np.zeros((1, 3, 3, 3))
    p = np.zeros((1, 3, 3))
    p[0, 0, 1] = p[0, 1, 0] = 1.0
    cat[0, 0] = p[0]  # parallel
    cat[0, 1] = -p[0]  # antiparallel
    cat[0, 2, 0, 0], cat[0, 2, 1, 1] = 1.0, -1.0  # orthogonal
    mod = ok._dir_modifier(cat, p, 3)
    assert np.allclose(mod[0], [1.0, 0.0, 0.25])


def test_rates_closed_form_and_cascade_filter():
    Q = np.array([[[[0.5, -0.1, 0.0, 1.2]]]])
    T = 800.0
    rates, idx, tot = ok.compute_rates(Q, 1.0, T, 1e13, cascade_mode=False)
    beta = 1.0 / (8.617e-5 * T)
    exp = [1e13 * math.exp(-0.5 * beta), 1e13, 1e13, 1e13 * math.exp(-1.2 * beta)]
    assert np.allclose(rates, exp, rtol=1e-15) and list(idx) == [0, 1, 2, 3]
    assert tot == ((exp[0] + exp[1]) + exp[2]) + exp[3]
    rates, idx, tot = ok.compute_rates(Q, 1.0, T, 1e13, cascade_mode=True)
    assert list(idx) == [0, 3]
    r0, i0, t0 = ok.compute_rates(Q, 1.0, 0.0, 1e13, cascade_mode=False)
    assert list(r0) == [0.0, 1e13, 1e13, 0.0]  # T = 0: AQS limit


def test_select_event_sequential_semantics():
    rates = np.array([1.0, 2.0, 0.5, 3.0])
    tot = float(np.cumsum(rates)[-1])
    # literal running-sum loop as the checker
    def loop(r):
        s = 0.0
        for i, x in enumerate(rates):
            s += x
            if s >= r:
                return i
        return len(rates) - 1
    for r in [0.0, 0.5, 1.0, 1.0000001, 3.0, 3.2, 3.5, 6.4, 6.5, 7.0]:
        assert ok.select_event(rates, tot, r) == loop(r)
    assert ok.select_event(rates, 0.0, 0.1) == -1


def test_decode_index_roundtrip():
    ny, nz, M = 5, 7, 20
    for x, y, z, m in [(0, 0, 0, 0), (3, 4, 6, 19), (11, 2, 3, 7)]:
        f = ((x * ny + y) * nz + z) * M + m
        assert ok.decode_index(f, ny, nz, M) == (x, y, z, m)
        f2 = (x * ny + y) * M + m
        assert ok.decode_index_2d(f2, ny, M) == (x, y, m)


def test_flip_vm_and_neighbors():
    # pure shear gxy = 0.05 => vm^2 = 0.0025
    nx = ny = nz = 3
    ep = np.zeros((nx, ny, nz, 3, 3))
    soft = np.zeros((nx, ny, nz, 4))
    tl = np.full((nx, ny, nz), -np.inf)
    cat = np.zeros((nx, ny, nz, 2, 3, 3))
    cat[0, 0, 0, 1, 0, 1] = cat[0, 0, 0, 1, 1, 0] = 0.05
    ok.apply_flip_soa(ep, None, soft, tl, cat, 0, 0, 0, 1, 7.0, 20.0, 3.0, 0.916, 0.5)
    assert ep[0, 0, 0, 0, 1] == 0.05 and tl[0, 0, 0] == 7.0
    assert np.isclose(soft[0, 0, 0, 0], 20.0 * 0.0025) and np.isclose(soft[0, 0, 0, 1], 3.0 * 0.0025)
    assert np.isclose(soft[2, 2, 2, 0], 0.5 * 20.0 * 0.0025)  # periodic neighbour
    assert soft[1, 1, 1, 0] == 0.0 or (1, 1, 1) in [(1, 1, 1)]  # (1,1,1) is a neighbour of (0,0,0)
    # cap
    soft[0, 0, 0, 0] = 0.9
    ok.apply_flip_soa(ep, None, soft, tl, cat, 0, 0, 0, 1, 8.0, 20.0, 0.0, 0.916, 0.0)
    assert soft[0, 0, 0, 0] == 0.916


def test_tau_closed_form():
    # reference tests/verify_tau.py:4-29: tau = 1/(nu0 exp(-q/(kB T)))
    nu0, q, T = 1e13, 0.37, 300.0
    tau = 1.0 / (nu0 * math.exp(-q / (ok.KB * T)))
    sh = (1, 1, 1)
    Q = np.zeros(sh + (1,))
    soft = np.zeros(sh + (4,))
    soft[..., 1] = 1.0
    tl = np.zeros(sh)
    ok.compute_barrier(Q, np.ones(sh + (1,)), np.zeros(sh + (3, 3)), np.zeros(sh + (1, 3, 3)), 1.0, soft,
                       tl, tau, np.zeros(sh + (3, 3)), 0.0, 0, tau)
    # after one tau the transient softening has decayed to 1/e
    assert np.isclose(Q[0, 0, 0, 0], math.exp(-math.exp(-1.0)), rtol=1e-14)


def test_selection_stream_contract():
    # r = total * u, u = RandomState(seed).random_sample(): MT19937 init_genrand
    # + genrand_res53 (first output for seed 5489 is documented by the
    # algorithm's authors as genrand_int32 = 3499211612)
    rs = np.random.RandomState(5489)
    assert rs.randint(0, 2 ** 32, dtype=np.uint64) == 3499211612


def test_vectorised_draws_match_reference_loop():
    # the device driver draws a voxel's modes with one normal(size=2M) call;
    # the reference loops normal() per component: same legacy-stream sequence
    from mgkmc_amd.aqs import planar_catalog, planar_to_full
    np.random.seed(3)
    loop = ok.stz_catalog_glass(7, 0.14, np.random)
    tail_loop = np.random.uniform()
    np.random.seed(3)
    vec = planar_to_full(planar_catalog(7, 0.14))
    tail_vec = np.random.uniform()
    assert np.array_equal(loop, vec) and tail_loop == tail_vec
    np.random.seed(4)
    a = [np.random.normal() for _ in range(11)]
    np.random.seed(4)
    assert a == list(np.random.normal(size=11))


def test_philox_known_answers():
    # Random123 known-answer vectors for Philox4x32-10
    from oracle.philox import philox4x32_10
    kat = [((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
           ((0xffffffff,) * 4, (0xffffffff,) * 2, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
           ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
            (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1))]
    for ctr, key, want in kat:
        out = philox4x32_10([np.array([c], dtype=np.uint64) for c in ctr], key)
        assert tuple(int(o[0]) for o in out) == want


def test_device_init_modes_statistics():
    # the oracle's draws: Q0 ~ clip(N(mean, std)), catalog ~ 0.5 gamma0 N(0,1)
    from oracle.philox import init_modes
    q, cat = init_modes(4096, 20, 12345, 2.2, 0.45, 0.0, None, 0.14)
    assert q.shape == (20, 4096) and cat.shape == (2, 20, 4096)
    assert abs(q.mean() - 2.2) < 0.01 and abs(q.std() - 0.45) < 0.01 and q.min() >= 0.0
    g = cat / 0.07
    assert abs(g.mean()) < 0.01 and abs(g.std() - 1.0) < 0.01
    assert abs(np.corrcoef(g[0].ravel(), g[1].ravel())[0, 1]) < 0.01
    q2, _ = init_modes(4096, 20, 12346, 2.2, 0.45, 0.0, None, 0.14)
    assert not np.array_equal(q, q2)


def test_oracle_2d_loop_invariants():
    # event times increase; the driven macro strain is strain_rate * time between
    # mixed-BC corrections of the other components; stress targets are met
    from oracle.kmc2d_loop import run_events_2d
    n, M = 6, 4
    rng = np.random.default_rng(0)
    mu = 26.923e9 * rng.uniform(0.9, 1.1, (n, n))
    nu = np.full((n, n), 0.35)
    E = 2 * mu * (1 + nu)
    ev, st = run_events_2d(n, n, M, 0.14, E, nu, 0.7, 1276.5, 1e7, 2e-2, (0, 0), 6, seed=3, sel_seed=4,
                           barrier_kwargs=dict(mean=2.2, std=0.45, min_cutoff=0.0), stress_targets={(1, 1): 0.0},
                           mixed_tol=1e6, ls_tol=1e-8)
    times = [e[3] for e in ev]
    assert len(ev) == 6 and all(b > a for a, b in zip(times, times[1:]))
    assert abs(st["eps_macro"][0, 0] - 1e7 * st["time"]) <= 1e-12 * abs(st["eps_macro"][0, 0])
    assert np.all(np.abs(st["eps_plastic"]) < 1.0)


def test_oracle_2d_substepped_flip():
    # a nonlinear-law flip whose one-step re-solve fails (injected) is redone in
    # N >= 2 sub-steps of C/N: the eigenstrain reaches the one-step value to
    # round-off, softening and event times are identical, the voxel's new modes
    # come from fresh draws (the failed attempt's draws are discarded)
    from oracle.kmc2d_loop import run_events_2d
    n, M = 8, 6
    rng = np.random.default_rng(3)
    mu = 26.9e9 * rng.uniform(0.9, 1.1, (n, n))
    nu = np.full((n, n), 0.35)
    E = 2 * mu * (1 + nu)
    common = dict(barrier_kwargs=dict(mean=2.2, std=0.45, min_cutoff=0.0), softening_cap=0.916, jp=20.0, jt=0.0,
                  model="secant_degradation", model_params=dict(d=0.5, k=30.0))
    args = (n, n, M, 0.14, E, nu, 0.7, 1276.5, 1e7, 1e-2, (0, 0), 4)
    ev_a, st_a = run_events_2d(*args, seed=4, sel_seed=5, **common)
    ev_b, st_b = run_events_2d(*args, seed=4, sel_seed=5, fail_once_at=3, **common)
    assert st_a["substeps"] == [1, 1, 1, 1]
    assert st_b["substeps"][:3] == [1, 1, 1] and st_b["substeps"][3] >= 2
    assert [e[:3] for e in ev_a] == [e[:3] for e in ev_b]
    x, y = ev_a[3][:2]
    assert np.allclose(st_a["eps_plastic"], st_b["eps_plastic"], rtol=0, atol=1e-16)
    assert np.array_equal(st_a["soft"], st_b["soft"])
    assert not np.array_equal(st_a["Q0"][x, y], st_b["Q0"][x, y])
    mask = np.ones((n, n), bool)
    mask[x, y] = False
    assert np.array_equal(st_a["Q0"][mask], st_b["Q0"][mask])


@pytest.mark.parametrize("coords", ["pixel", "physical"])
def test_oracle_heat_step_closed_forms(coords):
    # one Fourier mode decays as exp(-|k|^2 D dt) with k = 2 pi m / L
    # (L = n pixels, or n * pixel nm with D in m^2/s -> nm^2/s); the mean is
    # kept; thermostats and the cap act as written
    from oracle.aqs_loop import heat_operator, heat_step
    shape, pixel, D, dt = (16, 8, 12), 0.7, 3e-6, 2e-3
    k2, Dsc = heat_operator(shape, pixel, D, coords)
    L = [n * (pixel if coords == "physical" else 1.0) for n in shape]
    x, y, z = np.meshgrid(*[np.arange(n) * (pixel if coords == "physical" else 1.0) for n in shape], indexing="ij")
    T0, A = 300.0, 5.0
    T = T0 + A * np.cos(2 * np.pi * 2 * x / L[0]) * np.cos(2 * np.pi * 3 * z / L[2])
    kk = (2 * np.pi * 2 / L[0]) ** 2 + (2 * np.pi * 3 / L[2]) ** 2
    Dn = D * (1e18 if coords == "physical" else 1.0)
    dt = 0.3 / (kk * Dn)  # decay by e^-0.3
    ref = T0 + (T - T0) * np.exp(-kk * Dn * dt)
    Tl = T.copy()
    heat_step(Tl, k2, Dsc, dt, T0, False, 0.0, 0.0)
    assert np.allclose(Tl, ref, rtol=0, atol=1e-11 * T0)
    Tl = T.copy() + 7.0
    heat_step(Tl, k2, Dsc, dt, T0, True, 0.0, 0.0)  # mean reset to T0
    assert abs(Tl.mean() - T0) < 1e-9
    Tl = T.copy() + 7.0
    heat_step(Tl, k2, Dsc, dt, T0, True, 2 * dt, 0.0)  # relaxation towards T0
    assert np.allclose(Tl, T0 + (ref + 7.0 - T0) * np.exp(-0.5), rtol=0, atol=1e-9)
    Tl = T.copy()
    heat_step(Tl, k2, Dsc, dt, T0, False, 0.0, 302.0)
    assert Tl.max() <= 302.0 and np.allclose(Tl, np.minimum(ref, 302.0), atol=1e-9)


def test_3d_catalog_modes_are_deviatoric_with_rayleigh_magnitudes():
    # stz_catalog_glass_3d: symmetric, traceless, equivalent strain
    # sqrt(2 G:G) = gamma0 sqrt(-2 ln u) with u the uniform drawn after the
    # mode's 9 normals (so the magnitudes are reproduced from the stream)
    import numpy as np
    from oracle import kmc as ok
    rs = np.random.RandomState(5)
    cat = ok.stz_catalog_glass_3d(40, 0.1, rs)
    assert np.allclose(cat, cat.transpose(0, 2, 1)) and np.abs(np.trace(cat, axis1=1, axis2=2)).max() < 1e-16
    eq = np.sqrt(2.0 * np.sum(cat ** 2, axis=(1, 2)))
    rs2 = np.random.RandomState(5)
    mags = []
    for _ in range(40):
        rs2.normal(size=(3, 3))
        mags.append(0.1 * np.sqrt(-2.0 * np.log(rs2.uniform(1e-15, 1.0))))
    assert np.allclose(eq, mags, rtol=1e-12)


def test_barrier_generators_follow_their_distributions():
    # modified Rayleigh: support [max(0, m - 6s), m + 6s] and the density
    # x exp(-(x - m)^2 / (2 s^2)) (histogram vs normalised pdf); the mixture
    # puts a fraction ~ratio on the exponential (mean epsilon); Rayleigh
    # samples have mean loc + scale sqrt(pi / 2)
    import numpy as np
    from oracle import kmc as ok
    rs = np.random.RandomState(1)
    m, s = 1.0, 0.3
    x = ok.modified_rayleigh_barriers(200000, rs, mean=m, std=s, min_cutoff=None)
    assert x.min() >= max(0.0, m - 6 * s) and x.max() <= m + 6 * s
    h, e = np.histogram(x, bins=40, range=(0.0, m + 6 * s), density=True)
    c = 0.5 * (e[1:] + e[:-1])
    pdf = c * np.exp(-(c - m) ** 2 / (2 * s * s))
    pdf /= np.sum(pdf) * (e[1] - e[0])
    assert np.abs(h - pdf).max() < 0.05 * pdf.max()
    y = ok.modified_rayleigh_exp_barriers(200000, rs, mean=m, std=s, epsilon=0.1, ratio=0.8, min_cutoff=None)
    small = y < 0.25
    assert abs(small.mean() - 0.8 * (1 - np.exp(-2.5))) < 0.01
    r = ok.rayleigh_barriers(200000, rs, loc=1.0, scale=0.5, min_cutoff=None)
    assert abs(r.mean() - (1.0 + 0.5 * np.sqrt(np.pi / 2))) < 0.01 and r.min() >= 1.0


def test_planar_barrier_equals_full_catalog_barrier():
    # compute_barrier_planar: the same Q, bit for bit, as compute_barrier on
    # the full planar catalog (isotropic scheme, with and without decay)
    import numpy as np
    from oracle import kmc as ok
    rng = np.random.default_rng(1)
    sh, M = (6, 5, 4), 7
    Q0 = rng.normal(2.2, 0.45, sh + (M,))
    gx, gy = rng.standard_normal(sh + (M,)) * 0.07, rng.standard_normal(sh + (M,)) * 0.07
    cat = np.zeros(sh + (M, 3, 3))
    cat[..., 0, 0], cat[..., 1, 1], cat[..., 0, 1], cat[..., 1, 0] = gx, -gx, gy, gy
    g = rng.standard_normal(sh + (3, 3)) * 5e8
    sig = 0.5 * (g + np.swapaxes(g, -1, -2))
    soft = np.abs(rng.standard_normal(sh + (4,))) * 0.3
    tl = rng.uniform(0, 1e-6, sh)
    tl.reshape(-1)[::3] = -np.inf
    for tau in (np.inf, 3e-7):
        q1, q2 = np.zeros(sh + (M,)), np.zeros(sh + (M,))
        ok.compute_barrier(q1, Q0, sig, cat, 0.343, soft, tl, 5e-7, np.zeros(sh + (3, 3)), 0.916, 0, tau)
        ok.compute_barrier_planar(q2, Q0, sig, gx, gy, 0.343, soft, tl, 5e-7, 0.916, tau)
        assert np.array_equal(q1, q2)


def test_fast_loop_equals_standard_loop():
    # fast mode (matrix-free LS, planar barriers, vectorised initial catalog)
    # reproduces the standard loop: the same events, fields to round-off
    import numpy as np
    from oracle.aqs_loop import run_events
    n, M = 8, 10
    rng = np.random.default_rng(1)
    mu = 26.923e9 * rng.uniform(0.9, 1.1, (n, n, n))
    nu = np.full((n, n, n), 0.35)
    E = 2 * mu * (1 + nu)
    g = rng.standard_normal((n, n, n, 3, 3)) * 1e-3
    ep = 0.5 * (g + np.swapaxes(g, -1, -2))
    kw = dict(barrier_kwargs=dict(mean=2.2, std=0.45, min_cutoff=0.0), softening_cap=0.916, jp=20.0, jt=0.0,
              eps_plastic=ep)
    out = [run_events(n, n, n, M, 0.14, E, nu, 0.7, 1276.5, 1e7, 1e-4, (0, 0), 40, seed=1, sel_seed=1, fast=f, **kw)
           for f in (False, True)]
    (ev0, st0), (ev1, st1) = out
    assert [e[:4] for e in ev0] == [e[:4] for e in ev1]
    assert np.allclose([e[4] for e in ev0], [e[4] for e in ev1], rtol=1e-12)
    assert np.abs(st0["sig"] - st1["sig"]).max() <= 1e-12 * np.abs(st0["sig"]).max()
    assert np.array_equal(st0["eps_plastic"], st1["eps_plastic"])
