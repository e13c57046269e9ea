"""GPU parity of the KMC kernels with the oracle: barriers and rates to the
last bits of exp(), identical selections, compaction and flips, and an
identical event sequence of the device-resident KMC loop."""
import numpy as np
import pytest

from oracle import kmc as ok

pytestmark = pytest.mark.gpu


def _state(shape, M, seed, d=3, general=False):
    rng = np.random.default_rng(seed)
    Q0 = rng.normal(2.2, 0.45, shape + (M,))
    g = rng.standard_normal(shape + (d, d)) * 5e8
    sig = 0.5 * (g + np.swapaxes(g, -1, -2))
    if general:
        cat = rng.standard_normal(shape + (M, d, d)) * 0.07
    else:
        gx = rng.standard_normal(shape + (M,)) * 0.07
        gy = rng.standard_normal(shape + (M,)) * 0.07
        cat = np.zeros(shape + (M, d, d))
        cat[..., 0, 0], cat[..., 1, 1], cat[..., 0, 1], cat[..., 1, 0] = gx, -gx, gy, gy
    soft = np.abs(rng.standard_normal(shape + (4 if d == 3 else 2,))) * 0.3
    tl = rng.uniform(0, 1e-6, shape)
    tl.reshape(-1)[::3] = -np.inf
    prev = rng.standard_normal(shape + (d, d)) * 0.07
    return Q0, sig, cat, soft, tl, prev


def _close(a, b, rtol):
    scale = np.maximum(np.abs(b), 1e-300)
    return np.all(np.abs(a - b) <= rtol * scale)


@pytest.mark.parametrize("d,scheme,tau,general", [(3, 0, np.inf, False), (3, 1, 3e-7, False), (3, 1, np.inf, True),
                                                  (2, 0, np.inf, False), (2, 1, 5e-7, True)])
def test_barrier_parity(gpu, d, scheme, tau, general):
    from mgkmc_amd.stz.barriers import compute_barrier
    from mgkmc_amd.kmc_simulator_functions import compute_barrier_2d
    shape = (6, 5, 4) if d == 3 else (9, 7)
    M = 20
    Q0, sig, cat, soft, tl, prev = _state(shape, M, 1, d, general)
    Qr = np.zeros(shape + (M,))
    Qg = np.zeros(shape + (M,))
    args = (0.343, soft, tl, 5e-7, prev, 0.916, scheme, tau)
    if d == 3:
        ok.compute_barrier(Qr, Q0, sig, cat, *args)
        compute_barrier(Qg, Q0, sig, cat, *args)
    else:
        ok.compute_barrier_2d(Qr, Q0, sig, cat, *args)
        compute_barrier_2d(Qg, Q0, sig, cat, *args)
    # identical op order; only exp() may differ by an ulp
    assert np.abs(Qg - Qr).max() <= 4e-16 * np.abs(Qr).max() * 8, np.abs(Qg - Qr).max()


@pytest.mark.parametrize("cascade", [False, True])
def test_rates_parity(gpu, cascade):
    from mgkmc_amd.stz.kmc import compute_rates
    rng = np.random.default_rng(2)
    Q = rng.normal(0.8, 0.6, (7, 6, 5, 20))
    r0, i0, t0 = ok.compute_rates(Q, 1.0, 1276.5, 1e13, cascade_mode=cascade)
    r1, i1, t1 = compute_rates(Q, 1.0, 1276.5, 1e13, cascade_mode=cascade)
    assert np.array_equal(i0, i1)
    assert _close(r1, r0, 4e-15)
    assert abs(t1 - t0) <= 1e-12 * t0


def test_select_array_parity(gpu):
    from mgkmc_amd.kmc_device import select_host
    rng = np.random.default_rng(3)
    rates = np.exp(rng.normal(0, 3, 50000))
    rates[::7] = 0.0
    tot = float(np.cumsum(rates)[-1])
    us = np.concatenate([rng.random(200), [0.0, 1e-12, 0.5, 1 - 1e-15]])
    for u in us:
        r = tot * u
        assert select_host(rates, r) == ok.select_event(rates, tot, r), u


def test_find_unstable_parity(gpu):
    from mgkmc_amd.stz.cascade import find_unstable
    rng = np.random.default_rng(4)
    Q = rng.normal(0.5, 0.5, (9, 8, 7, 20))
    assert np.array_equal(find_unstable(Q, 0.0), ok.find_unstable(Q, 0.0))
    assert find_unstable(Q, -100.0).shape == (0, 4)


@pytest.mark.parametrize("jn", [0.0, 0.3])
def test_flip_parity(gpu, jn):
    from mgkmc_amd.stz.cascade import apply_flip_soa
    shape = (5, 4, 3)
    M = 6
    Q0, sig, cat, soft, tl, prev = _state(shape, M, 5)
    ep = np.zeros(shape + (3, 3))
    a = (ep.copy(), None, soft.copy(), tl.copy())
    b = (ep.copy(), None, soft.copy(), tl.copy())
    ok.apply_flip_soa(*a, cat, 4, 0, 2, 3, 1e-6, 20.0, 3.0, 0.916, jn)
    apply_flip_soa(*b, cat, 4, 0, 2, 3, 1e-6, 20.0, 3.0, 0.916, jn)
    assert np.array_equal(a[0], b[0])
    assert np.allclose(a[2][..., :2], b[2][..., :2], rtol=1e-15, atol=0)
    assert np.array_equal(a[3], b[3])


def test_fused_select_matches_oracle(gpu):
    # device barrier+rate pass and two-level selection vs the oracle's
    # compute_barrier -> compute_rates -> select_event on the same draws
    import torch
    from mgkmc_amd.kmc_device import KmcState, params, SIG_GEN64
    from mgkmc_amd import device as dev
    shape = (10, 9, 8)
    M = 20
    Q0, sig, cat, soft, tl, prev = _state(shape, M, 6)
    st = KmcState(shape, M, Q0, cat, soft_prop=soft, last_event_time=tl, prev_strain_dir=prev, directional=True)
    sig_d = dev.upload(np.ascontiguousarray(sig.reshape(-1, 9).T), 0)
    T = 1276.5
    p = params(volume=0.343, temperature=T, cap=0.916, tau=3e-7, time=5e-7, scheme=1)
    total = st.barrier_rates(sig_d, SIG_GEN64, p)
    Q = np.zeros(shape + (M,))
    ok.compute_barrier(Q, Q0, sig, cat, 0.343, soft, tl, 5e-7, prev, 0.916, 1, 3e-7)
    rates, idx, tot = ok.compute_rates(Q, 1.0, T, 1e13, cascade_mode=False)
    assert abs(total - tot) <= 1e-12 * tot
    rng = np.random.default_rng(7)
    for u in list(rng.random(300)) + [0.0, 0.999999999]:
        v, m = st.select(u, sig_d, SIG_GEN64, p)
        i = ok.select_event(rates, tot, tot * u)
        assert (v * M + m) == idx[i], (u, v, m, idx[i])
        qref = Q.reshape(-1, M)[v, m]
        assert abs(st.sel.q - qref) <= 1e-14 * max(1.0, abs(qref))


def test_kmc_event_sequence_matches_oracle(gpu):
    # device-resident loop vs the oracle loop: same event list for fixed seeds
    from mgkmc_amd.aqs import ThermalSimulation
    from mgkmc_amd.stz import kmc as kmod
    from oracle.aqs_loop import run_events
    n = 8
    M = 20
    rng = np.random.default_rng(1)
    mu = 26.923e9 * rng.uniform(0.9, 1.1, (n, n, n))
    nu = np.full((n, n, n), 0.35)
    E = 2 * mu * (1 + nu)
    T = 1276.5
    kw = dict(mean=2.2, std=0.45, min_cutoff=0.0)
    ev_ref, st_ref = run_events(n, n, n, M, 0.14, E, nu, 0.7, T, 1e7, 1e-4, (0, 0), 40, seed=11, sel_seed=12,
                                barrier_kwargs=kw, softening_cap=0.916, jp=20.0, jt=0.0)
    np.random.seed(11)
    kmod.seed_selection(12)
    sim = ThermalSimulation(n, n, n, M, 0.14, E, nu, pixel=0.7, barrier_generator="gaussian", barrier_kwargs=kw,
                            softening_cap=0.916, jp=20.0, jt=0.0, temperature=T, strain_rate=1e7)
    ev = sim.run_simulation(10 ** 9, 1e-4, component=(0, 0), max_events=40)
    got = [e[:4] for e in ev]
    ref = [e[:4] for e in ev_ref]
    print("events:", got[:10], flush=True)
    assert got == ref
    assert np.allclose([e[4] for e in ev], [e[4] for e in ev_ref], rtol=1e-10)
    ep = sim.eps_plastic
    assert np.allclose(ep, st_ref["eps_plastic"], rtol=0, atol=1e-15)
    sig = sim.sig_field
    assert np.abs(sig - st_ref["sig"]).max() <= 1e-10 * np.abs(st_ref["sig"]).max()


@pytest.mark.parametrize("scheme,plane_mode", [("isotropic", "plane_strain"), ("directional", "plane_stress")])
def test_kmc2d_event_sequence_matches_oracle(gpu, scheme, plane_mode):
    # 2D device loop (KmcSimulation2D) vs the oracle 2D loop, with the mixed-BC
    # secant loop active between increments
    from mgkmc_amd.kmc_simulator import KmcSimulation2D
    from oracle.kmc2d_loop import run_events_2d
    print(f"[kmc2d] {scheme} {plane_mode}", flush=True)
    n, M = 16, 20
    rng = np.random.default_rng(2)
    mu = 26.923e9 * rng.uniform(0.9, 1.1, (n, n))
    nu = np.full((n, n), 0.35)
    E = 2 * mu * (1 + nu)
    kw = dict(mean=2.2, std=0.45, min_cutoff=0.0)
    args = dict(softening_scheme=scheme, softening_cap=0.916, jp=20.0, jt=0.0)
    ev_ref, st_ref = run_events_2d(n, n, M, 0.14, E, nu, 0.7, 1276.5, 1e7, 1e-2, (0, 0), 30, seed=21, sel_seed=22,
                                   barrier_kwargs=kw, plane_mode=plane_mode, stress_targets={(1, 1): 0.0},
                                   mixed_tol=1e6, **args)
    sim = KmcSimulation2D(n, n, M, 0.14, E, nu, pixel=0.7, barrier_generator="gaussian", barrier_kwargs=kw,
                          temperature=1276.5, strain_rate=1e7, plane_mode=plane_mode,
                          rng=np.random.RandomState(21), selection_stream=np.random.RandomState(22), **args)
    # the oracle seeds the global stream; the simulator was given an equal private one
    ev = sim.run_simulation(10 ** 9, 1e-2, component=(0, 0), stress_targets={(1, 1): 0.0}, mixed_tol=1e6,
                            max_events=30)
    got = [e[:3] for e in ev]
    ref = [e[:3] for e in ev_ref]
    print("events:", got[:8], flush=True)
    assert got == ref
    assert np.allclose([e[3] for e in ev], [e[3] for e in ev_ref], rtol=1e-10)
    assert [e[4] for e in ev] == [e[4] for e in ev_ref]
    assert np.allclose(sim.eps_plastic, st_ref["eps_plastic"], rtol=0, atol=1e-15)
    assert np.abs(sim.eps_macro - st_ref["eps_macro"]).max() <= 1e-10 * np.abs(st_ref["eps_macro"]).max()
    sig = sim.sig_field
    assert np.abs(sig - st_ref["sig"]).max() <= 1e-10 * np.abs(st_ref["sig"]).max()


def test_device_init_modes_match_oracle(gpu):
    from mgkmc_amd.kmc_device import KmcState
    from mgkmc_amd.kmc_engine import device_init_modes
    from oracle.philox import init_modes
    shape, M = (12, 10, 9), 7
    st = KmcState(shape, M, None, None)
    prm = dict(mean=2.2, std=0.45, min_cutoff=0.0, max_cutoff=None)
    device_init_modes(shape, M, 0.14, prm, 987654321987, st.Q0, st.cat)
    q, cat = init_modes(int(np.prod(shape)), M, 987654321987, 2.2, 0.45, 0.0, None, 0.14)
    gq = st.Q0.cpu().numpy()
    gc = st.cat.cpu().numpy()
    assert np.abs(gq - q).max() <= 1e-12 * np.abs(q).max()
    assert np.abs(gc - cat).max() <= 1e-12 * np.abs(cat).max()


def test_kmc_refined_ls_event_sequence_matches_fp64(gpu):
    # the FP32 + FP64 refinement LS changes sigma by ~tol; the event sequence of
    # the FP64 path is reproduced (selection differs only for draws within
    # ~1e-8 relative of a boundary)
    from mgkmc_amd.aqs import ThermalSimulation
    n, M = 16, 20
    rng = np.random.default_rng(5)
    mu = 26.923e9 * rng.uniform(0.9, 1.1, (n, n, n))
    nu = np.full((n, n, n), 0.35)
    E = 2 * mu * (1 + nu)
    kw = dict(mean=2.2, std=0.45, min_cutoff=0.0)
    out = {}
    for prec in ("fp64", "mixed"):
        sim = ThermalSimulation(n, n, n, M, 0.14, E, nu, pixel=0.7, barrier_generator="gaussian", barrier_kwargs=kw,
                                softening_cap=0.916, jp=20.0, jt=0.0, temperature=1276.5, strain_rate=1e7,
                                ls_precision=prec, rng=np.random.RandomState(31),
                                selection_stream=np.random.RandomState(32))
        out[prec] = (sim.run_simulation(10 ** 9, 1e-4, component=(0, 0), max_events=25), sim.sig_field)
    assert [e[:4] for e in out["fp64"][0]] == [e[:4] for e in out["mixed"][0]]
    s64, smx = out["fp64"][1], out["mixed"][1]
    assert np.abs(smx - s64).max() <= 1e-5 * np.abs(s64).max()


def test_kmc2d_athermal_aqs_with_eigenstrain_matches_oracle(gpu):
    # the 2D plain-method recipe at reduced size: T = 0 (only q <= 0 modes are
    # events), initial zero-mean Gaussian eigenstrain, LS tol 1e-8, strain
    # increments along xx with sig_yy = 0
    from mgkmc_amd.kmc_simulator import KmcSimulation2D
    from oracle.kmc2d_loop import run_events_2d
    n, M = 16, 20
    rng = np.random.default_rng(0)
    mu = 26.923e9 * rng.uniform(0.9, 1.1, (n, n))
    nu = np.full((n, n), 0.35)
    E = 2 * mu * (1 + nu)
    g = rng.standard_normal((3, n, n)) * 1e-3
    g -= g.mean(axis=(1, 2), keepdims=True)
    ep = np.zeros((n, n, 2, 2))
    ep[..., 0, 0], ep[..., 1, 1], ep[..., 0, 1] = g[0], g[1], g[2]
    ep[..., 1, 0] = g[2]
    kw = dict(mean=0.55, std=0.15, min_cutoff=0.0)  # low barriers: ~60 athermal events in 8 increments
    args = dict(softening_cap=0.916, jp=20.0, jt=0.0)
    ev_ref, st_ref = run_events_2d(n, n, M, 0.14, E, nu, 0.7, 0.0, 1e7, 1e-3, (0, 0), 10 ** 9, seed=5, sel_seed=6,
                                   barrier_kwargs=kw, stress_targets={(1, 1): 0.0}, mixed_tol=1e6, ls_tol=1e-8,
                                   n_global_steps=8, eps_plastic=ep, **args)
    sim = KmcSimulation2D(n, n, M, 0.14, E, nu, pixel=0.7, barrier_generator="gaussian", barrier_kwargs=kw,
                          temperature=0.0, strain_rate=1e7, rng=np.random.RandomState(5),
                          selection_stream=np.random.RandomState(6), **args)
    sim.eps_plastic = ep
    sim.solver_args = {"tol": 1e-8}
    ev = sim.run_simulation(8, 1e-3, component=(0, 0), stress_targets={(1, 1): 0.0}, mixed_tol=1e6)
    print(f"[kmc2d aqs] events={len(ev)} (oracle {len(ev_ref)})", flush=True)
    assert len(ev_ref) > 0
    assert [e[:3] for e in ev] == [e[:3] for e in ev_ref]
    assert np.allclose(sim.eps_plastic, st_ref["eps_plastic"], rtol=0, atol=1e-15)
    assert np.abs(sim.eps_macro - st_ref["eps_macro"]).max() <= 1e-10 * np.abs(st_ref["eps_macro"]).max()
    assert np.abs(sim.sig_field - st_ref["sig"]).max() <= 1e-10 * np.abs(st_ref["sig"]).max()


_NL_LANDAU = dict(v1=-236.5e9, v2=-27.4e9, v3=11.8e9, g1=-5640.7e9, g2=2207.1e9, g3=-332.5e9, g4=-305.2e9)


@pytest.mark.parametrize("model,plane_mode,mp,fail", [
    ("secant_degradation", "plane_strain", dict(d=0.5, k=30.0), None),
    ("secant_degradation", "plane_stress", dict(d=0.5, k=30.0), 5),
    ("rose", "plane_stress", dict(eta=25.0, mu_floor_fraction=0.1), None),
    ("landau", "plane_strain", _NL_LANDAU, None),
    ("landau", "plane_stress", _NL_LANDAU, 3),
])
def test_kmc2d_nonlinear_law_sequence_matches_oracle(gpu, model, plane_mode, mp, fail):
    # KmcSimulation2D with a nonlinear small-strain law: the same 20 events as
    # the oracle loop; with ``fail`` the one-step re-solve after that event's
    # flip is made to fail (host-side injection, on both sides) and the flip is
    # redone in sub-steps
    from mgkmc_amd.kmc_simulator import KmcSimulation2D
    from oracle.kmc2d_loop import run_events_2d
    print(f"[kmc2d nonlinear] {model} {plane_mode} fail={fail}", flush=True)
    n, M = 16, 20
    rng = np.random.default_rng(2)
    mu = 26.923e9 * rng.uniform(0.9, 1.1, (n, n))
    nu = np.full((n, n), 0.35)
    E = 2 * mu * (1 + nu)
    kw = dict(mean=2.2, std=0.45, min_cutoff=0.0)
    args = dict(softening_cap=0.916, jp=20.0, jt=0.0)
    ev_ref, st_ref = run_events_2d(n, n, M, 0.14, E, nu, 0.7, 1276.5, 1e7, 1e-2, (0, 0), 20, seed=21, sel_seed=22,
                                   barrier_kwargs=kw, plane_mode=plane_mode, stress_targets={(1, 1): 0.0},
                                   mixed_tol=1e6, model=model, model_params=mp, fail_once_at=fail, ls_tol=1e-8,
                                   **args)
    sim = KmcSimulation2D(n, n, M, 0.14, E, nu, pixel=0.7, barrier_generator="gaussian", barrier_kwargs=kw,
                          temperature=1276.5, strain_rate=1e7, plane_mode=plane_mode, hyperelastic_model=model,
                          rng=np.random.RandomState(21), selection_stream=np.random.RandomState(22), **mp, **args)
    sim.solver_args = {"tol": 1e-8}
    # the failure injection wraps the Python elastic_run: those cases run the Python
    # sub-stepping; the others run the C++ event loop with the library's sub-stepping
    sim.native_loop = fail is None
    assert sim._native_ok() == (fail is None)
    armed = [False]
    substeps = []
    real_run, real_fs = sim.elastic_run, sim._flip_substepped

    def elastic_run(eps_macro):
        if armed[0]:
            armed[0] = False
            raise ValueError("injected solver failure")
        return real_run(eps_macro)

    def flip_substepped(flips):
        if fail is not None and len(sim.kmc_event_history) == fail:
            armed[0] = True
        real_fs(flips)
        substeps.append(sim.last_substeps)
    sim.elastic_run = elastic_run
    sim._flip_substepped = flip_substepped
    ev = sim.run_simulation(10 ** 9, 1e-2, component=(0, 0), stress_targets={(1, 1): 0.0}, mixed_tol=1e6,
                            max_events=20)
    if fail is None:
        assert substeps == [] and len(sim.kmc_substeps) == len(ev)
        substeps = sim.kmc_substeps
    print("events:", [e[:3] for e in ev[:8]], "substeps:", substeps, flush=True)
    assert substeps == st_ref["substeps"]
    assert [e[:3] for e in ev] == [e[:3] for e in ev_ref]
    assert np.allclose([e[3] for e in ev], [e[3] for e in ev_ref], rtol=1e-10)
    assert np.allclose(sim.eps_plastic, st_ref["eps_plastic"], rtol=0, atol=1e-15)
    assert np.abs(sim.eps_macro - st_ref["eps_macro"]).max() <= 1e-10 * np.abs(st_ref["eps_macro"]).max()
    assert np.abs(sim.sig_field - st_ref["sig"]).max() <= 1e-10 * np.abs(st_ref["sig"]).max()


def test_kmc3d_landau_sequence_matches_linear_limit_and_runs(gpu):
    # ThermalSimulation with hyperelastic_model="landau": with all Landau
    # coefficients zero the law is linear, so the event sequence equals the
    # linear simulation's; with the coefficients on, events run and the stress
    # stays finite
    from mgkmc_amd.aqs import ThermalSimulation
    n, M = 12, 8
    rng = np.random.default_rng(4)
    E = 70e9 * rng.uniform(0.9, 1.1, (n, n, n))
    nu = np.full((n, n, n), 0.3)
    kw = dict(mean=2.0, std=0.5, min_cutoff=0.0)
    base = dict(pixel=0.7, barrier_generator="gaussian", barrier_kwargs=kw, temperature=1200.0, strain_rate=1e7)
    zero = {k: 0.0 for k in _NL_LANDAU}
    runs = []
    for model, prm in (("svk", zero), ("landau", zero)):
        sim = ThermalSimulation(n, n, n, M, 0.1, E, nu, hyperelastic_model=model, rng=np.random.RandomState(3),
                                selection_stream=np.random.RandomState(4), **prm, **base)
        sim.solver_args = {"tol": 1e-10}
        runs.append([e[:4] for e in sim.run_simulation(10 ** 9, 1e-3, max_events=15)])
    print("[kmc3d landau] linear limit events:", runs[1][:5], flush=True)
    assert runs[0] == runs[1]
    sims = []
    for native in (True, False):  # the C++ event loop (cold Landau solves) against the Python one
        sim = ThermalSimulation(n, n, n, M, 0.1, E, nu, hyperelastic_model="landau", rng=np.random.RandomState(3),
                                selection_stream=np.random.RandomState(4), **_NL_LANDAU, **base)
        sim.native_loop = native
        assert sim._native_ok() == native
        ev = sim.run_simulation(10 ** 9, 1e-3, max_events=10)
        assert len(ev) == 10 and np.all(np.isfinite(sim.sig_mean()))
        sims.append(sim)
    a, b = sims
    assert a.kmc_event_history == b.kmc_event_history
    assert np.array_equal(a.eps_plastic, b.eps_plastic) and np.array_equal(a.sig_field, b.sig_field)
    sa, sb = a.rng.get_state(), b.rng.get_state()
    assert np.array_equal(sa[1], sb[1]) and sa[2:] == sb[2:]


@pytest.mark.parametrize("T,mean,steps,model,mp", [
    (400.0, 0.9, 4, "linear", {}),
    (0.0, 0.55, 6, "linear", {}),
    (400.0, 0.8, 4, "secant_degradation", dict(d=0.3, k=20.0)),
])
def test_kmc2d_cascade_mode_matches_oracle(gpu, T, mean, steps, model, mp):
    # cascade_mode: the cascade batches (order of flips included), the KMC
    # events between them (rates over Q > 0) and the final fields equal the
    # oracle loop's
    from mgkmc_amd.kmc_simulator import KmcSimulation2D
    from oracle.kmc2d_loop import run_events_2d
    print(f"[kmc2d cascade] T={T} mean={mean} {model}", flush=True)
    n, M = 16, 20
    rng = np.random.default_rng(0)
    mu = 26.923e9 * rng.uniform(0.9, 1.1, (n, n))
    nu = np.full((n, n), 0.35)
    E = 2 * mu * (1 + nu)
    kw = dict(mean=mean, std=0.15, min_cutoff=0.0)
    args = dict(softening_cap=0.916, jp=20.0, jt=0.0)
    ev_ref, st_ref = run_events_2d(n, n, M, 0.14, E, nu, 0.7, T, 1e7, 1e-3, (0, 0), 10 ** 9, seed=5, sel_seed=6,
                                   barrier_kwargs=kw, stress_targets={(1, 1): 0.0}, mixed_tol=1e6, ls_tol=1e-8,
                                   n_global_steps=steps, cascade_mode=True, max_cascade_steps_pct=0.02,
                                   model=model, model_params=mp, **args)
    hm = model if model != "linear" else "svk"
    sim = KmcSimulation2D(n, n, M, 0.14, E, nu, pixel=0.7, barrier_generator="gaussian", barrier_kwargs=kw,
                          temperature=T, strain_rate=1e7, cascade_mode=True, hyperelastic_model=hm,
                          rng=np.random.RandomState(5), selection_stream=np.random.RandomState(6), **mp, **args)
    sim.solver_args = {"tol": 1e-8}
    ev = sim.run_simulation(steps, 1e-3, component=(0, 0), stress_targets={(1, 1): 0.0}, mixed_tol=1e6,
                            max_cascade_steps_pct=0.02)
    print(f"[kmc2d cascade] events {len(ev)} (oracle {len(ev_ref)}), cascade flips {len(sim.cascade_history)} "
          f"(oracle {len(st_ref['cascades'])})", flush=True)
    assert len(st_ref["cascades"]) > 0
    assert sim.cascade_history == st_ref["cascades"]
    assert [e[:3] for e in ev] == [e[:3] for e in ev_ref]
    assert np.allclose(sim.eps_plastic, st_ref["eps_plastic"], rtol=0, atol=1e-15)
    assert np.abs(sim.sig_field - st_ref["sig"]).max() <= 1e-10 * np.abs(st_ref["sig"]).max()


@pytest.mark.parametrize("T,step_size,n_events,steps,mixed_tol", [
    (800.0, 1e-3, 20, 5, 1e6),      # cascades between KMC events
    (0.0, 2e-3, None, 4, 1e5),      # athermal: cascades inside the mixed-BC loop
    (300.0, 2e-3, None, 4, 1e5),
])
def test_kmc3d_cascade_mode_matches_oracle(gpu, T, step_size, n_events, steps, mixed_tol):
    from mgkmc_amd.aqs import ThermalSimulation
    from oracle.aqs_loop import run_events
    print(f"[kmc3d cascade] T={T} mixed_tol={mixed_tol}", flush=True)
    n, M = 8, 10
    rng = np.random.default_rng(1)
    E = 70e9 * rng.uniform(0.9, 1.1, (n, n, n))
    nu = np.full((n, n, n), 0.3)
    kw = dict(mean=0.5, std=0.15, min_cutoff=0.0)
    args = dict(softening_cap=0.916, jp=20.0, jt=0.0)
    ev_ref, st_ref = run_events(n, n, n, M, 0.14, E, nu, 0.7, T, 1e7, step_size, (0, 0),
                                10 ** 9 if n_events is None else n_events, seed=7, sel_seed=8, barrier_kwargs=kw,
                                mixed_tol=mixed_tol, ls_tol=1e-8, n_global_steps=steps, cascade_mode=True, **args)
    sim = ThermalSimulation(n, n, n, M, 0.14, E, nu, pixel=0.7, barrier_generator="gaussian", barrier_kwargs=kw,
                            temperature=T, strain_rate=1e7, cascade_mode=True, rng=np.random.RandomState(7),
                            selection_stream=np.random.RandomState(8), **args)
    sim.solver_args = {"tol": 1e-8}
    ev = sim.run_simulation(steps, step_size, component=(0, 0), mixed_tol=mixed_tol, max_events=n_events)
    print(f"[kmc3d cascade] events {len(ev)} (oracle {len(ev_ref)}), cascade flips {len(sim.cascade_history)} "
          f"(oracle {len(st_ref['cascades'])})", flush=True)
    assert len(st_ref["cascades"]) > 0
    assert sim.cascade_history == st_ref["cascades"]
    assert [e[:4] for e in ev] == [e[:4] for e in ev_ref]
    assert np.allclose(sim.eps_plastic, st_ref["eps_plastic"], rtol=0, atol=1e-15)
    assert np.abs(sim.sig_field - st_ref["sig"]).max() <= 1e-10 * np.abs(st_ref["sig"]).max()


@pytest.mark.parametrize("d,prec,scheme", [(3, "fp64", "isotropic"), (3, "mixed", "directional"),
                                           (2, "fp32", "directional"), (2, "mixed", "isotropic")])
def test_native_loop_matches_python_loop(gpu, d, prec, scheme):
    # the library's C++ event loop (mgk_kmc_run) against the per-event Python
    # loop on the same seeds: same events, times, fields and stream states after
    from mgkmc_amd.aqs import ThermalSimulation
    from mgkmc_amd.kmc_simulator import KmcSimulation2D
    print(f"[native loop] {d}D {prec} {scheme}", flush=True)
    n, M = (12, 10) if d == 3 else (32, 20)
    shape = (n,) * d
    rng = np.random.default_rng(9)
    E = 70e9 * rng.uniform(0.9, 1.1, shape)
    nu = np.full(shape, 0.3)
    kw = dict(mean=1.0, std=0.3, min_cutoff=0.0)
    sims = []
    for native in (True, False):
        args = dict(pixel=0.7, barrier_generator="gaussian", barrier_kwargs=kw, temperature=900.0, strain_rate=1e7,
                    softening_scheme=scheme, softening_cap=0.916, jp=20.0, jt=5.0, neighbor_softening_fraction=0.1,
                    ls_precision=prec, rng=np.random.RandomState(31), selection_stream=np.random.RandomState(32))
        sim = ThermalSimulation(n, n, n, M, 0.14, E, nu, **args) if d == 3 else KmcSimulation2D(n, n, M, 0.14, E, nu,
                                                                                               **args)
        sim.native_loop = native
        assert sim._native_ok() == native
        sim.run_simulation(3, 2e-3, component=(0, 0), mixed_tol=1e6, max_events=40)
        sims.append(sim)
    a, b = sims
    print(f"[native loop] {len(a.kmc_event_history)} events, first {a.kmc_event_history[:3]}", flush=True)
    assert len(a.kmc_event_history) == 40
    assert [e[:-2] for e in a.kmc_event_history] == [e[:-2] for e in b.kmc_event_history]
    assert [e[-1] for e in a.kmc_event_history] == [e[-1] for e in b.kmc_event_history]
    assert np.allclose([e[-2] for e in a.kmc_event_history], [e[-2] for e in b.kmc_event_history], rtol=1e-12)
    assert a.ls_iters_total == b.ls_iters_total
    assert np.array_equal(a.eps_plastic, b.eps_plastic)
    assert np.abs(a.sig_field - b.sig_field).max() <= 1e-10 * np.abs(b.sig_field).max()
    sa, sb = a.rng.get_state(), b.rng.get_state()
    assert np.array_equal(sa[1], sb[1]) and sa[2:] == sb[2:]
    assert a.selection_stream.get_state()[2] == b.selection_stream.get_state()[2]


@pytest.mark.parametrize("d,prec,solver,model,cascade", [
    (3, "fp64", "newton_cg", "svk", False),
    (3, "mixed", "newton_cg", "svk", False),
    (3, "fp64", "dbfft", "svk", False),
    (2, "fp64", "newton_cg", "landau", False),
    (2, "fp64", "dbfft", "svk", False),
    (2, "fp64", "newton_cg", "svk", True),
])
def test_native_finite_flip_matches_python_flip(gpu, d, prec, solver, model, cascade):
    # mgk_kmc_fs_flip (the sub-stepped finite-strain flip in the library) against the
    # per-sub-step Python path on the same seeds: the same events, sub-step counts,
    # cascades and host stream states; F_p to 1e-14 and, in FP64, F, F_macro and sigma
    # to 1e-10 relative (expm(C/N) is rounded differently: scipy's Pade form against a
    # Taylor form, both exact to rounding)
    from mgkmc_amd.aqs import ThermalSimulation
    from mgkmc_amd.kmc_simulator import KmcSimulation2D
    print(f"[native fs flip] {d}D {prec} {solver} {model} cascade={cascade}", flush=True)
    # the settings of the finite-strain oracle tests (the cascade case: 2D cascade_mode's)
    n, M = (8, 6) if d == 3 else ((12, 10) if cascade else (12, 6))
    shape = (n,) * d
    E = 70e9 * np.random.default_rng(0 if cascade else 3).uniform(0.9, 1.1, shape)
    nu = np.full(shape, 0.3)
    kw = dict(mean=0.3, std=0.15, min_cutoff=0.0) if cascade else dict(mean=1.0, std=0.3, min_cutoff=0.0)
    mp = dict(v1=-236.5e9, v2=-27.4e9, v3=11.8e9, g1=-5640.7e9, g2=2207.1e9, g3=-332.5e9, g4=-305.2e9) \
        if model == "landau" else {}
    g0, seeds = (0.05, (5, 6)) if cascade else (0.03, (7, 8))
    sims, used = [], []
    for native in (True, False):
        args = dict(pixel=0.7, barrier_generator="gaussian", barrier_kwargs=kw, temperature=900.0 if d == 3 else 400.0,
                    strain_rate=1e7, softening_cap=0.916, jp=20.0, jt=0.0 if cascade else 5.0,
                    strain_assumption="finite_strain", solver=solver, ls_precision=prec, cascade_mode=cascade,
                    rng=np.random.RandomState(seeds[0]), selection_stream=np.random.RandomState(seeds[1]))
        if d == 3:
            sim = ThermalSimulation(n, n, n, M, g0, E, nu, **args)
        else:
            sim = KmcSimulation2D(n, n, M, g0, E, nu, hyperelastic_model=model, **mp, **args)
        sim.native_loop = native
        assert sim._native_fs_ok() == native
        # KMC-mode runs: whole events in the C++ loop (mgk_kmc_run calling the flip);
        # cascade mode: per-event Python with the flips (events and batches) in C++
        assert sim._native_ok() == (native and not cascade)
        calls = [0]
        real_native = sim._flip_finite_native

        def counted(pairs, real_native=real_native, calls=calls):
            calls[0] += 1
            return real_native(pairs)
        sim._flip_finite_native = counted
        kw_run = dict(component=(0, 0), max_events=10 if cascade else 8)
        if d == 2:
            kw_run["stress_targets"] = {(1, 1): 0.0}
        if cascade:
            kw_run["max_cascade_steps_pct"] = 0.05
        sim.run_simulation(4 if cascade else 3, 2e-3, **kw_run)
        sims.append(sim)
        used.append(calls[0])
    a, b = sims
    print(f"[native fs flip] {len(a.kmc_event_history)} events, flips through the wrapper {used}, "
          f"substeps {a.kmc_substeps}", flush=True)
    assert used[1] == 0 and (used[0] > 0 if cascade else used[0] == 0)
    assert len(a.kmc_substeps) > 0 and a.kmc_substeps == b.kmc_substeps
    assert [e[:-2] for e in a.kmc_event_history] == [e[:-2] for e in b.kmc_event_history]
    assert np.allclose([e[-2] for e in a.kmc_event_history], [e[-2] for e in b.kmc_event_history], rtol=1e-12)
    if cascade:
        assert len(a.cascade_history) == len(b.cascade_history) and len(a.cascade_history) > 0
    # the FP32-Krylov form turns the rounding difference into FP32-level ones (its
    # oracle test allows 1e-5 in F)
    tol, stol = (1e-10, 1e-10) if prec == "fp64" else (1e-6, 1e-3)
    assert np.abs(a.F_plastic - b.F_plastic).max() <= 1e-14
    assert np.abs(a.F_field - b.F_field).max() <= tol * np.abs(b.F_field).max()
    assert np.abs(a.F_macro - b.F_macro).max() <= tol
    assert np.abs(a.sig_field - b.sig_field).max() <= stol * np.abs(b.sig_field).max()
    sa, sb = a.rng.get_state(), b.rng.get_state()
    assert np.array_equal(sa[1], sb[1]) and sa[2:] == sb[2:]


@pytest.mark.parametrize("model,mp", [("secant_degradation", dict(d=0.5, k=30.0)), ("landau", _NL_LANDAU)])
def test_native_law_flip_retry_schedule_matches_python(gpu, model, mp):
    # mgk_kmc_law_flip's retry schedule against the Python sub-stepping: a macro strain
    # whose mean stress exceeds 20 GPa makes the one-step solve and every one of the 20
    # sub-stepped attempts fail (forced from the host), so both raise the reference's
    # ValueError after drawing the same redraws (21 attempts) and leave the same state
    from mgkmc_amd.kmc_simulator import KmcSimulation2D
    print(f"[native law retry] {model}", flush=True)
    n, M = 16, 8
    rng = np.random.default_rng(2)
    mu = 26.923e9 * rng.uniform(0.9, 1.1, (n, n))
    nu = np.full((n, n), 0.35)
    E = 2 * mu * (1 + nu)
    out = []
    for native in (True, False):
        sim = KmcSimulation2D(n, n, M, 0.14, E, nu, pixel=0.7, barrier_generator="gaussian",
                              barrier_kwargs=dict(mean=2.2, std=0.45, min_cutoff=0.0), temperature=1276.5,
                              strain_rate=1e7, plane_mode="plane_strain", hyperelastic_model=model,
                              rng=np.random.RandomState(21), selection_stream=np.random.RandomState(22),
                              softening_cap=0.916, jp=20.0, jt=0.0, **mp)
        sim.native_loop = native
        assert sim._native_law_ok() == native
        sim.eps_macro[...] = np.diag([0.4, 0.0])  # <sigma_xx> far above 20 GPa
        with pytest.raises(ValueError, match="sub-stepping"):
            sim._flip_substepped([(37, 3), (200, 5)])
        out.append(sim)
    a, b = out
    sa, sb = a.rng.get_state(), b.rng.get_state()
    assert np.array_equal(sa[1], sb[1]) and sa[2:] == sb[2:]
    for name in ("eps_p", "soft", "tlast", "cat", "Q0"):
        xa, xb = getattr(a.kmc, name), getattr(b.kmc, name)
        assert np.array_equal(xa.cpu().numpy(), xb.cpu().numpy()), name
    print(f"[native law retry] last N {a.last_substeps if hasattr(a, 'last_substeps') else None}", flush=True)


@pytest.mark.parametrize("d,prec,T,pct,records", [
    (3, "fp64", 0.0, 0.3, None),        # athermal: cascades inside the mixed-BC loop, logged per batch
    (3, "mixed", 0.0, 0.002, None),     # batch limit: the cascade is truncated
    (3, "fp64", 300.0, 0.3, (3, 2)),    # small records: each cascade continues over several calls
    (2, "fp64", 400.0, 0.02, None),
    (2, "fp32", 0.0, 0.02, (2, 1)),
])
def test_native_cascade_matches_python_cascade(gpu, tmp_path, monkeypatch, d, prec, T, pct, records):
    # the library's C++ cascade loop (mgk_kmc_cascade) against the per-batch
    # Python loop on the same seeds: same batches and flips, per-batch logs,
    # events, fields and stream state
    from mgkmc_amd.aqs import ThermalSimulation
    from mgkmc_amd.kmc_simulator import KmcSimulation2D
    print(f"[native cascade] {d}D {prec} T={T} pct={pct} records={records}", flush=True)
    monkeypatch.chdir(tmp_path)
    n, M = (8, 10) if d == 3 else (16, 20)
    shape = (n,) * d
    rng = np.random.default_rng(1)
    E = 70e9 * rng.uniform(0.9, 1.1, shape)
    nu = np.full(shape, 0.3)
    kw = dict(mean=0.5 if d == 3 else (0.55 if T == 0 else 0.6), std=0.15, min_cutoff=0.0)
    sims = []
    for native in (True, False):
        args = dict(pixel=0.7, barrier_generator="gaussian", barrier_kwargs=kw, temperature=T, strain_rate=1e7,
                    cascade_mode=True, softening_cap=0.916, jp=20.0, jt=0.0, ls_precision=prec,
                    rng=np.random.RandomState(7), selection_stream=np.random.RandomState(8))
        if d == 3:
            sim = ThermalSimulation(n, n, n, M, 0.14, E, nu, **args)
        else:
            sim = KmcSimulation2D(n, n, M, 0.14, E, nu, **args)
        sim.solver_args = {"tol": 1e-8}
        sim.native_loop = native
        if records is not None:
            sim.cascade_records = records
        assert sim._native_cascade_ok() == native
        if d == 3:
            sim.run_simulation(4, 2e-3, component=(0, 0), mixed_tol=1e5 if T == 0 else 1e6,
                               max_cascade_steps_pct=pct, max_events=None if T == 0 else 20)
        else:
            sim.run_simulation(6 if T == 0 else 4, 1e-3, component=(0, 0), stress_targets={(1, 1): 0.0}, mixed_tol=1e6,
                               max_cascade_steps_pct=pct, max_events=20)
        sims.append(sim)
    a, b = sims
    print(f"[native cascade] flips {len(a.cascade_history)} (python {len(b.cascade_history)}), events "
          f"{len(a.kmc_event_history)}, logged {len(a.history_global)}", flush=True)
    assert len(b.cascade_history) > 0
    assert a.cascade_history == b.cascade_history
    assert a.kmc_event_history == b.kmc_event_history
    assert a.history_global == b.history_global
    assert np.array_equal(a.eps_plastic, b.eps_plastic)
    assert np.array_equal(a.sig_field, b.sig_field)
    sa, sb = a.rng.get_state(), b.rng.get_state()
    assert np.array_equal(sa[1], sb[1]) and sa[2:] == sb[2:]


@pytest.mark.parametrize("prec,solver", [("fp64", "newton_cg"), ("mixed", "newton_cg"), ("fp64", "dbfft")])
def test_kmc3d_finite_strain_matches_oracle(gpu, prec, solver):
    # strain_assumption="finite_strain" (svk, newton_cg or dbfft): every flip enters F_p
    # as expm(C/N) over N sub-steps, each followed by a Newton-Krylov step with
    # the stress targets; events, sub-step counts, F, F_p and F_macro equal
    # the oracle loop's (FP64; the FP32-Krylov form within its Newton tolerance)
    from mgkmc_amd.aqs import ThermalSimulation
    from oracle.aqs_loop import run_events
    print(f"[kmc3d finite] {prec} {solver}", flush=True)
    n, M = 8, 6
    rng = np.random.default_rng(3)
    E = 70e9 * rng.uniform(0.9, 1.1, (n, n, n))
    nu = np.full((n, n, n), 0.3)
    kw = dict(mean=1.0, std=0.3, min_cutoff=0.0)
    args = dict(softening_cap=0.916, jp=20.0, jt=5.0)
    ev_ref, st_ref = run_events(n, n, n, M, 0.03, E, nu, 0.7, 900.0, 1e7, 2e-3, (0, 0), 8, seed=7, sel_seed=8,
                                barrier_kwargs=kw, n_global_steps=3, strain_assumption="finite_strain", solver=solver,
                                **args)
    sim = ThermalSimulation(n, n, n, M, 0.03, E, nu, pixel=0.7, barrier_generator="gaussian", barrier_kwargs=kw,
                            temperature=900.0, strain_rate=1e7, strain_assumption="finite_strain", solver=solver,
                            ls_precision=prec, rng=np.random.RandomState(7), selection_stream=np.random.RandomState(8),
                            **args)
    ev = sim.run_simulation(3, 2e-3, component=(0, 0), max_events=8)
    subs = sim.kmc_substeps  # N of every KMC event's flip (the C++ event loop's or the Python one's)
    print(f"[kmc3d finite] events {[e[:4] for e in ev]} substeps {subs} (oracle {st_ref['substeps']})", flush=True)
    assert [e[:4] for e in ev] == [e[:4] for e in ev_ref]
    assert subs == st_ref["substeps"] and max(subs) >= 3
    # FP64: fields to 1e-10 relative; FP32 Krylov: within its Newton tolerance
    tol, stol = (1e-10, 1e-10) if prec == "fp64" else (1e-5, 1e-2)
    Fr = np.einsum("ijxyz->xyzij", st_ref["F"])
    eF = np.abs(sim.F_field - Fr).max() / np.abs(Fr).max()
    eS = np.abs(sim.sig_field - st_ref["sig"]).max() / np.abs(st_ref["sig"]).max()
    print(f"[kmc3d finite] relF={eF:.2e} relSig={eS:.2e}", flush=True)
    assert eF <= tol
    assert np.abs(sim.F_plastic - np.einsum("ijxyz->xyzij", st_ref["F_plastic"])).max() <= 1e-14
    assert np.abs(sim.F_macro - st_ref["F_macro"]).max() <= tol
    assert eS <= stol


@pytest.mark.parametrize("d,thermostat,tau_bath,cascade,coords", [
    (3, False, 0.0, False, "pixel"),
    (3, True, 0.0, False, "physical"),
    (2, True, 5e-13, False, "pixel"),
    (2, False, 0.0, True, "pixel"),
])
def test_kmc_thermal_matches_oracle(gpu, d, thermostat, tau_bath, cascade, coords):
    # enable_thermal: local temperatures in the rates, spectral heat conduction
    # over every waiting time (device mgk_heat_step), flip heating, thermostat
    # and cap: events and the final temperature field equal the oracle loop's
    from mgkmc_amd.aqs import ThermalSimulation
    from mgkmc_amd.kmc_simulator import KmcSimulation2D
    from oracle.aqs_loop import run_events
    from oracle.kmc2d_loop import run_events_2d
    print(f"[kmc thermal] {d}D thermostat={thermostat} tau_bath={tau_bath} cascade={cascade} {coords}", flush=True)
    n, M = (8, 8) if d == 3 else (16, 12)
    shape = (n,) * d
    rng = np.random.default_rng(6)
    E = 70e9 * rng.uniform(0.9, 1.1, shape)
    nu = np.full(shape, 0.3)
    kw = dict(mean=0.8 if cascade else 1.0, std=0.3, min_cutoff=0.0)
    th = dict(enable_thermal=True, Cp=420.0, rho=6125.0, thermal_diffusivity=3e-6 if coords == "physical" else 2e11,
              thermal_coords=coords, temperature_cap=1500.0, thermostat=thermostat, tau_bath=tau_bath)
    args = dict(softening_cap=0.916, jp=20.0, jt=5.0)
    T, steps, step_size = 700.0, 3, 2e-3
    if d == 3:
        ev_ref, st_ref = run_events(n, n, n, M, 0.14, E, nu, 0.7, T, 1e7, step_size, (0, 0), 25, seed=3, sel_seed=4,
                                    barrier_kwargs=kw, mixed_tol=1e6, ls_tol=1e-8, n_global_steps=steps,
                                    cascade_mode=cascade, **args, **th)
        sim = ThermalSimulation(n, n, n, M, 0.14, E, nu, pixel=0.7, barrier_generator="gaussian", barrier_kwargs=kw,
                                temperature=T, strain_rate=1e7, cascade_mode=cascade, rng=np.random.RandomState(3),
                                selection_stream=np.random.RandomState(4), **args, **th)
    else:
        ev_ref, st_ref = run_events_2d(n, n, M, 0.14, E, nu, 0.7, T, 1e7, step_size, (0, 0), 25, seed=3, sel_seed=4,
                                       barrier_kwargs=kw, stress_targets={(1, 1): 0.0}, mixed_tol=1e6, ls_tol=1e-8,
                                       n_global_steps=steps, cascade_mode=cascade, **args, **th)
        sim = KmcSimulation2D(n, n, M, 0.14, E, nu, pixel=0.7, barrier_generator="gaussian", barrier_kwargs=kw,
                              temperature=T, strain_rate=1e7, cascade_mode=cascade, rng=np.random.RandomState(3),
                              selection_stream=np.random.RandomState(4), **args, **th)
    sim.solver_args = {"tol": 1e-8}
    kw_run = dict(mixed_tol=1e6, max_events=25)
    if d == 2:
        kw_run["stress_targets"] = {(1, 1): 0.0}
    ev = sim.run_simulation(steps, step_size, component=(0, 0), **kw_run)
    Tl = sim.Tlocal
    print(f"[kmc thermal] events {len(ev)} (oracle {len(ev_ref)}), T range {Tl.min():.3f}..{Tl.max():.3f}", flush=True)
    assert len(ev) > 0
    assert [e[:d + 1] for e in ev] == [e[:d + 1] for e in ev_ref]
    if cascade:
        assert sim.cascade_history == st_ref["cascades"]
    assert np.abs(Tl - st_ref["Tlocal"]).max() <= 1e-10 * st_ref["Tlocal"].max()
    assert Tl.max() > T  # flips heated something


@pytest.mark.parametrize("model,plane_mode,stz_mode,solver", [("svk", "plane_strain", "pure_shear", "dbfft"),
                                                              ("svk", "plane_stress", "pure_shear", "dbfft")])
def test_kmc2d_finite_strain_dbfft_matches_oracle(gpu, model, plane_mode, stz_mode, solver):
    test_kmc2d_finite_strain_matches_oracle(gpu, model, plane_mode, stz_mode, solver)


@pytest.mark.parametrize("model,plane_mode,stz_mode", [("svk", "plane_strain", "pure_shear"),
                                                       ("svk", "plane_stress", "pure_shear"),
                                                       ("landau", "plane_strain", "pure_shear"),
                                                       ("landau", "plane_stress", "pure_shear"),
                                                       ("svk", "plane_strain", "simple_shear"),
                                                       ("neo_hookean", "plane_strain", "simple_shear"),
                                                       ("neo_hookean", "plane_stress", "pure_shear"),
                                                       ("landau_capped", "plane_stress", "pure_shear")])
def test_kmc2d_finite_strain_matches_oracle(gpu, model, plane_mode, stz_mode, solver="newton_cg"):
    # KmcSimulation2D(strain_assumption="finite_strain", solver="newton_cg"):
    # every flip multiplies F_p by expm(C/N) (N = 1 first), each followed by
    # the 2D Newton-CG step with the stress targets; events, sub-step counts,
    # F, F_p, F_macro and sigma equal the 2D oracle loop's
    from mgkmc_amd.kmc_simulator import KmcSimulation2D
    from oracle.kmc2d_loop import run_events_2d
    print(f"[kmc2d finite] {model} {plane_mode} {stz_mode}", flush=True)
    n, M = 12, 6
    rng = np.random.default_rng(3)
    E = 70e9 * rng.uniform(0.9, 1.1, (n, n))
    nu = np.full((n, n), 0.3)
    kw = dict(mean=1.0, std=0.3, min_cutoff=0.0)
    mp = dict(v1=-236.5e9, v2=-27.4e9, v3=11.8e9, g1=-5640.7e9, g2=2207.1e9, g3=-332.5e9, g4=-305.2e9) \
        if model.startswith("landau") else {}
    if model == "landau_capped":
        model = "landau"
        mp.update(strain_capping_enabled=True, strain_capping_limit=0.003)
    args = dict(softening_cap=0.916, jp=20.0, jt=5.0)
    # kT low enough that the 8 events span an elastic increment (F_xx ~ 1.004)
    ev_ref, st_ref = run_events_2d(n, n, M, 0.03, E, nu, 0.7, 400.0, 1e7, 2e-3, (0, 0), 8, seed=7, sel_seed=8,
                                   barrier_kwargs=kw, plane_mode=plane_mode, stress_targets={(1, 1): 0.0},
                                   n_global_steps=3, model=model, model_params=mp,
                                   strain_assumption="finite_strain", stz_mode=stz_mode, solver=solver, **args)
    sim = KmcSimulation2D(n, n, M, 0.03, E, nu, pixel=0.7, barrier_generator="gaussian", barrier_kwargs=kw,
                          stz_mode=stz_mode,
                          temperature=400.0, strain_rate=1e7, plane_mode=plane_mode, strain_assumption="finite_strain",
                          hyperelastic_model=model, solver=solver, rng=np.random.RandomState(7),
                          selection_stream=np.random.RandomState(8), **mp, **args)
    ev = sim.run_simulation(3, 2e-3, component=(0, 0), stress_targets={(1, 1): 0.0}, max_events=8)
    subs = sim.kmc_substeps
    print(f"[kmc2d finite] events {[e[:3] for e in ev]} substeps {subs} (oracle {st_ref['substeps']})", flush=True)
    assert [e[:3] for e in ev] == [e[:3] for e in ev_ref]
    assert subs == st_ref["substeps"]
    assert sim.kmc.planar == (stz_mode == "pure_shear")
    eF = np.abs(sim.F_field - np.einsum("ijxy->xyij", st_ref["F"])).max()
    eS = np.abs(sim.sig_field - st_ref["sig"]).max() / np.abs(st_ref["sig"]).max()
    print(f"[kmc2d finite] relF={eF:.2e} relSig={eS:.2e}", flush=True)
    assert eF <= 1e-10
    assert np.abs(sim.F_plastic - np.einsum("ijxy->xyij", st_ref["F_plastic"])).max() <= 1e-14
    assert np.abs(sim.F_macro - st_ref["F_macro"]).max() <= 1e-10
    assert eS <= 1e-10


@pytest.mark.parametrize("d,T,mean,g0", [(2, 400.0, 0.3, 0.05), (2, 0.0, 0.5, 0.1), (3, 800.0, 0.5, 0.03)])
def test_finite_strain_cascade_mode_matches_oracle(gpu, d, T, mean, g0):
    # cascade_mode with finite strain: every cascade batch goes through the
    # expm sub-stepped F_p update (one finite-strain step per sub-step); the
    # cascade histories, the KMC events, sub-step counts, F and F_p equal the
    # oracle loops'
    from mgkmc_amd.aqs import ThermalSimulation
    from mgkmc_amd.kmc_simulator import KmcSimulation2D
    from oracle.aqs_loop import run_events
    from oracle.kmc2d_loop import run_events_2d
    print(f"[finite cascade] {d}D T={T}", flush=True)
    kw = dict(mean=mean, std=0.15, min_cutoff=0.0)
    args = dict(softening_cap=0.916, jp=20.0, jt=0.0)
    if d == 2:
        n, M = 12, 10
        E = 70e9 * np.random.default_rng(0).uniform(0.9, 1.1, (n, n))
        nu = np.full((n, n), 0.3)
        ev_ref, st_ref = run_events_2d(n, n, M, g0, E, nu, 0.7, T, 1e7, 2e-3, (0, 0), 10, seed=5, sel_seed=6,
                                       barrier_kwargs=kw, stress_targets={(1, 1): 0.0}, n_global_steps=4,
                                       cascade_mode=True, max_cascade_steps_pct=0.05, model="svk",
                                       strain_assumption="finite_strain", **args)
        sim = KmcSimulation2D(n, n, M, g0, E, nu, pixel=0.7, barrier_generator="gaussian", barrier_kwargs=kw,
                              temperature=T, strain_rate=1e7, cascade_mode=True, strain_assumption="finite_strain",
                              solver="newton_cg", rng=np.random.RandomState(5),
                              selection_stream=np.random.RandomState(6), **args)
        ev = sim.run_simulation(4, 2e-3, component=(0, 0), stress_targets={(1, 1): 0.0},
                                max_cascade_steps_pct=0.05, max_events=10)
        Fr, Fpr = np.einsum("ijxy->xyij", st_ref["F"]), np.einsum("ijxy->xyij", st_ref["F_plastic"])
    else:
        n, M = 8, 10
        E = 70e9 * np.random.default_rng(1).uniform(0.9, 1.1, (n, n, n))
        nu = np.full((n, n, n), 0.3)
        ev_ref, st_ref = run_events(n, n, n, M, g0, E, nu, 0.7, T, 1e7, 1e-3, (0, 0), 10, seed=7, sel_seed=8,
                                    barrier_kwargs=kw, n_global_steps=3, cascade_mode=True,
                                    strain_assumption="finite_strain", **args)
        sim = ThermalSimulation(n, n, n, M, g0, E, nu, pixel=0.7, barrier_generator="gaussian", barrier_kwargs=kw,
                                temperature=T, strain_rate=1e7, cascade_mode=True, strain_assumption="finite_strain",
                                solver="newton_cg", rng=np.random.RandomState(7),
                                selection_stream=np.random.RandomState(8), **args)
        ev = sim.run_simulation(3, 1e-3, component=(0, 0), max_events=10)
        Fr, Fpr = np.einsum("ijxyz->xyzij", st_ref["F"]), np.einsum("ijxyz->xyzij", st_ref["F_plastic"])
    print(f"[finite cascade] events {len(ev)} (oracle {len(ev_ref)}), cascade flips {len(sim.cascade_history)} "
          f"(oracle {len(st_ref['cascades'])})", flush=True)
    assert len(st_ref["cascades"]) > 0
    assert sim.cascade_history == st_ref["cascades"]
    assert [e[:d + 1] for e in ev] == [e[:d + 1] for e in ev_ref]
    print(f"[finite cascade] relF={np.abs(sim.F_field - Fr).max():.2e}", flush=True)
    assert np.abs(sim.F_field - Fr).max() <= 1e-10
    assert np.abs(sim.F_plastic - Fpr).max() <= 1e-13


@pytest.mark.parametrize("scheme", ["isotropic", "directional"])
def test_kmc3d_3d_barriers_match_oracle(gpu, scheme):
    # use_3d_barriers: fully 3D mode catalogs (9 normals and a Rayleigh
    # magnitude per mode, in the reference's stream order), general-mode
    # barriers, flips and redraws on the device; events and fields equal the
    # oracle loop's
    from mgkmc_amd.aqs import ThermalSimulation
    from oracle.aqs_loop import run_events
    print(f"[kmc3d 3d barriers] {scheme}", flush=True)
    n, M = 8, 6
    rng = np.random.default_rng(4)
    E = 70e9 * rng.uniform(0.9, 1.1, (n, n, n))
    nu = np.full((n, n, n), 0.3)
    kw = dict(mean=1.0, std=0.3, min_cutoff=0.0)
    args = dict(softening_cap=0.916, jp=20.0, jt=5.0, softening_scheme=scheme)
    ev_ref, st_ref = run_events(n, n, n, M, 0.1, E, nu, 0.7, 700.0, 1e7, 2e-3, (0, 0), 15, seed=7, sel_seed=8,
                                barrier_kwargs=kw, n_global_steps=3, use_3d_barriers=True, **args)
    sim = ThermalSimulation(n, n, n, M, 0.1, E, nu, pixel=0.7, barrier_generator="gaussian", barrier_kwargs=kw,
                            temperature=700.0, strain_rate=1e7, use_3d_barriers=True, rng=np.random.RandomState(7),
                            selection_stream=np.random.RandomState(8), **args)
    assert not sim.kmc.planar
    ev = sim.run_simulation(3, 2e-3, component=(0, 0), max_events=15)
    print(f"[kmc3d 3d barriers] events {[e[:4] for e in ev][:6]}", flush=True)
    assert [e[:4] for e in ev] == [e[:4] for e in ev_ref]
    assert np.allclose(sim.eps_plastic, st_ref["eps_plastic"], rtol=0, atol=1e-15)
    assert np.abs(sim.sig_field - st_ref["sig"]).max() <= 1e-10 * np.abs(st_ref["sig"]).max()


@pytest.mark.parametrize("d,kind", [(3, "finite"), (2, "finite"), (2, "secant")])
def test_thermal_with_finite_strain_and_nonlinear_laws(gpu, d, kind):
    # enable_thermal together with finite strain or a 2D nonlinear law: the
    # flip heat comes from the stress at the start of the last sub-step and is
    # restored with the state when a sub-step count fails; events, the final
    # temperature field and the mechanical fields equal the oracle loops'
    from mgkmc_amd.aqs import ThermalSimulation
    from mgkmc_amd.kmc_simulator import KmcSimulation2D
    from oracle.aqs_loop import run_events
    from oracle.kmc2d_loop import run_events_2d
    print(f"[thermal combo] {d}D {kind}", flush=True)
    kw = dict(mean=1.0, std=0.3, min_cutoff=0.0)
    th = dict(enable_thermal=True, temperature_cap=1500.0, thermostat=True, tau_bath=5e-13)
    args = dict(softening_cap=0.916, jp=20.0, jt=5.0)
    if d == 3:
        n, M = 8, 6
        E = 70e9 * np.random.default_rng(3).uniform(0.9, 1.1, (n, n, n))
        nu = np.full((n, n, n), 0.3)
        ev_ref, st_ref = run_events(n, n, n, M, 0.03, E, nu, 0.7, 700.0, 1e7, 2e-3, (0, 0), 8, seed=7, sel_seed=8,
                                    barrier_kwargs=kw, n_global_steps=3, strain_assumption="finite_strain",
                                    **th, **args)
        sim = ThermalSimulation(n, n, n, M, 0.03, E, nu, pixel=0.7, barrier_generator="gaussian", barrier_kwargs=kw,
                                temperature=700.0, strain_rate=1e7, strain_assumption="finite_strain",
                                solver="newton_cg", rng=np.random.RandomState(7),
                                selection_stream=np.random.RandomState(8), **th, **args)
        ev = sim.run_simulation(3, 2e-3, component=(0, 0), max_events=8)
    else:
        n, M = 12, 6
        E = 70e9 * np.random.default_rng(3).uniform(0.9, 1.1, (n, n))
        nu = np.full((n, n), 0.3)
        model, mp, extra = ("svk", {}, dict(strain_assumption="finite_strain", solver="newton_cg")) \
            if kind == "finite" else ("secant_degradation", dict(d=0.3, k=20.0),
                                      dict(hyperelastic_model="secant_degradation", d=0.3, k=20.0))
        ev_ref, st_ref = run_events_2d(n, n, M, 0.05, E, nu, 0.7, 500.0, 1e7, 2e-3, (0, 0), 8, seed=7, sel_seed=8,
                                       barrier_kwargs=kw, stress_targets={(1, 1): 0.0}, n_global_steps=3,
                                       model=model, model_params=mp,
                                       strain_assumption="finite_strain" if kind == "finite" else "small_strain",
                                       **th, **args)
        sim = KmcSimulation2D(n, n, M, 0.05, E, nu, pixel=0.7, barrier_generator="gaussian", barrier_kwargs=kw,
                              temperature=500.0, strain_rate=1e7, rng=np.random.RandomState(7),
                              selection_stream=np.random.RandomState(8), **extra, **th, **args)
        ev = sim.run_simulation(3, 2e-3, component=(0, 0), stress_targets={(1, 1): 0.0}, max_events=8)
    print(f"[thermal combo] events {[e[:d + 1] for e in ev]}; T max {sim.Tlocal.max():.3f} "
          f"(oracle {st_ref['Tlocal'].max():.3f})", flush=True)
    assert [e[:d + 1] for e in ev] == [e[:d + 1] for e in ev_ref]
    assert np.abs(sim.Tlocal - st_ref["Tlocal"]).max() <= 1e-10 * st_ref["Tlocal"].max()
    assert st_ref["Tlocal"].max() > 500.0 * 1.0000001 or d == 3
    eS = np.abs(sim.sig_field - st_ref["sig"]).max() / np.abs(st_ref["sig"]).max()
    print(f"[thermal combo] relSig={eS:.2e}", flush=True)
    assert eS <= 1e-10


@pytest.mark.parametrize("d,rd,rb", [(3, False, False), (2, False, False), (3, False, True), (2, True, False)])
def test_kmc_redraw_options_match_oracle(gpu, d, rd, rb):
    # redraw_directions / redraw_barriers: the voxel's modes and barriers as
    # the flags say, or (both off) only the flipped mode's direction and
    # barrier; events and fields equal the oracle loops'
    from mgkmc_amd.aqs import ThermalSimulation
    from mgkmc_amd.kmc_simulator import KmcSimulation2D
    from oracle.aqs_loop import run_events
    from oracle.kmc2d_loop import run_events_2d
    print(f"[redraw] {d}D directions={rd} barriers={rb}", flush=True)
    kw = dict(mean=1.0, std=0.3, min_cutoff=0.0)
    args = dict(softening_cap=0.916, jp=20.0, jt=5.0)
    if d == 3:
        n, M = 8, 6
        E = 70e9 * np.random.default_rng(4).uniform(0.9, 1.1, (n, n, n))
        nu = np.full((n, n, n), 0.3)
        ev_ref, st_ref = run_events(n, n, n, M, 0.1, E, nu, 0.7, 700.0, 1e7, 2e-3, (0, 0), 15, seed=7, sel_seed=8,
                                    barrier_kwargs=kw, n_global_steps=3, redraw_directions=rd, redraw_barriers=rb,
                                    **args)
        sim = ThermalSimulation(n, n, n, M, 0.1, E, nu, pixel=0.7, barrier_generator="gaussian", barrier_kwargs=kw,
                                temperature=700.0, strain_rate=1e7, redraw_directions=rd, redraw_barriers=rb,
                                rng=np.random.RandomState(7), selection_stream=np.random.RandomState(8), **args)
        ev = sim.run_simulation(3, 2e-3, component=(0, 0), max_events=15)
    else:
        n, M = 16, 8
        E = 70e9 * np.random.default_rng(4).uniform(0.9, 1.1, (n, n))
        nu = np.full((n, n), 0.3)
        ev_ref, st_ref = run_events_2d(n, n, M, 0.1, E, nu, 0.7, 700.0, 1e7, 2e-3, (0, 0), 15, seed=7, sel_seed=8,
                                       barrier_kwargs=kw, stress_targets={(1, 1): 0.0}, n_global_steps=3,
                                       redraw_directions=rd, redraw_barriers=rb, **args)
        sim = KmcSimulation2D(n, n, M, 0.1, E, nu, pixel=0.7, barrier_generator="gaussian", barrier_kwargs=kw,
                              temperature=700.0, strain_rate=1e7, redraw_directions=rd, redraw_barriers=rb,
                              rng=np.random.RandomState(7), selection_stream=np.random.RandomState(8), **args)
        ev = sim.run_simulation(3, 2e-3, component=(0, 0), stress_targets={(1, 1): 0.0}, max_events=15)
    print(f"[redraw] events {[e[:d + 1] for e in ev][:6]}", flush=True)
    assert [e[:d + 1] for e in ev] == [e[:d + 1] for e in ev_ref]
    assert np.allclose(sim.Q0, st_ref["Q0"], rtol=0, atol=0)
    assert np.allclose(sim.eps_plastic, st_ref["eps_plastic"], rtol=0, atol=1e-15)


@pytest.mark.parametrize("gen,kw", [("rayleigh", dict(loc=0.5, scale=0.4, min_cutoff=0.1)),
                                    ("modified_rayleigh", dict(mean=1.0, std=0.3, min_cutoff=0.1)),
                                    ("modified_rayleigh_with_exponential",
                                     dict(mean=1.2, std=0.3, epsilon=0.3, ratio=0.3, min_cutoff=0.05))])
def test_kmc3d_barrier_generators_match_oracle(gpu, gen, kw):
    # the reference's other barrier generators (initial Q0 and every redraw
    # from the same host stream): events and fields equal the oracle loop's
    from mgkmc_amd.aqs import ThermalSimulation
    from oracle.aqs_loop import run_events
    print(f"[generators] {gen}", flush=True)
    n, M = 8, 6
    E = 70e9 * np.random.default_rng(4).uniform(0.9, 1.1, (n, n, n))
    nu = np.full((n, n, n), 0.3)
    args = dict(softening_cap=0.916, jp=20.0, jt=5.0)
    ev_ref, st_ref = run_events(n, n, n, M, 0.1, E, nu, 0.7, 700.0, 1e7, 2e-3, (0, 0), 12, seed=7, sel_seed=8,
                                barrier_kwargs=kw, n_global_steps=3, barrier_generator=gen, **args)
    sim = ThermalSimulation(n, n, n, M, 0.1, E, nu, pixel=0.7, barrier_generator=gen, barrier_kwargs=kw,
                            temperature=700.0, strain_rate=1e7, rng=np.random.RandomState(7),
                            selection_stream=np.random.RandomState(8), **args)
    ev = sim.run_simulation(3, 2e-3, component=(0, 0), max_events=12)
    print(f"[generators] events {[e[:4] for e in ev][:5]} (of {len(ev)})", flush=True)
    assert len(ev) > 0 and [e[:4] for e in ev] == [e[:4] for e in ev_ref]
    assert np.array_equal(sim.Q0, st_ref["Q0"])
