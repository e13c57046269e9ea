"""GPU parity of fast patching (the reference's predictor-corrector stress
update, mgkmc/aqs.py:196-256, :1021-1054; mgkmc/kmc_simulator.py:229-261,
:764-792, :1044-1098) with the oracle loops: identical events and cascade
flips, patched stress fields and kernels to 1e-10."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _moduli(shape, seed):
    rng = np.random.default_rng(seed)
    mu = 26.923e9 * rng.uniform(0.9, 1.1, shape)
    nu = np.full(shape, 0.35)
    return 2 * mu * (1 + nu), nu


def test_patch_kernels_3d_match_oracle(gpu):
    from mgkmc_amd.aqs import ThermalSimulation
    from oracle import patching as opt
    shape = (9, 8, 10)
    E, nu = _moduli(shape, 1)
    sim = ThermalSimulation(*shape, 4, 0.14, E, nu, pixel=0.7, temperature=300.0,
                            fast_patching=dict(enabled=True, patch_radius=3), rng=np.random.RandomState(1),
                            selection_stream=np.random.RandomState(2))
    K, mm = opt.patch_kernels_3d(E, nu, 3, pixel=0.7)
    assert sim.fast_patching_enabled and sim.patch_radius == 3 and sim.sync_interval == 100
    for a, b in zip(sim.patch_kernels, K):
        assert a.shape == b.shape
        assert np.abs(a - b).max() <= 1e-10 * np.abs(b).max()
    for a, b in zip(sim.patch_missing_mean, mm):
        assert np.abs(a - b).max() <= 1e-10 * np.abs(K[0]).max()


def test_patch_kernels_2d_match_oracle_flat_and_wrapped(gpu):
    # 2D, and a radius whose box wraps the grid (2R + 1 > n: voxels covered twice)
    from mgkmc_amd.kmc_simulator import KmcSimulation2D
    from oracle import patching as opt
    for shape, R, pm in [((16, 12), 5, "plane_strain"), ((9, 5), 5, "plane_stress")]:
        E, nu = _moduli(shape, 2)
        sim = KmcSimulation2D(*shape, 4, 0.14, E, nu, pixel=0.7, temperature=300.0, plane_mode=pm,
                              fast_patching=dict(enabled=True), rng=np.random.RandomState(1),
                              selection_stream=np.random.RandomState(2))
        assert sim.patch_radius == R
        K, mm = opt.patch_kernels_2d(E, nu, R, pixel=0.7, plane_mode=pm)
        for a, b in zip(sim.patch_kernels, K):
            assert np.abs(a - b).max() <= 1e-10 * np.abs(b).max()
        for a, b in zip(sim.patch_missing_mean, mm):
            assert np.abs(a - b).max() <= 1e-10 * np.abs(K[0]).max()
        # one stamp on the device equals the oracle's stamp on the same field
        sig0 = sim.sig_field.copy()
        C = np.array([[0.031, -0.012], [-0.012, -0.031]])
        v = 5 * shape[1] + (8 % shape[1])
        sim._patch_stamp(v, C)
        ref = sig0.copy()
        opt.stamp(ref, (5, 8 % shape[1]), opt.coefficients_2d(C), sim.patch_kernels, sim.patch_missing_mean, R)
        got = sim.sig_field
        assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
        assert np.allclose(sim.sig_mean(), ref.mean(axis=(0, 1)), rtol=1e-12, atol=0)


@pytest.mark.parametrize("sync,T", [(7, 1276.5), (1000, 900.0)])
def test_kmc3d_fast_patching_matches_oracle(gpu, sync, T):
    from mgkmc_amd.aqs import ThermalSimulation
    from oracle.aqs_loop import run_events
    n, M = 8, 20
    E, nu = _moduli((n, n, n), 1)
    kw = dict(mean=2.2, std=0.45, min_cutoff=0.0)
    fp = dict(enabled=True, patch_radius=2, sync_interval=sync)
    args = dict(softening_cap=0.916, jp=20.0, jt=0.0)
    ev_ref, st_ref = run_events(n, n, n, M, 0.14, E, nu, 0.7, T, 1e7, 1e-4, (0, 0), 30, seed=11, sel_seed=12,
                                barrier_kwargs=kw, fast_patching=fp, **args)
    sim = ThermalSimulation(n, n, n, M, 0.14, E, nu, pixel=0.7, barrier_generator="gaussian", barrier_kwargs=kw,
                            temperature=T, strain_rate=1e7, fast_patching=fp, rng=np.random.RandomState(11),
                            selection_stream=np.random.RandomState(12), **args)
    ev = sim.run_simulation(10 ** 9, 1e-4, component=(0, 0), max_events=30)
    print(f"[patch3d] sync {sync}: {len(ev)} events, flips since sync {sim.flips_since_sync}", flush=True)
    assert [e[:4] for e in ev] == [e[:4] for e in ev_ref]
    assert np.allclose([e[4] for e in ev], [e[4] for e in ev_ref], rtol=1e-10)
    assert np.allclose(sim.eps_plastic, st_ref["eps_plastic"], rtol=0, atol=1e-15)
    assert np.abs(sim.sig_field - st_ref["sig"]).max() <= 1e-10 * np.abs(st_ref["sig"]).max()
    # the logged driven stress of every event is the mean of the patched field
    hx = np.array([h[1] for h in sim.history_global]) * 1e9
    hr = np.array(st_ref["stress_history"])[1:]  # the oracle's first entry is the initial state
    assert hx.shape == hr.shape
    assert np.abs(hx - hr).max() <= 1e-10 * np.abs(hr).max()
    assert sim.flips_since_sync == len(ev) % sync


@pytest.mark.parametrize("plane_mode,sync", [("plane_strain", 6), ("plane_stress", 1000)])
def test_kmc2d_fast_patching_matches_oracle(gpu, plane_mode, sync):
    # KMC events with the macro-unit stress over waiting times and remaining
    # increment times, and the mixed-BC loop between increments
    from mgkmc_amd.kmc_simulator import KmcSimulation2D
    from oracle.kmc2d_loop import run_events_2d
    n, M = 16, 20
    E, nu = _moduli((n, n), 2)
    kw = dict(mean=2.2, std=0.45, min_cutoff=0.0)
    fp = dict(enabled=True, sync_interval=sync)
    args = dict(softening_cap=0.916, jp=20.0, jt=0.0)
    ev_ref, st_ref = run_events_2d(n, n, M, 0.14, E, nu, 0.7, 1276.5, 1e7, 2e-4, (0, 0), 40, seed=21, sel_seed=22,
                                   barrier_kwargs=kw, plane_mode=plane_mode, stress_targets={(1, 1): 0.0},
                                   mixed_tol=1e6, fast_patching=fp, **args)
    sim = KmcSimulation2D(n, n, M, 0.14, E, nu, pixel=0.7, barrier_generator="gaussian", barrier_kwargs=kw,
                          temperature=1276.5, strain_rate=1e7, plane_mode=plane_mode, fast_patching=fp,
                          rng=np.random.RandomState(21), selection_stream=np.random.RandomState(22), **args)
    ev = sim.run_simulation(10 ** 9, 2e-4, component=(0, 0), stress_targets={(1, 1): 0.0}, mixed_tol=1e6,
                            max_events=40)
    print(f"[patch2d] {plane_mode}: {len(ev)} events, increments {len(st_ref['increments'])}", flush=True)
    assert len(st_ref["increments"]) > 0
    assert [e[:3] for e in ev] == [e[:3] for e in ev_ref]
    assert np.allclose([e[3] for e in ev], [e[3] for e in ev_ref], rtol=1e-10)
    assert np.allclose(sim.eps_plastic, st_ref["eps_plastic"], rtol=0, atol=1e-15)
    assert np.abs(sim.eps_macro - st_ref["eps_macro"]).max() <= 1e-10 * np.abs(st_ref["eps_macro"]).max()
    assert np.abs(sim.sig_field - st_ref["sig"]).max() <= 1e-10 * np.abs(st_ref["sig"]).max()


@pytest.mark.parametrize("T,mean", [(0.0, 0.55), (400.0, 0.8)])
def test_kmc2d_cascade_fast_patching_matches_oracle(gpu, T, mean):
    # cascade flips patched (no direction redraws, so the catalog entry the
    # reference reads after the redraw is the flipped mode), one re-solve after
    # each cascade that flipped anything, KMC events between cascades
    from mgkmc_amd.kmc_simulator import KmcSimulation2D
    from oracle.kmc2d_loop import run_events_2d
    n, M = 16, 20
    E, nu = _moduli((n, n), 0)
    kw = dict(mean=mean, std=0.15, min_cutoff=0.0)
    fp = dict(enabled=True, patch_radius=4, sync_interval=5)
    args = dict(softening_cap=0.916, jp=20.0, jt=0.0)
    ev_ref, st_ref = run_events_2d(n, n, M, 0.14, E, nu, 0.7, T, 1e7, 1e-3, (0, 0), 10 ** 9, seed=5, sel_seed=6,
                                   barrier_kwargs=kw, stress_targets={(1, 1): 0.0}, mixed_tol=1e6, ls_tol=1e-8,
                                   n_global_steps=4, cascade_mode=True, max_cascade_steps_pct=0.02,
                                   fast_patching=fp, redraw_directions=False, **args)
    sim = KmcSimulation2D(n, n, M, 0.14, E, nu, pixel=0.7, barrier_generator="gaussian", barrier_kwargs=kw,
                          temperature=T, strain_rate=1e7, cascade_mode=True, fast_patching=fp,
                          redraw_directions=False, rng=np.random.RandomState(5),
                          selection_stream=np.random.RandomState(6), **args)
    sim.solver_args = {"tol": 1e-8}
    ev = sim.run_simulation(4, 1e-3, component=(0, 0), stress_targets={(1, 1): 0.0}, mixed_tol=1e6,
                            max_cascade_steps_pct=0.02)
    print(f"[patch2d cascade] T={T}: events {len(ev)}, cascade flips {len(sim.cascade_history)} "
          f"(oracle {len(st_ref['cascades'])})", flush=True)
    assert len(st_ref["cascades"]) > 0
    assert sim.cascade_history == st_ref["cascades"]
    assert [e[:3] for e in ev] == [e[:3] for e in ev_ref]
    assert np.allclose(sim.eps_plastic, st_ref["eps_plastic"], rtol=0, atol=1e-15)
    assert np.abs(sim.sig_field - st_ref["sig"]).max() <= 1e-10 * np.abs(st_ref["sig"]).max()


def test_kmc2d_cascade_fast_patching_redrawn_catalog_matches_oracle(gpu):
    # with direction redraws the reference patches a cascade flip with the
    # voxel's redrawn catalog entry, not the flipped mode; the predicted
    # stresses then drive the cascade until the corrector's re-solve finds an
    # unstable stress.  Device and oracle flip the same modes in the same
    # batches and both raise there.
    from mgkmc_amd.kmc_simulator import KmcSimulation2D
    from oracle.kmc2d_loop import run_events_2d
    n, M = 16, 20
    E, nu = _moduli((n, n), 0)
    kw = dict(mean=0.56, std=0.15, min_cutoff=0.0)
    fp = dict(enabled=True, patch_radius=4, sync_interval=5)
    args = dict(softening_cap=0.916, jp=20.0, jt=0.0)
    rec = {}
    with pytest.raises(ValueError):
        run_events_2d(n, n, M, 0.14, E, nu, 0.7, 0.0, 1e7, 1e-3, (0, 0), 10 ** 9, seed=5, sel_seed=6,
                      barrier_kwargs=kw, stress_targets={(1, 1): 0.0}, mixed_tol=1e6, ls_tol=1e-8, n_global_steps=4,
                      cascade_mode=True, max_cascade_steps_pct=0.02, fast_patching=fp, record=rec, **args)
    sim = KmcSimulation2D(n, n, M, 0.14, E, nu, pixel=0.7, barrier_generator="gaussian", barrier_kwargs=kw,
                          temperature=0.0, strain_rate=1e7, cascade_mode=True, fast_patching=fp,
                          rng=np.random.RandomState(5), selection_stream=np.random.RandomState(6), **args)
    sim.solver_args = {"tol": 1e-8}
    with pytest.raises(ValueError):
        sim.run_simulation(4, 1e-3, component=(0, 0), stress_targets={(1, 1): 0.0}, mixed_tol=1e6,
                           max_cascade_steps_pct=0.02)
    print(f"[patch2d cascade redrawn] cascade flips {len(sim.cascade_history)} (oracle {len(rec['cascades'])})",
          flush=True)
    assert len(rec["cascades"]) > 0
    assert sim.cascade_history == rec["cascades"]


def test_patch_fp32_stress_field_matches_fp64(gpu):
    # the FP32 LS form patches its FP32 stress field: same arithmetic in FP64,
    # one rounding per stored value
    from mgkmc_amd.aqs import ThermalSimulation
    n, M = 12, 4
    E, nu = _moduli((n, n, n), 4)
    fp = dict(enabled=True, patch_radius=3)
    sims = [ThermalSimulation(n, n, n, M, 0.14, E, nu, pixel=0.7, temperature=300.0, strain_rate=1e7,
                              fast_patching=fp, ls_precision=p, rng=np.random.RandomState(1),
                              selection_stream=np.random.RandomState(2)) for p in ("fp64", "fp32")]
    C = np.array([[0.02, 0.01, -0.005], [0.01, 0.004, 0.003], [-0.005, 0.003, -0.024]])
    for s in sims:
        s.update_stresses()
        s._patch_after_flip(5 * n * n + 7 * n + 2, C, 3e-10, np.diag([1.0, 0, 0]))
    a, b = sims[0].sig_field, sims[1].sig_field
    assert np.abs(b - a).max() <= 2e-6 * np.abs(a).max()
    assert np.allclose(sims[1].sig_mean(), sims[0].sig_mean(), rtol=0, atol=2e-6 * np.abs(a).max())


@pytest.mark.parametrize("d,prec", [(3, "fp64"), (2, "fp64"), (3, "mixed"), (2, "fp32")])
def test_native_fast_patching_equals_python_loop(gpu, d, prec):
    # the C++ event loop with the predictor (mgk_kmc_run with a patch
    # configuration) against the per-event Python loop: events, fields, sync
    # counters and stream states bit for bit
    from mgkmc_amd.aqs import ThermalSimulation
    from mgkmc_amd.kmc_simulator import KmcSimulation2D
    n, M = (10, 20) if d == 3 else (24, 20)
    shape = (n,) * d
    E, nu = _moduli(shape, 9)
    kw = dict(mean=2.2, std=0.45, min_cutoff=0.0)
    fp = dict(enabled=True, patch_radius=3, sync_interval=7)
    out = []
    for native in (True, False):
        args = dict(pixel=0.7, barrier_generator="gaussian", barrier_kwargs=kw, temperature=1276.5, strain_rate=1e7,
                    fast_patching=fp, ls_precision=prec, rng=np.random.RandomState(41),
                    selection_stream=np.random.RandomState(42), softening_cap=0.916, jp=20.0, jt=0.0)
        if d == 3:
            sim = ThermalSimulation(n, n, n, M, 0.14, E, nu, **args)
            sim.native_loop = native
            ev = sim.run_simulation(10 ** 9, 1e-4, component=(0, 0), max_events=25)
        else:
            sim = KmcSimulation2D(n, n, M, 0.14, E, nu, **args)
            sim.native_loop = native
            ev = sim.run_simulation(10 ** 9, 2e-4, component=(0, 0), stress_targets={(1, 1): 0.0}, mixed_tol=1e6,
                                    max_events=25)
        out.append((ev, sim.sig_field, sim.eps_plastic, sim.flips_since_sync, sim.rng.get_state()[2],
                    sim.selection_stream.get_state()[2], sim.sig_mean(), list(sim.history_global)))
    a, b = out
    print(f"[patch native] d={d} {prec}: {len(a[0])} events, flips since sync {a[3]}", flush=True)
    assert [e[:-2] for e in a[0]] == [e[:-2] for e in b[0]]
    assert [e[-2] for e in a[0]] == [e[-2] for e in b[0]]
    assert np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])
    assert a[3] == b[3] and a[4] == b[4] and a[5] == b[5]
    assert np.array_equal(a[6], b[6])
    assert a[7] == b[7]
