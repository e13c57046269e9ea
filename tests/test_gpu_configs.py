"""The five headline workloads on the device at their own sizes.

cfg0 (2D 64^2 athermal AQS, 200 increments, LS tol 1e-8) and cfg1 (3D 128^3,
500 KMC events at kT/E0 = 0.05) are replayed against oracle fixtures
(tests/golden/make_fixtures.py, inputs tests/golden/cfg_inputs.py): the same
events and the macro and field values to 1e-10 relative.  cfg2 (256^3
finite-strain Landau, simple shear to 5 % in 50 increments), cfg3 (512^3
mixed-BC KMC) and cfg4 (the 512-request serving stream) are beyond the
oracle's reach and are checked through size-independent properties: the
imposed mean deformation, the equilibrium residual, the stress targets, the
event count, completion of every request and solo-vs-mix equality.
"""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, GOLDEN)


def _fixture(name):
    path = os.path.join(GOLDEN, name)
    if not os.path.exists(path):
        pytest.skip(f"{name} not generated (python tests/golden/make_fixtures.py)")
    return np.load(path, allow_pickle=False)


def _rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(np.asarray(b)).max())


def test_cfg0_aqs_64_matches_fixture(gpu, tmp_path):
    import cfg_inputs as ci
    from mgkmc_amd.kmc_simulator import KmcSimulation2D
    f = _fixture("cfg0_aqs_64.npz")
    c = ci.CFG0
    E, nu, ep = ci.cfg0_inputs(c["n"])
    sim = KmcSimulation2D(c["n"], c["n"], c["M"], c["gamma0"], E, nu, pixel=c["pixel"], barrier_generator="gaussian",
                          barrier_kwargs=ci.BARRIERS, temperature=c["temperature"], strain_rate=c["strain_rate"],
                          output_dir=str(tmp_path), rng=np.random.RandomState(c["seed"]),
                          selection_stream=np.random.RandomState(c["sel_seed"]), **ci.STZ)
    sim.eps_plastic = ep
    sim.solver_args = {"tol": c["ls_tol"]}
    ev = sim.run_simulation(c["steps"], c["step_size"], component=(0, 0), stress_targets={(1, 1): 0.0},
                            mixed_tol=c["mixed_tol"], enable_console_log=False)
    inc_eps = np.array([e for e, _ in sim.increment_history])
    inc_sig = np.array([s for _, s in sim.increment_history])
    sig = sim.sig_field
    got_sig = np.stack([sig[..., 0, 0], sig[..., 1, 1], sig[..., 0, 1]])
    print(f"[cfg0] events {len(ev)} (oracle {len(f['events'])}), increments {len(inc_eps)}; "
          f"macro eps {_rel(inc_eps, f['inc_eps']):.2e} sig {_rel(inc_sig, f['inc_sig']):.2e}, "
          f"field sig {_rel(got_sig, f['sig']):.2e}", flush=True)
    assert [e[:3] for e in ev] == [tuple(int(v) for v in r) for r in f["events"]]
    assert len(inc_eps) == c["steps"] == len(f["inc_eps"])
    assert _rel(inc_eps, f["inc_eps"]) <= 1e-10 and _rel(inc_sig, f["inc_sig"]) <= 1e-10
    assert _rel(got_sig, f["sig"]) <= 1e-10
    epf = sim.eps_plastic
    assert np.abs(np.stack([epf[..., 0, 0], epf[..., 1, 1], epf[..., 0, 1]]) - f["eps_plastic"]).max() <= 1e-15
    assert abs(sim.eps_macro[0, 0] - 0.02) < 1e-12


def _cfg1_run(prec, out_dir, max_events):
    import cfg_inputs as ci
    from mgkmc_amd.aqs import ThermalSimulation
    c = ci.CFG1
    n = c["n"]
    E, nu, ep = ci.cfg1_inputs(n)
    sim = ThermalSimulation(n, n, n, c["M"], c["gamma0"], E, nu, pixel=c["pixel"], barrier_generator="gaussian",
                            barrier_kwargs=ci.BARRIERS, temperature=c["temperature"], strain_rate=c["strain_rate"],
                            output_dir=str(out_dir), ls_precision=prec, rng=np.random.RandomState(c["seed"]),
                            selection_stream=np.random.RandomState(c["sel_seed"]), **ci.STZ)
    sim.eps_plastic = ep
    sim.solver_args = {"tol": c["ls_tol"]}
    ev = sim.run_simulation(10 ** 9, c["step_size"], component=(0, 0), mixed_tol=c["mixed_tol"],
                            enable_console_log=False, enable_summary_log=False, enable_global_log=False,
                            enable_cascade_log=False, max_events=max_events)
    return sim, ev


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_cfg1_kmc_128_matches_fixture(gpu, tmp_path, prec):
    # FP64: the oracle's 500-event sequence exactly, fields to 1e-10.  The FP32 +
    # FP64-refinement form stops each solve at the FP64 test with tol 1e-6 while
    # the fixture's solves run to 1e-8, so its fields differ by O(tol) and a
    # selection whose draw falls within that of a boundary can differ; from
    # there the sequences part.  Stated for it: the events it shares with the
    # oracle (at least 50) and its field error against the FP64 path over them.
    f = _fixture("cfg1_kmc_128.npz")
    nev = len(f["events"])
    sim, ev = _cfg1_run(prec, tmp_path, nev)
    got = [e[:4] for e in ev]
    ref = [tuple(int(v) for v in r) for r in f["events"]]
    first_diff = next((i for i, (a, b) in enumerate(zip(got, ref)) if a != b), None)
    sig = sim.sig_field
    flat = sig.reshape(-1, 3, 3)
    e_samp = _rel(flat[f["sample_index"]], f["sample_sig"])
    e_mean = _rel(sig.mean(axis=(0, 1, 2)), f["sig_mean"])
    print(f"[cfg1 {prec}] {len(ev)} events (oracle {nev}), first difference at {first_diff}; "
          f"sampled sigma {e_samp:.2e}, mean sigma {e_mean:.2e}, |sigma| {abs(np.sqrt(np.sum(sig ** 2)) / f['sig_norm'] - 1):.2e}",
          flush=True)
    assert len(ev) == nev
    if prec == "fp64":
        assert got == ref
        assert np.allclose([e[4] for e in ev], f["times"], rtol=1e-10, atol=0)
        assert [e[5] for e in ev] == list(f["instab"])
        assert e_samp <= 1e-10 and e_mean <= 1e-10
        assert abs(np.sqrt(np.sum(sig ** 2)) / f["sig_norm"] - 1) <= 1e-10
        assert np.abs(sim.eps_macro - f["eps_macro"]).max() <= 1e-12 * max(1e-30, np.abs(f["eps_macro"]).max())
        assert np.abs(sim.eps_plastic.sum(axis=(0, 1, 2)) - f["eps_plastic_sum"]).max() <= 1e-9
    else:
        shared = nev if first_diff is None else first_diff
        assert shared >= 50
        k = min(shared, 100)
        del sim
        s64, ev64 = _cfg1_run("fp64", tmp_path, k)
        smx, evmx = _cfg1_run("mixed", tmp_path, k)
        assert [e[:4] for e in ev64] == [e[:4] for e in evmx] == ref[:k]
        err = _rel(smx.sig_field, s64.sig_field)
        print(f"[cfg1 mixed] field error vs the FP64 path after {k} shared events: sigma {err:.2e}", flush=True)
        assert err <= 1e-5


def _fs_residual(P):
    """Equilibrium residual of a first Piola-Kirchhoff field P (3, 3, nx, ny, nz)
    (torch, FP64): sqrt(sum_k |P_hat(k) q|^2 / |q|^2) over k != 0 off even-axis
    Nyquist planes, relative to the norm of the whole field (mean included,
    as the reference's residuals, finite_strain_simulator.py:1431-1436,
    :2130-2131, divide by the norm of P)."""
    import torch
    shape = P.shape[2:]
    q = [torch.fft.fftfreq(shape[0], dtype=torch.float64, device=P.device),
         torch.fft.fftfreq(shape[1], dtype=torch.float64, device=P.device),
         torch.fft.rfftfreq(shape[2], dtype=torch.float64, device=P.device)]
    g = torch.meshgrid(*q, indexing="ij")
    mask = torch.ones(g[0].shape, dtype=torch.bool, device=P.device)
    mask[0, 0, 0] = False
    for a, n in enumerate(shape):
        if n % 2 == 0:
            idx = [slice(None)] * 3
            idx[a] = n // 2
            mask[tuple(idx)] = False
    w = torch.full(g[0].shape, 2.0, dtype=torch.float64, device=P.device)  # half spectrum: interior bins twice
    w[:, :, 0] = 1.0
    if shape[2] % 2 == 0:
        w[:, :, -1] = 1.0
    k2 = g[0] ** 2 + g[1] ** 2 + g[2] ** 2
    k2 = torch.where(mask, k2, torch.ones_like(k2))
    num = torch.zeros((), dtype=torch.float64, device=P.device)
    den = torch.zeros((), dtype=torch.float64, device=P.device)
    for i in range(3):
        r = torch.zeros(g[0].shape, dtype=torch.complex128, device=P.device)
        for j in range(3):
            ph = torch.fft.rfftn(P[i, j].to(torch.float64))
            den += (w * ph.abs() ** 2).sum()
            r += ph * mask * g[j]
        num += (w * r.abs() ** 2 / k2).sum()
    return float(torch.sqrt(num / den))


def test_cfg2_finite_strain_256_simple_shear_properties(gpu):
    # 256^3 Landau + smooth capping, mu ~ mu_unit lognormal(sigma 0.1), kappa = 2.5 mu;
    # simple shear F_xy to 5 % in 50 increments (tol_NW 1e-5; tol_CG 1e-8 FP64, 1e-6
    # refined).  Each increment stops on the Newton update norm after two
    # iterations, so the residual left by the last update accumulates over the
    # ramp (the oracle shows the same growth at 16^3: 5e-10 after one increment,
    # 5e-8 after ten); it stays below tol_NW relative to |P|.
    import torch
    import cfg_inputs as ci
    from mgkmc_amd.finite_strain_simulator import FsSolver, material, MIXED
    from bench import LANDAU
    n = 256
    rng = np.random.default_rng(2)
    mu = ci.MU_UNIT * np.exp(rng.normal(-0.005, 0.1, (n, n, n)))
    lam = (2.5 - 2.0 / 3.0) * mu
    mat = material("landau", **LANDAU, strain_capping_enabled=True, strain_capping_type="smooth",
                   strain_capping_limit=0.10, mu_mean=float(mu.mean()))
    out = {}
    for prec, tcg in ((0, 1e-8), (MIXED, 1e-6)):
        s = FsSolver(lam, mu, precision=prec)
        newton = krylov = 0
        for k in range(1, 51):
            Fb = np.eye(3)
            Fb[0, 1] = 0.05 * k / 50
            inf = s.step(Fb, mat, (0, 1), tol_NW=1e-5, tol_CG=tcg, max_NW=20)
            newton += inf.newton_iters
            krylov += inf.krylov_iters
        Fm = s.F.mean(dim=(2, 3, 4)).cpu().numpy()
        res = _fs_residual(s.P)
        out[prec] = (s.F.clone(), res)
        print(f"[cfg2 {'fp64' if prec == 0 else 'refined'}] 50 increments: Newton {newton}, Krylov {krylov}; "
              f"<F> - F_bar {np.abs(Fm - Fb).max():.2e}; equilibrium residual {res:.2e}", flush=True)
        assert np.abs(Fm - Fb).max() <= 1e-12
        assert res <= 1e-5
        del s
        torch.cuda.empty_cache()
    eF = float((out[MIXED][0] - out[0][0]).abs().max() / out[0][0].abs().max())
    print(f"[cfg2] refined vs FP64 F: {eF:.2e}", flush=True)
    assert eF <= 1e-6


def test_cfg3_mixed_bc_512_kmc_properties(gpu, tmp_path):
    # 512^3, refined LS, 5 % of voxels directionally softened, kT/E0 = 0.03:
    # the events run, then one mixed-BC relaxation brings sigma_yy, sigma_zz
    # within 1 MPa of 0 at an FP64 equilibrium residual of the solve's tolerance
    import torch
    import cfg_inputs as ci
    from bench import eq_residual
    from mgkmc_amd.aqs import ThermalSimulation
    n, events = 512, 40
    rng = np.random.default_rng(3)
    mu = ci.MU_UNIT * rng.uniform(0.8, 1.2, (n, n, n))
    nu = np.full((n, n, n), 0.35)
    E = 2.0 * mu * (1.0 + nu)
    del mu
    sim = ThermalSimulation(n, n, n, 20, 0.14, E, nu, pixel=0.7, barrier_generator="gaussian",
                            barrier_kwargs=ci.BARRIERS, softening_scheme="directional", temperature=0.03 * 2.2 / ci.KB,
                            strain_rate=1e7, ls_precision="mixed", output_dir=str(tmp_path),
                            rng=np.random.RandomState(3), selection_stream=np.random.RandomState(4),
                            init_modes=("device", 3), **ci.STZ)
    del E, nu
    NV = n ** 3
    idx = torch.from_numpy(rng.choice(NV, NV // 20, replace=False)).cuda()
    th = torch.from_numpy(rng.uniform(0.0, np.pi, NV // 20)).cuda()
    sim.kmc.soft[0, idx] = float(np.log(2.0))
    c2, s2 = torch.cos(2 * th).double(), torch.sin(2 * th).double()
    prev = sim.kmc.prev
    prev[0, idx], prev[1, idx], prev[3, idx], prev[4, idx] = c2, s2, s2, -c2
    targets = {(1, 1): 0.0, (2, 2): 0.0}
    ev = sim.run_simulation(10 ** 9, 1e-4, component=(0, 0), stress_targets=targets, mixed_tol=1e6,
                            enable_console_log=False, enable_summary_log=False, enable_global_log=False,
                            enable_cascade_log=False, max_events=events)
    conv, err = sim.mixed_bc(targets, (0, 0), 1e6, 50)
    sm = sim.sig_mean()
    res = eq_residual(sim.ls.sig)
    print(f"[cfg3] {len(ev)} events, eps_xx {sim.eps_macro[0, 0]:.3e}, sigma_yy {sm[1, 1]:.3e} sigma_zz "
          f"{sm[2, 2]:.3e} Pa, equilibrium residual {res:.2e}, {sim.footprint_bytes() / 1e9:.1f} GB", flush=True)
    assert len(ev) == events
    assert conv and abs(sm[1, 1]) < 1e6 and abs(sm[2, 2]) < 1e6
    assert res <= 1e-6


def test_cfg4_serving_stream_properties(gpu):
    # the 512-request stream (Poisson arrivals 4/s, seed 4; 2D and 3D sizes;
    # 100 events each, refined LS), arrivals compressed 8x so the run measures
    # admission under load: every request completes, the admitted footprint
    # stays within the budget, and sampled results are those of a solo run
    from mgkmc_amd.serving import SimServer, request_stream, run_request
    reqs = request_stream(512, 4.0, 4, 100)
    srv = SimServer()
    res = srv.serve(reqs, time_scale=0.125)
    by_id = {r.rid: r for r in res}
    lat = np.array([r.latency for r in res])
    print(f"[cfg4] {len(res)} requests in {srv.wall:.1f} s, p50 {np.percentile(lat, 50):.3f} s, "
          f"p99 {np.percentile(lat, 99):.3f} s, peak {srv.peak_active} active, {srv.peak_used / 1e9:.1f} GB", flush=True)
    assert len(res) == 512 and not any("error" in r.extra for r in res)
    assert srv.peak_used <= srv.hbm_budget
    for rq in (reqs[0], reqs[1], max(reqs, key=lambda r: np.prod(r.shape))):
        solo = run_request(rq)
        mix = by_id[rq.rid]
        assert solo.events == mix.events
        assert np.array_equal(solo.eps_macro, mix.eps_macro) and np.array_equal(solo.sig_mean, mix.sig_mean)
