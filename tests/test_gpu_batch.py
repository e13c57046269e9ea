"""Batched launches of the serving form: several simulations of one grid shape
solved (mgk_ls_solve_batch, mgk_ls_solve_refined_batch) and stepped
(mgk_kmc_run_batch) together give each member exactly what it gets alone --
the same fields bit for bit, the same iteration counts, events and random
stream states -- including members that stop early."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _solvers(shape, precision, n, seed):
    from mgkmc_amd.linear_elastic_simulator import LsSolver
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        E = 70e9 * rng.uniform(0.7 + 0.1 * i, 1.3, shape)
        nu = np.full(shape, 0.25 + 0.05 * i)
        out.append(LsSolver(E, nu, precision=precision))
    return out


def _inputs(shape, n, seed, dtype):
    import torch
    d = len(shape)
    C = 6 if d == 3 else 3
    rng = np.random.default_rng(seed)
    ebs, eps = [], []
    for i in range(n):
        eb = np.zeros((d, d))
        eb[0, 0] = 1e-3 * (i + 1)
        eb[0, 1] = 2e-4 * i  # an antisymmetric part for one member
        ebs.append(eb)
        ep = rng.standard_normal((C,) + shape) * 1e-3 * (i + 1)
        eps.append(torch.from_numpy(ep).to(device="cuda", dtype=dtype))
    return ebs, eps


def _same_info(a, b):
    return (a.iters == b.iters and a.status == b.status and a.diff == b.diff
            and list(a.sig_mean) == list(b.sig_mean) and list(a.eps_mean) == list(b.eps_mean))


@pytest.mark.parametrize("shape,prec", [((32, 24, 20), 0), ((16, 16, 16), 1), ((64, 48), 0), ((128, 128), 1)])
def test_ls_batch_bitwise(gpu, shape, prec):
    import torch
    from mgkmc_amd.linear_elastic_simulator import solve_batch, FP64, FP32
    p = FP64 if prec == 0 else FP32
    dt = torch.float64 if prec == 0 else torch.float32
    n = 3
    ebs, eps = _inputs(shape, n, 5, dt)
    solo = _solvers(shape, p, n, 3)
    infos = [s.solve(eb, ep, tol=1e-8 if prec == 0 else 1e-5) for s, eb, ep in zip(solo, ebs, eps)]
    solo_f = [(s.eps.clone(), s.sig.clone()) for s in solo]
    iters = [i.iters for i in infos]
    solo_i = [_copy_info(i) for i in infos]
    bat = _solvers(shape, p, n, 3)
    got = solve_batch(bat, ebs, eps, tol=1e-8 if prec == 0 else 1e-5)
    print(f"[ls batch] shape={shape} prec={prec} iters solo {iters} batch {[g.iters for g in got]}", flush=True)
    assert len(set(iters)) > 1  # the members stop at different iterations
    for (e, s), b, i0, i1 in zip(solo_f, bat, solo_i, got):
        assert torch.equal(e, b.eps) and torch.equal(s, b.sig)
        assert _same_info(i0, i1)
    # warm restart from the solved state, with a changed macro strain
    ebs2 = [eb * 1.5 for eb in ebs]
    w_solo = [s.solve(eb, ep, warm=True, tol=1e-8 if prec == 0 else 1e-5) for s, eb, ep in zip(solo, ebs2, eps)]
    w_solo = [_copy_info(i) for i in w_solo]
    w_bat = solve_batch(bat, ebs2, eps, warm=True, tol=1e-8 if prec == 0 else 1e-5)
    for s, b, i0, i1 in zip(solo, bat, w_solo, w_bat):
        assert torch.equal(s.eps, b.eps) and torch.equal(s.sig, b.sig) and _same_info(i0, i1)


def _copy_info(i):
    import ctypes
    from mgkmc_amd import _lib
    c = _lib.MgkLsInfo()
    ctypes.memmove(ctypes.byref(c), ctypes.byref(i), ctypes.sizeof(i))
    return c


@pytest.mark.parametrize("shape", [(32, 32, 32), (96, 64)])
def test_ls_refined_batch_bitwise(gpu, shape):
    import torch
    from mgkmc_amd.linear_elastic_simulator import solve_batch, MIXED
    n = 4
    ebs, eps = _inputs(shape, n, 9, torch.float64)
    solo = _solvers(shape, MIXED, n, 4)
    infos = [_copy_info(s.solve(eb, ep, tol=1e-8)) for s, eb, ep in zip(solo, ebs, eps)]
    it32 = [s.iters32 for s in solo]
    bat = _solvers(shape, MIXED, n, 4)
    got = solve_batch(bat, ebs, eps, tol=1e-8)
    print(f"[refined batch] shape={shape} iters32 {it32} iters64 {[i.iters for i in infos]}", flush=True)
    for s, b, i0, i1 in zip(solo, bat, infos, got):
        assert torch.equal(s.eps, b.eps) and torch.equal(s.sig, b.sig) and torch.equal(s.eps32, b.eps32)
        assert _same_info(i0, i1) and s.iters32 == b.iters32


def test_ls_batch_rejects_mixed_shapes(gpu):
    from mgkmc_amd.linear_elastic_simulator import solve_batch, FP64
    a = _solvers((16, 16, 16), FP64, 1, 1)[0]
    b = _solvers((16, 16, 8), FP64, 1, 1)[0]
    with pytest.raises(ValueError):
        solve_batch([a, b], [np.eye(3) * 1e-3] * 2)
    with pytest.raises(ValueError):
        solve_batch([a] * 9, [np.eye(3) * 1e-3] * 9)


def _state(sim):
    import torch
    k = sim.kmc
    return dict(events=list(sim.kmc_event_history), time=sim.time, eps_macro=sim.eps_macro.copy(),
                eps_p=k.eps_p.clone(), eps=sim.ls.eps.clone(), sig=sim.ls.sig.clone(), Q0=k.Q0.clone(),
                cat=k.cat.clone(), soft=k.soft.clone(), rng=sim.rng.get_state()[1:],
                sel=sim.selection_stream.get_state()[1:], iters=sim.ls_iters_total,
                sm=np.array(sim.ls.info.sig_mean), dt=torch.float64)


def _equal_state(a, b):
    import torch
    same = (a["events"] == b["events"] and a["time"] == b["time"] and np.array_equal(a["eps_macro"], b["eps_macro"])
            and a["iters"] == b["iters"] and np.array_equal(a["sm"], b["sm"]))
    for k in ("eps_p", "eps", "sig", "Q0", "cat", "soft"):
        same = same and torch.equal(a[k], b[k])
    for k in ("rng", "sel"):
        same = same and np.array_equal(a[k][0], b[k][0]) and a[k][1:] == b[k][1:]
    return same


@pytest.mark.parametrize("shape", [(24, 24, 24), (64, 64)])
def test_kmc_batch_bitwise(gpu, shape):
    # the serving recipe (refined LS, gaussian barriers drawn on the device);
    # member 1 stops after 7 events, the others run 12 in chunks of 5
    from mgkmc_amd.serving import SimRequest, build_simulation
    from mgkmc_amd.kmc_engine import kmc_events_batch
    reqs = [SimRequest(i, 0.0, shape, 0.02 + 0.02 * i, 1000 + 17 * i, events=12) for i in range(3)]
    want = [12, 7, 12]
    d = len(shape)
    unit = np.zeros((d, d))
    unit[0, 0] = 1.0
    solo = []
    for r, w in zip(reqs, want):
        sim = build_simulation(r)
        sim.update_stresses()
        sim.kmc_events(1e30, unit, w)
        solo.append(_state(sim))
        del sim
    sims = [build_simulation(r) for r in reqs]
    for s in sims:
        s.update_stresses()
    left = list(want)
    rem = [1e30] * 3
    seq = [0] * 3
    live = [0, 1, 2]
    while live:
        outs = kmc_events_batch([sims[i] for i in live], [rem[i] for i in live], [unit] * len(live),
                                [min(5, left[i]) for i in live], seqs=[seq[i] for i in live])
        nxt = []
        for i, (n, r, reason, sq, err) in zip(live, outs):
            assert err is None and reason == "count"
            left[i] -= n
            rem[i], seq[i] = r, sq
            if left[i] > 0:
                nxt.append(i)
        live = nxt
    for i, s in enumerate(sims):
        got = _state(s)
        print(f"[kmc batch] shape={shape} member {i}: {len(got['events'])} events, solo==batch "
              f"{_equal_state(solo[i], got)}", flush=True)
        assert len(got["events"]) == want[i]
        assert _equal_state(solo[i], got)


def test_ls_batch_mixed_radix_bitwise(gpu):
    # a 96^3 batch runs the compile-time radix-3 sweeps with a member dimension
    import torch
    from mgkmc_amd.linear_elastic_simulator import solve_batch, MIXED
    shape = (96, 96, 96)
    ebs, eps = _inputs(shape, 2, 13, torch.float64)
    solo = _solvers(shape, MIXED, 2, 6)
    infos = [_copy_info(s.solve(eb, ep, tol=1e-8)) for s, eb, ep in zip(solo, ebs, eps)]
    bat = _solvers(shape, MIXED, 2, 6)
    got = solve_batch(bat, ebs, eps, tol=1e-8)
    for s, b, i0, i1 in zip(solo, bat, infos, got):
        assert torch.equal(s.eps, b.eps) and torch.equal(s.sig, b.sig) and _same_info(i0, i1)
