"""Serving loop: a request's result is the same alone and in the mix (bitwise:
every kernel reduction has a fixed order and each request owns its random
streams), admission keeps the estimated footprint within the budget, and the
estimate covers the state a simulation actually holds."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SIZES = ((32, 32), (64, 48), (128, 128), (16, 16, 16), (32, 24, 20), (48, 48, 48))


def _same(a, b):
    return (a.events == b.events and np.array_equal(a.eps_macro, b.eps_macro) and a.eps_p_sum == b.eps_p_sum
            and np.array_equal(a.sig_mean, b.sig_mean) and a.time == b.time)


def test_solo_vs_mix_identical(gpu):
    # two shapes only, so requests of one shape share a batch and join it
    # while it runs (chunks of 4 events); every result equals a solo run
    from mgkmc_amd.serving import SimServer, request_stream
    reqs = request_stream(12, rate=40.0, seed=7, events=15, sizes=((32, 24, 20), (64, 48)))
    print("[serving] mix of", [r.shape for r in reqs], flush=True)
    srv = SimServer(max_active=8, chunk_events=4)
    mix = {r.rid: r for r in srv.serve(reqs)}
    print(f"[serving] peak active {srv.peak_active}, peak batch {srv.peak_batch}", flush=True)
    assert len(mix) == len(reqs) and srv.peak_active >= 2 and srv.peak_batch >= 2
    for q in reqs[:6]:
        solo = SimServer(max_active=1, batch=False).serve([q])[0]
        m = mix[q.rid]
        print(f"[serving] rid={q.rid} shape={q.shape} events={len(m.events)} solo==mix: {_same(solo, m)}",
              flush=True)
        assert len(m.events) == q.events and "error" not in m.extra
        assert _same(solo, m)


def test_admission_by_footprint(gpu):
    from mgkmc_amd.serving import SimServer, SimRequest, footprint_bytes, build_simulation
    reqs = [SimRequest(i, 0.0, (48, 48, 48), 0.05, 100 + i, events=3) for i in range(6)]
    fp = footprint_bytes(reqs[0])
    sim = build_simulation(reqs[0])
    assert sim.footprint_bytes() <= fp
    del sim
    srv = SimServer(hbm_budget=int(2.5 * fp), max_active=8)
    res = srv.serve(reqs)
    assert len(res) == 6
    assert srv.peak_used <= srv.hbm_budget and srv.peak_active == 2


def test_unbatched_mode_matches_batched(gpu):
    from mgkmc_amd.serving import SimServer, request_stream
    reqs = request_stream(6, rate=40.0, seed=11, events=8, sizes=SIZES)
    a = {r.rid: r for r in SimServer(max_active=6).serve(reqs)}
    b = {r.rid: r for r in SimServer(max_active=6, batch=False).serve(reqs)}
    assert set(a) == set(b)
    for k in a:
        assert _same(a[k], b[k])
