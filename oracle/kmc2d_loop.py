"""Oracle: the 2D KMC event loop (CPU restatement, small grids only).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Restates mgkmc/kmc_simulator.py:14-176 (initial draws: Q0 from the gaussian
generator, then the pure-shear catalog voxel by voxel), :387-525 (elastic_run,
small-strain linear branch: spectral_solver_2d cold start + unstable-stress
check), :796-1101 (KMC mode inner loop: barriers, rates, waiting time from
np.random.rand, selection, flip, prev_strain_dir, redraw of directions then
barriers, re-solve) and :1103-1143 (mixed-BC loop with the compliance step and
the per-component secant update).

Random streams: numpy's global legacy stream as the reference uses it, and a
separate RandomState for the selection draw r = total * u (the reference's
numba-internal draw, made seedable).
"""
import numpy as np

from . import kmc as ok
from . import ls as ol

KB = 8.617e-5


def run_events_2d(nx, ny, M, gamma0, E, nu, pixel, temperature, strain_rate, step_size, component, n_events,
                  seed, sel_seed, barrier_kwargs=None, softening_scheme="isotropic", softening_cap=2.0, jp=10.0,
                  jt=30.0, jn=0.0, nu0=1e13, q_act_temp=0.37, plane_mode="plane_strain", ls_tol=1e-6,
                  stress_targets=None, mixed_tol=1e-4, mixed_max_iter=50, n_global_steps=10 ** 9,
                  eps_plastic=None):
    """Returns (events [(x, y, m, time, is_instab)], final state dict)."""
    np.random.seed(seed)
    sel = np.random.RandomState(sel_seed)
    bk = dict(mean=2.0, std=0.6, min_cutoff=0.1) if barrier_kwargs is None else dict(barrier_kwargs)

    def gen(shape):
        return ok.gaussian_barriers(shape, np.random, mean=bk.get("mean", 2.0), std=bk.get("std", 0.6),
                                    min_cutoff=bk.get("min_cutoff", 0.1), max_cutoff=bk.get("max_cutoff", None))
    E = np.asarray(E, float)
    E = E * 1e9 if E.mean() < 1e6 else E
    nu = np.asarray(nu, float)
    volume = pixel ** 3
    tau = 1.0 / (nu0 * np.exp(-q_act_temp / (KB * temperature))) if temperature > 0 else np.inf
    Q0 = gen((nx, ny, M))
    cat = np.zeros((nx, ny, M, 2, 2))
    for x in range(nx):
        for y in range(ny):
            cat[x, y] = ok.stz_catalog_glass_2d(M, gamma0, np.random)
    Q = np.zeros((nx, ny, M))
    soft = np.zeros((nx, ny, 2))
    tl = np.full((nx, ny), -np.inf)
    prev = np.zeros((nx, ny, 2, 2))
    ep = np.zeros((nx, ny, 2, 2)) if eps_plastic is None else np.array(eps_plastic, float)
    eps_macro = np.zeros((2, 2))
    t = 0.0
    scheme = 1 if softening_scheme == "directional" else 0
    state = {}
    targets = dict(stress_targets or {})

    def elastic_run():
        e, s, _, _ = ol.spectral_solver_2d(E, nu, eps_macro, eps_plastic=ep, pixel=pixel, plane_mode=plane_mode,
                                           tol=ls_tol)
        state["sig"] = s
        sm = s.mean(axis=(0, 1))
        if np.any(np.isnan(s)) or np.max(np.abs(sm)) / 1e9 > 20.0:
            raise ValueError("Unstable stress in small-strain elastic_run")
        return sm

    unit = np.zeros((2, 2))
    unit[component] = 1.0
    elastic_run()
    dt_step = abs(step_size) / strain_rate if strain_rate > 0 else 1.0
    events = []
    remaining = 0.0
    done = 0
    E_avg, nu_avg = E.mean(), nu.mean()

    def result():
        return events, dict(eps_plastic=ep, Q0=Q0, soft=soft, sig=state["sig"], eps_macro=eps_macro.copy(), time=t)

    while done < n_global_steps:
        remaining += dt_step
        while remaining > 0:
            ok.compute_barrier_2d(Q, Q0, state["sig"], cat, volume, soft, tl, t, prev, softening_cap, scheme, tau)
            rates, idx, total = ok.compute_rates_2d(Q, 1.0, temperature, nu0, cascade_mode=False)
            if total > 0:
                tw = -np.log(np.random.rand()) / total
                if tw < remaining:
                    t += tw
                    eps_macro += unit * (strain_rate * tw)
                    remaining -= tw
                    i = ok.select_event(rates, total, total * sel.random_sample())
                    x, y, m = ok.decode_index_2d(idx[i], ny, M)
                    inst = Q[x, y, m] <= 0.0
                    C = cat[x, y, m].copy()
                    ok.apply_flip_soa_2d(ep, soft, tl, cat, x, y, m, t, jp, jt, softening_cap, jn)
                    prev[x, y] = C
                    cat[x, y] = ok.stz_catalog_glass_2d(M, gamma0, np.random)
                    Q0[x, y] = gen((M,))
                    elastic_run()
                    events.append((int(x), int(y), int(m), t, bool(inst)))
                    if len(events) >= n_events:
                        return result()
                    continue
            eps_macro += unit * (strain_rate * remaining)
            t += remaining
            remaining = 0.0
        prev_e, prev_s = {}, {}
        for _ in range(mixed_max_iter):
            sm = elastic_run()
            err = np.zeros((2, 2))
            mx = 0.0
            for k, v in targets.items():
                if k[0] < 2 and k[1] < 2:
                    err[k] = v - sm[k]
                    mx = max(mx, abs(err[k]))
            if mx < mixed_tol:
                break
            d_eps = (err - nu_avg * np.trace(err) * np.eye(2)) / E_avg
            for k in targets:
                if k[0] < 2 and k[1] < 2:
                    st = d_eps[k]
                    if k in prev_e:
                        de = eps_macro[k] - prev_e[k]
                        ds = sm[k] - prev_s[k]
                        if abs(de) > 1e-14 and np.isfinite(ds):
                            slope = ds / de
                            if np.isfinite(slope) and slope > 0.01 * E_avg:
                                st = err[k] / slope
                    prev_e[k] = eps_macro[k]
                    prev_s[k] = sm[k]
                    eps_macro[k] += st
        done += 1
    return result()
