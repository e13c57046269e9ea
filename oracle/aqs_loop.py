"""Oracle: the 3D KMC event loop (CPU restatement, small grids only).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Restates mgkmc/aqs.py:16-215 (initial draws), :273-351 (update_stresses via
update_stress_fft_full, cold start), :848-1087 (KMC mode inner loop) and
:1105-1142 (mixed-BC relaxation), for small-strain linear elasticity,
KMC mode, planar catalogs, redraw of directions and barriers.

cascade_mode (:480-680 small-strain branch, :820-870, :1103-1142): the
initial relaxation is a cascade (no solve before it: barriers from the zero
initial stress); each pass of the inner loop, and each iteration of the
mixed-BC loop, first runs a cascade when any Q < stability_threshold:
repeat {barriers; unstable modes sorted by Q (np.argsort over the C-order
scan); first mode per voxel; flip, prev and redraw each in that order; one
re-solve} until none is unstable, stopping the run when the number of batches
exceeds int(max_cascade_steps_pct * voxels); rates then count Q > 0 only.
Cascade flips are recorded as (step, batch, x, y, z, m) with the reference's
global step (1 at the start, +1 per KMC event and per elastic increment).

Random streams: numpy's global legacy stream exactly as the reference uses it
(Q0, then the catalog voxel by voxel with stz_catalog_glass, then per event
one uniform for the waiting time, 2M normals for the new modes and M normals
for the new Q0), and a separate RandomState for the selection draw
r = total * u (the reference's numba-internal draw, made seedable).
"""
import numpy as np

from . import kmc as ok
from . import ls as ol

KB = 8.617e-5


def run_events(nx, ny, nz, M, gamma0, E, nu, pixel, temperature, strain_rate, step_size, component,
               n_events, seed, sel_seed, barrier_kwargs=None, softening_scheme="isotropic", softening_cap=2.0,
               jp=10.0, jt=30.0, jn=0.0, nu0=1e13, q_act_temp=0.37, ls_tol=1e-6, stress_targets=None,
               mixed_tol=1e-4, mixed_max_iter=50, n_global_steps=10 ** 9, cascade_mode=False,
               stability_threshold=0.0, max_cascade_steps_pct=0.3, eps_plastic=None):
    """Returns the list of events (x, y, z, m, time, is_instab) and final state."""
    np.random.seed(seed)
    sel = np.random.RandomState(sel_seed)
    bk = dict(mean=2.0, std=0.6, min_cutoff=0.1) if barrier_kwargs is None else dict(barrier_kwargs)
    gen = lambda shape: ok.gaussian_barriers(shape, np.random, **{"mean": bk.get("mean", 2.0),
                                                                  "std": bk.get("std", 0.6),
                                                                  "min_cutoff": bk.get("min_cutoff", 0.1),
                                                                  "max_cutoff": bk.get("max_cutoff", None)})
    E = np.asarray(E, float)
    E = E * 1e9 if E.mean() < 1e6 else E
    nu = np.asarray(nu, float)
    volume = pixel ** 3
    tau = 1.0 / (nu0 * np.exp(-q_act_temp / (KB * temperature))) if temperature > 0 else np.inf
    Q0 = gen((nx, ny, nz, M))
    cat = np.zeros((nx, ny, nz, M, 3, 3))
    for x in range(nx):
        for y in range(ny):
            for z in range(nz):
                cat[x, y, z] = ok.stz_catalog_glass(M, gamma0, np.random)
    Q = np.zeros((nx, ny, nz, M))
    soft = np.zeros((nx, ny, nz, 4))
    tl = np.full((nx, ny, nz), -np.inf)
    prev = np.zeros((nx, ny, nz, 3, 3))
    ep = np.zeros((nx, ny, nz, 3, 3)) if eps_plastic is None else np.array(eps_plastic, float)
    eps_macro = np.zeros((3, 3))
    t = 0.0
    scheme = 0 if softening_scheme == "isotropic" else 1
    state = {}

    def update():
        e, s, em, sm = ol.update_stress_fft_full(ep, eps_macro, E, nu, pixel=pixel, tol=ls_tol)
        state["sig"] = s
        state["sm"] = sm

    state["sig"] = np.zeros((nx, ny, nz, 3, 3))
    state["sm"] = np.zeros((3, 3))
    cascades = []

    def run_cascade(step):
        local = 0
        while True:
            ok.compute_barrier(Q, Q0, state["sig"], cat, volume, soft, tl, t, prev, softening_cap, scheme, tau)
            un = ok.find_unstable(Q, stability_threshold)
            if len(un) == 0:
                break
            qv = Q[un[:, 0], un[:, 1], un[:, 2], un[:, 3]]
            un = un[np.argsort(qv)]
            seen = set()
            flips = []
            for x, y, z, m in un:
                if (x, y, z) in seen:
                    continue
                seen.add((x, y, z))
                flips.append((x, y, z, m))
            for x, y, z, m in flips:
                ok.apply_flip_soa(ep, None, soft, tl, cat, x, y, z, m, t, jp, jt, softening_cap, jn)
                prev[x, y, z] = cat[x, y, z, m].copy()
                cat[x, y, z] = ok.stz_catalog_glass(M, gamma0, np.random)
                Q0[x, y, z] = gen((M,))
                cascades.append((step, local, int(x), int(y), int(z), int(m)))
            update()
            local += 1
            if local > int(max_cascade_steps_pct * nx * ny * nz):
                return True
        return False

    def result():
        return events, dict(eps_plastic=ep, Q=Q, Q0=Q0, soft=soft, sig=state["sig"], eps_macro=eps_macro, time=t,
                            cascades=cascades)

    targets = dict(stress_targets or {})
    if not targets:
        if component == (0, 0):
            targets = {(1, 1): 0.0, (2, 2): 0.0}
        elif component == (0, 1):
            targets = {(0, 0): 0.0, (1, 1): 0.0, (2, 2): 0.0}
    unit = np.zeros((3, 3))
    unit[component] = 1.0
    events = []
    if cascade_mode:
        if run_cascade(0):
            return result()
    else:
        update()
    dt_step = abs(step_size) / strain_rate if strain_rate > 0 else 1.0
    remaining = 0.0
    done = 0
    step = 1
    E_avg, nu_avg = np.mean(E), np.mean(nu)
    while done < n_global_steps:
        remaining += dt_step
        while remaining > 0:
            ok.compute_barrier(Q, Q0, state["sig"], cat, volume, soft, tl, t, prev, softening_cap, scheme, tau)
            if cascade_mode and len(ok.find_unstable(Q, stability_threshold)) > 0:
                if run_cascade(step):
                    return result()
                continue
            rates, idx, total = ok.compute_rates(Q, 1.0, temperature, nu0, cascade_mode=cascade_mode)
            if total > 0:
                u = np.random.uniform()
                tw = -np.log(u) / total
                if tw < remaining:
                    t += tw
                    eps_macro += unit * (strain_rate * tw)
                    remaining -= tw
                    i = ok.select_event(rates, total, total * sel.random_sample())
                    x, y, z, m = ok.decode_index(idx[i], ny, nz, M)
                    inst = Q[x, y, z, m] <= stability_threshold
                    C = cat[x, y, z, m].copy()
                    ok.apply_flip_soa(ep, None, soft, tl, cat, x, y, z, m, t, jp, jt, softening_cap, jn)
                    prev[x, y, z] = C
                    cat[x, y, z] = ok.stz_catalog_glass(M, gamma0, np.random)
                    Q0[x, y, z] = gen((M,))
                    update()
                    events.append((int(x), int(y), int(z), int(m), t, bool(inst)))
                    if len(events) >= n_events:
                        return result()
                    step += 1
                    continue
            eps_macro += unit * (strain_rate * remaining)
            t += remaining
            remaining = 0.0
        update()
        for _ in range(mixed_max_iter):
            if cascade_mode and run_cascade(step):
                return result()
            sc = state["sm"]
            err = np.zeros((3, 3))
            mx = 0.0
            for k, v in targets.items():
                err[k] = v - sc[k]
                mx = max(mx, abs(err[k]))
            if mx < mixed_tol:
                break
            corr = (err - nu_avg * np.trace(err) * np.eye(3)) / E_avg
            corr[component] = 0.0
            eps_macro += corr
            update()
        done += 1
        step += 1
    return result()
