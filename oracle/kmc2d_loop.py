"""Oracle: the 2D KMC event loop (CPU restatement, small grids only).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Restates mgkmc/kmc_simulator.py:14-176 (initial draws: Q0 from the gaussian
generator, then the pure-shear catalog voxel by voxel), :387-525 (elastic_run,
small-strain linear branch: spectral_solver_2d cold start + unstable-stress
check), :796-1101 (KMC mode inner loop: barriers, rates, waiting time from
np.random.rand, selection, flip, prev_strain_dir, redraw of directions then
barriers, re-solve) and :1103-1143 (mixed-BC loop with the compliance step and
the per-component secant update).

Nonlinear small-strain models (``model`` = "secant_degradation", "rose",
"landau"): elastic_run calls the matching LS solver (:449-512: secant with the
Lame pair of the plane mode, Rose and Landau with the plane-strain pair;
Landau warm-started from the current field with one e33 state for the run),
and a flip goes through the sub-stepping of :268-385 (first one step; if the
re-solve raises ValueError or FloatingPointError, the state is restored and the
flip is retried in N = max(N_h, 2) .. +19 sub-steps of C/N each, softening,
prev and redraws at the last, N_h = min(max(2, ceil(sqrt(C:C/2)/0.02)), 40)).
cascade_mode (:577-794, :870-888): after the initial solve, each pass of the
inner loop first runs a cascade when any Q < stability_threshold: repeat
{barriers; unstable modes sorted by Q; first mode per voxel; flip, prev and
redraw each (sub-stepped together under a nonlinear law); one re-solve} until
none is unstable or the batch count exceeds nx ny; the run stops when more than
int(max_cascade_steps_pct nx ny) consecutive cascades flipped something; rates
then count Q > 0 only.  Cascade flips are recorded as (step, batch, x, y, m)
with the reference's global step (1 at the start, +1 per KMC event and per
elastic increment).

enable_thermal (:60-75, :559-575, :749-753, :896, :1029-1033, :1099): a local temperature
field sets the rates, diffuses over each waiting time and each remaining
increment time (oracle/aqs_loop.py heat_step), and each flip heats its voxel
by |sigma:C| / (rho Cp), capped (linear law).

strain_assumption="finite_strain" (:156-168, :389-445, :895-1000, :1103-1104):
elastic_run is the 2D Newton-CG step (oracle/finite_strain2d.py solver_step_2d
with the reference's default tolerances; ``model`` = "svk", "landau" or
"murnaghan", constants in ``model_params``) from the current F with F_plastic,
macro BC from the driven strain with F_macro as the initial F_bar, F_macro and
the macro strain from the returned F_bar, ValueError on NaN or |<sigma>| > 20
GPa; a flip multiplies F_p by expm(C/N) N times (N = 1, then N_h .. N_h + 19
on ValueError from the restored state), softening, prev and redraws at the
last, a step after each; elastic increments end with one elastic_run (no
mixed loop).

``fail_once_at`` (test hook) makes the first re-solve after that event's flip
raise ValueError, to drive the sub-stepped path.

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
                  eps_plastic=None, model="linear", model_params=None, fail_once_at=None, cascade_mode=False,
                  stability_threshold=0.0, max_cascade_steps_pct=0.3, enable_thermal=False, Cp=420.0, rho=6125.0,
                  thermal_diffusivity=3.0e-6, thermal_coords="pixel", temperature_cap=1000.0, thermostat=False,
                  tau_bath=0.0, strain_assumption="small_strain", stz_mode="pure_shear", redraw_directions=True,
                  redraw_barriers=True, solver="newton_cg", fast_patching=None, record=None):
    """Returns (events [(x, y, m, time, is_instab)], final state dict; its
    "increments" holds (eps_macro, <sigma>) after each elastic increment).

    ``fast_patching`` ({"enabled", "patch_radius" (5), "sync_interval" (100)},
    mgkmc/kmc_simulator.py:76-78, :154-171, :229-261, :764-792, :1044-1098;
    linear small strain only): KMC events and cascade flips predict the stress
    with oracle/patching.py instead of re-solving (events: the macro-unit
    stress over the waiting time first, weights from the flipped mode; cascade
    flips: weights from the catalog entry after its redraw), a re-solve every
    sync_interval event flips and after a cascade that flipped anything; once
    the macro-unit stress exists the remaining time of an increment adds it
    too.

    ``record`` (a dict, test hook): its "cascades" list is the run's cascade
    history as it grows, so a run that raises still shows how far it got."""
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
            cat[x, y] = ok.stz_catalog_glass_2d(M, gamma0, np.random, stz_mode)
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

    mp = dict(model_params or {})
    finite = strain_assumption == "finite_strain"
    if finite:
        from . import finite_strain2d as f2
        from scipy.linalg import expm
        fI2, fI4, fI4rt, fI4s, fII = f2.identity_tensors_2d(nx, ny)
        fC4 = f2.build_C4_2d(E, nu, fI4s, fII, plane_mode)
        fG = f2.build_ghat4_2d(nx, ny, nx * pixel, ny * pixel)
        state["F"] = fI2.copy()
        state["Fp"] = fI2.copy()
        state["F_macro"] = np.eye(2)
    elif model == "secant_degradation":
        lam, mu = ol.compute_lame_2d(E, nu, plane_mode)
    elif model in ("rose", "landau"):
        lam, mu = ol.compute_lame_2d(E, nu, "plane_strain")
    elif model != "linear":
        raise ValueError(model)
    state["eps"] = np.zeros((nx, ny, 2, 2))
    e33_state = {}
    inject = {"armed": False}

    def elastic_run_finite():
        Fb, Pt, Pm = f2.build_bc_2d(component, eps_macro[component], targets, state["F_macro"])
        F, P, Sig, _, Fb = f2.solver_step_2d(state["F"], Fb, fG, fC4, fI2, fI4, fI4rt, Fp=state["Fp"],
                                             driving_component=component, P_target=Pt, P_mask=Pm, E_avg=E.mean(),
                                             nu_avg=nu.mean(), model_type=model, plane_mode=plane_mode,
                                             solver=solver, pixel=pixel, **mp)
        state["F"] = F
        state["sig"] = np.einsum("ijxy->xyij", Sig)
        state["F_macro"] = Fb
        for i in range(2):
            for j in range(2):
                eps_macro[i, j] = Fb[i, j] - 1.0 if i == j else Fb[i, j]
        sm = state["sig"].mean(axis=(0, 1))
        if np.any(np.isnan(state["sig"])) or np.any(np.isnan(F)) or np.max(np.abs(sm)) / 1e9 > 20.0:
            raise ValueError("Simulation stopped due to unstable stress")
        return sm

    def elastic_run():
        if inject["armed"]:
            inject["armed"] = False
            raise ValueError("injected solver failure")
        if finite:
            return elastic_run_finite()
        if model == "linear":
            e, s, _, _ = ol.spectral_solver_2d(E, nu, eps_macro, eps_plastic=ep, pixel=pixel, plane_mode=plane_mode,
                                               tol=ls_tol)
        elif model == "secant_degradation":
            e, s, _, _ = ol.spectral_solver_secant_2d(lam, mu, mp["d"], mp["k"], eps_macro, eps_plastic=ep,
                                                      tol=ls_tol, pixel=pixel, plane_mode=plane_mode)
        elif model == "rose":
            e, s, _, _ = ol.spectral_solver_rose_2d(lam, mu, mp["eta"], mp["mu_floor_fraction"], eps_macro,
                                                    eps_plastic=ep, tol=ls_tol, pixel=pixel, plane_mode=plane_mode)
        else:
            lp = {k: mp[k] for k in ("v1", "v2", "v3", "g1", "g2", "g3", "g4")}
            cap = {k: v for k, v in mp.items() if k.startswith("strain_capping")}
            e, s, _, _ = ol.spectral_solver_landau_2d(lam, mu, **lp, eps_bar=eps_macro, eps_plastic=ep, tol=ls_tol,
                                                      pixel=pixel, plane_mode=plane_mode, e33_state=e33_state,
                                                      eps_init=state["eps"], **cap)
        state["eps"] = e
        state["sig"] = s
        sm = s.mean(axis=(0, 1))
        if np.any(np.isnan(s)) or np.max(np.abs(sm)) / 1e9 > 20.0:
            raise ValueError("Unstable stress in small-strain elastic_run")
        return sm

    fp = dict(fast_patching or {})
    fp_on = bool(fp.get("enabled", False)) and not finite and model == "linear"
    if fp_on:
        from . import patching as opt
        fp_R = int(fp.get("patch_radius", 5))
        fp_sync = int(fp.get("sync_interval", 100))
        # at construction, with the default solver arguments (kmc_simulator.py:134, :171)
        fp_K, fp_mm = opt.patch_kernels_2d(E, nu, fp_R, pixel=pixel, plane_mode=plane_mode)
        state["fp_flips"] = 0
        state["fp_unit"] = None

    def patch(x, y, C):
        opt.stamp(state["sig"], (x, y), opt.coefficients_2d(C), fp_K, fp_mm, fp_R)

    def redraw(x, y, m):
        # the voxel's modes and barriers, or only the flipped mode's (kmc_simulator.py:1000-1008)
        if redraw_directions or redraw_barriers:
            if redraw_directions:
                cat[x, y] = ok.stz_catalog_glass_2d(M, gamma0, np.random, stz_mode)
            if redraw_barriers:
                Q0[x, y] = gen((M,))
        else:
            cat[x, y, m] = ok.stz_catalog_glass_2d(1, gamma0, np.random, stz_mode)[0]
            Q0[x, y, m] = gen((1,))[0]

    def soften(x, y, c):
        sq = ok.vm_sq_2d(c)
        gp = soft[x, y, 0] + jp * sq
        if softening_cap > 0 and gp > softening_cap:
            gp = softening_cap
        soft[x, y, 0] = gp
        soft[x, y, 1] = jt * sq
        tl[x, y] = t
        if jn > 0.0:
            for dx in (-1, 0, 1):
                for dy in (-1, 0, 1):
                    if dx == 0 and dy == 0:
                        continue
                    a, b = (x + dx + nx) % nx, (y + dy + ny) % ny
                    g = soft[a, b, 0] + jn * jp * sq
                    if softening_cap > 0 and g > softening_cap:
                        g = softening_cap
                    soft[a, b, 0] = g
                    soft[a, b, 1] += jn * jt * sq

    def flip_substepped(flips):
        arrays = (ep, soft, tl, prev, cat, Q0, Tl)
        backup = [a.copy() for a in arrays]
        b_eps, b_sig = state["eps"].copy(), state["sig"].copy()
        nh = min(max(max(2, int(np.ceil(np.sqrt(0.5 * np.sum(cat[x, y, m] ** 2)) / 0.02))) for x, y, m in flips), 40)
        for N in [1] + list(range(max(nh, 2), max(nh, 2) + 20)):
            for a, b in zip(arrays, backup):
                a[...] = b
            state["eps"], state["sig"] = b_eps.copy(), b_sig.copy()
            try:
                for k in range(1, N + 1):
                    for x, y, m in flips:
                        c = cat[x, y, m].copy()
                        ep[x, y] += c / N
                        if k == N:
                            soften(x, y, c)
                            prev[x, y] = c
                            flip_heat(x, y, c)
                            redraw(x, y, m)
                    elastic_run()
                state.setdefault("substeps", []).append(N)
                return
            except (ValueError, FloatingPointError):
                continue
        raise ValueError("Mechanical solver failed to converge under small-strain sub-stepping.")

    def flip_finite(flips):
        # one KMC event or one cascade batch (:605-700, :895-1000)
        Cs = [cat[x, y, m].copy() for x, y, m in flips]
        arrays = (soft, tl, prev, cat, Q0, Tl)
        backup = [a.copy() for a in arrays]
        b = {k: state[k].copy() for k in ("F", "Fp", "sig", "F_macro")}
        b_em = eps_macro.copy()
        nh = min(max(max(2, int(np.ceil(np.sqrt(0.5 * np.sum(C ** 2)) / 0.02))) for C in Cs), 40)
        for N in [1] + list(range(max(nh, 2), max(nh, 2) + 20)):
            for a, bb in zip(arrays, backup):
                a[...] = bb
            for k in b:
                state[k] = b[k].copy()
            eps_macro[...] = b_em
            try:
                for k in range(1, N + 1):
                    for (x, y, m), C in zip(flips, Cs):
                        state["Fp"][:, :, x, y] = expm(C / N) @ state["Fp"][:, :, x, y]
                        if k == N:
                            soften(x, y, C)
                            prev[x, y] = C
                            flip_heat(x, y, C)
                            redraw(x, y, m)
                    elastic_run()
                state.setdefault("substeps", []).append(N)
                return
            except ValueError:
                continue
        raise ValueError("Mechanical solver failed to converge under sub-stepping.")

    # thermal extension (:60-75, :559-575, :749-753, :1029-1033): as oracle/aqs_loop.py
    from .aqs_loop import heat_operator, heat_step
    Tl = np.full((nx, ny), float(temperature))
    if enable_thermal:
        k2, Dsc = heat_operator((nx, ny), pixel, thermal_diffusivity, thermal_coords)
    Tr = Tl if enable_thermal else temperature

    def heat(dt):
        if enable_thermal:
            heat_step(Tl, k2, Dsc, dt, temperature, thermostat, tau_bath, temperature_cap)

    def flip_heat(x, y, C):
        if enable_thermal:
            Tl[x, y] += abs(np.sum(state["sig"][x, y] * C)) / (rho * Cp)
            if temperature_cap > 0:
                Tl[x, y] = min(Tl[x, y], temperature_cap)

    cascades = [] if record is None else record.setdefault("cascades", [])

    def run_cascade(step):
        local = total = 0
        while True:
            ok.compute_barrier_2d(Q, Q0, state["sig"], cat, volume, soft, tl, t, prev, softening_cap, scheme, tau)
            un = ok.find_unstable_2d(Q, stability_threshold)
            if len(un) == 0:
                break
            un = un[np.argsort(Q[un[:, 0], un[:, 1], un[:, 2]])]
            seen = set()
            flips = []
            for x, y, m in un:
                if (x, y) in seen:
                    continue
                seen.add((x, y))
                flips.append((x, y, m))
            if finite:
                flip_finite(flips)
            elif model == "linear":
                for x, y, m in flips:
                    ok.apply_flip_soa_2d(ep, soft, tl, cat, x, y, m, t, jp, jt, softening_cap, jn)
                    prev[x, y] = cat[x, y, m].copy()
                    flip_heat(x, y, prev[x, y])
                    redraw(x, y, m)
                    if fp_on:
                        patch(x, y, cat[x, y, m])  # the entry after its redraw (kmc_simulator.py:764-765)
                if not fp_on:
                    elastic_run()
            else:
                flip_substepped(flips)
            cascades.extend((step, local, int(x), int(y), int(m)) for x, y, m in flips)
            total += len(flips)
            local += 1
            if local > nx * ny:
                break
        if fp_on and total > 0:
            elastic_run()  # kmc_simulator.py:790-792
        return total

    unit = np.zeros((2, 2))
    unit[component] = 1.0
    elastic_run()
    n_casc = 0
    step = 1
    dt_step = abs(step_size) / strain_rate if strain_rate > 0 else 1.0
    events = []
    increments = []  # (eps_macro, <sigma>) after each elastic increment
    remaining = 0.0
    done = 0
    E_avg, nu_avg = E.mean(), nu.mean()

    def result():
        return events, dict(eps_plastic=ep, Q0=Q0, soft=soft, sig=state["sig"], eps_macro=eps_macro.copy(), time=t,
                            increments=increments,
                            substeps=state.get("substeps"), cascades=cascades, Tlocal=Tl, F=state.get("F"),
                            F_plastic=state.get("Fp"), F_macro=state.get("F_macro"))

    while done < n_global_steps:
        remaining += dt_step
        while remaining > 0:
            ok.compute_barrier_2d(Q, Q0, state["sig"], cat, volume, soft, tl, t, prev, softening_cap, scheme, tau)
            if cascade_mode:
                if len(ok.find_unstable_2d(Q, stability_threshold)) > 0:
                    if run_cascade(step) > 0:
                        n_casc += 1
                    if n_casc > int(max_cascade_steps_pct * nx * ny):
                        return result()
                    continue
                n_casc = 0
            rates, idx, total = ok.compute_rates_2d(Q, 1.0, Tr, nu0, cascade_mode=cascade_mode)
            if total > 0:
                tw = -np.log(np.random.rand()) / total
                if tw < remaining:
                    t += tw
                    heat(tw)
                    eps_macro += unit * (strain_rate * tw)
                    remaining -= tw
                    i = ok.select_event(rates, total, total * sel.random_sample())
                    x, y, m = ok.decode_index_2d(idx[i], ny, M)
                    inst = Q[x, y, m] <= stability_threshold
                    if fail_once_at is not None and len(events) == fail_once_at:
                        inject["armed"] = True
                    if finite:
                        flip_finite([(x, y, m)])
                    elif model == "linear":
                        C = cat[x, y, m].copy()
                        ok.apply_flip_soa_2d(ep, soft, tl, cat, x, y, m, t, jp, jt, softening_cap, jn)
                        prev[x, y] = C
                        flip_heat(x, y, C)
                        redraw(x, y, m)
                        if fp_on:  # kmc_simulator.py:1044-1078
                            if state["fp_unit"] is None:
                                _, state["fp_unit"], _, _ = ol.spectral_solver_2d(
                                    E, nu, unit, eps_plastic=np.zeros_like(ep), pixel=pixel, plane_mode=plane_mode,
                                    tol=ls_tol)
                            if tw > 0:
                                state["sig"] += state["fp_unit"] * (strain_rate * tw)
                            patch(x, y, C)
                            state["fp_flips"] += 1
                            if state["fp_flips"] >= fp_sync:
                                elastic_run()
                                state["fp_flips"] = 0
                        else:
                            elastic_run()
                    else:
                        flip_substepped([(x, y, m)])
                    events.append((int(x), int(y), int(m), t, bool(inst)))
                    if len(events) >= n_events:
                        return result()
                    step += 1
                    continue
            if fp_on and state["fp_unit"] is not None:
                state["sig"] += state["fp_unit"] * (strain_rate * remaining)  # kmc_simulator.py:1094-1096
            eps_macro += unit * (strain_rate * remaining)
            heat(remaining)
            t += remaining
            remaining = 0.0
        if finite:
            elastic_run()
        prev_e, prev_s = {}, {}
        for _ in range(0 if finite else mixed_max_iter):
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
        increments.append((eps_macro.copy(), state["sig"].mean(axis=(0, 1))))
        done += 1
        step += 1
    return result()
