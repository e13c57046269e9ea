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

strain_assumption="finite_strain" (:273-330, :887-996, :1088-1102; svk
and newton_cg): update() is the finite-strain step (oracle/finite_strain.py)
from the current F with F_plastic, the macro BC built from the driven strain
with F_macro as the initial F_bar; F_macro is the returned F_bar and the macro
strain is read from it (F_ii - 1, F_ij); after each elastic increment's step
the reference runs a cascade (cascade_mode) or a second step from the
converged state (:1089-1102); barriers use the Cauchy stress; a
NaN field or |<sigma>| > 20 GPa raises ValueError.  A flip is applied in N
sub-steps F_p <- expm(C/N) F_p (N from max(2, ceil(sqrt(C:C/2) / 0.005)), at
most 40, then N + 1 .. on ValueError, state restored), with softening, prev
and redraws at the last sub-step and a finite-strain step after each; no
mixed-BC loop after an increment (the step enforces the stress targets).

enable_thermal (:83-99, :462-478, :877, :1000-1005; also in cascades,
:630-635): a local temperature field (initially the temperature) sets the
rates; after a waiting time dt (and the remaining time of an increment) it
diffuses spectrally, T <- ifft(fft(T) exp(-|k|^2 D dt)), with k of the box n
(pixel coordinates) or n * pixel (physical, D in m^2/s -> nm^2/s), then the
thermostat (T0 + (T - T0) exp(-dt/tau_bath), or T - (<T> - T0) for tau_bath
<= 0) and the cap; a flip heats its voxel by |sigma:C| / (rho Cp) (stress
before the re-solve), capped.  The rates of an event are those computed
before its heat step.

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


def heat_operator(shape, pixel, thermal_diffusivity, thermal_coords="pixel"):
    """|k|^2 on the periodic box (n per axis in pixel coordinates, n * pixel in
    physical ones) and the diffusivity in those units (m^2/s -> nm^2/s for
    physical), as the simulators build them (mgkmc/aqs.py:93-99,
    kmc_simulator.py:66-75)."""
    phys = thermal_coords.lower() != "pixel"
    Ls = [n * pixel if phys else float(n) for n in shape]
    ks = np.meshgrid(*[2 * np.pi * np.fft.fftfreq(n, d=L / n) for n, L in zip(shape, Ls)], indexing="ij")
    k2 = ks[0] ** 2 + ks[1] ** 2
    if len(shape) == 3:
        k2 = k2 + ks[2] ** 2
    return k2, thermal_diffusivity * (1e18 if phys else 1.0)


def heat_step(Tl, k2, Dsc, dt, temperature, thermostat, tau_bath, temperature_cap):
    """heat_conducting_3d / _2d in place (mgkmc/aqs.py:462-478)."""
    if dt <= 0:
        return
    Th = ol.fft_field(Tl)
    Th *= np.exp(-k2 * Dsc * dt)
    Tl[...] = ol.ifft_field(Th)
    if thermostat:
        if tau_bath > 0:
            Tl[...] = temperature + (Tl - temperature) * np.exp(-dt / tau_bath)
        else:
            Tl[...] = Tl - (np.mean(Tl) - temperature)
    if temperature_cap > 0:
        Tl[...] = np.clip(Tl, a_min=None, a_max=temperature_cap)


class _PlanarCatalog:
    """Planar catalogs [[gxx, gxy, 0], [gxy, -gxx, 0], [0, 0, 0]] stored as
    their two parameters (..., M, 2) with the indexing of the full
    (..., M, 3, 3) array the loop uses (fast mode: 128^3 x 20 modes would
    take 3 GB as full tensors)."""

    def __init__(self, shape, M):
        self.p = np.zeros(tuple(shape) + (M, 2))

    def __getitem__(self, k):
        g = self.p[k]
        out = np.zeros(g.shape[:-1] + (3, 3))
        out[..., 0, 0], out[..., 1, 1] = g[..., 0], -g[..., 0]
        out[..., 0, 1] = out[..., 1, 0] = g[..., 1]
        return out

    def __setitem__(self, k, v):
        v = np.asarray(v)
        self.p[k] = np.stack([v[..., 0, 0], v[..., 0, 1]], axis=-1)


def run_events(nx, ny, nz, M, gamma0, E, nu, pixel, temperature, strain_rate, step_size, component,
               n_events, seed, sel_seed, barrier_kwargs=None, softening_scheme="isotropic", softening_cap=2.0,
               jp=10.0, jt=30.0, jn=0.0, nu0=1e13, q_act_temp=0.37, ls_tol=1e-6, stress_targets=None,
               mixed_tol=1e-4, mixed_max_iter=50, n_global_steps=10 ** 9, cascade_mode=False,
               stability_threshold=0.0, max_cascade_steps_pct=0.3, eps_plastic=None,
               strain_assumption="small_strain", enable_thermal=False, Cp=420.0, rho=6125.0,
               thermal_diffusivity=3.0e-6, thermal_coords="pixel", temperature_cap=1000.0, thermostat=False,
               tau_bath=0.0, solver="newton_cg", use_3d_barriers=False, redraw_directions=True,
               redraw_barriers=True, barrier_generator="gaussian", stop_on_stress_drop=None,
               stress_drop_component=None, stop_post_drop_steps=20, ignore_drop_steps=0, stress_drop_lookback=1,
               eps_target=None, fast=False, progress=None, fast_patching=None):
    """Returns the list of events (x, y, z, m, time, is_instab) and final state.
    ``solver``: the finite-strain step's solver ("newton_cg", "al", "dbfft").
    Stress-drop stop (aqs.py:834, :1070, :1149-1181): the driven stress
    history holds the mean stress after the initial relaxation, after each
    event and after each increment; after an increment with step >
    ignore_drop_steps a drop (|s_ref| - |s|) / |s_ref| against the entry
    stress_drop_lookback back above stop_on_stress_drop arms the stop, and the
    run ends after stop_post_drop_steps further increments (state
    "drop_step", "stress_history").  eps_target (aqs.py:660-666): a cascade
    batch after which eps_macro[component] >= eps_target ends the run.

    ``fast`` (large grids; small strain, KMC mode, isotropic softening, planar
    catalogs, gaussian barriers): the LS solves run matrix-free
    (ol.spectral_solver_3d_mf, equal to the materialised solver to round-off
    with the same iteration counts), the barriers from the two planar
    parameters (ok.compute_barrier_planar, bit-identical), and the initial
    catalog is drawn in one vectorised call that consumes the stream like the
    per-voxel loop; the event sequence is the standard mode's
    (tests/test_oracle_kmc.py).  ``progress(n_events)`` is called after each
    event.

    ``fast_patching`` ({"enabled", "patch_radius" (3), "sync_interval" (100)},
    mgkmc/aqs.py:196-218, :1021-1054; small strain only): after a KMC event's
    flip, prev, heat and redraw the stress field is predicted by
    oracle/patching.py (macro-unit stress over the waiting time, the patch
    around the voxel, the missing mean) instead of re-solved; every
    sync_interval such flips a full re-solve.  Cascades and elastic
    increments re-solve as without it.  The driven-stress history and the
    returned "sm" are the mean of the current (patched) field."""
    np.random.seed(seed)
    sel = np.random.RandomState(sel_seed)
    bk = dict(mean=2.0, std=0.6, min_cutoff=0.1) if barrier_kwargs is None else dict(barrier_kwargs)
    if barrier_generator == "gaussian":
        gen = lambda shape: ok.gaussian_barriers(shape, np.random, **{"mean": bk.get("mean", 2.0),
                                                                      "std": bk.get("std", 0.6),
                                                                      "min_cutoff": bk.get("min_cutoff", 0.1),
                                                                      "max_cutoff": bk.get("max_cutoff", None)})
    else:  # barrier_generators.py:32-170 on numpy's global stream
        fn = {"rayleigh": ok.rayleigh_barriers, "modified_rayleigh": ok.modified_rayleigh_barriers,
              "modified_rayleigh_with_exponential": ok.modified_rayleigh_exp_barriers}[barrier_generator]
        gen = lambda shape: fn(shape, np.random.mtrand._rand, **bk)
    E = np.asarray(E, float)
    E = E * 1e9 if E.mean() < 1e6 else E
    nu = np.asarray(nu, float)
    volume = pixel ** 3
    tau = 1.0 / (nu0 * np.exp(-q_act_temp / (KB * temperature))) if temperature > 0 else np.inf
    def catalog(m):
        # stz_catalog_glass, or stz_catalog_glass_3d with use_3d_barriers (aqs.py:162-168)
        return (ok.stz_catalog_glass_3d if use_3d_barriers else ok.stz_catalog_glass)(m, gamma0, np.random)

    def redraw(x, y, z, m):
        # the voxel's modes and barriers, or only the flipped mode's (aqs.py:594-606)
        if redraw_directions or redraw_barriers:
            if redraw_directions:
                cat[x, y, z] = catalog(M)
            if redraw_barriers:
                Q0[x, y, z] = gen((M,))
        else:
            cat[x, y, z, m] = catalog(1)[0]
            Q0[x, y, z, m] = gen((1,))[0]

    Q0 = gen((nx, ny, nz, M))
    if fast:
        if (strain_assumption != "small_strain" or cascade_mode or enable_thermal or use_3d_barriers
                or softening_scheme != "isotropic" or barrier_generator != "gaussian"):
            raise ValueError("fast mode covers small-strain KMC runs with isotropic softening and planar catalogs")
        cat = _PlanarCatalog((nx, ny, nz), M)
        # one call: the legacy stream's normals in the per-mode order (gxx, gxy) of stz_catalog_glass
        cat.p[...] = 0.5 * gamma0 * np.random.normal(size=nx * ny * nz * M * 2).reshape(nx, ny, nz, M, 2)
    else:
        cat = np.zeros((nx, ny, nz, M, 3, 3))
        for x in range(nx):
            for y in range(ny):
                for z in range(nz):
                    cat[x, y, z] = catalog(M)
    Q = np.zeros((nx, ny, nz, M))
    soft = np.zeros((nx, ny, nz, 4))
    tl = np.full((nx, ny, nz), -np.inf)
    prev = np.zeros((nx, ny, nz, 3, 3))
    ep = np.zeros((nx, ny, nz, 3, 3)) if eps_plastic is None else np.array(eps_plastic, float)
    eps_macro = np.zeros((3, 3))
    t = 0.0
    scheme = 0 if softening_scheme == "isotropic" else 1
    state = {}

    Tl = np.full((nx, ny, nz), float(temperature))
    if enable_thermal:
        k2, Dsc = heat_operator((nx, ny, nz), pixel, thermal_diffusivity, thermal_coords)

    def heat(dt):
        if enable_thermal:
            heat_step(Tl, k2, Dsc, dt, temperature, thermostat, tau_bath, temperature_cap)

    def flip_heat(x, y, z, C):
        if enable_thermal:
            Tl[x, y, z] += abs(np.sum(state["sig"][x, y, z] * C)) / (rho * Cp)
            if temperature_cap > 0:
                Tl[x, y, z] = min(Tl[x, y, z], temperature_cap)

    Tr = Tl if enable_thermal else temperature
    finite = strain_assumption == "finite_strain"
    if finite:
        from scipy.linalg import expm
        from . import finite_strain as ofs
        I2, I4, I4rt, I4s, II = ofs.identity_tensors(nx, ny, nz)
        Gh = ofs.build_ghat4(nx, ny, nz, nx * pixel, ny * pixel, nz * pixel)
        C4 = ofs.build_C4(E, nu, I4s, II)
        Ff = I2.copy()
        Fp = I2.copy()
        state["F_macro"] = np.eye(3)

    def update():
        if fast:
            e, s, em, sm = ol.spectral_solver_3d_mf(E, nu, eps_macro, eps_plastic=ep, pixel=pixel, tol=ls_tol)
            state["sig"] = s
            state["sm"] = sm
            return
        if not finite:
            e, s, em, sm = ol.update_stress_fft_full(ep, eps_macro, E, nu, pixel=pixel, tol=ls_tol)
            state["sig"] = s
            state["sm"] = sm
            return
        Fb, _, Pt, Pm = ofs.build_bc(component, eps_macro[component], targets, F_bar_initial=state["F_macro"])
        Fo, Po, So, _, Fb = ofs.solver_step(Ff, Fb, Gh, C4, I2, I4, I4rt, Fp=Fp, driving_component=component,
                                            P_target=Pt, P_mask=Pm, E_avg=E.mean(), nu_avg=nu.mean(),
                                            solver=solver, pixel=pixel)
        Ff[...] = Fo
        state["F_macro"] = Fb
        for i in range(3):
            for j in range(3):
                eps_macro[i, j] = Fb[i, j] - 1.0 if i == j else Fb[i, j]
        state["sig"] = np.einsum("ijxyz->xyzij", So)
        state["sm"] = state["sig"].mean(axis=(0, 1, 2))
        if np.any(np.isnan(So)) or np.any(np.isnan(Fo)) or np.max(np.abs(state["sm"])) / 1e9 > 20.0:
            raise ValueError("Simulation stopped due to unstable stress")

    fp = dict(fast_patching or {})
    fp_on = bool(fp.get("enabled", False)) and not finite
    if fp_on:
        from . import patching as opt
        fp_R = int(fp.get("patch_radius", 3))
        fp_sync = int(fp.get("sync_interval", 100))
        # at construction, with the default solver arguments (aqs.py:175, :217)
        fp_K, fp_mm = opt.patch_kernels_3d(E, nu, fp_R, pixel=pixel)
        state["fp_flips"] = 0
        state["fp_unit"] = None

    def patch_step(x, y, z, C, tw):
        # mgkmc/aqs.py:1021-1054
        if state["fp_unit"] is None:
            _, state["fp_unit"], _, _ = ol.update_stress_fft_full(np.zeros_like(ep), unit, E, nu, pixel=pixel,
                                                                  tol=ls_tol)
        if tw > 0:
            state["sig"] += state["fp_unit"] * (strain_rate * tw)
        opt.stamp(state["sig"], (x, y, z), opt.coefficients_3d(C), fp_K, fp_mm, fp_R)
        state["sm"] = state["sig"].mean(axis=(0, 1, 2))
        state["fp_flips"] += 1
        if state["fp_flips"] >= fp_sync:
            update()
            state["fp_flips"] = 0

    def flip_finite(flips):
        # one KMC event or one cascade batch (:512-615, :887-996)
        Cs = [cat[x, y, z, m].copy() for x, y, z, m in flips]
        arrays = (Ff, Fp, soft, tl, prev, cat, Q0, Tl)
        backup = [a.copy() for a in arrays]
        b_sig, b_sm, b_fm, b_em = state["sig"].copy(), state["sm"].copy(), state["F_macro"].copy(), eps_macro.copy()
        n0 = min(max(max(2, int(np.ceil(np.sqrt(0.5 * np.sum(C ** 2)) / 0.005))) for C in Cs), 40)
        for N in range(n0, n0 + 20):
            for a, b in zip(arrays, backup):
                a[...] = b
            state["sig"], state["sm"], state["F_macro"] = b_sig.copy(), b_sm.copy(), b_fm.copy()
            eps_macro[...] = b_em
            try:
                for k in range(1, N + 1):
                  for (x, y, z, m), C in zip(flips, Cs):
                    dF = expm(C / N)
                    Fp[:, :, x, y, z] = np.dot(dF, Fp[:, :, x, y, z])
                    if k == N:
                        sq = ok.vm_sq_3d(C)
                        gp = soft[x, y, z, 0] + jp * sq
                        if softening_cap > 0 and gp > softening_cap:
                            gp = softening_cap
                        soft[x, y, z, 0] = gp
                        soft[x, y, z, 1] = jt * sq
                        tl[x, y, z] = t
                        if jn > 0.0:
                            for dx in (-1, 0, 1):
                                for dy in (-1, 0, 1):
                                    for dz in (-1, 0, 1):
                                        if dx == 0 and dy == 0 and dz == 0:
                                            continue
                                        a, b, c = (x + dx + nx) % nx, (y + dy + ny) % ny, (z + dz + nz) % nz
                                        g = soft[a, b, c, 0] + jn * jp * sq
                                        if softening_cap > 0 and g > softening_cap:
                                            g = softening_cap
                                        soft[a, b, c, 0] = g
                                        soft[a, b, c, 1] += jn * jt * sq
                        prev[x, y, z] = C
                        flip_heat(x, y, z, C)
                        redraw(x, y, z, m)
                  update()
                state.setdefault("substeps", []).append(N)
                return
            except ValueError:
                continue
        raise ValueError("Mechanical solver failed to converge under sub-stepping.")

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
            if finite:
                flip_finite(flips)
                cascades.extend((step, local, int(x), int(y), int(z), int(m)) for x, y, z, m in flips)
            else:
                for x, y, z, m in flips:
                    ok.apply_flip_soa(ep, None, soft, tl, cat, x, y, z, m, t, jp, jt, softening_cap, jn)
                    prev[x, y, z] = cat[x, y, z, m].copy()
                    flip_heat(x, y, z, prev[x, y, z])
                    redraw(x, y, z, m)
                    cascades.append((step, local, int(x), int(y), int(z), int(m)))
                update()
            if eps_target is not None and eps_macro[component] >= eps_target:
                state["cascade_stop"] = "eps_target"
                return True
            local += 1
            if local > int(max_cascade_steps_pct * nx * ny * nz):
                state["cascade_stop"] = "limit"
                return True
        return False

    def result():
        out = dict(eps_plastic=ep, Q=Q, Q0=Q0, soft=soft, sig=state["sig"], eps_macro=eps_macro, time=t,
                   cascades=cascades, substeps=state.get("substeps"), Tlocal=Tl,
                   stress_history=state.get("stress_history"), drop_step=state.get("drop_step"),
                   last_step=state.get("last_step"), cascade_stop=state.get("cascade_stop"))
        if finite:
            out.update(F=Ff, F_plastic=Fp, F_macro=state["F_macro"])
        return events, out

    targets = dict(stress_targets or {})
    if not targets:
        if component == (0, 0):
            targets = {(1, 1): 0.0, (2, 2): 0.0}
        elif component == (0, 1):
            targets = {(0, 0): 0.0, (1, 1): 0.0, (2, 2): 0.0}
    unit = np.zeros((3, 3))
    unit[component] = 1.0
    events = []
    sdc = component if stress_drop_component is None else tuple(stress_drop_component)
    if cascade_mode:
        if run_cascade(0):
            return result()
    else:
        update()
    hist = [state["sm"][sdc]]
    state["stress_history"] = hist
    state["drop_step"] = None
    countdown = stop_post_drop_steps
    dt_step = abs(step_size) / strain_rate if strain_rate > 0 else 1.0
    remaining = 0.0
    done = 0
    step = 1
    E_avg, nu_avg = np.mean(E), np.mean(nu)
    while done < n_global_steps:
        remaining += dt_step
        while remaining > 0:
            if fast:
                ok.compute_barrier_planar(Q, Q0, state["sig"], cat.p[..., 0], cat.p[..., 1], volume, soft, tl, t,
                                          softening_cap, tau)
            else:
                ok.compute_barrier(Q, Q0, state["sig"], cat, volume, soft, tl, t, prev, softening_cap, scheme, tau)
            if cascade_mode and len(ok.find_unstable(Q, stability_threshold)) > 0:
                if run_cascade(step):
                    return result()
                continue
            rates, idx, total = ok.compute_rates(Q, 1.0, Tr, nu0, cascade_mode=cascade_mode)
            if total > 0:
                u = np.random.uniform()
                tw = -np.log(u) / total
                if tw < remaining:
                    t += tw
                    heat(tw)
                    eps_macro += unit * (strain_rate * tw)
                    remaining -= tw
                    i = ok.select_event(rates, total, total * sel.random_sample())
                    x, y, z, m = ok.decode_index(idx[i], ny, nz, M)
                    inst = Q[x, y, z, m] <= stability_threshold
                    if finite:
                        flip_finite([(x, y, z, m)])
                    else:
                        C = cat[x, y, z, m].copy()
                        ok.apply_flip_soa(ep, None, soft, tl, cat, x, y, z, m, t, jp, jt, softening_cap, jn)
                        prev[x, y, z] = C
                        flip_heat(x, y, z, C)
                        redraw(x, y, z, m)
                        if fp_on:
                            patch_step(x, y, z, C, tw)
                        else:
                            update()
                    events.append((int(x), int(y), int(z), int(m), t, bool(inst)))
                    hist.append(state["sm"][sdc])
                    if progress is not None:
                        progress(len(events))
                    if len(events) >= n_events:
                        return result()
                    step += 1
                    continue
            eps_macro += unit * (strain_rate * remaining)
            heat(remaining)
            t += remaining
            remaining = 0.0
        update()
        if finite:
            # aqs.py:1089-1102: a cascade, or a second finite-strain step from the converged state
            if cascade_mode:
                if run_cascade(step):
                    return result()
            else:
                update()
        for _ in range(0 if finite else mixed_max_iter):
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
        cur = state["sm"][sdc]
        hist.append(cur)
        if stop_on_stress_drop is not None and state["drop_step"] is None and step > ignore_drop_steps:
            lb = max(1, stress_drop_lookback)
            if len(hist) >= lb:
                ref = hist[-lb]
                drop = (abs(ref) - abs(cur)) / abs(ref) if abs(ref) > 1e-6 else 0.0
                if drop > stop_on_stress_drop:
                    state["drop_step"] = step
        if state["drop_step"] is not None:
            if countdown > 0:
                countdown -= 1
            else:
                break
        step += 1
    state["last_step"] = step
    return result()
