"""Device-resident 3D STZ/KMC simulation (drop-in for mgkmc/aqs.py ThermalSimulation).

Same constructor and run_simulation signatures as the reference
(mgkmc/aqs.py:16-53, :682-705); the event loop itself is
``mgkmc_amd.kmc_engine.DeviceKmc`` (per-event work on the GPU, random streams
on the host in the reference's order).  Extensions (keyword-only in use):
``ls_precision`` ("fp64", "fp32", "mixed" = FP32 + FP64 refinement),
``ls_warm_start``, ``rng`` / ``selection_stream`` (isolated random streams),
``init_modes=("device", seed)`` (initial modes drawn on the GPU).
"""
import numpy as np

from .kmc_engine import (DeviceKmc, LS_PRECISIONS, gaussian_generator, planar_catalog,  # noqa: F401
                         planar_to_full, KB)


class ThermalSimulation(DeviceKmc):
    def __init__(self, nx, ny, nz, M, gamma0, E_field, nu_field, pixel=1.0, barrier_generator=None,
                 barrier_kwargs={}, softening_scheme="isotropic", softening_cap=2.0, jp=10.0, jt=30.0,
                 neighbor_softening_fraction=0.0, output_dir="output", temperature=0.0, strain_rate=1.0,
                 stability_threshold=0.0, redraw_directions=True, redraw_barriers=True,
                 max_cascade_steps_pct=0.3, nu0=1e13, q_act_temp=0.37, cascade_mode=False,
                 scale_rate_by_volume=False, fast_patching=None, enable_thermal=False, Cp=420.0, rho=6125.0,
                 thermal_diffusivity=3.0e-6, thermal_coords="pixel", temperature_cap=1000.0, thermostat=False,
                 tau_bath=0.0, strain_assumption="small_strain", use_3d_barriers=False, hyperelastic_model="svk",
                 A_m=0.0, B_m=0.0, C_m=0.0, solver="al", v1=0.0, v2=0.0, v3=0.0, g1=0.0, g2=0.0, g3=0.0, g4=0.0,
                 strain_capping_enabled=False, strain_capping_limit=None, strain_capping_tangent_ratio=0.1,
                 strain_capping_type="piecewise", strain_capping_smooth_power=1.0,
                 ls_precision="fp64", ls_warm_start=False, rng=None, selection_stream=None, init_modes=None):
        # finite strain: the Newton-Krylov, augmented-Lagrangian or DBFFT step, KMC mode
        fs_solvers = ("newton_cg", "al", "dbfft")
        self._finite_ok = strain_assumption == "finite_strain" and solver in fs_solvers and not cascade_mode
        extra = ["use_3d_barriers"] if use_3d_barriers else []
        if strain_assumption == "finite_strain" and solver not in fs_solvers:
            extra.append(f"finite-strain solver={solver!r} (the device step is 'newton_cg', 'al' or 'dbfft')")
        self._unsupported(cascade_mode, fast_patching, enable_thermal, strain_assumption, hyperelastic_model, extra)
        self.output_dir = output_dir
        self.strain_assumption = strain_assumption
        self.hyperelastic_model = hyperelastic_model
        self.v1, self.v2, self.v3, self.g1, self.g2, self.g3, self.g4 = v1, v2, v3, g1, g2, g3, g4
        cap = dict(strain_capping_enabled=strain_capping_enabled, strain_capping_limit=strain_capping_limit,
                   strain_capping_tangent_ratio=strain_capping_tangent_ratio,
                   strain_capping_type=strain_capping_type, strain_capping_smooth_power=strain_capping_smooth_power)
        for key, val in cap.items():
            setattr(self, key, val)
        # small strain: only "landau" leaves the linear solve (mgkmc/aqs.py:330-352)
        law = ("landau", (v1, v2, v3, g1, g2, g3, g4), cap) if hyperelastic_model == "landau" else None
        finite = None
        self.solver, self.A_m, self.B_m, self.C_m = solver, A_m, B_m, C_m
        if strain_assumption == "finite_strain":
            law = None
            finite = dict(model_type=hyperelastic_model, A_m=A_m, B_m=B_m, C_m=C_m, v1=v1, v2=v2, v3=v3, g1=g1,
                          g2=g2, g3=g3, g4=g4, solver=solver, **cap)
        self._setup((nx, ny, nz), M, gamma0, E_field, nu_field, pixel, barrier_generator, barrier_kwargs,
                    softening_scheme, softening_cap, jp, jt, neighbor_softening_fraction, temperature, strain_rate,
                    stability_threshold, redraw_directions, redraw_barriers, nu0, q_act_temp, cascade_mode,
                    scale_rate_by_volume, "plane_strain", ls_precision, ls_warm_start, rng, selection_stream,
                    init_modes, dict(mean=2.0, std=0.6), law=law, finite=finite)
        self.max_cascade_steps_pct = max_cascade_steps_pct
        if enable_thermal:
            self._setup_thermal(Cp, rho, thermal_diffusivity, thermal_coords, temperature_cap, thermostat, tau_bath)

    def update_stresses(self):
        """LS re-equilibration with the current eps_plastic and eps_macro
        (mgkmc/aqs.py:330-351: update_stress_fft_full, cold start, or the
        Landau solver for hyperelastic_model="landau")."""
        return self._solve(self.eps_macro)

    def mixed_bc(self, stress_targets, component, mixed_tol, mixed_max_iter, step=0):
        """Elastic relaxation loop with mixed BCs (mgkmc/aqs.py:1103-1142): in
        cascade_mode each iteration first runs a cascade (None is returned
        when it is truncated); the corrections re-solve with the linear LS
        (update_stress_fft_full) whatever the law, as the reference does."""
        E_avg, nu_avg = np.mean(self.E), np.mean(self.nu)
        converged = False
        for _ in range(mixed_max_iter):
            if self.cascade_mode and self._run_cascade(step)[2]:
                return None
            sig_curr = self.sig_mean()
            err = np.zeros((3, 3))
            max_err = 0.0
            for idx_t, target in stress_targets.items():
                e = target - sig_curr[idx_t]
                err[idx_t] = e
                max_err = max(max_err, abs(e))
            if max_err < mixed_tol:
                converged = True
                break
            corr = (err - nu_avg * np.trace(err) * np.eye(3)) / E_avg
            corr[component] = 0.0
            self.eps_macro += corr
            self._solve_linear(self.eps_macro)
        return converged

    def run_simulation(self, n_global_steps, step_size, component=(0, 1), stress_targets={}, mixed_tol=1e-4,
                       mixed_max_iter=50, max_kmc_steps_pct=0.3, max_cascade_steps_pct=0.3, max_events=None,
                       **kwargs):
        """KMC + elastic increments (mgkmc/aqs.py:682-1150).  Logging,
        checkpoint, VTK and stress-drop options of the reference are accepted
        and ignored.  In cascade_mode the initial relaxation is a cascade (no
        solve before it) and every inner pass first runs a cascade while any
        mode is unstable; a truncated cascade ends the run.  Cascade flips go
        to ``cascade_history``.  ``max_events`` (extension) stops after that
        many KMC events."""
        if max_cascade_steps_pct is not None:
            self.max_cascade_steps_pct = max_cascade_steps_pct
        targets = dict(stress_targets)
        self.driving_component = component
        if not targets:
            if component == (0, 0):
                targets = {(1, 1): 0.0, (2, 2): 0.0}
            elif component == (0, 1):
                targets = {(0, 0): 0.0, (1, 1): 0.0, (2, 2): 0.0}
        self.stress_targets = targets
        unit = np.zeros((3, 3))
        unit[component] = 1.0
        if self.cascade_mode:
            if self._run_cascade(0)[2]:
                return self.flip_event_history
        else:
            self.update_stresses()
        dt_step = abs(step_size) / self.strain_rate if self.strain_rate > 0 else 1.0
        max_seq = int(max_kmc_steps_pct * self.total_voxels)
        seq = 0
        remaining = 0.0
        done = 0
        events = 0
        step = 1  # the reference's global step: +1 per KMC event and per elastic increment
        while done < n_global_steps:
            remaining += dt_step
            while remaining > 0:
                if self.cascade_mode and self._cascade_pairs():
                    if self._run_cascade(step)[2]:
                        return self.flip_event_history
                    continue
                # one event at a time in cascade mode (a cascade check follows each)
                limit = 1 if self.cascade_mode else (None if max_events is None else max_events - events)
                n, remaining, reason, seq = self.kmc_events(remaining, unit, limit, max_seq, seq)
                events += n
                step += n
                if reason == "seq" or (max_events is not None and events >= max_events):
                    return self.flip_event_history
                if reason == "time":
                    break
            self.update_stresses()
            # finite strain: the step itself enforces the stress targets (mgkmc/aqs.py:1089-1102)
            if not self.finite and self.mixed_bc(targets, component, mixed_tol, mixed_max_iter, step) is None:
                return self.flip_event_history
            done += 1
            step += 1
        return self.flip_event_history
