"""Device-resident 2D STZ/KMC simulation (drop-in for mgkmc/kmc_simulator.py
KmcSimulation2D).

Same constructor and run_simulation signatures as the reference
(mgkmc/kmc_simulator.py:14-35, :796-803).  The per-event work of its KMC loop
(:871-1101, small strain) runs on the GPU through
``mgkmc_amd.kmc_engine.DeviceKmc``; elastic_run (:387-525, the small-strain
branches: linear, secant degradation, Rose, Landau) is the device LS solve of
that law with the reference's plane mode and its unstable-stress check, and
flips under a nonlinear law are sub-stepped on solver failure (:263-385); the mixed-BC loop (:1106-1143) keeps the reference's
compliance-then-secant update on the host.  strain_assumption="finite_strain"
(:156-168, :389-445, :895-1000) runs the 2D Newton-CG step (solver="newton_cg")
with F_p updated by expm(C/N) sub-steps.  Extensions as in
``mgkmc_amd.aqs.ThermalSimulation``.
"""
import os

import numpy as np

from . import _lib
from .kmc_engine import DeviceKmc


class KmcSimulation2D(DeviceKmc):
    def __init__(self, nx, ny, M, gamma0, E_field, nu_field, pixel=1.0, barrier_generator=None, barrier_kwargs={},
                 softening_scheme="isotropic", softening_cap=2.0, jp=10.0, jt=30.0, neighbor_softening_fraction=0.0,
                 q_act_temp=0.37, output_dir="output", temperature=0.0, strain_rate=1.0, stability_threshold=0.0,
                 nu0=1e13, plane_mode="plane_strain", fast_patching=None, cascade_mode=False,
                 scale_rate_by_volume=False, redraw_directions=True, redraw_barriers=True, enable_thermal=False,
                 Cp=420.0, rho=6125.0, thermal_diffusivity=3.0e-6, thermal_coords="pixel", temperature_cap=1000.0,
                 thermostat=False, tau_bath=0.0, strain_assumption="small_strain", hyperelastic_model="svk",
                 A_m=0.0, B_m=0.0, C_m=0.0, solver="al", stz_mode="pure_shear", d=0.0, k=0.0, eta=5.0,
                 mu_floor_fraction=0.1, v1=0.0, v2=0.0, v3=0.0, g1=0.0, g2=0.0, g3=0.0, g4=0.0,
                 strain_capping_enabled=False, strain_capping_limit=None, strain_capping_tangent_ratio=0.1,
                 strain_capping_type="piecewise", strain_capping_smooth_power=1.0,
                 ls_precision="fp64", ls_warm_start=False, rng=None, selection_stream=None, init_modes=None):
        # finite strain: the Newton-CG step of the plane problem, KMC or cascade mode
        self._finite_ok = strain_assumption == "finite_strain" and solver in ("newton_cg", "dbfft", "al")
        # simple-shear modes are non-symmetric: finite strain only (expm(C/N) F_p)
        extra = [f"stz_mode={stz_mode} with small strain"] if stz_mode != "pure_shear" and \
            strain_assumption != "finite_strain" else []
        self.stz_mode = stz_mode
        if strain_assumption == "finite_strain" and solver not in ("newton_cg", "dbfft", "al"):
            extra.append(f"2D finite-strain solver={solver!r} (the device steps are 'newton_cg', 'al', 'dbfft')")
        self._unsupported(cascade_mode, fast_patching, enable_thermal, strain_assumption, hyperelastic_model, extra)
        self.nx, self.ny = nx, ny
        self.output_dir = output_dir
        os.makedirs(output_dir, exist_ok=True)  # as the reference's constructor
        self.strain_assumption = strain_assumption
        self.hyperelastic_model = hyperelastic_model
        self.stz_mode = stz_mode
        # the reference's names: d, k (secant degradation), eta, mu_floor_fraction (Rose), v*, g* (Landau)
        self.d, self.k, self.eta, self.mu_floor_fraction = d, k, eta, mu_floor_fraction
        self.v1, self.v2, self.v3, self.g1, self.g2, self.g3, self.g4 = v1, v2, v3, g1, g2, g3, g4
        cap = dict(strain_capping_enabled=strain_capping_enabled, strain_capping_limit=strain_capping_limit,
                   strain_capping_tangent_ratio=strain_capping_tangent_ratio,
                   strain_capping_type=strain_capping_type, strain_capping_smooth_power=strain_capping_smooth_power)
        for key, val in cap.items():
            setattr(self, key, val)
        law = {"secant_degradation": ("secant", d, k), "rose": ("rose", eta, mu_floor_fraction),
               "landau": ("landau", (v1, v2, v3, g1, g2, g3, g4), cap)}.get(hyperelastic_model)
        finite = None
        self.solver, self.A_m, self.B_m, self.C_m = solver, A_m, B_m, C_m
        if strain_assumption == "finite_strain":
            law = None
            finite = dict(model_type=hyperelastic_model, A_m=A_m, B_m=B_m, C_m=C_m, v1=v1, v2=v2, v3=v3, g1=g1,
                          g2=g2, g3=g3, g4=g4, solver=solver, **cap)
        self._setup((nx, ny), M, gamma0, E_field, nu_field, pixel, barrier_generator, barrier_kwargs,
                    softening_scheme, softening_cap, jp, jt, neighbor_softening_fraction, temperature, strain_rate,
                    stability_threshold, redraw_directions, redraw_barriers, nu0, q_act_temp, cascade_mode,
                    scale_rate_by_volume, plane_mode, ls_precision, ls_warm_start, rng, selection_stream,
                    init_modes, dict(mean=2.0, std=0.6, min_cutoff=0.1), law=law, finite=finite)
        if enable_thermal:
            self._setup_thermal(Cp, rho, thermal_diffusivity, thermal_coords, temperature_cap, thermostat, tau_bath)
        # fast patching (mgkmc/kmc_simulator.py:76-78, 154-171): off for finite strain and the nonlinear laws
        self._setup_fast_patching(fast_patching, 5, strain_assumption == "finite_strain" or
                                  hyperelastic_model in ("secant_degradation", "landau", "murnaghan", "rose"))

    def elastic_run(self, eps_macro):
        """Small-strain equilibrium (mgkmc/kmc_simulator.py:449-525): LS solve
        of the linear or nonlinear law, then ValueError on a NaN stress or
        |<sigma>| > 20 GPa.  Returns <sigma> (2, 2)."""
        inf = self._solve(eps_macro)
        sig_mean = self.sig_mean()
        if inf.status == _lib.MGK_ERR_NONFINITE or not np.all(np.isfinite(sig_mean)) or \
                np.max(np.abs(sig_mean)) / 1e9 > 20.0:
            raise ValueError(f"Unstable stress in small-strain elastic_run: {sig_mean / 1e9} GPa")
        return sig_mean

    def update_stresses(self):
        return self.elastic_run(self.eps_macro)

    def mixed_bc(self, stress_targets, mixed_tol, mixed_max_iter):
        """Macro loop after an increment (mgkmc/kmc_simulator.py:1106-1143):
        solve, stop if every targeted stress is within mixed_tol, else the
        Poisson-corrected compliance step, replaced per component by the secant
        slope once two iterates exist (and the slope exceeds 0.01 E_avg)."""
        if self._native_mixed_ok():
            self._run_mixed_native(stress_targets, (0, 0), mixed_tol, mixed_max_iter)
            return
        E_avg = self.E_field.mean()
        nu_avg = self.nu_field.mean()
        prev_eps, prev_sig = {}, {}
        for _ in range(mixed_max_iter):
            sigM = self.elastic_run(self.eps_macro)
            err = np.zeros((2, 2))
            err_max = 0.0
            for idx, target in stress_targets.items():
                if idx[0] < 2 and idx[1] < 2:
                    e = target - sigM[idx]
                    err[idx] = e
                    err_max = max(err_max, abs(e))
            if err_max < mixed_tol:
                break
            d_eps = (err - nu_avg * np.trace(err) * np.eye(2)) / E_avg
            for idx in stress_targets.keys():
                if idx[0] < 2 and idx[1] < 2:
                    step = d_eps[idx]
                    if idx in prev_eps:
                        d_e = self.eps_macro[idx] - prev_eps[idx]
                        d_s = sigM[idx] - prev_sig[idx]
                        if abs(d_e) > 1e-14 and np.isfinite(d_s):
                            slope = d_s / d_e
                            if np.isfinite(slope) and slope > 0.01 * E_avg:
                                step = err[idx] / slope
                    prev_eps[idx] = self.eps_macro[idx]
                    prev_sig[idx] = sigM[idx]
                    self.eps_macro[idx] += step

    def run_simulation(self, n_global_steps, step_size, component=(0, 0), stress_targets={}, mixed_tol=1e-4,
                       mixed_max_iter=50, checkpoint_interval="none", checkpoint_path="checkpoint",
                       vtk_interval="none", vtk_elastic_only=True, track_cascades=False, enable_console_log=True,
                       summary_filename="summary_log.txt", enable_summary_log=True, enable_global_log=True,
                       max_kmc_steps_pct=0.3, max_cascade_steps_pct=0.3, enable_kmc_log=False, max_events=None,
                       **kwargs):
        """KMC + elastic increments (mgkmc/kmc_simulator.py:796-1160), the
        reference's parameters in its order; ``max_events`` (extension, last)
        stops after that many KMC events.  Logs as the reference writes them:
        summary (summary_filename, enable_summary_log), global_log.txt
        (enable_global_log), flipped_voxels_log.txt (enable_kmc_log: one line
        per flipped mode, its softening and barrier at selection), console
        lines (enable_console_log).  checkpoint_interval / checkpoint_path and
        track_cascades are accepted without effect, as in the reference's 2D
        loop; VTK output (vtk_interval other than "none") raises
        NotImplementedError.  In cascade_mode every inner pass first runs a
        cascade while any mode is unstable; the run ends when more than
        int(max_cascade_steps_pct nx ny) consecutive cascades flipped
        something.  Cascade flips go to ``cascade_history``.  Returns
        kmc_event_history (x, y, m, time, is_instab per KMC event)."""
        from .runlog import RunLog, FlipLog
        if vtk_interval not in (None, "none"):
            raise NotImplementedError("VTK export (mgkmc/analysis/vtk.py) is not part of the device build; "
                                      "pass vtk_interval='none'")
        self.driving_component = component
        self.stress_targets = stress_targets
        log = RunLog(self.output_dir, 2, summary_filename, enable_summary_log, enable_global_log,
                     enable_console_log=enable_console_log, thermal=self.enable_thermal,
                     t_avg=lambda: float(np.mean(self.Tlocal)), plane_mode=self.plane_mode,
                     nu_avg=float(self.nu_field.mean()))
        self._flip_log = FlipLog(self.output_dir) if enable_kmc_log else None
        c = dict(elastic=0, kmc=0, step=1, events=0, casc=0)

        def do_log(step, kind, cid, flips, eps=None, sig=None):
            eps = self.eps_macro.copy() if eps is None else eps
            sig = self.sig_mean() if sig is None else sig
            log.log_global(step, c["elastic"], c["kmc"], self.time, eps, sig, cid, flips)
            log.summary(step, kind, eps[component], sig[component], c["kmc"], cid, flips, self.time,
                        first_2d=step == 0)

        try:
            unit = np.zeros((2, 2))
            unit[component] = 1.0
            self.elastic_run(self.eps_macro)
            do_log(0, "INIT", 0, 0)
            dt_step = abs(step_size) / self.strain_rate if self.strain_rate > 0 else 1.0
            max_seq = int(max_kmc_steps_pct * self.nx * self.ny)
            max_casc = int(max_cascade_steps_pct * self.nx * self.ny)
            seq = 0
            remaining = 0.0

            def on_event(ev, eps, sig, over):
                # kmc_simulator.py:1081-1092: unstable-sequence stop before the log line
                if over:
                    print(f"\n[TERMINATE] KMC instability sequence limit reached! {max_seq + 1} steps.")
                    return
                c["kmc"] += 1
                c["events"] += 1
                do_log(c["step"], "KMC_INSTAB" if ev[-1] else "KMC", c["casc"] if self.cascade_mode else 0, 1,
                       eps, sig)
                c["step"] += 1
                self._flip_log_step = c["step"]

            self._flip_log_step = c["step"]
            while c["elastic"] < n_global_steps:
                remaining += dt_step
                while remaining > 0:
                    if self.cascade_mode:
                        if self._cascade_pairs():
                            if self._run_cascade(c["step"])[1] > 0:
                                c["casc"] += 1
                            if c["casc"] > max_casc:
                                print(f"\n[TERMINATE] Cascade instability sequence limit reached! "
                                      f"{c['casc']} batches.")
                                return self.kmc_event_history
                            continue
                        c["casc"] = 0
                    # one event at a time in cascade mode (a cascade check follows each)
                    limit = 1 if self.cascade_mode else (None if max_events is None else max_events - c["events"])
                    n, remaining, reason, seq = self.kmc_events(remaining, unit, limit, max_seq, seq, on_event)
                    if reason == "seq" or (max_events is not None and c["events"] >= max_events):
                        return self.kmc_event_history
                    if reason == "time":
                        break
                if self.finite:  # the step enforces the stress targets itself (kmc_simulator.py:1103-1104)
                    self.elastic_run(self.eps_macro)
                else:
                    self.mixed_bc(stress_targets, mixed_tol, mixed_max_iter)
                c["elastic"] += 1
                self.increment_history.append((self.eps_macro.copy(), self.sig_mean().copy()))
                do_log(c["step"], "ELAST", c["casc"], 0)
                c["step"] += 1
                self._flip_log_step = c["step"]
            log.finish()
        finally:
            log.close()
            self._flip_log = None
        return self.kmc_event_history
