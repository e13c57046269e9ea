"""Device-resident 3D STZ/KMC simulation (drop-in for mgkmc/aqs.py ThermalSimulation).

Same constructor and run_simulation signatures as the reference
(mgkmc/aqs.py:16-53, :682-705).  The per-event work of the KMC loop
(mgkmc/aqs.py:848-1087) stays on the GPU: fused barrier + rate pass, FP64
rate total, two-level event selection, flip, mode redraw upload, and the LS
re-equilibration.  The host keeps the scalar state (time, macro strain,
remaining time) and the random streams:

  * numpy's global legacy stream (np.random), exactly as the reference uses
    it: initial Q0 (generator), initial catalog (2 normals per mode, voxel
    order), then per event one uniform (waiting time), 2M normals (new modes)
    and M normals (new Q0);
  * the selection stream ``mgkmc_amd.stz.kmc.SELECTION_STREAM``
    (r = total * u), the reference's numba-internal draw made seedable.

Supported: small strain, linear elasticity, KMC mode (cascade_mode=False),
isotropic or directional softening, redraw of directions and barriers,
planar catalogs.  Options the device path does not cover raise
NotImplementedError instead of falling back to a CPU path.
"""
import math

import numpy as np

from . import device as dev
from .kmc_device import KmcState, params as kmc_params, SIG_SYM64, SIG_SYM32
from .linear_elastic_simulator import LsSolver, FP64, FP32, MIXED
from .stz import kmc as _kmcmod

KB = 8.617e-5
# "mixed": FP32 iterations, then FP64 refinement to the FP64 stop test
LS_PRECISIONS = {"fp64": FP64, "fp32": FP32, "mixed": MIXED}


def gaussian_generator(mean=2.0, std=0.6, min_cutoff=0.1, max_cutoff=None):
    """np.random.normal + clip (mgkmc/stz/barrier_generators.py:4-30)."""
    def gen(shape):
        b = np.random.normal(loc=mean, scale=std, size=shape)
        if min_cutoff is not None or max_cutoff is not None:
            b = np.clip(b, a_min=min_cutoff, a_max=max_cutoff)
        return b
    return gen


def planar_catalog(n_modes_total, gamma0):
    """n planar modes drawn as the reference's per-mode loop does
    (gxx then gxy, mgkmc/stz/catalog.py:7-40): returns (n, 2) (gxx, gxy).
    One vectorised call consumes the legacy stream exactly like the loop."""
    g = np.random.normal(size=2 * n_modes_total).reshape(n_modes_total, 2)
    return 0.5 * gamma0 * g


def planar_to_full(p, d=3):
    out = np.zeros(p.shape[:-1] + (d, d))
    out[..., 0, 0] = p[..., 0]
    out[..., 1, 1] = -p[..., 0]
    out[..., 0, 1] = p[..., 1]
    out[..., 1, 0] = p[..., 1]
    return out


class ThermalSimulation:
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
                 ls_precision="fp64", ls_warm_start=False):
        unsupported = []
        if cascade_mode:
            unsupported.append("cascade_mode")
        if fast_patching and fast_patching.get("enabled", False):
            unsupported.append("fast_patching")
        if enable_thermal:
            unsupported.append("enable_thermal")
        if strain_assumption != "small_strain":
            unsupported.append("finite_strain KMC")
        if hyperelastic_model in ("landau", "secant_degradation", "rose"):
            unsupported.append(f"hyperelastic_model={hyperelastic_model}")
        if use_3d_barriers:
            unsupported.append("use_3d_barriers")
        if unsupported:
            raise NotImplementedError("device KMC path does not cover: " + ", ".join(unsupported))
        self.output_dir = output_dir
        self.M, self.gamma0 = M, gamma0
        self.volume = pixel ** 3
        self.pixel = pixel
        self.softening_scheme = softening_scheme
        self.softening_cap = softening_cap
        self.jp, self.jt = jp, jt
        self.neighbor_softening_fraction = neighbor_softening_fraction
        self.temperature = temperature
        self.strain_rate = strain_rate
        self.stability_threshold = stability_threshold
        self.redraw_directions, self.redraw_barriers = redraw_directions, redraw_barriers
        self.nu0 = nu0
        self.q_act_temp = q_act_temp
        self.cascade_mode = cascade_mode
        self.scale_rate_by_volume = scale_rate_by_volume
        if self.temperature > 0:
            self.tau = 1.0 / (self.nu0 * np.exp(-self.q_act_temp / (KB * self.temperature)))
        else:
            self.tau = np.inf
        self.grid_shape = (nx, ny, nz)
        self.total_voxels = nx * ny * nz
        E = np.asarray(E_field, dtype=np.float64)
        self.E_field = E * 1e9 if E.mean() < 1e6 else E
        self.nu_field = np.asarray(nu_field, dtype=np.float64)
        self.E, self.nu = self.E_field, self.nu_field
        if barrier_generator is None:
            self.barrier_generator = gaussian_generator(mean=2.0, std=0.6)
        elif isinstance(barrier_generator, str):
            if barrier_generator.lower() != "gaussian":
                raise NotImplementedError(f"barrier generator {barrier_generator!r}")
            self.barrier_generator = gaussian_generator(**barrier_kwargs)
        else:
            self.barrier_generator = barrier_generator
        # initial draws in the reference's order: Q0, then the catalog
        Q0 = self.barrier_generator((nx, ny, nz, M))
        cat = planar_to_full(planar_catalog(self.total_voxels * M, gamma0)).reshape(nx, ny, nz, M, 3, 3)
        if ls_precision not in LS_PRECISIONS:
            raise ValueError(f"ls_precision must be one of {sorted(LS_PRECISIONS)}")
        self.precision = LS_PRECISIONS[ls_precision]
        self.sig_kind = SIG_SYM32 if self.precision == FP32 else SIG_SYM64
        self.ls = LsSolver(self.E_field, self.nu_field, precision=self.precision)
        self.kmc = KmcState((nx, ny, nz), M, Q0, cat, directional=(softening_scheme == "directional"),
                            eps_precision=FP32 if self.precision == FP32 else FP64,
                            eps_shadow32=self.precision == MIXED)
        self.ls_warm_start = ls_warm_start
        self._solved_once = False
        self.eps_macro = np.zeros((3, 3))
        self.time = 0.0
        self.solver_args = {}
        self.flip_event_history = []
        self.strain_assumption = strain_assumption
        self.ls_info = None

    # ----- state mirrors (synchronised on access) --------------------------
    @property
    def sig_field(self):
        return dev.from_sym_soa(dev.download(self.ls.sig), 3)

    @property
    def eps_field(self):
        return dev.from_sym_soa(dev.download(self.ls.eps), 3)

    @property
    def eps_plastic(self):
        return dev.from_sym_soa(dev.download(self.kmc.eps_p), 3)

    @property
    def Q(self):
        """Barriers of the current state (evaluated on access; the event loop
        does not store them)."""
        from .kmc_device import modes_from_device
        self.kmc.barrier_rates(self.ls.sig, self.sig_kind, self._params(), write_q=True)
        return modes_from_device(self.kmc.Q, self.grid_shape)

    @property
    def Q0(self):
        from .kmc_device import modes_from_device
        return modes_from_device(self.kmc.Q0, self.grid_shape)

    @property
    def soft_prop(self):
        s = dev.download(self.kmc.soft)
        out = np.zeros(self.grid_shape + (4,))
        out[..., 0] = s[0].reshape(self.grid_shape)
        out[..., 1] = s[1].reshape(self.grid_shape)
        return out

    # ----- seams -----------------------------------------------------------
    def update_stresses(self):
        """LS re-equilibration with the current eps_plastic and eps_macro
        (mgkmc/aqs.py:345-351: update_stress_fft_full, cold start)."""
        warm = self.ls_warm_start and self._solved_once
        args = dict(max_iter=200, tol=1e-6)
        args.update(self.solver_args)
        self.ls_info = self.ls.solve(self.eps_macro, self.kmc.eps_p, warm=warm,
                                     max_iter=args["max_iter"], tol=args["tol"], eps_plastic32=self.kmc.eps_p32)
        self._solved_once = True
        return self.ls_info

    def sig_mean(self):
        return np.array(self.ls_info.sig_mean[:9]).reshape(3, 3)

    def _params(self):
        eff_volume = self.volume if self.scale_rate_by_volume else 1.0
        return kmc_params(volume=self.volume, rate_volume=eff_volume, nu0=self.nu0, temperature=self.temperature,
                          cap=self.softening_cap, tau=self.tau, time=self.time,
                          scheme=1 if self.softening_scheme == "directional" else 0, cascade=self.cascade_mode)

    def kmc_event(self, remaining_time, strain_unit_tensor):
        """One pass of the inner KMC loop body (mgkmc/aqs.py:852-1087).
        Returns (event or None, remaining_time)."""
        p = self._params()
        # barriers are not stored: the selection recomputes the chosen voxel's
        total = self.kmc.barrier_rates(self.ls.sig, self.sig_kind, p, write_q=False)
        if total > 0:
            u = np.random.uniform()
            t_wait = -np.log(u) / total
            if t_wait < remaining_time:
                self.time += t_wait
                self.eps_macro += strain_unit_tensor * (self.strain_rate * t_wait)
                remaining_time -= t_wait
                us = _kmcmod.SELECTION_STREAM.random_sample()
                v, m = self.kmc.select(us, self.ls.sig, self.sig_kind, p)
                x, y, z = (int(c) for c in self.kmc.voxel_coords(v))
                is_instab = self.kmc.sel.q <= self.stability_threshold
                self.kmc.flip(v, m, self.time, self.jp, self.jt, self.softening_cap, self.neighbor_softening_fraction)
                newcat = newq0 = None
                if self.redraw_directions or self.redraw_barriers:
                    if self.redraw_directions:
                        newcat = planar_to_full(planar_catalog(self.M, self.gamma0))
                    if self.redraw_barriers:
                        newq0 = self.barrier_generator((self.M,))
                    self.kmc.set_voxel(v, newcat, newq0)
                else:
                    raise NotImplementedError("single-mode redraw")
                self.update_stresses()
                ev = (x, y, z, m, self.time, is_instab)
                self.flip_event_history.append(ev)
                return ev, remaining_time
            self.eps_macro += strain_unit_tensor * (self.strain_rate * remaining_time)
            self.time += remaining_time
            return None, 0.0
        self.eps_macro += strain_unit_tensor * (self.strain_rate * remaining_time)
        self.time += remaining_time
        return None, 0.0

    def mixed_bc(self, stress_targets, component, mixed_tol, mixed_max_iter):
        """Elastic relaxation loop with mixed BCs (mgkmc/aqs.py:1107-1142)."""
        E_avg, nu_avg = np.mean(self.E), np.mean(self.nu)
        converged = False
        for _ in range(mixed_max_iter):
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
            self.update_stresses()
        return converged

    def run_simulation(self, n_global_steps, step_size, component=(0, 1), stress_targets={}, mixed_tol=1e-4,
                       mixed_max_iter=50, max_kmc_steps_pct=0.3, max_events=None, **kwargs):
        """KMC + elastic increments (mgkmc/aqs.py:682-1150, KMC mode).  Logging,
        checkpoint, VTK and stress-drop options of the reference are accepted
        and ignored.  ``max_events`` (extension) stops after that many events."""
        targets = dict(stress_targets)
        if not targets:
            if component == (0, 0):
                targets = {(1, 1): 0.0, (2, 2): 0.0}
            elif component == (0, 1):
                targets = {(0, 0): 0.0, (1, 1): 0.0, (2, 2): 0.0}
        unit = np.zeros((3, 3))
        unit[component] = 1.0
        self.update_stresses()
        dt_step = abs(step_size) / self.strain_rate if self.strain_rate > 0 else 1.0
        max_seq = int(max_kmc_steps_pct * self.total_voxels)
        seq = 0
        remaining = 0.0
        done = 0
        events = 0
        while done < n_global_steps:
            remaining += dt_step
            while remaining > 0:
                ev, remaining = self.kmc_event(remaining, unit)
                if ev is None:
                    break
                events += 1
                seq = seq + 1 if ev[5] else 0
                if seq > max_seq or (max_events is not None and events >= max_events):
                    return self.flip_event_history
            self.update_stresses()
            self.mixed_bc(targets, component, mixed_tol, mixed_max_iter)
            done += 1
        return self.flip_event_history
