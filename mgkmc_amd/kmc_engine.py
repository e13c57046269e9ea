"""Device-resident STZ/KMC event loop shared by the 2D and 3D simulators.

The per-event work of the reference's KMC loops (mgkmc/aqs.py:848-1087 in 3D,
mgkmc/kmc_simulator.py:871-1101 in 2D) stays on the GPU: fused barrier + rate
pass, FP64 rate total, two-level event selection (the chosen voxel's barriers
are recomputed, Q is not stored), flip, mode redraw upload and the LS
re-equilibration.  The host keeps the scalar state (time, macro strain) and
the random streams:

  * ``rng``: numpy's legacy stream (the np.random module by default, exactly
    as the reference uses it; a RandomState for an isolated simulation):
    initial Q0 (generator), initial catalog (2 normals per mode, voxel order),
    then per event one uniform (waiting time), 2M normals (new modes) and M
    normals (new Q0);
  * ``selection_stream``: r = total * u for the event selection, the
    reference's numba-internal draw made seedable (default: the module-level
    ``mgkmc_amd.stz.kmc.SELECTION_STREAM``).

``init_modes=("device", seed)`` draws the initial Q0 and catalog on the GPU
with a counter-based generator instead (no reference counterpart: used by the
serving loop, where host draws of M modes per voxel would dominate start-up).

Supported: small strain, KMC mode and cascade mode, isotropic or
directional softening, redraw of directions and barriers, planar (pure-shear)
catalogs; linear elasticity, and the nonlinear laws of the reference's small-
strain branches (3D Landau; 2D Landau, secant degradation and Rose, whose
flips are sub-stepped on solver failure as the reference does).  Options the device path does not cover raise
NotImplementedError instead of falling back to a CPU path.
"""
import ctypes

import numpy as np

from . import _lib
from . import device as dev
from .kmc_device import KmcState, params as kmc_params, SIG_SYM64, SIG_SYM32, modes_from_device
from .linear_elastic_simulator import (LsSolver, FP64, FP32, MIXED, LandauSolver, LandauSolver2D, SecantSolver,
                                       SECANT_EXP, SECANT_ROSE, landau_params)
from .kmc_device import SIG_GEN64, CAT_PLANAR, CAT_GENERAL, is_planar, catalog_to_device, modes_to_device
from .elasticity import compute_lame_2d, compute_lame_3d
from .stz import kmc as _kmcmod

KB = 8.617e-5
# "mixed": FP32 iterations, then FP64 refinement to the FP64 stop test
LS_PRECISIONS = {"fp64": FP64, "fp32": FP32, "mixed": MIXED}


class _SigView:
    """The in-plane 2x2 Cauchy field (4, NV) of a 2D problem solved on an
    (nx, 1, ny) grid, the layout the 2D barrier kernels read."""

    sig = None

    def sync(self, sig3):
        v = sig3[[0, 2]][:, [0, 2]].reshape(4, -1)
        if self.sig is None or self.sig.shape != v.shape:
            self.sig = v.contiguous()
        else:  # in place: the native finite-strain flip writes this buffer
            self.sig.copy_(v)


def gaussian_generator(mean=2.0, std=0.6, min_cutoff=0.1, max_cutoff=None, rng=None):
    """normal + clip on the legacy stream (mgkmc/stz/barrier_generators.py:4-30,
    kmc_simulator_functions.py:222-229)."""
    def gen(shape):
        r = np.random if rng is None else rng
        b = r.normal(loc=mean, scale=std, size=shape)
        if min_cutoff is not None or max_cutoff is not None:
            b = np.clip(b, a_min=min_cutoff, a_max=max_cutoff)
        return b
    gen.params = dict(mean=mean, std=std, min_cutoff=min_cutoff, max_cutoff=max_cutoff)
    gen.rng = rng
    return gen


def planar_catalog(n_modes_total, gamma0, rng=None):
    """n planar pure-shear modes drawn as the reference's per-mode loops do
    (gxx then gxy; mgkmc/stz/catalog.py:7-40, kmc_simulator_functions.py:78-98):
    returns (n, 2) (gxx, gxy).  One vectorised call consumes the legacy stream
    exactly like the loop."""
    r = np.random if rng is None else rng
    g = r.normal(size=2 * n_modes_total).reshape(n_modes_total, 2)
    return 0.5 * gamma0 * g


def _clip(b, lo, hi):
    return np.clip(b, a_min=lo, a_max=hi) if (lo is not None or hi is not None) else b


def rayleigh_generator(loc=1.0, scale=0.5, min_cutoff=0.1, max_cutoff=None, rng=None):
    """scipy.stats.rayleigh(loc, scale).rvs + clip (mgkmc/stz/barrier_generators.py:32-58),
    on the given stream (numpy's global one by default)."""
    from scipy.stats import rayleigh

    def gen(shape):
        r = np.random.mtrand._rand if rng is None else rng
        return _clip(rayleigh(loc=loc, scale=scale).rvs(size=shape, random_state=r), min_cutoff, max_cutoff)
    gen.rng = rng
    return gen


def _modified_rayleigh(n, mean, std, r):
    """Rejection sampling of f(x) ~ x exp(-(x - mean)^2 / (2 std^2)) on
    [max(0, mean - 6 std), mean + 6 std] in batches of max(3 n_needed, 100)
    proposals (x then u uniform), first accepted kept in order
    (barrier_generators.py:60-107)."""
    out = np.empty(n)
    got = 0
    a, b = max(0.0, mean - 6 * std), mean + 6 * std
    x_max = (mean + np.sqrt(mean ** 2 + 4 * std ** 2)) / 2.0
    c_max = x_max * np.exp(-(x_max - mean) ** 2 / (2 * std ** 2))
    while got < n:
        need = n - got
        m = max(int(need * 3), 100)
        x = r.uniform(a, b, m)
        u = r.uniform(0, c_max, m)
        acc = x[u < x * np.exp(-(x - mean) ** 2 / (2 * std ** 2))]
        take = min(len(acc), need)
        out[got:got + take] = acc[:take]
        got += take
    return out


def modified_rayleigh_generator(mean=2.0, std=0.6, min_cutoff=0.1, max_cutoff=None, rng=None):
    """barrier_generators.py:60-107 on the given stream."""
    def gen(shape):
        r = np.random if rng is None else rng
        n = shape if isinstance(shape, int) else int(np.prod(shape))
        return _clip(_modified_rayleigh(n, mean, std, r).reshape(shape), min_cutoff, max_cutoff)
    gen.rng = rng
    return gen


def modified_rayleigh_exp_generator(mean=2.0, std=0.6, epsilon=0.1, ratio=0.8, min_cutoff=0.1, max_cutoff=None,
                                    rng=None):
    """Mixture of modified Rayleigh and exponential(epsilon) barriers
    (barrier_generators.py:111-170): one uniform per sample picks the
    exponential when below ratio; the Rayleigh samples are drawn first, then
    the exponential ones."""
    def gen(shape):
        r = np.random if rng is None else rng
        n = shape if isinstance(shape, int) else int(np.prod(shape))
        out = np.empty(n)
        exp_mask = r.uniform(0.0, 1.0, n) < ratio
        n_ray = int(np.sum(~exp_mask))
        if n_ray > 0:
            out[~exp_mask] = _modified_rayleigh(n_ray, mean, std, r)
        if n - n_ray > 0:
            out[exp_mask] = r.exponential(scale=epsilon, size=n - n_ray)
        return _clip(out.reshape(shape), min_cutoff, max_cutoff)
    gen.rng = rng
    return gen


BARRIER_GENERATORS = {"gaussian": gaussian_generator, "rayleigh": rayleigh_generator,
                      "modified_rayleigh": modified_rayleigh_generator,
                      "modified_rayleigh_with_exponential": modified_rayleigh_exp_generator}


def catalog_3d(n_modes, gamma0, rng=None):
    """n fully 3D modes in the stream order of stz_catalog_glass_3d
    (mgkmc/stz/catalog.py:47-82): per mode 9 normals (row-major 3x3), the
    symmetric deviatoric part G scaled to equivalent strain 1 (1 / (sqrt 2 |G|);
    a zero mode and no further draw when |G| < 1e-12), times the Rayleigh
    magnitude gamma0 sqrt(-2 ln u), u = uniform(1e-15, 1).  Mode by mode: the
    normals' cached second value carries over between modes."""
    r = np.random if rng is None else rng
    out = np.zeros((n_modes, 3, 3))
    for i in range(n_modes):
        a = r.normal(size=(3, 3))
        sym = 0.5 * (a + a.T)
        g = sym - (1.0 / 3.0) * np.trace(sym) * np.eye(3)
        nrm = np.sqrt(np.sum(g ** 2))
        if nrm < 1e-12:
            continue
        g = g * (1.0 / (np.sqrt(2.0) * nrm))
        out[i] = g * (gamma0 * np.sqrt(-2.0 * np.log(r.uniform(1e-15, 1.0))))
    return out


def simple_shear_catalog(n_modes, gamma0, rng=None):
    """n 2D simple-shear modes in the stream order of stz_catalog_glass_2d
    with stz_mode != "pure_shear" (mgkmc/kmc_simulator_functions.py:78-98):
    per mode gxx, gxy (normals), then the sign s = +1 if rand() > 0.5 else -1;
    [[gxx, gxy + s R], [gxy - s R, -gxx]] with R = sqrt(gxx^2 + gxy^2)."""
    r = np.random if rng is None else rng
    out = np.zeros((n_modes, 2, 2))
    for i in range(n_modes):
        gxx = 0.5 * gamma0 * r.normal()
        gxy = 0.5 * gamma0 * r.normal()
        R = np.sqrt(gxx ** 2 + gxy ** 2)
        sgn = 1.0 if r.rand() > 0.5 else -1.0
        out[i] = ((gxx, gxy + sgn * R), (gxy - sgn * R, -gxx))
    return out


def planar_to_full(p, d=3):
    out = np.zeros(p.shape[:-1] + (d, d))
    out[..., 0, 0] = p[..., 0]
    out[..., 1, 1] = -p[..., 0]
    out[..., 0, 1] = p[..., 1]
    out[..., 1, 0] = p[..., 1]
    return out


def device_init_modes(shape, M, gamma0, gen_params, seed, Q0, cat, stream=None):
    """Counter-based (Philox4x32-10) initial Q0 ~ clip(N(mean, std)) and planar
    catalog gxx, gxy ~ 0.5 gamma0 N(0, 1), written into the device arrays Q0
    [m][v] and cat [k][m][v].  Deterministic in (seed, voxel, mode)."""
    nv = int(np.prod(shape))
    lo = gen_params.get("min_cutoff")
    hi = gen_params.get("max_cutoff")
    rc = _lib.load().mgk_kmc_init_modes(
        nv, int(M), ctypes.c_uint64(int(seed) & (2 ** 64 - 1)), float(gen_params.get("mean", 2.0)),
        float(gen_params.get("std", 0.6)), 1 if lo is not None else 0, float(lo if lo is not None else 0.0),
        1 if hi is not None else 0, float(hi if hi is not None else 0.0), float(gamma0), _lib.dptr(Q0),
        _lib.dptr(cat), ctypes.c_void_p(dev.stream_ptr() if stream is None else stream))
    _lib.check(rc, "mgk_kmc_init_modes")


# trace-free basis eigenstrains of the fast-patching kernels (mgkmc/aqs.py:224-231,
# kmc_simulator.py:235-238)
_PATCH_BASES_3D = [np.array([[1.0, 0, 0], [0, -1.0, 0], [0, 0, 0]]), np.array([[0, 0, 0], [0, 1.0, 0], [0, 0, -1.0]]),
                   np.array([[0, 1.0, 0], [1.0, 0, 0], [0, 0, 0]]), np.array([[0, 0, 1.0], [0, 0, 0], [1.0, 0, 0]]),
                   np.array([[0, 0, 0], [0, 0, 1.0], [0, 1.0, 0]])]
_PATCH_BASES_2D = [np.array([[1.0, 0], [0, -1.0]]), np.array([[0, 1.0], [1.0, 0]])]


class DeviceKmc:
    """Grid-dimension-independent state and per-event step."""

    def _unsupported(self, cascade_mode, fast_patching, enable_thermal, strain_assumption, hyperelastic_model,
                     extra=()):
        bad = list(extra)
        if strain_assumption != "small_strain" and not getattr(self, "_finite_ok", False):
            bad.append("finite_strain KMC")
        if bad:
            raise NotImplementedError("device KMC path does not cover: " + ", ".join(bad))

    def _setup(self, shape, M, gamma0, E_field, nu_field, pixel, barrier_generator, barrier_kwargs,
               softening_scheme, softening_cap, jp, jt, neighbor_softening_fraction, temperature, strain_rate,
               stability_threshold, redraw_directions, redraw_barriers, nu0, q_act_temp, cascade_mode,
               scale_rate_by_volume, plane_mode, ls_precision, ls_warm_start, rng, selection_stream, init_modes,
               barrier_defaults, law=None, finite=None):
        self.rng = np.random if rng is None else rng
        self.selection_stream = _kmcmod.SELECTION_STREAM if selection_stream is None else selection_stream
        self.M, self.gamma0 = M, gamma0
        self.pixel = pixel
        self.volume = pixel ** 3
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
        self.plane_mode = plane_mode
        if self.temperature > 0:
            self.tau = 1.0 / (self.nu0 * np.exp(-self.q_act_temp / (KB * self.temperature)))
        else:
            self.tau = np.inf
        self.grid_shape = tuple(int(s) for s in shape)
        self.ndim = len(self.grid_shape)
        self.total_voxels = int(np.prod(self.grid_shape))
        d = self.ndim
        E = np.asarray(E_field, dtype=np.float64)
        if E.ndim == 0:
            E = np.full(self.grid_shape, float(E))
        E = E * 1e9 if E.mean() < 1e6 else E
        nu = np.asarray(nu_field, dtype=np.float64)
        if nu.ndim == 0:
            nu = np.full(self.grid_shape, float(nu))
        self.E_field, self.nu_field = E, nu
        self.E, self.nu = E, nu
        if barrier_generator is None:
            self.barrier_generator = gaussian_generator(**barrier_defaults, rng=rng)
        elif isinstance(barrier_generator, str):
            make = BARRIER_GENERATORS.get(barrier_generator.lower())
            if make is None:
                raise ValueError(f"Unknown barrier generator type: {barrier_generator}")
            self.barrier_generator = make(**barrier_kwargs, rng=rng)
        else:
            self.barrier_generator = barrier_generator
        if ls_precision not in LS_PRECISIONS:
            raise ValueError(f"ls_precision must be one of {sorted(LS_PRECISIONS)}")
        self.precision = LS_PRECISIONS[ls_precision]
        self.sig_kind = SIG_SYM32 if self.precision == FP32 else SIG_SYM64
        self.law = law
        self.finite = finite is not None
        if self.finite:
            # de Geus finite-strain state (mgkmc/aqs.py:205-215, kmc_simulator.py:156-168):
            # F, F_plastic = I; the step's own macro BC.  2D: the plane problem on an
            # (nx, 1, ny) grid of the 3D step, the lam of the plane mode
            from .finite_strain_simulator import FsSolver, MIXED as FS_MIXED, material, plane_law_supported
            if self.precision == FP32:
                raise NotImplementedError("finite-strain KMC runs with ls_precision 'fp64' or 'mixed'")
            finite = dict(finite)
            self._fs_solver = finite.pop("solver", "newton_cg")
            if d == 3:
                lam, mu = compute_lame_3d(E, nu)
                fs_grid = self.grid_shape
            else:
                if self._fs_solver not in ("newton_cg", "dbfft", "al") or self.precision != FP64:
                    raise NotImplementedError("2D finite-strain KMC runs solver='newton_cg', 'al' or 'dbfft' in FP64")
                lam, mu = compute_lame_2d(E, nu, plane_mode=plane_mode)
                fs_grid = (self.grid_shape[0], 1, self.grid_shape[1])
                lam, mu = lam.reshape(fs_grid), mu.reshape(fs_grid)
            eye = np.broadcast_to(np.eye(3).reshape(3, 3, 1, 1, 1), (3, 3) + fs_grid)
            self.fs = FsSolver(lam, mu, precision=FS_MIXED if self.precision == MIXED else 0, Fp=eye)
            self._fs_mat = material(**finite, mu_mean=float(np.mean(mu)), ndim=d)
            if d == 2:
                model = finite.get("model_type", "svk")
                if not plane_law_supported(model, plane_mode, finite.get("strain_capping_enabled", False)):
                    raise NotImplementedError(f"2D finite-strain {model} in {plane_mode}")
                self._fs_mat.plane = 2 if plane_mode == "plane_stress" else 1
                self._fs_sig2 = _SigView()
                self._fs_sig2.sync(self.fs.sig)
            self.F_macro = np.eye(d)
            self.sig_kind = SIG_GEN64
            self.ls = None
        elif law is None:
            self.ls = LsSolver(self.E_field, self.nu_field, plane_mode=plane_mode, precision=self.precision)
        else:
            if self.precision != FP64:
                raise NotImplementedError("the nonlinear small-strain laws run with ls_precision='fp64'")
            self.ls = self._law_solver(law)
        directional = softening_scheme == "directional"
        eps_prec = FP32 if self.precision == FP32 else FP64
        if init_modes is not None and init_modes[0] == "device":
            params = getattr(self.barrier_generator, "params", None)
            if params is None:
                raise NotImplementedError("device mode initialisation needs a gaussian barrier generator")
            self.kmc = KmcState(self.grid_shape, M, None, None, directional=directional, eps_precision=eps_prec,
                                eps_shadow32=self.precision == MIXED)
            device_init_modes(self.grid_shape, M, gamma0, params, init_modes[1], self.kmc.Q0, self.kmc.cat)
        else:
            # the reference's order: Q0, then the catalog voxel by voxel
            Q0 = self.barrier_generator(self.grid_shape + (M,))
            cat = self._draw_modes(self.total_voxels * M).reshape(self.grid_shape + (M, d, d))
            self.kmc = KmcState(self.grid_shape, M, Q0, cat, directional=directional, eps_precision=eps_prec,
                                eps_shadow32=self.precision == MIXED)
        self.ls_warm_start = ls_warm_start
        self._lin = None  # linear solver of the 3D mixed-BC corrections under a nonlinear law
        self._sig_owner = None  # solver whose stress field is current (None: self.ls)
        self.native_loop = True  # run eligible event sequences in the library's C++ loop
        self.ls_iters_total = 0  # LS iterations (FP64 + FP32) of the events' re-solves
        self._redraw_dev = None
        self._solved_once = False
        self.eps_macro = np.zeros((d, d))
        self.time = 0.0
        self.solver_args = {}
        self.enable_thermal = False
        self.Tfield = None
        self.kmc_event_history = []  # (x, y[, z], m, time, is_instab) per KMC event (extension)
        self.kmc_substeps = []  # finite strain: the N of each KMC event's sub-stepped flip (extension)
        self.flip_event_history = []  # tracked cascade flips (global step, batch, x, y, z, m), as the reference's
        self.cascade_history = []  # (global step, batch, x, y[, z], m) per cascade flip
        self.history_global = []  # (eps_xx, sig_xx / 1e9) per logged step (mgkmc/aqs.py:760)
        self.increment_history = []  # (eps_macro, <sigma>) after each elastic increment (extension)
        self._flip_log = None  # 2D flipped_voxels_log writer during a run
        self.max_cascade_steps_pct = 0.3
        self.ls_info = None
        self.fast_patching_enabled = False
        self._patched_mean = None  # <sigma> after a fast-patching update (None after a solve)

    # ----- thermal extension (local temperature, heat conduction) -----------
    def _setup_thermal(self, Cp, rho, thermal_diffusivity, thermal_coords, temperature_cap, thermostat, tau_bath):
        """Local temperature field on the device (mgkmc/aqs.py:83-99,
        kmc_simulator.py:60-75): T = temperature everywhere; heat conduction
        on the periodic box of n (pixel coordinates) or n * pixel (physical,
        diffusivity m^2/s -> nm^2/s) per axis."""
        self.enable_thermal = True
        self.Cp, self.rho = float(Cp), float(rho)
        self.thermal_diffusivity = float(thermal_diffusivity)
        self.thermal_coords = thermal_coords.lower()
        self.temperature_cap, self.thermostat, self.tau_bath = float(temperature_cap), bool(thermostat), float(tau_bath)
        torch = dev.torch
        self.Tfield = torch.full(self.grid_shape, float(self.temperature), dtype=torch.float64, device="cuda")
        phys = self.thermal_coords != "pixel"
        self._heat_L = (ctypes.c_double * 3)(*[float(n) * (self.pixel if phys else 1.0) for n in self.grid_shape])
        self.diffusivity_scaled = self.thermal_diffusivity * (1e18 if phys else 1.0)
        nb = ctypes.c_size_t()
        self._heat_shape = _lib.shape_arg(self.grid_shape)
        _lib.check(_lib.load().mgk_heat_workspace_bytes(self.ndim, self._heat_shape, ctypes.byref(nb)),
                   "mgk_heat_workspace_bytes")
        self._heat_ws = torch.empty(int(nb.value), dtype=torch.uint8, device="cuda")

    @property
    def Tlocal(self):
        return dev.download(self.Tfield) if self.enable_thermal else None

    @Tlocal.setter
    def Tlocal(self, value):
        if not self.enable_thermal:
            raise ValueError("Tlocal is only stored with enable_thermal=True")
        self.Tfield.copy_(dev.upload(np.array(np.broadcast_to(np.asarray(value, dtype=np.float64), self.grid_shape)), 0))

    def heat_conducting(self, dt):
        """heat_conducting_3d / _2d (mgkmc/aqs.py:462-478, kmc_simulator.py:559-575)
        on the device: spectral diffusion over dt, thermostat, cap."""
        if not self.enable_thermal or dt <= 0:
            return
        mode = 0 if not self.thermostat else (1 if self.tau_bath > 0 else 2)
        f = float(np.exp(-dt / self.tau_bath)) if mode == 1 else 1.0
        rc = _lib.load().mgk_heat_step(self.ndim, self._heat_shape, _lib.dptr(self.Tfield), self._heat_L,
                                       self.diffusivity_scaled * dt, float(self.temperature), mode, f,
                                       self.temperature_cap, _lib.dptr(self._heat_ws), self._heat_ws.numel(),
                                       ctypes.c_void_p(dev.stream_ptr()))
        _lib.check(rc, "mgk_heat_step")

    def _flip_heat(self, v, C):
        """Heat of a flip at voxel v (mgkmc/aqs.py:1000-1005): T += |sigma:C| /
        (rho Cp) with the stress before the re-solve, capped."""
        if not self.enable_thermal:
            return
        d = self.ndim
        coords = tuple(int(c) for c in self.kmc.voxel_coords(v))
        if self.finite:  # Cauchy stress of the finite-strain state (2D: its in-plane block)
            sig = dev.download(self.fs.sig[(slice(None), slice(None)) + coords]) if d == 3 else \
                dev.download(self._fs_sig2.sig[:, int(v)]).reshape(2, 2)
        else:
            s = dev.download(self._stress.sig[(slice(None),) + coords])
            sig = dev.from_sym_soa(s.reshape((-1,) + (1,) * d), d).reshape(d, d) if self.sig_kind != 2 else \
                s.reshape(d, d)
        dT = abs(np.sum(sig * C)) / (self.rho * self.Cp)
        flat = self.Tfield.view(-1)
        t = float(flat[int(v)]) + dT
        if self.temperature_cap > 0:
            t = min(t, self.temperature_cap)
        flat[int(v)] = t

    # ----- state mirrors (synchronised on access) --------------------------
    @property
    def _stress(self):
        if self.finite:
            return self.fs if self.ndim == 3 else self._fs_sig2
        return self.ls if self._sig_owner is None else self._sig_owner

    def _fs_field(self, t):
        """(3, 3, *fs grid) device field -> host (*grid, d, d)."""
        a = dev.download(t)
        if self.ndim == 2:
            from .finite_strain_simulator import _extract2
            a = _extract2(a)
            return np.einsum("ijxy->xyij", a)
        return np.einsum("ijxyz->xyzij", a)

    @property
    def sig_field(self):
        if self.finite:
            return self._fs_field(self.fs.sig)
        return dev.from_sym_soa(dev.download(self._stress.sig), self.ndim)

    @sig_field.setter
    def sig_field(self, value):
        self._set_sym_field(self._stress.sig, value)

    def _set_sym_field(self, t, value):
        """Store a host (*grid, d, d) field into a small-strain solver's
        symmetric-component array (its symmetric part)."""
        if self.finite:
            raise NotImplementedError("finite-strain states are set through F_field / F_plastic")
        sym, _ = dev.to_sym_soa(np.asarray(value, dtype=np.float64), self.ndim)
        t.copy_(dev.upload(sym, FP64).to(t.dtype).reshape(t.shape))

    @property
    def eps_field(self):
        if self.finite:  # Green-Lagrange strain of F, as the reference stores it
            F = self._fs_field(self.fs.F)
            return 0.5 * (np.einsum("...ki,...kj->...ij", F, F) - np.eye(self.ndim))
        return dev.from_sym_soa(dev.download(self._stress.eps), self.ndim)

    @eps_field.setter
    def eps_field(self, value):
        self._set_sym_field(self._stress.eps, value)

    @property
    def F_field(self):
        return self._fs_field(self.fs.F) if self.finite else None

    @property
    def F_plastic(self):
        return self._fs_field(self.fs.Fp) if self.finite else None

    @property
    def eps_plastic(self):
        return dev.from_sym_soa(dev.download(self.kmc.eps_p), self.ndim)

    @eps_plastic.setter
    def eps_plastic(self, value):
        """Set the eigenstrain field from a host (..., d, d) array (its
        symmetric part is stored)."""
        sym, _ = dev.to_sym_soa(np.asarray(value, dtype=np.float64), self.ndim)
        self.kmc.eps_p.copy_(dev.upload(sym, FP64).to(self.kmc.eps_p.dtype))
        if self.kmc.eps_p32 is not None:
            self.kmc.eps_p32.copy_(self.kmc.eps_p.to(dev.torch.float32))

    @property
    def Q(self):
        """Barriers of the current state (evaluated on access; the event loop
        does not store them)."""
        self.kmc.barrier_rates(self._stress.sig, self.sig_kind, self._params(), write_q=True, Tfield=self.Tfield)
        return modes_from_device(self.kmc.Q, self.grid_shape)

    @property
    def Q0(self):
        return modes_from_device(self.kmc.Q0, self.grid_shape)

    @Q0.setter
    def Q0(self, value):
        self.kmc.Q0.copy_(modes_to_device(value))

    @property
    def catalog(self):
        """(*grid, M, d, d) strain modes of every voxel."""
        d = self.ndim
        c = dev.download(self.kmc.cat).transpose(2, 1, 0)  # (NV, M, CK)
        full = planar_to_full(c, d) if self.kmc.planar else c.reshape(c.shape[:2] + (d, d))
        return np.ascontiguousarray(full).reshape(self.grid_shape + (self.M, d, d))

    @catalog.setter
    def catalog(self, value):
        d = self.ndim
        planar = is_planar(value, d)
        if planar == self.kmc.planar:
            self.kmc.cat.copy_(catalog_to_device(value, d, planar))
        else:
            self.kmc.cat = catalog_to_device(value, d, planar)
            self.kmc.planar = planar
            self.kmc.cat_kind = CAT_PLANAR if planar else CAT_GENERAL
            self.kmc.CK = 2 if planar else d * d

    @property
    def last_event_time(self):
        return dev.download(self.kmc.tlast).reshape(self.grid_shape)

    @last_event_time.setter
    def last_event_time(self, value):
        self.kmc.tlast.copy_(dev.upload(np.asarray(value, dtype=np.float64).reshape(-1), 0))

    @property
    def soft_prop(self):
        s = dev.download(self.kmc.soft)
        out = np.zeros(self.grid_shape + ((4,) if self.ndim == 3 else (2,)))
        out[..., 0] = s[0].reshape(self.grid_shape)
        out[..., 1] = s[1].reshape(self.grid_shape)
        return out

    @soft_prop.setter
    def soft_prop(self, value):
        """g_p, g_t from columns 0 and 1 (the others are unused)."""
        v = np.asarray(value, dtype=np.float64).reshape(self.total_voxels, -1)
        self.kmc.soft.copy_(dev.upload(np.ascontiguousarray(v[:, :2].T), 0))

    def _law_solver(self, law):
        """Device solver of a nonlinear small-strain law (the reference's
        elastic_run / update_stresses branches, mgkmc/kmc_simulator.py:449-512,
        mgkmc/aqs.py:331-348): ("landau", params) in 3D or 2D, ("secant", d, k),
        ("rose", eta, floor).  Secant takes the Lame pair of the plane mode,
        Rose and Landau the plane-strain (3D) pair."""
        kind = law[0]
        if self.ndim == 3:
            if kind != "landau":
                raise NotImplementedError(f"3D hyperelastic_model {kind!r} on the small-strain KMC path")
            lam, mu = compute_lame_3d(self.E_field, self.nu_field)
            return LandauSolver(lam, mu, FP64)
        if kind == "secant":
            lam, mu = compute_lame_2d(self.E_field, self.nu_field, plane_mode=self.plane_mode)
            return SecantSolver(lam, mu, SECANT_EXP, law[1], law[2], self.plane_mode, FP64)
        lam, mu = compute_lame_2d(self.E_field, self.nu_field, plane_mode="plane_strain")
        if kind == "rose":
            return SecantSolver(lam, mu, SECANT_ROSE, law[1], law[2], self.plane_mode, FP64)
        return LandauSolver2D(lam, mu, self.plane_mode, 5)

    # ----- fast patching (the reference's predictor-corrector) -------------
    def _setup_fast_patching(self, fast_patching, default_radius, disabled):
        """mgkmc/aqs.py:196-218, kmc_simulator.py:76-78, 154-171: the options
        under the reference's names; off for finite strain and the nonlinear
        laws (``disabled``); the patch kernels are computed at construction."""
        fp = fast_patching
        self.fast_patching_enabled = fp.get("enabled", False) if fp else False
        self.patch_radius = fp.get("patch_radius", default_radius) if fp else default_radius
        self.sync_interval = fp.get("sync_interval", 100) if fp else 100
        self.flips_since_sync = 0
        self.sigma_macro_unit = None
        self._su_live = False  # a patched flip has used sigma_macro_unit (2D remaining-time term)
        if disabled:
            self.fast_patching_enabled = False
        if self.fast_patching_enabled:
            self._precompute_patch_kernels()

    def _precompute_patch_kernels(self):
        """Truncated point solutions (mgkmc/aqs.py:220-256, kmc_simulator.py:229-261):
        per trace-free basis mode (five in 3D, two in 2D) a cold FP64 LS solve with
        that eigenstrain at the grid centre and the default solver arguments, the
        stress cropped to the periodic box of radius patch_radius around the
        centre (3D: one plane on a flat z axis), and the missing mean <sigma> -
        (box sum) / voxels.  Kept on the device as K [b][c][wx][wy][wz]; the
        host lists patch_kernels / patch_missing_mean hold the reference's
        (2R+1, 2R+1[, W], d, d) crops and (d, d) means."""
        torch = dev.torch
        d, R = self.ndim, int(self.patch_radius)
        if R < 0 or R > 64:
            raise ValueError("patch_radius must lie in 0..64 on the device path")
        shape = self.grid_shape
        pairs = dev.sym_pairs(d)
        C = len(pairs)
        bases = _PATCH_BASES_3D if d == 3 else _PATCH_BASES_2D
        self._patch_ls = LsSolver(self.E_field, self.nu_field, plane_mode=self.plane_mode, precision=FP64)
        radii = [R, R, R if shape[2] > 1 else 0] if d == 3 else [R, R]
        idx = [torch.tensor([(n // 2 + k) % n for k in range(-r, r + 1)], device="cuda")
               for n, r in zip(shape, radii)]
        nv = self.total_voxels
        ep = torch.zeros((C,) + shape, dtype=torch.float64, device="cuda")
        center = tuple(n // 2 for n in shape)
        Ks, self.patch_kernels, self.patch_missing_mean = [], [], []
        for P in bases:
            ep.zero_()
            for c, (i, j) in enumerate(pairs):
                ep[(c,) + center] = float(P[i, j])
            info = self._patch_ls.solve(np.zeros((d, d)), ep, warm=False, max_iter=200, tol=1e-6)
            sig = self._patch_ls.sig
            if d == 3:
                crop = sig[:, idx[0][:, None, None], idx[1][None, :, None], idx[2][None, None, :]]
            else:
                crop = sig[:, idx[0][:, None], idx[1][None, :]]
            crop = crop.contiguous()
            Ks.append(crop)
            host = dev.from_sym_soa(dev.download(crop), d)  # (wx, wy[, wz], d, d)
            full_mean = np.array(info.sig_mean[: d * d]).reshape(d, d)
            self.patch_kernels.append(host)
            self.patch_missing_mean.append(full_mean - np.sum(host, axis=tuple(range(d))) / nv)
        self._patch_K = torch.stack(Ks).contiguous()
        self._patch_radius_c = (ctypes.c_int * d)(*radii[:d]) if d == 2 else (ctypes.c_int * 3)(*radii)
        need = ctypes.c_size_t()
        _lib.check(_lib.load().mgk_kmc_patch_workspace_bytes(ctypes.byref(need)), "mgk_kmc_patch_workspace_bytes")
        self._patch_ws = torch.empty(int(need.value), dtype=torch.uint8, device="cuda")
        self._patch_shape = _lib.shape_arg(shape)

    def _patch_coefs(self, C):
        # basis weights (mgkmc/aqs.py:1030, kmc_simulator.py:765, :1055)
        if self.ndim == 3:
            return C[0, 0], -C[2, 2], C[0, 1], C[0, 2], C[1, 2]
        return C[0, 0], C[0, 1]

    def _patch_call(self, v, coefs, shift, unit, a):
        d = self.ndim
        pairs = dev.sym_pairs(d)
        sh = (ctypes.c_double * 6)(*[float(shift[i, j]) for i, j in pairs])
        cf = (ctypes.c_double * 5)(*[float(c) for c in coefs])
        sm = (ctypes.c_double * 9)()
        sig = self._stress.sig
        rc = _lib.load().mgk_kmc_patch(
            d, self._patch_shape, SIG_SYM64 if sig.dtype == dev.torch.float64 else SIG_SYM32, _lib.dptr(sig),
            _lib.dptr(unit), float(a), int(v), self._patch_radius_c, len(coefs), _lib.dptr(self._patch_K), cf, sh,
            _lib.dptr(self._patch_ws), self._patch_ws.numel(), ctypes.c_void_p(dev.stream_ptr()), sm)
        _lib.check(rc, "mgk_kmc_patch")
        self._patched_mean = np.array(sm[: d * d]).reshape(d, d)

    def _patch_stamp(self, v, C):
        """The patch of mode C around voxel v and the missing mean
        (mgkmc/aqs.py:1030-1047, kmc_simulator.py:764-774)."""
        coefs = self._patch_coefs(C)
        mm = self.patch_missing_mean
        shift = coefs[0] * mm[0]
        for c, m in zip(coefs[1:], mm[1:]):
            shift = shift + c * m
        self._patch_call(v, coefs, shift, None, 0.0)

    def _make_sigma_macro_unit(self, strain_unit_tensor):
        """sigma_macro_unit (mgkmc/aqs.py:1022-1026, kmc_simulator.py:1045-1050): the
        stress of the unit macro strain with no eigenstrain, a cold LS solve with the
        current solver arguments, made once."""
        if self.sigma_macro_unit is not None:
            return
        args = dict(max_iter=200, tol=1e-6)
        args.update(self.solver_args)
        self._patch_ls.solve(np.asarray(strain_unit_tensor, float), None, warm=False, max_iter=args["max_iter"],
                             tol=args["tol"])
        self.sigma_macro_unit = self._patch_ls.sig.clone()
        self._patch_ls = None  # set-up done: its fields are not needed again

    def _patch_after_flip(self, v, C, t_wait, strain_unit_tensor):
        """A KMC event's predictor (mgkmc/aqs.py:1021-1054, kmc_simulator.py:
        1044-1078): the macro-unit stress (computed at the first flip with the
        current solver arguments) scaled by strain_rate * t_wait, the patch and
        the missing mean in one pass; every sync_interval flips the full
        re-solve.  Returns True when it re-solved."""
        self._make_sigma_macro_unit(strain_unit_tensor)
        self._su_live = True
        coefs = self._patch_coefs(C)
        mm = self.patch_missing_mean
        shift = coefs[0] * mm[0]
        for c, m in zip(coefs[1:], mm[1:]):
            shift = shift + c * m
        self._patch_call(v, coefs, shift, self.sigma_macro_unit if t_wait > 0 else None,
                         self.strain_rate * t_wait)
        self.flips_since_sync += 1
        if self.flips_since_sync >= self.sync_interval:
            self.update_stresses()
            self.flips_since_sync = 0
            return True
        return False

    # ----- seams -----------------------------------------------------------
    def _solve(self, eps_macro):
        self._sig_owner = None
        self._patched_mean = None
        if self.finite:
            return self._solve_finite()
        if self.law is not None:
            return self._solve_law(eps_macro)
        warm = self.ls_warm_start and self._solved_once
        args = dict(max_iter=200, tol=1e-6)
        args.update(self.solver_args)
        self.ls_info = self.ls.solve(eps_macro, self.kmc.eps_p, warm=warm, max_iter=args["max_iter"],
                                     tol=args["tol"], eps_plastic32=self.kmc.eps_p32)
        self._solved_once = True
        return self.ls_info

    def _solve_finite(self):
        """Finite-strain update_stresses (mgkmc/aqs.py:276-330): macro BC from
        the driven strain with F_macro as the initial F_bar, one Newton-Krylov
        step from the current F with F_plastic; F_macro and the macro strain
        (F_ii - 1, F_ij) follow the step; ValueError on a NaN field or
        |<sigma>| > 20 GPa."""
        from .finite_strain_simulator import build_finite_strain_bc_3d, _embed2, _extract2, _PLANE_AXES
        d = self.ndim
        drv = getattr(self, "driving_component", (0, 0))
        tg = getattr(self, "stress_targets", {})
        Fb, _, Pt, Pm = build_finite_strain_bc_3d(drv, self.eps_macro[drv], tg, ndim=d, F_bar_initial=self.F_macro)
        drv3 = drv
        if d == 2:  # kmc_simulator.py:389-445 through the (nx, 1, ny) embedding
            Fb, Pt, Pm = _embed2(Fb), _embed2(Pt, 0.0), _embed2(Pm.astype(float), 0.0).astype(bool)
            drv3 = (_PLANE_AXES[drv[0]], _PLANE_AXES[drv[1]])
        info = self.fs.step(Fb, self._fs_mat, drv3, Pt, Pm, E_avg=self.E_field.mean(), nu_avg=self.nu_field.mean(),
                            solver=self._fs_solver, pixel=self.pixel)
        Fm = np.array(info.F_bar[:]).reshape(3, 3)
        sm = np.array(info.sig_mean[:9]).reshape(3, 3)
        if d == 2:
            Fm, sm = _extract2(Fm), _extract2(sm)
            self._fs_sig2.sync(self.fs.sig)
        self.F_macro = Fm
        for i in range(d):
            for j in range(d):
                self.eps_macro[i, j] = self.F_macro[i, j] - 1.0 if i == j else self.F_macro[i, j]
        self.ls_info = info
        self._solved_once = True
        torch = dev.torch
        if bool(torch.isnan(self.fs.sig).any()) or bool(torch.isnan(self.fs.F).any()) or \
                np.max(np.abs(sm)) / 1e9 > 20.0:
            raise ValueError(f"Simulation stopped due to unstable stress: {sm / 1e9} GPa")
        return info

    def _flip_finite(self, v, m):
        """One KMC event's finite-strain flip (see _flip_finite_pairs)."""
        self._flip_finite_pairs([(v, m)])
        self.kmc_substeps.append(self.last_substeps)

    def _flip_finite_pairs(self, pairs):
        """Finite-strain flips of one event or one cascade batch (mgkmc/aqs.py:512-615,
        887-996; 2D kmc_simulator.py:605-700, 895-1000): F_p <- expm(C/N) F_p at
        every flipped voxel in N sub-steps, retried from the state before the
        flips on ValueError; 3D: N = n0 .. n0 + 19, n0 = min(max over the
        flips of max(2, ceil(sqrt(C:C/2) / 0.005)), 40); 2D: N = 1, then
        n0 .. n0 + 19 with the step 0.02.  Softening, prev and redraws of each
        voxel at the last sub-step, one finite-strain step after each sub-step."""
        if self._native_fs_ok():
            return self._flip_finite_native(pairs)
        from scipy.linalg import expm
        torch = dev.torch
        Cs = [self._mode_host(v, m) for v, m in pairs]
        k, fs = self.kmc, self.fs
        live = [fs.F, fs.P, fs.sig, fs.Fp, k.soft, k.tlast, k.prev, k.cat, k.Q0, self.Tfield]
        backup = [None if a is None else a.clone() for a in live]
        b_fm, b_em = self.F_macro.copy(), self.eps_macro.copy()
        step = 0.005 if self.ndim == 3 else 0.02
        n0 = min(max(max(2, int(np.ceil(np.sqrt(0.5 * np.sum(C ** 2)) / step))) for C in Cs), 40)
        if self.ndim == 3:
            schedule = list(range(n0, n0 + 20))
        else:
            n0 = max(n0, 2)
            schedule = [1] + list(range(n0, n0 + 20))
        sels = []
        for v, _ in pairs:
            c = tuple(int(q) for q in self.kmc.voxel_coords(v))
            # 2D: the in-plane block of the (nx, 1, ny) grid
            sels.append((slice(None), slice(None)) + c if self.ndim == 3 else ([[0], [2]], [0, 2], c[0], 0, c[1]))
        for N in schedule:
            for a, b in zip(live, backup):
                if a is not None:
                    a.copy_(b)
            self.F_macro = b_fm.copy()
            self.eps_macro[...] = b_em
            if self.ndim == 2:
                self._fs_sig2.sync(fs.sig)
            try:
                for i in range(1, N + 1):
                    for (v, m), C, sel in zip(pairs, Cs, sels):
                        dF = expm(C / N)
                        fpv = fs.Fp[sel].cpu().numpy()
                        fs.Fp[sel] = torch.from_numpy(np.dot(dF, fpv)).to(fs.Fp.device)
                        if i == N:
                            k.flip(v, m, self.time, self.jp, self.jt, self.softening_cap,
                                   self.neighbor_softening_fraction, 0)
                            self._flip_heat(v, C)  # from the stress of the last sub-step's start
                            self._redraw(v, m)
                    self.update_stresses()
                self.last_substeps = N
                return
            except ValueError:
                continue
        raise ValueError("Mechanical solver failed to converge under sub-stepping.")

    def _native_fs_ok(self):
        """Whether the sub-stepped finite-strain flip runs in the library
        (mgk_kmc_fs_flip): planar catalogs redrawn from the Gaussian barrier
        generator on the host stream, no thermal coupling or flip log, FP64
        fields."""
        gen = self.barrier_generator
        gp = getattr(gen, "params", None)
        same_stream = getattr(gen, "rng", 0) is (None if self.rng is np.random else self.rng)
        return (self.native_loop and self.finite and not self.enable_thermal and self.kmc.planar
                and self._flip_log is None and gp is not None and same_stream
                and (self.redraw_directions or self.redraw_barriers) and hasattr(self.rng, "get_state")
                and self.fs.precision == 0)

    def _native_law_ok(self):
        """Whether the 2D nonlinear-law sub-stepped flip runs in the library
        (mgk_kmc_law_flip; also inside mgk_kmc_run): Landau (fixed point),
        secant or Rose, planar catalogs from the Gaussian generator on the host
        stream, no thermal coupling or flip log."""
        gen = self.barrier_generator
        gp = getattr(gen, "params", None)
        same_stream = getattr(gen, "rng", 0) is (None if self.rng is np.random else self.rng)
        return (self.native_loop and self.law is not None and not self.finite
                and (self.law[0] in ("landau", "secant", "rose") if self.ndim == 2 else self.law[0] == "landau")
                and not self.enable_thermal and self.kmc.planar
                and self._flip_log is None and gp is not None and same_stream
                and (self.redraw_directions or self.redraw_barriers) and hasattr(self.rng, "get_state")
                and self.kmc.eps_precision == FP64)

    def _law_substep_cfg(self):
        """MgkKmcLawSubstep of the law solver: its fields and buffers, the
        reference's solve arguments, the live state and its backup."""
        ls, k = self.ls, self.kmc
        torch = dev.torch
        args = dict(max_iter=400, tol=1e-6)
        args.update(self.solver_args)
        st = _lib.MgkLawStepper()
        if self.ndim == 2:
            st.shape[0], st.shape[1] = self.grid_shape[0], self.grid_shape[1]
        st.lam, st.mu, st.lam0, st.mu0 = _lib.dptr(ls.lam), _lib.dptr(ls.mu), ls.lam0, ls.mu0
        keep = [st]
        if self.law[0] == "landau" and self.ndim == 3:  # cold solves, no sub-stepping
            if not hasattr(self, "_law_params"):
                self._law_params = landau_params(*self.law[1], mu=self.ls.mu_host, **self.law[2])
            st.kind, st.landau = 2, ctypes.addressof(self._law_params)
            keep += [None, self._law_params]
        elif self.law[0] == "landau":
            if not hasattr(self, "_law_params"):
                self._law_params = landau_params(*self.law[1], mu=self.ls.mu_host, **self.law[2])
            st.kind, st.landau, st.plane_stress = 0, ctypes.addressof(self._law_params), int(ls.plane_stress)
            st.e33 = _lib.dptr(ls.e33)
            flag = ctypes.c_int(int(ls.e33_have))
            st.e33_have = ctypes.addressof(flag)
            st.anderson_m, st.aa = ls.anderson_m, _lib.dptr(ls.aa)
            keep += [flag, self._law_params]
        else:
            st.kind, st.secant = 1, ctypes.addressof(ls.params)
        st.eps, st.sig, st.tau = _lib.dptr(ls.eps), _lib.dptr(ls.sig), _lib.dptr(ls.tau)
        st.ws, st.ws_bytes = _lib.dptr(ls.ws), ls.ws.numel()
        st.max_iter, st.tol = int(args["max_iter"]), float(args["tol"])
        c = _lib.MgkKmcLawSubstep()
        c.step = ctypes.addressof(st)
        live = [k.eps_p, k.soft, k.tlast, k.prev, k.cat, k.Q0, ls.eps, ls.sig]
        c.nlive = len(live)
        total = 0
        for q, a in enumerate(live):
            c.live[q] = _lib.dptr(a)
            c.live_bytes[q] = 0 if a is None else a.numel() * a.element_size()
            total += (c.live_bytes[q] + 255) // 256 * 256
        if getattr(self, "_law_backup", None) is None or self._law_backup.numel() < total:
            self._law_backup = torch.empty(total, dtype=torch.uint8, device="cuda")
        c.backup, c.backup_bytes = _lib.dptr(self._law_backup), self._law_backup.numel()
        c._keep = keep
        return c

    def _law_after(self, c):
        """Solver state after a native law flip: the persisting E33 flag and
        the last solve's record."""
        if self.law[0] == "landau" and self.ndim == 2:
            self.ls.e33_have = bool(c._keep[1].value)
        self.ls_info = self.ls.info
        self._solved_once = True

    def _flip_substepped_native(self, flips):
        """_flip_substepped through mgk_kmc_law_flip (same schedule, same
        host-stream redraws)."""
        d = self.ndim
        L = self._loop_struct(np.zeros((d, d)), patch=False)
        c = self._law_substep_cfg()
        n = len(flips)
        vv = (ctypes.c_longlong * n)(*[int(v) for v, _ in flips])
        mm = (ctypes.c_int * n)(*[int(m) for _, m in flips])
        rng = _lib.mt_from(self.rng)
        em = (ctypes.c_double * 9)(*self.eps_macro.reshape(-1))
        nsub = ctypes.c_int()
        try:
            rc = _lib.load().mgk_kmc_law_flip(ctypes.byref(L), ctypes.byref(c), n, vv, mm, float(self.time),
                                              ctypes.byref(rng), em, ctypes.byref(nsub), ctypes.byref(self.ls.info),
                                              ctypes.c_void_p(dev.stream_ptr()))
        finally:
            _lib.mt_to(self.rng, rng)
        self._law_after(c)
        if rc == _lib.MGK_ERR_SOLVER:
            raise ValueError("Mechanical solver failed to converge under small-strain sub-stepping.")
        _lib.check(rc, "mgk_kmc_law_flip")
        self.last_substeps = nsub.value

    def _flip_finite_native(self, pairs):
        """_flip_finite_pairs through mgk_kmc_fs_flip: the same sub-stepping
        schedule, redraws from the same host stream, F_macro / eps_macro and
        the step's record handed back."""
        d = self.ndim
        L = self._loop_struct(np.zeros((d, d)), patch=False)
        c = self._fs_substep_cfg(L)
        n = len(pairs)
        vv = (ctypes.c_longlong * n)(*[int(v) for v, _ in pairs])
        mm = (ctypes.c_int * n)(*[int(m) for _, m in pairs])
        rng = _lib.mt_from(self.rng)
        fm = (ctypes.c_double * 9)(*self.F_macro.reshape(-1))
        em = (ctypes.c_double * 9)(*self.eps_macro.reshape(-1))
        nsub = ctypes.c_int()
        info = _lib.MgkFsInfo()
        try:
            rc = _lib.load().mgk_kmc_fs_flip(ctypes.byref(c), n, vv, mm, float(self.time), ctypes.byref(rng), fm, em,
                                             ctypes.byref(nsub), ctypes.byref(info),
                                             ctypes.c_void_p(dev.stream_ptr()))
        finally:
            _lib.mt_to(self.rng, rng)
        self.F_macro = np.array(fm[: d * d]).reshape(d, d)
        self.eps_macro[...] = np.array(em[: d * d]).reshape(d, d)
        self.ls_info = info
        self._solved_once = True
        self._fs_raise(rc)
        self.last_substeps = nsub.value

    @staticmethod
    def _fs_raise(rc):
        """The reference's exceptions of a failed sub-stepped flip."""
        if rc == _lib.MGK_ERR_SOLVER:
            raise ValueError("Mechanical solver failed to converge under sub-stepping.")
        if rc == _lib.MGK_ERR_NONFINITE:
            raise FloatingPointError(_lib.last_error().split(" (/")[0])
        _lib.check(rc, "mgk_kmc_fs_flip")

    def _fs_substep_cfg(self, L):
        """MgkKmcFsSubstep for loop description L: the step (solver, fields,
        material, the reference's step defaults), the macro BC inputs, the
        live device state and its backup area (kept on self)."""
        d, fs, k = self.ndim, self.fs, self.kmc
        torch = dev.torch
        st = fs.stepper(self._fs_mat, E_avg=self.E_field.mean(), nu_avg=self.nu_field.mean(),
                        solver=self._fs_solver, pixel=self.pixel)
        c = _lib.MgkKmcFsSubstep()
        c.kmc, c.step = ctypes.addressof(L), ctypes.addressof(st)
        drv = getattr(self, "driving_component", (0, 0))
        c.drv_i, c.drv_j = int(drv[0]), int(drv[1])
        # the step's BC reads the diagonal targets P_kk only (build_finite_strain_bc_3d)
        tg = getattr(self, "stress_targets", {}) or {}
        diag = [(int(a), float(v)) for (a, b), v in tg.items() if a == b and 0 <= int(a) < d]
        c.n_targets = len(diag)
        for q, (kk, v) in enumerate(diag):
            c.target_k[q], c.target_val[q] = kk, v
        live = [fs.F, fs.P, fs.sig, fs.Fp, k.soft, k.tlast, k.prev, k.cat, k.Q0]
        c.nlive = len(live)
        total = 0
        for q, a in enumerate(live):  # buffers a scheme leaves out (prev) are None
            c.live[q] = _lib.dptr(a)
            c.live_bytes[q] = 0 if a is None else a.numel() * a.element_size()
            total += (c.live_bytes[q] + 255) // 256 * 256
        if getattr(self, "_fs_backup", None) is None or self._fs_backup.numel() < total:
            self._fs_backup = torch.empty(total, dtype=torch.uint8, device="cuda")
            self._fs_flag = torch.zeros(1, dtype=torch.int32, device="cuda")
        c.backup, c.backup_bytes = _lib.dptr(self._fs_backup), self._fs_backup.numel()
        c.flag = _lib.dptr(self._fs_flag)
        c.sig2 = _lib.dptr(self._fs_sig2.sig) if d == 2 else None
        c._keep = (st, L)
        return c

    def _solve_linear(self, eps_macro):
        """update_stress_fft_full itself (linear LS, cold start) whatever the
        law, as the reference's 3D mixed-BC loop calls it (mgkmc/aqs.py:1137-1139);
        under a nonlinear law its stress field becomes the current one."""
        if self.law is None:
            return self._solve(eps_macro)
        self._patched_mean = None
        if self._lin is None:
            self._lin = LsSolver(self.E_field, self.nu_field, plane_mode=self.plane_mode, precision=FP64)
        args = dict(max_iter=200, tol=1e-6)
        args.update(self.solver_args)
        self.ls_info = self._lin.solve(eps_macro, self.kmc.eps_p, warm=False, max_iter=args["max_iter"],
                                       tol=args["tol"])
        self._sig_owner = self._lin
        return self.ls_info

    def _solve_law(self, eps_macro):
        """One nonlinear solve with the reference's defaults (max_iter 400,
        tol 1e-6) overridden by solver_args: Landau 3D and the secant laws
        start cold, Landau 2D from the current field with one e33 state for
        the run; a non-finite Landau update raises FloatingPointError."""
        args = dict(max_iter=400, tol=1e-6)
        args.update(self.solver_args)
        kind = self.law[0]
        kw = dict(max_iter=args["max_iter"], tol=args["tol"])
        if kind == "landau":
            if not hasattr(self, "_law_params"):
                self._law_params = landau_params(*self.law[1], mu=self.ls.mu_host, **self.law[2])
            if self.ndim == 3:
                self.ls_info = self.ls.solve(self._law_params, eps_macro, self.kmc.eps_p, warm=False, **kw)
            else:
                self.ls_info = self.ls.solve(self._law_params, eps_macro, self.kmc.eps_p, warm=True, persist_e33=True,
                                             **kw)
        else:
            self.ls_info = self.ls.solve(eps_macro, self.kmc.eps_p, warm=False, **kw)
        self._solved_once = True
        return self.ls_info

    def sig_mean(self):
        d = self.ndim
        if self._patched_mean is not None:
            return self._patched_mean.copy()
        if self.ls_info is None:  # no solve yet (cascade mode before its first batch): the zero initial stress
            return np.zeros((d, d))
        if self.finite:
            sm = np.array(self.ls_info.sig_mean[:9]).reshape(3, 3)
            if d == 2:
                from .finite_strain_simulator import _extract2
                sm = _extract2(sm)
            return sm
        return np.array(self.ls_info.sig_mean[: d * d]).reshape(d, d)

    def _params(self):
        eff_volume = self.volume if self.scale_rate_by_volume else 1.0
        return kmc_params(volume=self.volume, rate_volume=eff_volume, nu0=self.nu0, temperature=self.temperature,
                          cap=self.softening_cap, tau=self.tau, time=self.time,
                          scheme=1 if self.softening_scheme == "directional" else 0, cascade=self.cascade_mode)

    def _after_flip_solve(self):
        self.update_stresses()

    def kmc_event(self, remaining_time, strain_unit_tensor):
        """One pass of the inner KMC loop body.  Returns (event or None,
        remaining_time); an event is (x, y[, z], m, time, is_instab)."""
        p = self._params()
        sig = self._stress.sig
        total = self.kmc.barrier_rates(sig, self.sig_kind, p, write_q=False, Tfield=self.Tfield)
        if total > 0:
            u = self.rng.random_sample()
            t_wait = -np.log(u) / total
            if t_wait < remaining_time:
                self.time += t_wait
                self.eps_macro += strain_unit_tensor * (self.strain_rate * t_wait)
                remaining_time -= t_wait
                us = self.selection_stream.random_sample()
                v, m = self.kmc.select(us, sig, self.sig_kind, p, Tfield=self.Tfield)
                coords = tuple(int(c) for c in self.kmc.voxel_coords(v))
                is_instab = self.kmc.sel.q <= self.stability_threshold
                if self._flip_log is not None:
                    self._log_flip(v, m, self.kmc.sel.q, "kmc")
                # the reference conducts heat over t_wait before selecting, from
                # rates computed before it: conducting after the selection is the same
                self.heat_conducting(t_wait)
                if self.enable_thermal:
                    self._flip_heat_mode = self._mode_host(v, m)
                if self.finite:
                    self._flip_finite(v, m)
                elif self.law is not None and self.ndim == 2:
                    self._flip_substepped([(v, m)])
                elif self.fast_patching_enabled:
                    C = self._mode_host(v, m)
                    self._flip(v, m, 1)
                    self._flip_heat(v, getattr(self, "_flip_heat_mode", None))
                    solved = self._patch_after_flip(v, C, t_wait, strain_unit_tensor)
                else:
                    self._flip(v, m, 1)
                    self._flip_heat(v, getattr(self, "_flip_heat_mode", None))
                    self._after_flip_solve()
                if self.ls_info is not None and not self.finite and (not self.fast_patching_enabled or solved):
                    self.ls_iters_total += self.ls_info.iters + getattr(self.ls, "iters32", 0)
                ev = coords + (m, self.time, is_instab)
                self.kmc_event_history.append(ev)
                return ev, remaining_time
        if self.fast_patching_enabled and self.ndim == 2 and self._su_live:
            # kmc_simulator.py:1094-1096 (the 3D loop re-solves at the increment's end without it)
            self._patch_call(0, (), np.zeros((2, 2)), self.sigma_macro_unit, self.strain_rate * remaining_time)
        self.eps_macro += strain_unit_tensor * (self.strain_rate * remaining_time)
        self.heat_conducting(remaining_time)
        self.time += remaining_time
        return None, 0.0

    def _flip(self, v, m, eps_div):
        """Flip (eps_div = 1), or the last sub-step of a sub-stepped flip
        (eps_div = N: mode / N), then the redraw of the voxel's modes."""
        self.kmc.flip(v, m, self.time, self.jp, self.jt, self.softening_cap, self.neighbor_softening_fraction,
                      eps_div)
        self._redraw(v, m)

    def _redraw(self, v, m):
        """New modes and barriers of voxel v from the host stream (directions
        first, then barriers, each when its flag is on); with both flags off
        only the flipped mode m gets a new direction and barrier
        (mgkmc/aqs.py:594-606, kmc_simulator.py:1000-1008)."""
        newcat = newq0 = None
        if self.redraw_directions or self.redraw_barriers:
            if self.redraw_directions:
                newcat = self._draw_modes(self.M)
            if self.redraw_barriers:
                newq0 = self.barrier_generator((self.M,))
        else:
            newcat = np.stack([self._mode_host(v, mm) for mm in range(self.M)])
            newcat[m] = self._draw_modes(1)[0]
            newq0 = dev.download(self.kmc.Q0[:, int(v)]).copy()
            newq0[m] = self.barrier_generator((1,))[0]
        self.kmc.set_voxel(v, newcat, newq0)

    def _draw_modes(self, n):
        """n catalog modes from the host stream: planar pure shear, or fully 3D
        modes with use_3d_barriers."""
        if getattr(self, "use_3d_barriers", False):
            return catalog_3d(n, self.gamma0, self.rng)
        if getattr(self, "stz_mode", "pure_shear") != "pure_shear":
            return simple_shear_catalog(n, self.gamma0, self.rng)
        return planar_to_full(planar_catalog(n, self.gamma0, self.rng), self.ndim)

    def _mode_host(self, v, m):
        c = dev.download(self.kmc.cat[:, m, v])
        if self.kmc.planar:
            return planar_to_full(c[None, :], self.ndim)[0]
        return c.reshape(self.ndim, self.ndim)

    def _flip_substepped(self, flips):
        """Flips [(v, m), ...] with the reference's small-strain sub-stepping
        (mgkmc/kmc_simulator.py:263-385): one step first; if the re-solve
        raises ValueError or FloatingPointError, restore the state (eigenstrain,
        softening, event times, prev, modes, fields) and retry with N sub-steps
        of mode / N per flip, N = max(N_h, 2) .. +19 (N_h = min(max over the
        flips of max(2, ceil(sqrt(C:C/2) / 0.02)), 40)), softening, prev and
        redraws at the last, one re-solve per sub-step.  Every attempt draws
        its redraws anew from the host stream."""
        if self._native_law_ok():
            return self._flip_substepped_native(flips)
        k = self.kmc
        live = [k.eps_p, k.soft, k.tlast, k.prev, k.cat, k.Q0, self.ls.eps, self.ls.sig, self.Tfield]
        backup = [None if a is None else a.clone() for a in live]
        schedule = [1]
        for N in schedule:
            if N > 1:
                for a, b in zip(live, backup):
                    if a is not None:
                        a.copy_(b)
            try:
                for i in range(1, N + 1):
                    for v, m in flips:
                        if i < N:
                            k.flip(v, m, self.time, self.jp, self.jt, self.softening_cap,
                                   self.neighbor_softening_fraction, -N)
                        else:
                            # heat at the last sub-step from the current stress (kmc_simulator.py:358-364)
                            C = self._mode_host(v, m) if self.enable_thermal else None
                            self._flip(v, m, N)
                            self._flip_heat(v, C)
                    self.update_stresses()
                self.last_substeps = N
                return
            except (ValueError, FloatingPointError):
                if N == 1:
                    for a, b in zip(live, backup):
                        if a is not None:
                            a.copy_(b)
                    nh = 2
                    for v, m in flips:
                        C = self._mode_host(v, m)
                        nh = max(nh, max(2, int(np.ceil(np.sqrt(0.5 * np.sum(C ** 2)) / 0.02))))
                    nh = min(nh, 40)
                    schedule.extend(range(max(nh, 2), max(nh, 2) + 20))
        raise ValueError("Mechanical solver failed to converge under small-strain sub-stepping.")

    # ----- cascades (cascade_mode) -------------------------------------------
    def _cascade_pairs(self):
        """Modes to flip in one cascade batch (mgkmc/aqs.py:489-510,
        kmc_simulator.py:581-597): barriers of the current state, modes with
        Q < stability_threshold sorted by Q (np.argsort over the voxel-major
        scan, as the reference sorts), the first (most unstable) per voxel."""
        self.kmc.barrier_rates(self._stress.sig, self.sig_kind, self._params(), write_q=True, Tfield=self.Tfield)
        v, m, q = self.kmc.unstable_modes(self.stability_threshold)
        pairs, seen = [], set()
        self._pair_q = {}
        for i in np.argsort(q):
            if v[i] in seen:
                continue
            seen.add(v[i])
            pairs.append((int(v[i]), int(m[i])))
            self._pair_q[pairs[-1]] = float(q[i])
        return pairs

    def _run_cascade(self, step, eps_target=None, component=(0, 0), log_cb=None, track=False):
        """Cascade batches until no mode is unstable (mgkmc/aqs.py:480-680,
        kmc_simulator.py:577-794, small strain): each batch flips its modes in
        order (flip, prev, redraw; sub-stepped together under a 2D nonlinear
        law), then one re-solve.  Returns (batches, flips, truncated,
        stopped_by_eps): 3D stops with truncated = True when the batch count
        exceeds int(max_cascade_steps_pct * voxels), and with stopped_by_eps =
        True once eps_macro[component] >= eps_target after a batch (checked
        after ``log_cb(batch, n_flips)``, aqs.py:660-666); ``track`` appends
        the flips to flip_event_history as the reference's track_cascades does;
        2D stops after nx * ny + 1 batches.  ``log_cb(batch, n_flips, sig)``
        gets the mean stress after the batch's re-solve.  Linear-law
        simulations with the gaussian barrier generator and redraws run in
        the library's C++ cascade loop (mgk_kmc_cascade; same streams, same
        batches)."""
        if self._native_cascade_ok():
            return self._run_cascade_native(step, eps_target, component, log_cb, track)
        local = total = 0
        while True:
            pairs = self._cascade_pairs()
            if not pairs:
                break
            flog = self._flip_log is not None and self.ndim == 2
            if self.finite:
                if flog:
                    for v, m in pairs:
                        self._log_flip(v, m, self._pair_q[(v, m)], "cascade")
                self._flip_finite_pairs(pairs)
            elif self.law is not None and self.ndim == 2:
                if flog:
                    for v, m in pairs:
                        self._log_flip(v, m, self._pair_q[(v, m)], "cascade")
                self._flip_substepped(pairs)
            else:
                fp2 = self.fast_patching_enabled and self.ndim == 2
                for v, m in pairs:
                    if flog:
                        self._log_flip(v, m, self._pair_q[(v, m)], "cascade")
                    C = self._mode_host(v, m) if self.enable_thermal else None
                    self._flip(v, m, 1)
                    self._flip_heat(v, C)
                    if fp2:
                        # the catalog entry after its redraw (kmc_simulator.py:764-774)
                        self._patch_stamp(v, self._mode_host(v, m))
                if not fp2:
                    self.update_stresses()
            for v, m in pairs:
                rec = (step, local) + tuple(int(c) for c in self.kmc.voxel_coords(v)) + (m,)
                self.cascade_history.append(rec)
                if track and self.ndim == 3:
                    self.flip_event_history.append(tuple(int(r) for r in rec))
            total += len(pairs)
            if log_cb is not None:
                log_cb(local, len(pairs), self.sig_mean())
            if self.ndim == 3 and eps_target is not None and self.eps_macro[component] >= eps_target:
                return local + 1, total, False, True
            local += 1
            if self.ndim == 3 and local > int(self.max_cascade_steps_pct * self.total_voxels):
                print(f"Cascade limit reached ({local} steps > {self.max_cascade_steps_pct * 100:.1f}% of "
                      f"{self.total_voxels} voxels)")
                return local, total, True, False
            if self.ndim == 2 and local > self.total_voxels:
                break
        if self.fast_patching_enabled and self.ndim == 2 and not self.finite and self.law is None and total > 0:
            self.update_stresses()  # the corrector after a cascade (kmc_simulator.py:790-792)
        return local, total, False, False

    # flip and batch records of one mgk_kmc_cascade call; a cascade that fills
    # them continues in further calls
    cascade_records = (1 << 16, 1024)

    def _native_cascade_ok(self):
        gen = self.barrier_generator
        gp = getattr(gen, "params", None)
        same_stream = getattr(gen, "rng", 0) is (None if self.rng is np.random else self.rng)
        return (self.native_loop and self.law is None and not self.finite and not self.enable_thermal
                and not (self.fast_patching_enabled and self.ndim == 2)
                and self.kmc.planar and self._flip_log is None and self._sig_owner is None
                and gp is not None and same_stream and (self.redraw_directions or self.redraw_barriers)
                and hasattr(self.rng, "get_state") and (self._solved_once or not self.ls_warm_start))

    def _run_cascade_native(self, step, eps_target, component, log_cb, track):
        """_run_cascade through mgk_kmc_cascade: the library runs whole
        batches (barriers, unstable list, sort, flips, redraws, re-solve) and
        returns each batch's flips and mean stress; histories and per-batch
        logs are written here in batch order.  A batch whose re-solve fails
        the 2D stability check raises ValueError as update_stresses does,
        with that batch left out of the histories and the log."""
        d = self.ndim
        lib = _lib.load()
        L = self._loop_struct(np.zeros((d, d)))
        if self.kmc.Q is None:
            self.kmc.Q = dev.torch.zeros_like(self.kmc.Q0)
        if d == 3:
            maxb, trunc = int(self.max_cascade_steps_pct * self.total_voxels), 1
            ec = component[0] * d + component[1] if eps_target is not None else -1
        else:
            maxb, trunc, ec = self.total_voxels, 0, -1
        et = float(eps_target) if eps_target is not None else 0.0
        rng = _lib.mt_from(self.rng)
        em = (ctypes.c_double * 9)(*self.eps_macro.reshape(-1))
        cap, bcap = min(self.total_voxels, self.cascade_records[0]), self.cascade_records[1]
        flips = (_lib.MgkCascadeFlip * cap)()
        sizes = (ctypes.c_int * bcap)()
        bsig = (ctypes.c_double * (9 * bcap))()
        nf, nb, why, it32 = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        local = total = 0
        ran = False
        try:
            while True:
                rc = lib.mgk_kmc_cascade(ctypes.byref(L), ctypes.byref(rng), _lib.dptr(self.kmc.Q), self.time, em,
                                         maxb - local, trunc, ec, et, flips, cap, ctypes.byref(nf), sizes, bsig,
                                         bcap, ctypes.byref(nb), ctypes.byref(why), ctypes.byref(self.ls.info),
                                         ctypes.byref(it32), ctypes.c_void_p(dev.stream_ptr()))
                if rc < 0:
                    _lib.check(rc, "mgk_kmc_cascade")
                ran = ran or nb.value > 0
                n_ok = nb.value - int(why.value == _lib.MGK_CASCADE_UNSTABLE)
                nrec = sum(sizes[:n_ok])
                rec = np.frombuffer((ctypes.c_char * ctypes.sizeof(flips)).from_buffer(flips),
                                    dtype=_CASCADE_FLIP)[:nrec]
                cols = [np.full(nrec, step), rec["batch"] + local, *np.unravel_index(rec["v"], self.kmc.shape),
                        rec["m"]]
                recs = list(map(tuple, np.stack(cols, axis=1).astype(np.int64).tolist()))
                self.cascade_history.extend(recs)
                if track and d == 3:
                    self.flip_event_history.extend(recs)
                for b in range(n_ok):
                    total += sizes[b]
                    if log_cb is not None:
                        log_cb(local, sizes[b], np.array(bsig[9 * b: 9 * b + d * d]).reshape(d, d))
                    local += 1
                if why.value == _lib.MGK_CASCADE_CAPACITY:
                    if nb.value == 0:  # one batch holds more flips than the records: room for every voxel
                        cap = self.total_voxels
                        flips = (_lib.MgkCascadeFlip * cap)()
                    continue
                break
        finally:
            _lib.mt_to(self.rng, rng)
            if ran:
                self._solved_once = True
                self.ls_info = self.ls.info
                self._patched_mean = None
                if self.precision == MIXED:
                    self.ls.iters32 = it32.value
        if why.value == _lib.MGK_CASCADE_UNSTABLE:
            sm = np.array(self.ls.info.sig_mean[: d * d]).reshape(d, d)
            raise ValueError(f"Unstable stress in small-strain elastic_run: {sm / 1e9} GPa")
        if why.value == _lib.MGK_CASCADE_TRUNCATED:
            print(f"Cascade limit reached ({local} steps > {self.max_cascade_steps_pct * 100:.1f}% of "
                  f"{self.total_voxels} voxels)")
            return local, total, True, False
        return local, total, False, why.value == _lib.MGK_CASCADE_EPS_TARGET

    def _log_flip(self, v, m, q, flip_type):
        """One flipped_voxels_log line (mgkmc/kmc_simulator.py:535-557): the
        voxel's softening g_p + g_t (decayed over the time since its last
        event, capped) and the mode's barrier at selection."""
        soft = dev.download(self.kmc.soft[:, int(v)])
        tl = float(dev.download(self.kmc.tlast[int(v)]))
        gp, gt = float(soft[0]), float(soft[1])
        gtc = 0.0
        if gt > 0:
            if self.tau == np.inf:
                gtc = gt
            elif tl != -np.inf:
                gtc = gt * np.exp(-max(self.time - tl, 0.0) / self.tau)
        sft = gp + gtc
        if self.softening_cap > 0 and sft > self.softening_cap:
            sft = self.softening_cap
        x, y = (int(c) for c in self.kmc.voxel_coords(v))
        self._flip_log.write(self._flip_log_step, x, y, int(m), sft, float(q), flip_type)

    # ----- event sequences ----------------------------------------------------
    def kmc_events(self, remaining_time, strain_unit_tensor, max_events=None, max_seq=None, seq=0, on_event=None):
        """Events until the remaining time holds no further event ("time": the
        time and macro strain then advance by what was left), max_events events
        ("count") or more than max_seq consecutive unstable-mode events ("seq").
        Returns (n events, remaining time, reason, seq).  ``on_event(event,
        eps_macro, sig_mean, over_seq)`` is called after each event with the
        macro strain and mean stress after its re-solve (over_seq: this event
        ends the run on the unstable-sequence limit).  Linear-law KMC-mode
        simulations with the gaussian barrier generator run in the library's
        C++ loop (mgk_kmc_run; same streams, same sequence); others event by
        event through kmc_event."""
        if self._native_ok():
            return self._run_native(remaining_time, strain_unit_tensor, max_events, max_seq, seq, on_event)
        n = 0
        while True:
            ev, remaining_time = self.kmc_event(remaining_time, strain_unit_tensor)
            if ev is None:
                return n, remaining_time, "time", seq
            n += 1
            seq = seq + 1 if ev[-1] else 0
            over = max_seq is not None and seq > max_seq
            if on_event is not None:
                on_event(ev, self.eps_macro.copy(), self.sig_mean(), over)
            if over:
                return n, remaining_time, "seq", seq
            if max_events is not None and n >= max_events:
                return n, remaining_time, "count", seq

    def _native_ok(self):
        gen = self.barrier_generator
        gp = getattr(gen, "params", None)
        same_stream = getattr(gen, "rng", 0) is (None if self.rng is np.random else self.rng)
        return (self.native_loop and (self.law is None or self._native_law_ok())
                and (not self.finite or self._native_fs_ok())
                and not self.cascade_mode and not self.enable_thermal and self.kmc.planar and self._flip_log is None
                and gp is not None and same_stream and (self.redraw_directions or self.redraw_barriers)
                and hasattr(self.rng, "get_state") and hasattr(self.selection_stream, "get_state")
                and (self._solved_once or not self.ls_warm_start))

    def _loop_struct(self, unit, patch=True):
        k, ls, d = self.kmc, self.ls, self.ndim
        L = _lib.MgkKmcLoop()
        L.ndim, L.M = d, self.M
        for i, n in enumerate(self.grid_shape):
            L.shape[i] = n
        if self._redraw_dev is None:
            self._redraw_dev = dev.torch.zeros(3 * self.M, dtype=dev.torch.float64, device="cuda")
        L.Q0, L.cat, L.soft, L.tlast, L.prev = (_lib.dptr(k.Q0), _lib.dptr(k.cat), _lib.dptr(k.soft),
                                                _lib.dptr(k.tlast), _lib.dptr(k.prev))
        L.epsp, L.eps_precision, L.epsp_shadow = _lib.dptr(k.eps_p), k.eps_precision, _lib.dptr(k.eps_p32)
        L.kmc_ws, L.redraw_dev = k.ws.ptr, _lib.dptr(self._redraw_dev)
        L.kp = self._params()
        L.jp, L.jt, L.jn = self.jp, self.jt, self.neighbor_softening_fraction
        L.stability_threshold = self.stability_threshold
        L.redraw_directions, L.redraw_barriers = int(bool(self.redraw_directions)), int(bool(self.redraw_barriers))
        gp = getattr(self.barrier_generator, "params", None) or {}
        L.gamma0, L.q_mean, L.q_std = self.gamma0, float(gp.get("mean", 2.0)), float(gp.get("std", 0.6))
        lo, hi = gp.get("min_cutoff"), gp.get("max_cutoff")
        L.q_has_min, L.q_min = int(lo is not None), float(lo if lo is not None else 0.0)
        L.q_has_max, L.q_max = int(hi is not None), float(hi if hi is not None else 0.0)
        L.ls_precision, L.sig_kind = {FP64: 0, FP32: 1, MIXED: 2}[self.precision], self.sig_kind
        if ls is None:  # finite strain: the KMC fields and the Cauchy field of the barrier pass
            L.sig = _lib.dptr(self.fs.sig if d == 3 else self._fs_sig2.sig)
            L.strain_rate = float(self.strain_rate)
            u = np.asarray(unit, dtype=np.float64).reshape(-1)
            for i in range(d * d):
                L.unit[i] = u[i]
            return L
        L.lam, L.mu, L.lam0, L.mu0 = _lib.dptr(ls.lam), _lib.dptr(ls.mu), ls.lam0, ls.mu0
        L.eps, L.sig, L.ws, L.ws_bytes = _lib.dptr(ls.eps), _lib.dptr(ls.sig), _lib.dptr(ls.ws), ls.ws.numel()
        if self.precision == MIXED:
            L.lam32, L.mu32, L.eps32 = _lib.dptr(ls.lam32), _lib.dptr(ls.mu32), _lib.dptr(ls.eps32)
            L.ws32, L.ws32_bytes = _lib.dptr(ls.ws32), ls.ws32.numel()
        args = dict(max_iter=200, tol=1e-6)
        args.update(self.solver_args)
        L.warm = int(bool(self.ls_warm_start))
        L.max_iter, L.tol = int(args["max_iter"]), float(args["tol"])
        L.tol32 = max(float(args["tol"]), 1e-6)
        L.check_unstable = int(d == 2)
        L.strain_rate = float(self.strain_rate)
        u = np.asarray(unit, dtype=np.float64).reshape(-1)
        for i in range(d * d):
            L.unit[i] = u[i]
        if patch and self.fast_patching_enabled:
            self._make_sigma_macro_unit(unit)
            P = self._patch_cfg = _lib.MgkKmcPatchCfg()
            for i in range(len(self._patch_radius_c)):
                P.radius[i] = self._patch_radius_c[i]
            P.nbasis = len(self.patch_missing_mean)
            P.kernels = self._patch_K.data_ptr()
            for b, mm in enumerate(self.patch_missing_mean):
                for c, (i, j) in enumerate(dev.sym_pairs(d)):
                    P.missing_mean[b][c] = float(mm[i, j])
            P.sig_unit = self.sigma_macro_unit.data_ptr()
            P.workspace, P.workspace_bytes = self._patch_ws.data_ptr(), self._patch_ws.numel()
            P.sync_interval = int(self.sync_interval)
            self._patch_io = (ctypes.c_int(int(self.flips_since_sync)), ctypes.c_int(int(self._su_live)),
                              (ctypes.c_double * 9)(), ctypes.c_int(int(self._patched_mean is not None)))
            P.flips_since_sync = ctypes.pointer(self._patch_io[0])
            P.unit_live = ctypes.pointer(self._patch_io[1])
            P.sig_mean = ctypes.cast(self._patch_io[2], ctypes.POINTER(ctypes.c_double))
            P.mean_from_patch = ctypes.pointer(self._patch_io[3])
            L.patch = ctypes.addressof(P)
        return L

    def _native_mixed_ok(self):
        """Whether the macro relaxation runs in the library (mgk_kmc_mixed_bc):
        linear law, small strain, no cascade interleaving, the solver's fields
        current."""
        return (self.native_loop and self.law is None and not self.finite
                and (not self.cascade_mode or self.ndim == 2)  # the 3D loop interleaves cascades
                and self._sig_owner is None and (self._solved_once or not self.ls_warm_start))

    def _run_mixed_native(self, stress_targets, component, mixed_tol, mixed_max_iter):
        """The macro relaxation after an increment through mgk_kmc_mixed_bc (3D:
        mgkmc/aqs.py:1105-1142; 2D: kmc_simulator.py:1106-1143).  Returns
        (converged, last max error)."""
        d = self.ndim
        lib = _lib.load()
        L = self._loop_struct(np.zeros((d, d)), patch=False)  # every iteration solves
        items = list(stress_targets.items())
        idx = (ctypes.c_int * max(1, 2 * len(items)))(*[int(i) for k, _ in items for i in k])
        val = (ctypes.c_double * max(1, len(items)))(*[float(v) for _, v in items])
        em = (ctypes.c_double * 9)(*self.eps_macro.reshape(-1))
        conv, mx, it, it32 = ctypes.c_int(), ctypes.c_double(), ctypes.c_int(), ctypes.c_int()
        if self.ls_info is None:  # no solve yet: the zero initial stress
            for k in range(9):
                self.ls.info.sig_mean[k] = 0.0
        rc = lib.mgk_kmc_mixed_bc(ctypes.byref(L), len(items), idx, val, int(component[0]), int(component[1]),
                                  float(np.mean(self.E_field)), float(np.mean(self.nu_field)), float(mixed_tol),
                                  int(mixed_max_iter), em, ctypes.byref(conv), ctypes.byref(mx), ctypes.byref(it),
                                  ctypes.byref(self.ls.info), ctypes.byref(it32), ctypes.c_void_p(dev.stream_ptr()))
        self.eps_macro[...] = np.array(em[: d * d]).reshape(d, d)
        if it.value:
            self._solved_once = True
            self._patched_mean = None
            self.ls_info = self.ls.info
            if self.precision == MIXED:
                self.ls.iters32 = it32.value
        if rc == _lib.MGK_KMC_RUN_UNSTABLE:
            sm = np.array(self.ls.info.sig_mean[: d * d]).reshape(d, d)
            raise ValueError(f"Unstable stress in small-strain elastic_run: {sm / 1e9} GPa")
        _lib.check(rc, "mgk_kmc_mixed_bc")
        return bool(conv.value), mx.value

    def _take_native_events(self, evs, rc, it32, on_event=None):
        """Record events returned by the native loop (history, LS iteration
        totals, on_event callbacks) for this simulation."""
        d = self.ndim
        for i, e in enumerate(evs):
            coords = tuple(int(c) for c in self.kmc.voxel_coords(e.v))
            ev = coords + (int(e.m), float(e.time), bool(e.is_instab))
            self.kmc_event_history.append(ev)
            self.ls_iters_total += e.ls_iters + e.ls_iters32
            if self.finite or (self.law is not None and d == 2):
                self.kmc_substeps.append(int(e.substeps))
                self.last_substeps = int(e.substeps)
            if on_event is not None:
                over = rc == _lib.MGK_KMC_RUN_SEQ and i == len(evs) - 1
                on_event(ev, np.array(e.eps_macro[: d * d]).reshape(d, d),
                         np.array(e.sig_mean[: d * d]).reshape(d, d), over)
        if len(evs):
            self._solved_once = True
            if self.ls is not None:
                self.ls_info = self.ls.info
                if self.precision == MIXED:
                    self.ls.iters32 = it32

    def _run_native(self, remaining, unit, max_events, max_seq, seq, on_event=None):
        d = self.ndim
        lib = _lib.load()
        L = self._loop_struct(unit)
        if self.law is not None:  # each event's flip is the law's sub-stepped flip
            lcfg = self._law_substep_cfg()
            L.law = ctypes.addressof(lcfg)
        if self.finite:  # each event's flip is the sub-stepped finite-strain flip
            cfg = self._fs_substep_cfg(L)
            fm = (ctypes.c_double * 9)(*self.F_macro.reshape(-1))
            fsi = _lib.MgkFsInfo()
            cfg.F_macro, cfg.info = ctypes.addressof(fm), ctypes.addressof(fsi)
            L.fs = ctypes.addressof(cfg)
        ls_info = self.ls.info if self.ls is not None else _lib.MgkLsInfo()
        rng, sel = _lib.mt_from(self.rng), _lib.mt_from(self.selection_stream)
        t = ctypes.c_double(self.time)
        rem = ctypes.c_double(remaining)
        em = (ctypes.c_double * 9)(*self.eps_macro.reshape(-1))
        sq = ctypes.c_int(int(seq))
        ms = 2 ** 31 - 1 if max_seq is None else int(min(max_seq, 2 ** 31 - 1))
        n_total = 0
        it32 = ctypes.c_int()
        nb = ctypes.c_int()
        reason = "time"
        try:
            while True:
                chunk = 4096 if max_events is None else min(4096, max_events - n_total)
                evs = (_lib.MgkKmcEvent * chunk)()
                rc = lib.mgk_kmc_run(ctypes.byref(L), ctypes.byref(rng), ctypes.byref(sel), ctypes.byref(t), em,
                                     ctypes.byref(rem), chunk, ms, ctypes.byref(sq), evs, ctypes.byref(nb),
                                     ctypes.byref(ls_info), ctypes.byref(it32),
                                     ctypes.c_void_p(dev.stream_ptr()))
                self._take_native_events(evs[: nb.value], rc, it32.value, on_event)
                n_total += nb.value
                if self.finite and (nb.value or rc < 0):
                    self.F_macro = np.array(fm[: d * d]).reshape(d, d)
                    self.ls_info = fsi
                if self.finite and rc < 0:
                    self._fs_raise(rc)
                if self.law is not None:
                    self._law_after(lcfg)
                    if rc == _lib.MGK_ERR_SOLVER:
                        raise ValueError("Mechanical solver failed to converge under small-strain sub-stepping.")
                    if rc == _lib.MGK_ERR_NONFINITE:  # 3D Landau (LandauSolver.solve's error)
                        raise FloatingPointError(_lib.last_error().split(" (/")[0])
                if rc == _lib.MGK_KMC_RUN_UNSTABLE:
                    sm = np.array(self.ls.info.sig_mean[: d * d]).reshape(d, d)
                    raise ValueError(f"Unstable stress in small-strain elastic_run: {sm / 1e9} GPa")
                if rc < 0:
                    _lib.check(rc, "mgk_kmc_run")
                if rc == _lib.MGK_KMC_RUN_COUNT and (max_events is None or n_total < max_events):
                    continue
                reason = {_lib.MGK_KMC_RUN_TIME: "time", _lib.MGK_KMC_RUN_COUNT: "count",
                          _lib.MGK_KMC_RUN_SEQ: "seq"}[rc]
                break
        finally:
            _lib.mt_to(self.rng, rng)
            _lib.mt_to(self.selection_stream, sel)
            self.time = t.value
            self.eps_macro[...] = np.array(em[: d * d]).reshape(d, d)
            if self.fast_patching_enabled:
                fl, live, sm, from_patch = self._patch_io
                self.flips_since_sync = fl.value
                self._su_live = bool(live.value)
                if from_patch.value:
                    self._patched_mean = np.array(sm[: d * d]).reshape(d, d)
                elif n_total:
                    self._patched_mean = None
        return n_total, rem.value, reason, sq.value

    def footprint_bytes(self):
        """HBM bytes held by this simulation's device state."""
        t = 0
        if self.finite:
            f = self.fs
            for x in (f.lam, f.mu, f.F, f.P, f.sig, f.Fp, f.ws, getattr(f, "lam32", None), getattr(f, "mu32", None),
                      getattr(f, "ws32", None), self.kmc.Q0, self.kmc.cat, self.kmc.soft, self.kmc.tlast,
                      self.kmc.prev, self.kmc.eps_p):
                if x is not None:
                    t += x.numel() * x.element_size()
            return t
        for x in (self.ls.lam, self.ls.mu, self.ls.eps, self.ls.sig, self.ls.ws,
                  getattr(self.ls, "lam32", None), getattr(self.ls, "mu32", None), getattr(self.ls, "eps32", None),
                  getattr(self.ls, "ws32", None), getattr(self.ls, "tau", None), getattr(self.ls, "e33", None),
                  getattr(self.ls, "aa", None), self.kmc.Q0, self.kmc.Q, self.kmc.cat, self.kmc.soft,
                  self.kmc.tlast, self.kmc.prev, self.kmc.eps_p, self.kmc.eps_p32, self.kmc.ws.buf):
            if x is not None:
                t += x.numel() * x.element_size()
        return t


_CASCADE_FLIP = np.dtype({"names": ["batch", "v", "m"], "formats": ["<i4", "<i8", "<i4"], "offsets": [0, 8, 16],
                          "itemsize": 24})  # mgk_cascade_flip


def kmc_events_batch(sims, remaining, units, max_events, max_seq=None, seqs=None, stream=None):
    """Events of several simulations of one grid shape and LS configuration in
    lockstep (mgk_kmc_run_batch, at most MGK_LS_MAX_BATCH): per step one
    synchronisation for all rate totals, one for all selections and one
    batched LS re-solve.  Member i runs until its own stop (no event within
    remaining[i], max_events[i] events, or more than max_seq consecutive
    unstable-mode events); its events, fields, random streams, time and macro
    strain come out bit-identical to sims[i].kmc_events(remaining[i],
    units[i], max_events[i], max_seq, seqs[i]).  Returns one (n events,
    remaining time, reason, seq, error) tuple per member: reason "time",
    "count" or "seq", or "error" with the exception the single-simulation
    call would raise (the others are unaffected by it)."""
    n = len(sims)
    if not 1 <= n <= _lib.MGK_LS_MAX_BATCH:
        raise ValueError(f"batch of {n} simulations (1..{_lib.MGK_LS_MAX_BATCH})")
    for sm in sims:
        if not sm._native_ok() or sm.finite:
            raise ValueError("batched events need simulations that run in the native loop "
                             "(linear law, small strain, KMC mode, gaussian barriers, planar catalog)")
    lib = _lib.load()
    seqs = [0] * n if seqs is None else list(seqs)
    cap = max(int(m) for m in max_events)
    loops = (_lib.MgkKmcLoop * n)()
    rngs = (_lib.MgkMtState * n)()
    sels = (_lib.MgkMtState * n)()
    t = (ctypes.c_double * n)()
    em = (ctypes.c_double * (9 * n))()
    rem = (ctypes.c_double * n)()
    mx = (ctypes.c_int * n)()
    sq = (ctypes.c_int * n)()
    for i, sm in enumerate(sims):
        loops[i] = sm._loop_struct(units[i])
        rngs[i] = _lib.mt_from(sm.rng)
        sels[i] = _lib.mt_from(sm.selection_stream)
        t[i] = sm.time
        e = sm.eps_macro.reshape(-1)
        for k in range(e.size):
            em[9 * i + k] = float(e[k])
        rem[i] = float(remaining[i])
        mx[i] = int(max_events[i])
        sq[i] = int(seqs[i])
    evs = (_lib.MgkKmcEvent * (n * cap))()
    nev = (ctypes.c_int * n)()
    status = (ctypes.c_int * n)()
    infos = (_lib.MgkLsInfo * n)()
    it32 = (ctypes.c_int * n)()
    ms = 2 ** 31 - 1 if max_seq is None else int(min(max_seq, 2 ** 31 - 1))
    s = dev.stream_ptr() if stream is None else stream
    rc = lib.mgk_kmc_run_batch(n, loops, rngs, sels, t, em, rem, mx, ms, sq, evs, cap, nev, status, infos, it32,
                               ctypes.c_void_p(s))
    out = []
    for i, sm in enumerate(sims):
        d = sm.ndim
        err = None
        _lib.mt_to(sm.rng, rngs[i])
        _lib.mt_to(sm.selection_stream, sels[i])
        sm.time = t[i]
        sm.eps_macro[...] = np.array(em[9 * i: 9 * i + d * d]).reshape(d, d)
        if nev[i] or status[i] == _lib.MGK_KMC_RUN_UNSTABLE:
            ctypes.memmove(ctypes.byref(sm.ls.info), ctypes.byref(infos[i]), ctypes.sizeof(infos[i]))
        sm._take_native_events(evs[i * cap: i * cap + nev[i]], status[i], int(it32[i]))
        st = status[i]
        if st == _lib.MGK_KMC_RUN_UNSTABLE:
            smn = np.array(sm.ls.info.sig_mean[: d * d]).reshape(d, d)
            err = ValueError(f"Unstable stress in small-strain elastic_run: {smn / 1e9} GPa")
            reason = "error"
        elif st < 0:
            err = RuntimeError(f"mgk_kmc_run_batch member {i} failed with status {st}: {_lib.last_error()}")
            reason = "error"
        else:
            reason = {_lib.MGK_KMC_RUN_TIME: "time", _lib.MGK_KMC_RUN_COUNT: "count",
                      _lib.MGK_KMC_RUN_SEQ: "seq"}[st]
        out.append((int(nev[i]), float(rem[i]), reason, int(sq[i]), err))
    if rc < 0:
        _lib.check(rc, "mgk_kmc_run_batch")
    return out
