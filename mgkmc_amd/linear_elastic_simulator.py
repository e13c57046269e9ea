"""Small-strain Lippmann-Schwinger solvers on the GPU (drop-ins).

``spectral_solver_2d`` / ``spectral_solver_3d`` keep the reference signatures
(mgkmc/linear_elastic_simulator.py:12-15 and :274-277) and return the same
(eps, sig, eps_macro, sig_macro) tuple of host arrays in the reference's
(nx, ny[, nz], d, d) layout.  The iteration itself (same initial guess, same
stopping test ||eps_new - eps||_F / (||eps||_F + 1e-20) < tol over the full
tensor, same silent stop at max_iter) runs in libmgkmc_hip.so.

``Gamma`` is accepted for signature compatibility and ignored: the Green
operator is evaluated from k in registers inside the FFT kernels, for the
reference medium (mean E, mean nu), which is what every caller of the
reference builds its Gamma from.

``LsSolver`` is the device-resident form used by the simulators and the
serving loop: moduli, eps, sig and eps_plastic stay in HBM between solves.
"""
import ctypes

import numpy as np

from . import _lib
from . import device as dev
from .elasticity import compute_lame_2d, compute_lame_3d

FP64 = 0
FP32 = 1
MIXED = 2  # FP32 iterations, then FP64 refinement iterations to the FP64 stop test


class LsSolver:
    """Device-resident LS solver for one grid.

    Fields (torch tensors on the GPU, SoA symmetric components):
      lam, mu : (nx, ny[, nz])
      eps, sig: (C, nx, ny[, nz]), C = 6 (3D) or 3 (2D)
    """

    def __init__(self, E, nu, plane_mode="plane_strain", precision=FP64, lam0=None, mu0=None):
        dev.require_gpu()
        E = np.asarray(E, dtype=np.float64)
        nu = np.asarray(nu, dtype=np.float64)
        self.shape = E.shape
        self.d = E.ndim
        if self.d not in (2, 3):
            raise ValueError("E must be a 2D or 3D field")
        self.C = 6 if self.d == 3 else 3
        self.precision = precision
        if self.d == 3:
            lam, mu = compute_lame_3d(E, nu)
            l0, m0 = compute_lame_3d(E.mean(), nu.mean())
        else:
            lam, mu = compute_lame_2d(E, nu, plane_mode)
            l0, m0 = compute_lame_2d(E.mean(), nu.mean(), plane_mode)
        self.lam0 = float(l0 if lam0 is None else lam0)
        self.mu0 = float(m0 if mu0 is None else mu0)
        self.mu_host = mu
        fp = FP64 if precision == MIXED else precision  # precision of lam, mu, eps, sig
        self.lam = dev.upload(lam, fp)
        self.mu = dev.upload(mu, fp)
        dt = dev.torch_dtype(fp)
        self.eps = dev.torch.zeros((self.C,) + self.shape, dtype=dt, device="cuda")
        self.sig = dev.torch.zeros((self.C,) + self.shape, dtype=dt, device="cuda")
        self._shape_c = _lib.shape_arg(self.shape)
        self.ws = self._workspace(fp)
        self.info = _lib.MgkLsInfo()
        self.iters32 = 0
        if precision == MIXED:
            self.lam32 = dev.upload(lam, FP32)
            self.mu32 = dev.upload(mu, FP32)
            self.eps32 = dev.torch.zeros((self.C,) + self.shape, dtype=dev.torch.float32, device="cuda")
            self.ws32 = self._workspace(FP32)
            self._iters32 = ctypes.c_int()

    def _workspace(self, precision):
        nbytes = ctypes.c_size_t()
        rc = _lib.load().mgk_ls_workspace_bytes(self.d, self._shape_c, precision, ctypes.byref(nbytes))
        _lib.check(rc, "mgk_ls_workspace_bytes")
        return dev.torch.empty(int(nbytes.value), dtype=dev.torch.uint8, device="cuda")

    @staticmethod
    def footprint_bytes(shape, precision):
        """HBM bytes a solver of this grid and precision holds (fields + workspaces)."""
        d = len(shape)
        C = 6 if d == 3 else 3
        NV = int(np.prod(shape))
        sc = _lib.shape_arg(shape)
        lib = _lib.load()

        def ws(p):
            nb = ctypes.c_size_t()
            _lib.check(lib.mgk_ls_workspace_bytes(d, sc, p, ctypes.byref(nb)), "mgk_ls_workspace_bytes")
            return int(nb.value)
        if precision == MIXED:
            return NV * 8 * (2 + 2 * C) + ws(FP64) + NV * 4 * (2 + C) + ws(FP32)
        s = 8 if precision == FP64 else 4
        return NV * s * (2 + 2 * C) + ws(precision)

    def solve(self, eps_bar, eps_plastic=None, warm=False, max_iter=200, tol=1e-6, stream=None,
              eps_plastic32=None, tol32=None):
        """Run the LS fixed point; eps/sig are updated in place on the device.
        eps_plastic: device tensor (C, ...) or None.  warm: start from self.eps
        (the reference's eps_init; in the MIXED form from self.eps32).
        MIXED form: FP32 iterations to tol32 (default max(tol, 1e-6)), then FP64
        iterations to tol; eps_plastic32 is the FP32 copy of eps_plastic (made
        here if not given).  Returns the mgk_ls_info struct of the final stage."""
        eb, ebp = _lib.host_dbl(np.asarray(eps_bar, dtype=np.float64).reshape(self.d, self.d))
        if eps_plastic is not None:
            if eps_plastic.dtype != self.eps.dtype or tuple(eps_plastic.shape) != tuple(self.eps.shape):
                raise ValueError("eps_plastic must be a device tensor of shape (C, ...) and solver dtype")
            eps_plastic = eps_plastic.contiguous()
        s = dev.stream_ptr() if stream is None else stream
        if self.precision == MIXED:
            if eps_plastic is not None and eps_plastic32 is None:
                eps_plastic32 = eps_plastic.to(dev.torch.float32)
            if eps_plastic32 is not None and (eps_plastic32.dtype != dev.torch.float32
                                              or tuple(eps_plastic32.shape) != tuple(self.eps.shape)):
                raise ValueError("eps_plastic32 must be an FP32 device tensor of shape (C, ...)")
            t32 = max(float(tol), 1e-6) if tol32 is None else float(tol32)
            rc = _lib.load().mgk_ls_solve_refined(
                self.d, self._shape_c, _lib.dptr(self.lam32), _lib.dptr(self.mu32), _lib.dptr(self.lam),
                _lib.dptr(self.mu), self.lam0, self.mu0, ebp, _lib.dptr(eps_plastic32), _lib.dptr(eps_plastic),
                _lib.dptr(self.eps32), _lib.dptr(self.eps), _lib.dptr(self.sig), 1 if warm else 0, int(max_iter),
                float(tol), t32, _lib.dptr(self.ws32), self.ws32.numel(), _lib.dptr(self.ws), self.ws.numel(),
                ctypes.c_void_p(s), ctypes.byref(self.info), ctypes.byref(self._iters32))
            _lib.check(rc, "mgk_ls_solve_refined")
            self.iters32 = self._iters32.value
            return self.info
        rc = _lib.load().mgk_ls_solve(
            self.d, self._shape_c, self.precision, _lib.dptr(self.lam), _lib.dptr(self.mu),
            self.lam0, self.mu0, ebp, _lib.dptr(eps_plastic), _lib.dptr(self.eps), _lib.dptr(self.sig),
            1 if warm else 0, int(max_iter), float(tol), _lib.dptr(self.ws), self.ws.numel(),
            ctypes.c_void_p(s), ctypes.byref(self.info))
        _lib.check(rc, "mgk_ls_solve")
        return self.info



def solve_batch(solvers, eps_bars, eps_plastics=None, warm=False, max_iter=200, tol=1e-6, stream=None,
                eps_plastics32=None, tol32=None):
    """Solve several LsSolvers of one grid shape and precision together: one
    launch per sweep for all of them (mgk_ls_solve_batch /
    mgk_ls_solve_refined_batch, at most MGK_LS_MAX_BATCH).  Each solver's
    fields, info and iteration counts come out bit-identical to its own
    LsSolver.solve with the same arguments.  Returns the list of info structs
    (copies)."""
    n = len(solvers)
    if not 1 <= n <= _lib.MGK_LS_MAX_BATCH:
        raise ValueError(f"batch of {n} solvers (1..{_lib.MGK_LS_MAX_BATCH})")
    s0 = solvers[0]
    if any(x.shape != s0.shape or x.precision != s0.precision for x in solvers):
        raise ValueError("a batch needs one grid shape and one precision")
    eps_plastics = [None] * n if eps_plastics is None else list(eps_plastics)
    eps_plastics32 = [None] * n if eps_plastics32 is None else list(eps_plastics32)
    mem = (_lib.MgkLsMember * n)()
    keep = []
    for i, (x, eb, ep) in enumerate(zip(solvers, eps_bars, eps_plastics)):
        _, ebp = _lib.host_dbl(np.asarray(eb, dtype=np.float64).reshape(x.d, x.d))
        keep.append(ebp)
        if ep is not None:
            if ep.dtype != x.eps.dtype or tuple(ep.shape) != tuple(x.eps.shape):
                raise ValueError("eps_plastic must be a device tensor of shape (C, ...) and solver dtype")
            ep = ep.contiguous()
            keep.append(ep)
        m = mem[i]
        m.lam, m.mu, m.lam0, m.mu0 = _lib.dptr(x.lam), _lib.dptr(x.mu), x.lam0, x.mu0
        m.eps_bar = ebp
        m.eps_plastic, m.eps, m.sig = _lib.dptr(ep), _lib.dptr(x.eps), _lib.dptr(x.sig)
        m.workspace, m.workspace_bytes = _lib.dptr(x.ws), x.ws.numel()
        if x.precision == MIXED:
            ep32 = eps_plastics32[i]
            if ep is not None and ep32 is None:
                ep32 = ep.to(dev.torch.float32)
            keep.append(ep32)
            m.lam32, m.mu32, m.eps_plastic32 = _lib.dptr(x.lam32), _lib.dptr(x.mu32), _lib.dptr(ep32)
            m.eps32, m.workspace32, m.workspace32_bytes = _lib.dptr(x.eps32), _lib.dptr(x.ws32), x.ws32.numel()
    s = dev.stream_ptr() if stream is None else stream
    infos = (_lib.MgkLsInfo * n)()
    lib = _lib.load()
    if s0.precision == MIXED:
        it32 = (ctypes.c_int * n)()
        t32 = max(float(tol), 1e-6) if tol32 is None else float(tol32)
        rc = lib.mgk_ls_solve_refined_batch(s0.d, s0._shape_c, n, mem, 1 if warm else 0, int(max_iter), float(tol),
                                            t32, ctypes.c_void_p(s), infos, it32)
        _lib.check(rc, "mgk_ls_solve_refined_batch")
        for x, v in zip(solvers, it32):
            x.iters32 = int(v)
    else:
        rc = lib.mgk_ls_solve_batch(s0.d, s0._shape_c, s0.precision, n, mem, 1 if warm else 0, int(max_iter),
                                    float(tol), ctypes.c_void_p(s), infos)
        _lib.check(rc, "mgk_ls_solve_batch")
    out = []
    for x, inf in zip(solvers, infos):
        ctypes.memmove(ctypes.byref(x.info), ctypes.byref(inf), ctypes.sizeof(inf))
        c = _lib.MgkLsInfo()
        ctypes.memmove(ctypes.byref(c), ctypes.byref(inf), ctypes.sizeof(inf))
        out.append(c)
    return out


_cache = {}


def _solver_for(E, nu, plane_mode):
    """Cache one device solver per (shape, moduli) so repeated calls from a
    driver reuse HBM buffers and the uploaded moduli."""
    key = (E.shape, plane_mode, E.__array_interface__["data"][0], nu.__array_interface__["data"][0])
    s = _cache.get(key)
    if s is not None and np.array_equal(s._E, E) and np.array_equal(s._nu, nu):
        return s
    s = LsSolver(E, nu, plane_mode=plane_mode, precision=FP64)
    s._E = E.copy()
    s._nu = nu.copy()
    _cache.clear()
    _cache[key] = s
    return s


def _drop_in(d, E, nu, eps_bar, eps_plastic, max_iter, tol, plane_mode, eps_init, info=None):
    E = np.ascontiguousarray(E, dtype=np.float64)
    nu = np.ascontiguousarray(nu, dtype=np.float64)
    if E.ndim != d:
        raise ValueError(f"E must be {d}D")
    eps_bar = np.asarray(eps_bar, dtype=np.float64)
    solver = _solver_for(E, nu, plane_mode)
    a_bar = 0.5 * (eps_bar - eps_bar.T)
    ep_dev = None
    a_p = None
    if eps_plastic is not None:
        sym, a_p = dev.to_sym_soa(eps_plastic, d)
        if not np.any(a_p):
            a_p = None
        ep_dev = dev.upload(sym, FP64)
    warm = eps_init is not None
    if warm:
        sym, a_i = dev.to_sym_soa(eps_init, d)
        if np.any(a_i != a_i.reshape(-1, d, d)[0]):
            raise ValueError("eps_init with a non-uniform antisymmetric part is not supported")
        solver.eps.copy_(dev.upload(sym, FP64))
    inf = solver.solve(eps_bar, ep_dev, warm=warm, max_iter=max_iter, tol=tol)
    if info is not None:
        info.update(iters=inf.iters, diff=inf.diff, status=inf.status)
    eps = dev.from_sym_soa(dev.download(solver.eps), d)
    sig = dev.from_sym_soa(dev.download(solver.sig), d)
    if np.any(a_bar):
        eps += a_bar
    eps_macro = np.array(inf.eps_mean[: d * d]).reshape(d, d)
    sig_macro = np.array(inf.sig_mean[: d * d]).reshape(d, d)
    if np.any(a_bar) or a_p is not None:
        # antisymmetric stress 2 mu (a_bar - a_p): pointwise, no solve involved
        anti = a_bar - (a_p if a_p is not None else 0.0)
        sig += 2.0 * solver.mu_host[..., None, None] * anti
        sig_macro = sig.mean(axis=tuple(range(d)))
    return eps, sig, eps_macro, sig_macro


def spectral_solver_2d(E, nu, eps_bar, eps_plastic=None, max_iter=200, tol=1e-6,
                       verbose=False, pixel=1.0, plane_mode="plane_strain",
                       Gamma=None, eps_init=None):
    """2D LS solver (mgkmc/linear_elastic_simulator.py:12-88) on the GPU."""
    return _drop_in(2, E, nu, eps_bar, eps_plastic, max_iter, tol, plane_mode, eps_init)


def spectral_solver_3d(E, nu, eps_bar, eps_plastic=None, max_iter=200, tol=1e-6,
                       verbose=False, pixel=1.0, Gamma=None, eps_init=None):
    """3D LS solver (mgkmc/linear_elastic_simulator.py:274-376) on the GPU."""
    return _drop_in(3, E, nu, eps_bar, eps_plastic, max_iter, tol, "plane_strain", eps_init)


def _simulation(d, E, nu, target_strain_mask, target_values, n_steps, pixel, tol_ls, max_iter_ls, tol_macro,
                max_iter_macro, store, plane_mode, precision=FP64, info=None):
    """Mixed stress/strain load stepping on one device solver (the reference's
    Gamma-once + warm-start structure): per step, macro iterations restart from
    the previous step's converged field; the stress-controlled entries of
    eps_bar get (err - nu_avg tr(err) I) / E_avg until max |err| < tol_macro."""
    E = np.ascontiguousarray(E, dtype=np.float64)
    nu = np.ascontiguousarray(nu, dtype=np.float64)
    if E.ndim != d:
        raise ValueError(f"E must be {d}D")
    solver = LsSolver(E, nu, plane_mode=plane_mode, precision=precision)
    torch = dev.torch
    E_avg, nu_avg = E.mean(), nu.mean()
    mu_mean = float(solver.mu_host.mean())
    mask = np.asarray(target_strain_mask, dtype=bool)
    smask = ~mask
    tv = np.asarray(target_values, dtype=np.float64)
    eb = np.zeros((d, d))
    warm = None  # device copy of the previous step's field
    warm32 = None
    em_l, sm_l, e_l, s_l, its = [], [], [], [], []
    for s in range(n_steps + 1):
        tgt = tv * (s / n_steps)
        max_err = 0.0
        for _ in range(max_iter_macro):
            eb[mask] = tgt[mask]
            if warm is not None:
                solver.eps.copy_(warm)
                if precision == MIXED:
                    solver.eps32.copy_(warm32)
            inf = solver.solve(eb, None, warm=warm is not None, max_iter=max_iter_ls, tol=tol_ls)
            its.append(inf.iters + solver.iters32)
            a_bar = 0.5 * (eb - eb.T)
            epsM = np.array(inf.eps_mean[: d * d]).reshape(d, d)
            sigM = np.array(inf.sig_mean[: d * d]).reshape(d, d) + 2.0 * mu_mean * a_bar
            err = np.zeros((d, d))
            err[smask] = tgt[smask] - sigM[smask]
            max_err = np.max(np.abs(err[smask])) if np.any(smask) else 0.0
            if max_err < tol_macro:
                break
            de = (err - nu_avg * np.trace(err) * np.eye(d)) / E_avg
            eb[smask] += de[smask]
        warm = solver.eps.clone()
        if precision == MIXED:
            warm32 = solver.eps32.clone()
        em_l.append(epsM)
        sm_l.append(sigM)
        if store:
            eps = dev.from_sym_soa(dev.download(solver.eps), d)
            sig = dev.from_sym_soa(dev.download(solver.sig), d)
            if np.any(a_bar):
                eps += a_bar
                sig += 2.0 * solver.mu_host[..., None, None] * a_bar
            e_l.append(eps)
            s_l.append(sig)
    if info is not None:
        info["ls_iters"] = its
    return np.array(em_l), np.array(sm_l), e_l, s_l


def linear_elastic_simulation_2d(E, nu, target_strain_mask, target_values, n_steps=20, pixel=1.0, tol_ls=1e-6,
                                 max_iter_ls=200, tol_macro=1e-4, max_iter_macro=20, store=True,
                                 plane_mode="plane_strain", log_path=None, global_log_path=None,
                                 driving_component=(0, 0), enable_console=True, vtk_path=None, vtk_interval="none",
                                 **solver_kw):
    """2D mixed-BC load stepping (mgkmc/linear_elastic_simulator.py:90-272) on
    the GPU.  Returns (eps_macro (n+1,2,2), sig_macro (n+1,2,2), eps list, sig
    list).  Log, VTK and checkpoint options are accepted and ignored;
    ``precision`` / ``info`` (extensions) select the LS form / receive the LS
    iteration counts."""
    return _simulation(2, E, nu, target_strain_mask, target_values, n_steps, pixel, tol_ls, max_iter_ls, tol_macro,
                       max_iter_macro, store, plane_mode, solver_kw.get("precision", FP64), solver_kw.get("info"))


def linear_elastic_simulation_3d(E, nu, target_strain_mask, target_values, n_steps=20, pixel=1.0, tol_ls=1e-6,
                                 max_iter_ls=200, tol_macro=1e-4, max_iter_macro=20, store=True, log_path=None,
                                 global_log_path=None, driving_component=(0, 0), enable_console=True, vtk_path=None,
                                 vtk_interval="none", **solver_kw):
    """3D mixed-BC load stepping (mgkmc/linear_elastic_simulator.py:378-566) on
    the GPU; as linear_elastic_simulation_2d."""
    return _simulation(3, E, nu, target_strain_mask, target_values, n_steps, pixel, tol_ls, max_iter_ls, tol_macro,
                       max_iter_macro, store, "plane_strain", solver_kw.get("precision", FP64),
                       solver_kw.get("info"))


def landau_params(v1, v2, v3, g1, g2, g3, g4, mu, strain_capping_enabled=False, strain_capping_limit=None,
                  strain_capping_tangent_ratio=0.1, strain_capping_type="piecewise", strain_capping_smooth_power=1.0):
    """mgk_landau_params with the reference's resolved capping strain
    (mgkmc/elasticity.py:312-319: sqrt(0.2 <mu> / |g4|) if g4 < 0 else 999 when
    no positive limit is given)."""
    if strain_capping_limit is None or strain_capping_limit <= 0:
        mu_val = float(np.mean(mu))
        ecap = float(np.sqrt(0.2 * mu_val / abs(g4))) if g4 < 0 else 999.0
    else:
        ecap = float(strain_capping_limit)
    return _lib.MgkLandauParams(float(v1), float(v2), float(v3), float(g1), float(g2), float(g3), float(g4),
                                1 if strain_capping_enabled else 0, 1 if strain_capping_type == "smooth" else 0,
                                ecap, float(strain_capping_smooth_power), float(strain_capping_tangent_ratio))


class LandauSolver:
    """Device-resident 3D Landau LS solver (moduli uploaded once)."""

    def __init__(self, lam, mu, precision=FP64):
        dev.require_gpu()
        lam = np.ascontiguousarray(lam, dtype=np.float64)
        mu = np.ascontiguousarray(mu, dtype=np.float64)
        if lam.ndim != 3 or lam.shape != mu.shape:
            raise ValueError("lam and mu must be 3D fields of one shape")
        self.shape = lam.shape
        self.precision = precision
        self.lam0, self.mu0 = float(lam.mean()), float(mu.mean())
        self.mu_host = mu
        self.lam = dev.upload(lam, precision)
        self.mu = dev.upload(mu, precision)
        dt = dev.torch_dtype(precision)
        self.eps = dev.torch.zeros((6,) + self.shape, dtype=dt, device="cuda")
        self.sig = dev.torch.zeros((6,) + self.shape, dtype=dt, device="cuda")
        self.tau = dev.torch.empty((6,) + self.shape, dtype=dt, device="cuda")
        self._shape_c = _lib.shape_arg(self.shape)
        nb = ctypes.c_size_t()
        _lib.check(_lib.load().mgk_ls_workspace_bytes(3, self._shape_c, precision, ctypes.byref(nb)),
                   "mgk_ls_workspace_bytes")
        self.ws = dev.torch.empty(int(nb.value), dtype=dev.torch.uint8, device="cuda")
        self.info = _lib.MgkLsInfo()

    def solve(self, params, eps_bar, eps_plastic=None, warm=False, max_iter=400, tol=1e-6, stream=None):
        eb, ebp = _lib.host_dbl(np.asarray(eps_bar, dtype=np.float64).reshape(3, 3))
        s = dev.stream_ptr() if stream is None else stream
        rc = _lib.load().mgk_ls_solve_landau(
            self._shape_c, self.precision, _lib.dptr(self.lam), _lib.dptr(self.mu), self.lam0, self.mu0,
            ctypes.byref(params), ebp, _lib.dptr(eps_plastic), _lib.dptr(self.eps), _lib.dptr(self.sig),
            _lib.dptr(self.tau), 1 if warm else 0, int(max_iter), float(tol), _lib.dptr(self.ws), self.ws.numel(),
            ctypes.c_void_p(s), ctypes.byref(self.info))
        _lib.check(rc, "mgk_ls_solve_landau")
        if rc == _lib.MGK_ERR_NONFINITE:
            raise FloatingPointError(
                "spectral_solver_landau_3d: non-finite strain update in LS loop (Landau energy likely unstable at "
                "this strain; consider strain capping or smaller load steps).")
        return self.info


def spectral_solver_landau_3d(lam, mu, v1, v2, v3, g1, g2, g3, g4, eps_bar, eps_plastic=None, max_iter=400, tol=1e-6,
                              verbose=False, pixel=1.0, Gamma=None, eps_init=None, strain_capping_enabled=False,
                              strain_capping_limit=None, strain_capping_tangent_ratio=0.1,
                              strain_capping_type="piecewise", strain_capping_smooth_power=1.0, solver="dbfft"):
    """Nonlinear LS with the 3D Landau stress on the GPU
    (mgkmc/linear_elastic_simulator.py:1231-1350); reference medium from
    (mean lam, mean mu); FloatingPointError on a non-finite update.  Strains
    must be symmetric (the simulators' are)."""
    if solver == "newton_cg":
        raise NotImplementedError("Newton-CG solver is not yet implemented for 3D small-strain. Please use 2D or "
                                  "default to dbfft solver.")
    lam = np.asarray(lam, dtype=np.float64)
    mu = np.asarray(mu, dtype=np.float64)
    eps_bar = np.asarray(eps_bar, dtype=np.float64)
    if np.any(eps_bar != eps_bar.T):
        raise ValueError("the device Landau solver takes symmetric eps_bar")
    solver_ = LandauSolver(lam, mu, FP64)
    ep = None
    if eps_plastic is not None:
        sym, a = dev.to_sym_soa(eps_plastic, 3)
        if np.any(a):
            raise ValueError("the device Landau solver takes a symmetric eps_plastic")
        ep = dev.upload(sym, FP64)
    warm = eps_init is not None
    if warm:
        sym, a = dev.to_sym_soa(eps_init, 3)
        if np.any(a):
            raise ValueError("the device Landau solver takes a symmetric eps_init")
        solver_.eps.copy_(dev.upload(sym, FP64))
    prm = landau_params(v1, v2, v3, g1, g2, g3, g4, mu, strain_capping_enabled, strain_capping_limit,
                        strain_capping_tangent_ratio, strain_capping_type, strain_capping_smooth_power)
    inf = solver_.solve(prm, eps_bar, ep, warm=warm, max_iter=max_iter, tol=tol)
    eps = dev.from_sym_soa(dev.download(solver_.eps), 3)
    sig = dev.from_sym_soa(dev.download(solver_.sig), 3)
    return eps, sig, np.array(inf.eps_mean[:9]).reshape(3, 3), np.array(inf.sig_mean[:9]).reshape(3, 3)


class LandauSolver2D:
    """Device-resident 2D Landau LS solver (FP64, plane strain or plane stress)."""

    def __init__(self, lam, mu, plane_mode="plane_strain", anderson_m=5):
        dev.require_gpu()
        lam = np.ascontiguousarray(lam, dtype=np.float64)
        mu = np.ascontiguousarray(mu, dtype=np.float64)
        if lam.ndim != 2 or lam.shape != mu.shape:
            raise ValueError("lam and mu must be 2D fields of one shape")
        if not 0 <= int(anderson_m) <= 5:
            raise ValueError("anderson_m must be in 0..5")
        self.shape = lam.shape
        NV = lam.size
        self.plane_stress = plane_mode == "plane_stress"
        self.anderson_m = int(anderson_m)
        self.lam0, self.mu0 = float(lam.mean()), float(mu.mean())
        self.mu_host = mu
        self.lam = dev.upload(lam, FP64)
        self.mu = dev.upload(mu, FP64)
        t = dev.torch
        self.eps = t.zeros((3,) + self.shape, dtype=t.float64, device="cuda")
        self.sig = t.zeros((3,) + self.shape, dtype=t.float64, device="cuda")
        self.tau = t.empty((3,) + self.shape, dtype=t.float64, device="cuda")
        self.e33 = t.zeros(self.shape, dtype=t.float64, device="cuda")
        self.e33_have = False
        m = self.anderson_m
        self.aa = t.empty((2 * (m + 1) + 1) * 3 * NV, dtype=t.float64, device="cuda") if m > 0 else None
        self._shape_c = _lib.shape_arg(self.shape)
        nb = ctypes.c_size_t()
        _lib.check(_lib.load().mgk_ls_workspace_bytes(2, self._shape_c, FP64, ctypes.byref(nb)),
                   "mgk_ls_workspace_bytes")
        self.ws = t.empty(int(nb.value), dtype=t.uint8, device="cuda")
        self.info = _lib.MgkLsInfo()

    def solve(self, params, eps_bar, eps_plastic=None, warm=False, max_iter=400, tol=1e-6, persist_e33=True,
              stream=None):
        eb, ebp = _lib.host_dbl(np.asarray(eps_bar, dtype=np.float64).reshape(2, 2))
        s = dev.stream_ptr() if stream is None else stream
        rc = _lib.load().mgk_ls_solve_landau2d(
            self._shape_c, _lib.dptr(self.lam), _lib.dptr(self.mu), self.lam0, self.mu0, ctypes.byref(params),
            1 if self.plane_stress else 0, ebp, _lib.dptr(eps_plastic), _lib.dptr(self.eps), _lib.dptr(self.sig),
            _lib.dptr(self.e33), 1 if persist_e33 else 0, 1 if (persist_e33 and self.e33_have) else 0,
            _lib.dptr(self.tau), self.anderson_m, _lib.dptr(self.aa), 1 if warm else 0, int(max_iter), float(tol),
            _lib.dptr(self.ws), self.ws.numel(), ctypes.c_void_p(s), ctypes.byref(self.info))
        _lib.check(rc, "mgk_ls_solve_landau2d")
        if persist_e33 and self.plane_stress:
            self.e33_have = True
        if rc == _lib.MGK_ERR_NONFINITE:
            raise FloatingPointError(
                "spectral_solver_landau_2d: non-finite update (Landau energy likely unstable at this strain; "
                "consider strain capping or smaller load steps).")
        return self.info

    def solve_ncg(self, params, eps_bar, eps_plastic=None, warm=False, max_iter=400, tol=1e-6, stream=None):
        """The reference's solver="newton_cg" (mgkmc/linear_elastic_simulator.py:1011-1117):
        Newton steps with BiCGSTAB on G C_h G (central differences of the Landau
        stress, h = 1e-7), stop when |dx| / |eps| < tol after the first step.
        info.iters = Newton steps; self.krylov_iters = BiCGSTAB iterations."""
        t = dev.torch
        lib = _lib.load()
        if getattr(self, "ws_ncg", None) is None:
            nb = ctypes.c_size_t()
            _lib.check(lib.mgk_ls_landau2d_ncg_workspace_bytes(self._shape_c, ctypes.byref(nb)),
                       "mgk_ls_landau2d_ncg_workspace_bytes")
            self.ws_ncg = t.empty(int(nb.value), dtype=t.uint8, device="cuda")
        eb, ebp = _lib.host_dbl(np.asarray(eps_bar, dtype=np.float64).reshape(2, 2))
        s = dev.stream_ptr() if stream is None else stream
        kit = ctypes.c_int()
        rc = lib.mgk_ls_solve_landau2d_ncg(
            self._shape_c, _lib.dptr(self.lam), _lib.dptr(self.mu), self.lam0, self.mu0, ctypes.byref(params),
            1 if self.plane_stress else 0, ebp, _lib.dptr(eps_plastic), _lib.dptr(self.eps), _lib.dptr(self.sig),
            1 if warm else 0, int(max_iter), float(tol), _lib.dptr(self.ws_ncg), self.ws_ncg.numel(),
            ctypes.c_void_p(s), ctypes.byref(self.info), ctypes.byref(kit))
        _lib.check(rc, "mgk_ls_solve_landau2d_ncg")
        self.krylov_iters = kit.value
        if rc == _lib.MGK_ERR_NONFINITE:
            raise FloatingPointError("spectral_solver_landau_2d (newton_cg): non-finite stress or update")
        return self.info


def spectral_solver_landau_2d(lam, mu, v1, v2, v3, g1, g2, g3, g4, eps_bar, eps_plastic=None, max_iter=400, tol=1e-6,
                              verbose=False, pixel=1.0, plane_mode="plane_strain", Gamma=None, eps_init=None,
                              strain_capping_enabled=False, strain_capping_limit=None,
                              strain_capping_tangent_ratio=0.1, strain_capping_type="piecewise",
                              strain_capping_smooth_power=1.0, solver="dbfft", e33_state=None, anderson_m=5):
    """2D nonlinear LS with the Landau stress and Anderson mixing on the GPU
    (mgkmc/linear_elastic_simulator.py:952-1228, "dbfft" path); e33_state
    (dict) carries the plane-stress e33 warm start across calls as in the
    reference.  solver="newton_cg" runs the reference's Newton-CG branch
    (:1011-1117; e33_state and anderson_m unused there, as in the reference).
    Strains must be symmetric."""
    if solver not in ("dbfft", "newton_cg"):
        raise ValueError(f"unknown solver {solver!r}")
    lam = np.asarray(lam, dtype=np.float64)
    mu = np.asarray(mu, dtype=np.float64)
    eps_bar = np.asarray(eps_bar, dtype=np.float64)
    if np.any(eps_bar != eps_bar.T):
        raise ValueError("the device Landau solver takes symmetric eps_bar")
    sol = LandauSolver2D(lam, mu, plane_mode, anderson_m)
    ep = None
    if eps_plastic is not None:
        sym, a = dev.to_sym_soa(eps_plastic, 2)
        if np.any(a):
            raise ValueError("the device Landau solver takes a symmetric eps_plastic")
        ep = dev.upload(sym, FP64)
    warm = eps_init is not None
    if warm:
        sym, a = dev.to_sym_soa(eps_init, 2)
        if np.any(a):
            raise ValueError("the device Landau solver takes a symmetric eps_init")
        sol.eps.copy_(dev.upload(sym, FP64))
    prm = landau_params(v1, v2, v3, g1, g2, g3, g4, mu, strain_capping_enabled, strain_capping_limit,
                        strain_capping_tangent_ratio, strain_capping_type, strain_capping_smooth_power)
    if solver == "newton_cg":
        inf = sol.solve_ncg(prm, eps_bar, ep, warm=warm, max_iter=max_iter, tol=tol)
        eps = dev.from_sym_soa(dev.download(sol.eps), 2)
        sig = dev.from_sym_soa(dev.download(sol.sig), 2)
        return eps, sig, np.array(inf.eps_mean[:4]).reshape(2, 2), np.array(inf.sig_mean[:4]).reshape(2, 2)
    persist = e33_state is not None
    if persist and e33_state.get("e33") is not None and np.shape(e33_state["e33"]) == lam.shape:
        sol.e33.copy_(dev.upload(np.ascontiguousarray(e33_state["e33"], dtype=np.float64), FP64))
        sol.e33_have = True
    inf = sol.solve(prm, eps_bar, ep, warm=warm, max_iter=max_iter, tol=tol, persist_e33=persist)
    if persist and sol.plane_stress:
        e33_state["e33"] = dev.download(sol.e33)
    eps = dev.from_sym_soa(dev.download(sol.eps), 2)
    sig = dev.from_sym_soa(dev.download(sol.sig), 2)
    return eps, sig, np.array(inf.eps_mean[:4]).reshape(2, 2), np.array(inf.sig_mean[:4]).reshape(2, 2)


# ---------------------------------------------------------------------------
# Secant-modulus laws: exponential degradation (2D/3D) and Rose (2D)
# ---------------------------------------------------------------------------
SECANT_EXP, SECANT_ROSE = 0, 1


class SecantSolver:
    """Device-resident nonlinear LS solver for the secant-modulus laws
    (mgkmc/linear_elastic_simulator.py:614-950): law SECANT_EXP with
    (p1, p2) = (d, k) in 2D or 3D, law SECANT_ROSE with (eta, floor
    fraction) in 2D.  The reference medium is (mean lam, mean mu), with
    mu_ref = (1 + floor) mean mu / 2 for the Rose law, as the reference sets it."""

    def __init__(self, lam, mu, law, p1, p2, plane_mode="plane_strain", precision=FP64):
        dev.require_gpu()
        lam = np.ascontiguousarray(lam, dtype=np.float64)
        mu = np.ascontiguousarray(mu, dtype=np.float64)
        if lam.shape != mu.shape or lam.ndim not in (2, 3):
            raise ValueError("lam and mu must be 2D or 3D fields of one shape")
        if law == SECANT_ROSE and lam.ndim != 2:
            raise ValueError("the Rose law is defined for 2D grids only")
        if precision not in (FP64, FP32):
            raise ValueError("precision must be FP64 or FP32")
        self.shape = lam.shape
        self.d = lam.ndim
        self.C = 6 if self.d == 3 else 3
        self.precision = precision
        self.params = _lib.MgkSecantParams(int(law), 1 if plane_mode == "plane_stress" else 0, float(p1), float(p2))
        self.lam0 = float(lam.mean())
        m = float(mu.mean())
        self.mu0 = 0.5 * (m + float(p2) * m) if law == SECANT_ROSE else m
        self.lam = dev.upload(lam, precision)
        self.mu = dev.upload(mu, precision)
        dt = dev.torch_dtype(precision)
        t = dev.torch
        self.eps = t.zeros((self.C,) + self.shape, dtype=dt, device="cuda")
        self.sig = t.zeros((self.C,) + self.shape, dtype=dt, device="cuda")
        self.tau = t.empty((self.C,) + self.shape, dtype=dt, device="cuda")
        self._shape_c = _lib.shape_arg(self.shape)
        nb = ctypes.c_size_t()
        _lib.check(_lib.load().mgk_ls_workspace_bytes(self.d, self._shape_c, precision, ctypes.byref(nb)),
                   "mgk_ls_workspace_bytes")
        self.ws = t.empty(int(nb.value), dtype=t.uint8, device="cuda")
        self.info = _lib.MgkLsInfo()

    def solve(self, eps_bar, eps_plastic=None, warm=False, max_iter=400, tol=1e-6, stream=None):
        """One solve; eps_plastic a device SoA field or None.  A non-finite
        update stops the loop with info.status == MGK_ERR_NONFINITE and
        non-finite fields (the reference iterates on to max_iter there)."""
        eb, ebp = _lib.host_dbl(np.asarray(eps_bar, dtype=np.float64).reshape(self.d, self.d))
        s = dev.stream_ptr() if stream is None else stream
        rc = _lib.load().mgk_ls_solve_secant(
            self.d, self._shape_c, self.precision, _lib.dptr(self.lam), _lib.dptr(self.mu), self.lam0, self.mu0,
            ctypes.byref(self.params), ebp, _lib.dptr(eps_plastic), _lib.dptr(self.eps), _lib.dptr(self.sig),
            _lib.dptr(self.tau), 1 if warm else 0, int(max_iter), float(tol), _lib.dptr(self.ws), self.ws.numel(),
            ctypes.c_void_p(s), ctypes.byref(self.info))
        _lib.check(rc, "mgk_ls_solve_secant")
        return self.info


def _secant_drop_in(d, lam, mu, law, p1, p2, eps_bar, eps_plastic, max_iter, tol, plane_mode, eps_init):
    lam = np.asarray(lam, dtype=np.float64)
    mu = np.asarray(mu, dtype=np.float64)
    shape = lam.shape if lam.ndim == d else mu.shape
    lam = np.array(np.broadcast_to(lam, shape))
    mu = np.array(np.broadcast_to(mu, shape))
    eps_bar = np.asarray(eps_bar, dtype=np.float64)
    if np.any(eps_bar != eps_bar.T):
        raise ValueError("the device secant solvers take a symmetric eps_bar")
    sol = SecantSolver(lam, mu, law, p1, p2, plane_mode, FP64)
    ep = None
    if eps_plastic is not None:
        sym, a = dev.to_sym_soa(eps_plastic, d)
        if np.any(a):
            raise ValueError("the device secant solvers take a symmetric eps_plastic")
        ep = dev.upload(sym, FP64)
    warm = eps_init is not None
    if warm:
        sym, a = dev.to_sym_soa(eps_init, d)
        if np.any(a):
            raise ValueError("the device secant solvers take a symmetric eps_init")
        sol.eps.copy_(dev.upload(sym, FP64))
    inf = sol.solve(eps_bar, ep, warm=warm, max_iter=max_iter, tol=tol)
    eps = dev.from_sym_soa(dev.download(sol.eps), d)
    sig = dev.from_sym_soa(dev.download(sol.sig), d)
    n = d * d
    return eps, sig, np.array(inf.eps_mean[:n]).reshape(d, d), np.array(inf.sig_mean[:n]).reshape(d, d)


def spectral_solver_secant_2d(lam, mu, d, k, eps_bar, eps_plastic=None, max_iter=400, tol=1e-6, verbose=False,
                              pixel=1.0, plane_mode="plane_strain", Gamma=None, eps_init=None):
    """2D LS with the exponential secant-degradation law on the GPU
    (mgkmc/linear_elastic_simulator.py:614-765); lam is used as given in
    either plane mode, as in the reference.  Gamma is ignored (the reference
    medium is (mean lam, mean mu))."""
    return _secant_drop_in(2, lam, mu, SECANT_EXP, d, k, eps_bar, eps_plastic, max_iter, tol, plane_mode, eps_init)


def spectral_solver_secant_3d(lam, mu, d, k, eps_bar, eps_plastic=None, max_iter=400, tol=1e-6, verbose=False,
                              pixel=1.0, Gamma=None, eps_init=None):
    """3D LS with the exponential secant-degradation law on the GPU
    (mgkmc/linear_elastic_simulator.py:768-849)."""
    return _secant_drop_in(3, lam, mu, SECANT_EXP, d, k, eps_bar, eps_plastic, max_iter, tol, "plane_strain",
                           eps_init)


def spectral_solver_rose_2d(lam, mu, eta, mu_floor_fraction=0.1, eps_bar=None, eps_plastic=None, max_iter=400,
                            tol=1e-6, verbose=False, pixel=1.0, plane_mode="plane_strain", Gamma=None,
                            eps_init=None):
    """2D LS with the Rose secant law on the GPU
    (mgkmc/linear_elastic_simulator.py:852-950); reference medium
    (mean lam, (1 + floor) mean mu / 2)."""
    return _secant_drop_in(2, lam, mu, SECANT_ROSE, eta, mu_floor_fraction, eps_bar, eps_plastic, max_iter, tol,
                           plane_mode, eps_init)


def _nl_stepping(d, solver, solve, E_avg, nu_avg, target_strain_mask, target_values, n_steps, tol_macro,
                 max_iter_macro, store, secant_slope, info):
    """Mixed stress/strain load stepping of a nonlinear law on one device
    solver (mgkmc/linear_elastic_simulator.py:1353-2516): per step, every macro
    iteration restarts from the previous step's converged field; the
    stress-controlled entries get (err - nu_avg tr(err) I) / E_avg with the
    compliance of the reference medium, or, with secant_slope (the Landau
    drivers, :2155-2205 and :2390-2438), err / slope for entries whose last two
    macro iterates give a finite slope above 0.01 E_avg and a strain change
    above 1e-14 (history restarted each step).  solve(eps_bar, warm) runs one
    solve on ``solver`` and returns its info."""
    mask = np.asarray(target_strain_mask, dtype=bool)
    smask = ~mask
    tv = np.asarray(target_values, dtype=np.float64)
    eb = np.zeros((d, d))
    warm = None
    em_l, sm_l, e_l, s_l, its = [], [], [], [], []
    n = d * d
    for s in range(n_steps + 1):
        tgt = tv * (s / n_steps)
        prev_e = prev_s = None
        for _ in range(max_iter_macro):
            eb[mask] = tgt[mask]
            if warm is not None:
                solver.eps.copy_(warm)
            inf = solve(eb, warm is not None)
            its.append(inf.iters)
            epsM = np.array(inf.eps_mean[:n]).reshape(d, d)
            sigM = np.array(inf.sig_mean[:n]).reshape(d, d)
            err = np.zeros((d, d))
            err[smask] = tgt[smask] - sigM[smask]
            max_err = np.max(np.abs(err[smask])) if np.any(smask) else 0.0
            if max_err < tol_macro:
                break
            de = (err - nu_avg * np.trace(err) * np.eye(d)) / E_avg
            dsc = de[smask]
            if secant_slope:
                e_sc = eb[smask].copy()
                s_sc = sigM[smask].copy()
                if prev_e is not None:
                    dh = e_sc - prev_e
                    with np.errstate(divide="ignore", invalid="ignore"):
                        slope = (s_sc - prev_s) / dh
                    ok = np.isfinite(slope) & (np.abs(dh) > 1e-14) & (slope > 0.01 * E_avg)
                    dsc[ok] = err[smask][ok] / slope[ok]
                prev_e, prev_s = e_sc, s_sc
            eb[smask] += dsc
        warm = solver.eps.clone()
        em_l.append(epsM)
        sm_l.append(sigM)
        if store:
            e_l.append(dev.from_sym_soa(dev.download(solver.eps), d))
            s_l.append(dev.from_sym_soa(dev.download(solver.sig), d))
    if info is not None:
        info["ls_iters"] = its
    return np.array(em_l), np.array(sm_l), e_l, s_l


def _medium_compliance(la, ma):
    """(E, nu) of the reference medium (lam0, mu0), the stress-correction compliance."""
    return ma * (3.0 * la + 2.0 * ma) / (la + ma), la / (2.0 * (la + ma))


def _secant_simulation(d, lam, mu, law, p1, p2, target_strain_mask, target_values, n_steps, tol_ls, max_iter_ls,
                       tol_macro, max_iter_macro, store, plane_mode, precision=FP64, info=None):
    lam = np.ascontiguousarray(lam, dtype=np.float64)
    mu = np.ascontiguousarray(mu, dtype=np.float64)
    if lam.ndim != d:
        raise ValueError(f"lam must be {d}D")
    solver = SecantSolver(lam, mu, law, p1, p2, plane_mode, precision)
    E_avg, nu_avg = _medium_compliance(solver.lam0, solver.mu0)
    return _nl_stepping(d, solver, lambda eb, warm: solver.solve(eb, None, warm=warm, max_iter=max_iter_ls,
                                                                 tol=tol_ls),
                        E_avg, nu_avg, target_strain_mask, target_values, n_steps, tol_macro, max_iter_macro, store,
                        False, info)


def secant_elastic_simulation_2d(lam, mu, d, k, target_strain_mask, target_values, n_steps=20, pixel=1.0, tol_ls=1e-6,
                                 max_iter_ls=400, tol_macro=1e-4, max_iter_macro=20, store=True,
                                 plane_mode="plane_strain", log_path=None, global_log_path=None,
                                 driving_component=(0, 0), enable_console=True, vtk_path=None, vtk_interval="none",
                                 **solver_kw):
    """2D mixed-BC load stepping with the secant-degradation law
    (mgkmc/linear_elastic_simulator.py:1353-1600) on the GPU; log, VTK and
    checkpoint options are accepted and ignored; ``precision`` / ``info``
    (extensions) as in linear_elastic_simulation_2d."""
    return _secant_simulation(2, lam, mu, SECANT_EXP, d, k, target_strain_mask, target_values, n_steps, tol_ls,
                              max_iter_ls, tol_macro, max_iter_macro, store, plane_mode,
                              solver_kw.get("precision", FP64), solver_kw.get("info"))


def secant_elastic_simulation_3d(lam, mu, d, k, target_strain_mask, target_values, n_steps=20, pixel=1.0, tol_ls=1e-6,
                                 max_iter_ls=400, tol_macro=1e-4, max_iter_macro=20, store=True, log_path=None,
                                 global_log_path=None, driving_component=(0, 0), enable_console=True, vtk_path=None,
                                 vtk_interval="none", **solver_kw):
    """3D mixed-BC load stepping with the secant-degradation law
    (mgkmc/linear_elastic_simulator.py:1857-2052) on the GPU."""
    return _secant_simulation(3, lam, mu, SECANT_EXP, d, k, target_strain_mask, target_values, n_steps, tol_ls,
                              max_iter_ls, tol_macro, max_iter_macro, store, "plane_strain",
                              solver_kw.get("precision", FP64), solver_kw.get("info"))


def rose_elastic_simulation_2d(lam, mu, eta, mu_floor_fraction=0.1, target_strain_mask=None, target_values=None,
                               n_steps=20, pixel=1.0, tol_ls=1e-6, max_iter_ls=400, tol_macro=1e-4, max_iter_macro=20,
                               store=True, plane_mode="plane_strain", log_path=None, global_log_path=None,
                               driving_component=(0, 0), enable_console=True, vtk_path=None, vtk_interval="none",
                               **solver_kw):
    """2D mixed-BC load stepping with the Rose law
    (mgkmc/linear_elastic_simulator.py:1603-1855) on the GPU."""
    return _secant_simulation(2, lam, mu, SECANT_ROSE, eta, mu_floor_fraction, target_strain_mask, target_values,
                              n_steps, tol_ls, max_iter_ls, tol_macro, max_iter_macro, store, plane_mode,
                              solver_kw.get("precision", FP64), solver_kw.get("info"))


def _landau_simulation(d, lam, mu, v, target_strain_mask, target_values, n_steps, tol_ls, max_iter_ls, tol_macro,
                       max_iter_macro, store, plane_mode, solver_kw):
    solver_name = solver_kw.get("solver", "dbfft")
    if solver_name == "newton_cg" and d == 3:
        raise NotImplementedError("Newton-CG solver is not yet implemented for 3D small-strain. Please use 2D or "
                                  "default to dbfft solver.")
    lam = np.ascontiguousarray(lam, dtype=np.float64)
    mu = np.ascontiguousarray(mu, dtype=np.float64)
    if lam.ndim != d:
        raise ValueError(f"lam must be {d}D")
    cap = {k: solver_kw[k] for k in ("strain_capping_enabled", "strain_capping_limit", "strain_capping_tangent_ratio",
                                     "strain_capping_type", "strain_capping_smooth_power") if k in solver_kw}
    prm = landau_params(*v, mu=mu, **cap)
    if d == 3:
        solver = LandauSolver(lam, mu, solver_kw.get("precision", FP64))
        solve = lambda eb, warm: solver.solve(prm, eb, None, warm=warm, max_iter=max_iter_ls, tol=tol_ls)
    elif solver_name == "newton_cg":
        solver = LandauSolver2D(lam, mu, plane_mode, 0)
        solve = lambda eb, warm: solver.solve_ncg(prm, eb, None, warm=warm, max_iter=max_iter_ls, tol=tol_ls)
    else:
        solver = LandauSolver2D(lam, mu, plane_mode, solver_kw.get("anderson_m", 5))
        # one e33 state for the whole run (the reference's e33_state dict)
        ps = plane_mode == "plane_stress"
        solve = lambda eb, warm: solver.solve(prm, eb, None, warm=warm, max_iter=max_iter_ls, tol=tol_ls,
                                              persist_e33=ps)
    E_avg, nu_avg = _medium_compliance(solver.lam0, solver.mu0)
    return _nl_stepping(d, solver, solve, E_avg, nu_avg, target_strain_mask, target_values, n_steps, tol_macro,
                        max_iter_macro, store, True, solver_kw.get("info"))


def landau_elastic_simulation_2d(lam, mu, v1, v2, v3, g1, g2, g3, g4, target_strain_mask, target_values, n_steps=20,
                                 pixel=1.0, tol_ls=1e-6, max_iter_ls=400, tol_macro=1e-4, max_iter_macro=20,
                                 store=True, plane_mode="plane_strain", log_path=None, global_log_path=None,
                                 driving_component=(0, 0), enable_console=True, vtk_path=None, vtk_interval="none",
                                 **solver_kw):
    """2D mixed-BC load stepping with the Landau law
    (mgkmc/linear_elastic_simulator.py:2055-2284) on the GPU: secant-slope
    macro corrections, Anderson depth 5, one plane-stress e33 state for the
    run.  Capping options and ``solver`` come through solver_kw as in the
    reference; ``info`` (extension) receives the LS iteration counts."""
    return _landau_simulation(2, lam, mu, (v1, v2, v3, g1, g2, g3, g4), target_strain_mask, target_values, n_steps,
                              tol_ls, max_iter_ls, tol_macro, max_iter_macro, store, plane_mode, solver_kw)


def landau_elastic_simulation_3d(lam, mu, v1, v2, v3, g1, g2, g3, g4, target_strain_mask, target_values, n_steps=20,
                                 pixel=1.0, tol_ls=1e-6, max_iter_ls=400, tol_macro=1e-4, max_iter_macro=20,
                                 store=True, log_path=None, global_log_path=None, driving_component=(0, 0),
                                 enable_console=True, vtk_path=None, vtk_interval="none", **solver_kw):
    """3D mixed-BC load stepping with the Landau law
    (mgkmc/linear_elastic_simulator.py:2287-2516) on the GPU; ``precision``
    (extension) selects FP64 or FP32 fields."""
    return _landau_simulation(3, lam, mu, (v1, v2, v3, g1, g2, g3, g4), target_strain_mask, target_values, n_steps,
                              tol_ls, max_iter_ls, tol_macro, max_iter_macro, store, "plane_strain", solver_kw)
