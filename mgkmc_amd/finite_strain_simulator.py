"""Finite-strain (de Geus) Newton-Krylov solve on the GPU (drop-ins).

Reference: mgkmc/finite_strain_simulator.py.  Field layout is the reference's
tensor-first (3, 3, nx, ny, nz), which is exactly the library's 9-component
SoA layout, so no transposition happens at the seam.

``finite_strain_solver_step_3d`` keeps the reference signature
(:2967-2982) for solver="newton_cg", "al" and "dbfft"; Ghat4, C4, I2, I4, I4rt are accepted for
compatibility and only lam = C4[0,0,1,1], mu = C4[0,1,0,1] are read (the
projection and the tangent are matrix-free on the device).  K4 is returned as
None, as no caller consumes it (mgkmc/aqs.py:289-303).

Errors follow the reference: ValueError for a failed line search, Newton
non-convergence and macro-BC non-convergence.
"""
import ctypes

import numpy as np

from . import _lib
from . import device as dev

MODELS = {"svk": 0, "landau": 1, "murnaghan": 2, "neo_hookean": 3}  # neo-Hookean: 2D plane strain only


def _scalar_const(x, name):
    a = np.asarray(x, dtype=np.float64)
    if a.ndim == 0:
        v = float(a)
    else:
        if not np.all(a == a.reshape(-1)[0]):
            raise NotImplementedError(f"per-voxel {name} is not supported on the device path")
        v = float(a.reshape(-1)[0])
    return v


def landau_pa(x, name):
    """GPa -> Pa autoconversion of |x| < 1e6, x != 0 (:1687-1690)."""
    v = _scalar_const(x, name)
    if abs(v) < 1e6 and v != 0.0:
        v *= 1e9
    return v


def material(model_type="svk", A_m=0.0, B_m=0.0, C_m=0.0, v1=0.0, v2=0.0, v3=0.0, g1=0.0, g2=0.0, g3=0.0,
             g4=0.0, strain_capping_enabled=False, strain_capping_limit=None, strain_capping_tangent_ratio=0.1,
             strain_capping_type="piecewise", strain_capping_smooth_power=1.0, mu_mean=None, ndim=3):
    """mgk_fs_material for the device kernels (constants resolved on the host).
    neo_hookean exists in the 2D law only; the 3D law treats it, like any
    other unknown name, as SVK."""
    m = _lib.MgkFsMaterial()
    m.model = MODELS.get(model_type, 0)  # any other name falls through to SVK, as the reference
    if model_type == "neo_hookean" and ndim == 3:
        m.model = 0
    if model_type == "landau":
        consts = [np.asarray(x, dtype=np.float64) for x in (v1, v2, v3, g1, g2, g3, g4)]
        if any(a.ndim > 0 for a in consts):
            # per-voxel constants (finite_strain_simulator.py:1678-1690): each element
            # converted from GPa as the reference converts them, kept on the device
            shape = np.broadcast_shapes(*[a.shape for a in consts])
            field = np.stack([np.broadcast_to(a, shape) for a in consts]).reshape(7, -1).copy()
            small = (np.abs(field) < 1e6) & (field != 0.0)
            field[small] *= 1e9
            m._landau_dev = dev.upload(field, 0)
            m.landau_field = m._landau_dev.data_ptr()
            m.v1, m.v2, m.v3, m.g1, m.g2, m.g3 = (float(field[k].mean()) for k in range(6))
            m.g4 = float(field[6].mean())  # the capping limit reads <g4> (:1700-1703)
        else:
            m.v1, m.v2, m.v3 = landau_pa(v1, "v1"), landau_pa(v2, "v2"), landau_pa(v3, "v3")
            m.g1, m.g2, m.g3, m.g4 = landau_pa(g1, "g1"), landau_pa(g2, "g2"), landau_pa(g3, "g3"), landau_pa(g4, "g4")
        m.capping = 1 if strain_capping_enabled else 0
        if strain_capping_enabled:
            if strain_capping_limit is None or strain_capping_limit <= 0:
                m.cap_limit = float(np.sqrt(0.2 * mu_mean / abs(m.g4))) if m.g4 < 0 else 999.0
            else:
                m.cap_limit = float(strain_capping_limit)
            m.cap_smooth = 1 if strain_capping_type == "smooth" else 0
            m.cap_power = float(strain_capping_smooth_power)
            m.cap_ratio = float(strain_capping_tangent_ratio)
    elif model_type == "murnaghan":
        m.Am, m.Bm, m.Cm = _scalar_const(A_m, "A_m"), _scalar_const(B_m, "B_m"), _scalar_const(C_m, "C_m")
    return m


MIXED = 2  # FP64 Newton loop, FP32 Krylov solves


class FsSolver:
    """Device-resident finite-strain state for one grid: lam, mu, F, P, Fp."""

    def __init__(self, lam, mu, precision=0, Fp=None):
        """precision: 0 FP64, 1 FP32, MIXED (2): FP64 Newton loop (F, P,
        residual, line search, macro BC) with the Krylov solves in FP32."""
        dev.require_gpu()
        lam = np.ascontiguousarray(lam, dtype=np.float64)
        self.shape = lam.shape
        self.NV = int(np.prod(self.shape))
        self.mixed = precision == MIXED
        self.precision = 0 if self.mixed else precision
        precision = self.precision
        self.lam = dev.upload(lam, precision)
        self.mu = dev.upload(np.ascontiguousarray(mu, dtype=np.float64), precision)
        if self.mixed:
            self.lam32 = dev.upload(lam, 1)
            self.mu32 = dev.upload(np.ascontiguousarray(mu, dtype=np.float64), 1)
        self.mu_mean = float(np.mean(mu))
        # reference medium of the augmented-Lagrangian solver: voxel means of
        # C4_iiii, C4_iijj, C4_ijij of the linear stiffness
        mu_h = np.asarray(mu, dtype=np.float64)
        self.L0 = (ctypes.c_double * 3)(float(np.mean(lam + 2.0 * mu_h)), float(np.mean(lam)), float(np.mean(mu_h)))
        dt = dev.torch_dtype(precision)
        eye = dev.torch.eye(3, dtype=dt, device="cuda").reshape(3, 3, 1, 1, 1)
        self.F = eye.expand((3, 3) + self.shape).contiguous()
        self.P = dev.torch.zeros((3, 3) + self.shape, dtype=dt, device="cuda")
        self.sig = dev.torch.zeros((3, 3) + self.shape, dtype=dt, device="cuda")
        self.Fp = None if Fp is None else dev.upload(np.ascontiguousarray(Fp, dtype=np.float64), precision)
        self._shape_c = _lib.shape_arg(self.shape)
        n = ctypes.c_size_t()
        _lib.check(_lib.load().mgk_fs_workspace_bytes(self._shape_c, precision, ctypes.byref(n)),
                   "mgk_fs_workspace_bytes")
        self.ws = dev.torch.empty(int(n.value), dtype=dev.torch.uint8, device="cuda")
        if self.mixed:
            _lib.check(_lib.load().mgk_fs_workspace_bytes(self._shape_c, 1, ctypes.byref(n)),
                       "mgk_fs_workspace_bytes")
            self.ws32 = dev.torch.empty(int(n.value), dtype=dev.torch.uint8, device="cuda")
        self.info = _lib.MgkFsInfo()

    def stepper(self, mat, E_avg=100e9, nu_avg=0.3, tol_NW=1e-5, tol_CG=1e-6, max_NW=50, tol_macro=1e6,
                max_iter_macro=20, solver="newton_cg", pixel=1.0):
        """The arguments of step() without the macro BC as an MgkFsStepper
        (for the library's native sub-stepped flip); keeps mat alive."""
        if solver not in ("newton_cg", "al", "dbfft"):
            raise NotImplementedError(f"solver={solver!r}: the device path implements 'newton_cg', 'al' and 'dbfft'")
        st = _lib.MgkFsStepper()
        st.solver = {"newton_cg": 0, "al": 1, "dbfft": 2}[solver]
        st.precision, st.mixed = self.precision, int(self.mixed and solver == "newton_cg")
        for i, n in enumerate(self.shape):
            st.shape[i] = n
        st.F, st.Fp, st.lam, st.mu = (_lib.dptr(self.F), _lib.dptr(self.Fp), _lib.dptr(self.lam),
                                      _lib.dptr(self.mu))
        if self.mixed:
            st.lam32, st.mu32 = _lib.dptr(self.lam32), _lib.dptr(self.mu32)
            st.ws32, st.ws32_bytes = _lib.dptr(self.ws32), self.ws32.numel()
        st.mat = ctypes.addressof(mat)
        st.P, st.sig, st.ws, st.ws_bytes = _lib.dptr(self.P), _lib.dptr(self.sig), _lib.dptr(self.ws), self.ws.numel()
        st.L0 = ctypes.addressof(self.L0)
        st.E_avg, st.nu_avg, st.tol_NW, st.tol_CG = float(E_avg), float(nu_avg), float(tol_NW), float(tol_CG)
        st.tol_macro, st.pixel, st.max_NW, st.max_iter_macro = float(tol_macro), float(pixel), int(max_NW), \
            int(max_iter_macro)
        st._keep = (mat, self)
        return st

    def step(self, F_bar, mat, driving_component=(0, 0), P_target=None, P_mask=None, E_avg=100e9, nu_avg=0.3,
             tol_NW=1e-5, tol_CG=1e-6, max_NW=50, tol_macro=1e6, max_iter_macro=20, solver="newton_cg", pixel=1.0):
        """One load step with solver "newton_cg" (FP64, FP32 or the FP32-Krylov
        form), "al" (augmented Lagrangian, FP64 or FP32) or "dbfft"
        (displacement-based FFT, FP64; ``pixel`` sets its wave vectors)."""
        Fb, Fbp = _lib.host_dbl(np.asarray(F_bar, dtype=np.float64).reshape(9))
        Pt, Ptp = _lib.host_dbl(np.zeros(9) if P_target is None else np.asarray(P_target, float).reshape(9))
        pm = np.zeros(9, dtype=np.int32) if P_mask is None else np.asarray(P_mask, bool).reshape(9).astype(np.int32)
        pm_c = (ctypes.c_int * 9)(*pm.tolist())
        if solver not in ("newton_cg", "al", "dbfft"):
            raise NotImplementedError(f"solver={solver!r}: the device path implements 'newton_cg', 'al' and 'dbfft'")
        if solver == "dbfft":
            rc = _lib.load().mgk_fs_solver_step_dbfft(
                self._shape_c, self.precision, _lib.dptr(self.F), _lib.dptr(self.Fp), _lib.dptr(self.lam),
                _lib.dptr(self.mu), ctypes.byref(mat), Fbp, Ptp, pm_c, int(driving_component[0]),
                int(driving_component[1]), float(E_avg), float(nu_avg), float(tol_NW), float(tol_CG), int(max_NW),
                float(tol_macro), int(max_iter_macro), float(pixel), _lib.dptr(self.P), _lib.dptr(self.sig),
                _lib.dptr(self.ws), self.ws.numel(), ctypes.c_void_p(dev.stream_ptr()), ctypes.byref(self.info))
        elif solver == "al":
            rc = _lib.load().mgk_fs_solver_step_al(
                self._shape_c, self.precision, _lib.dptr(self.F), _lib.dptr(self.Fp), _lib.dptr(self.lam),
                _lib.dptr(self.mu), ctypes.byref(mat), Fbp, Ptp, pm_c, int(driving_component[0]),
                int(driving_component[1]), float(E_avg), float(nu_avg), float(tol_NW), int(max_NW), float(tol_macro),
                int(max_iter_macro), self.L0, _lib.dptr(self.P), _lib.dptr(self.sig), _lib.dptr(self.ws),
                self.ws.numel(), ctypes.c_void_p(dev.stream_ptr()), ctypes.byref(self.info))
        elif self.mixed:
            rc = _lib.load().mgk_fs_solver_step_refined(
                self._shape_c, _lib.dptr(self.F), _lib.dptr(self.Fp), _lib.dptr(self.lam), _lib.dptr(self.mu),
                _lib.dptr(self.lam32), _lib.dptr(self.mu32), ctypes.byref(mat), Fbp, Ptp, pm_c,
                int(driving_component[0]), int(driving_component[1]), float(E_avg), float(nu_avg), float(tol_NW),
                float(tol_CG), int(max_NW), float(tol_macro), int(max_iter_macro), _lib.dptr(self.P),
                _lib.dptr(self.sig), _lib.dptr(self.ws), self.ws.numel(), _lib.dptr(self.ws32), self.ws32.numel(),
                ctypes.c_void_p(dev.stream_ptr()), ctypes.byref(self.info))
        else:
            rc = _lib.load().mgk_fs_solver_step(
                self._shape_c, self.precision, _lib.dptr(self.F), _lib.dptr(self.Fp), _lib.dptr(self.lam),
                _lib.dptr(self.mu), ctypes.byref(mat), Fbp, Ptp, pm_c, int(driving_component[0]),
                int(driving_component[1]), float(E_avg), float(nu_avg), float(tol_NW), float(tol_CG), int(max_NW),
                float(tol_macro), int(max_iter_macro), _lib.dptr(self.P), _lib.dptr(self.sig), _lib.dptr(self.ws),
                self.ws.numel(), ctypes.c_void_p(dev.stream_ptr()), ctypes.byref(self.info))
        if rc == _lib.MGK_ERR_SOLVER:
            raise ValueError(_lib.last_error().split(" (")[0])
        if rc == _lib.MGK_ERR_NONFINITE:  # the capped plane-stress E33 iteration (finite_strain_simulator.py:573-576)
            raise FloatingPointError(_lib.last_error().split(" (/")[0])
        _lib.check(rc, "mgk_fs_solver_step")
        return self.info


def finite_strain_solver_step_3d(F, F_bar, Ghat4, C4, I2, I4, I4rt, Fp=None, driving_component=(0, 0),
                                 P_target=None, P_mask=None, E_avg=100e9, nu_avg=0.3, tol_NW=1e-5, tol_CG=1e-6,
                                 max_NW=50, tol_macro=1e6, max_iter_macro=20, enable_console=True, model_type="svk",
                                 A_m=0.0, B_m=0.0, C_m=0.0, solver="al", pixel=1.0, v1=0.0, v2=0.0, v3=0.0, g1=0.0,
                                 g2=0.0, g3=0.0, g4=0.0, strain_capping_enabled=False, strain_capping_limit=None,
                                 strain_capping_tangent_ratio=0.1, strain_capping_type="piecewise",
                                 strain_capping_smooth_power=1.0):
    """One finite-strain load step (solver "newton_cg", "al" or "dbfft") -> (F, P, Sig, None, F_bar)."""
    if solver not in ("newton_cg", "al", "dbfft"):
        raise NotImplementedError(f"solver={solver!r}: the device path implements 'newton_cg', 'al' and 'dbfft'")
    C4 = np.asarray(C4)
    lam = C4[0, 0, 1, 1]
    mu = C4[0, 1, 0, 1]
    s = FsSolver(lam, mu, precision=0, Fp=Fp)
    s.F.copy_(dev.upload(np.ascontiguousarray(F, dtype=np.float64), 0))
    mat = material(model_type, A_m, B_m, C_m, v1, v2, v3, g1, g2, g3, g4, strain_capping_enabled,
                   strain_capping_limit, strain_capping_tangent_ratio, strain_capping_type,
                   strain_capping_smooth_power, mu_mean=float(np.mean(mu)))
    info = s.step(F_bar, mat, driving_component, P_target, P_mask, E_avg, nu_avg, tol_NW, tol_CG, max_NW, tol_macro,
                  max_iter_macro, solver=solver, pixel=pixel)
    finite_strain_solver_step_3d.last_info = info  # iteration counts of the last call (extension)
    Fo = dev.download(s.F)
    Po = dev.download(s.P)
    So = dev.download(s.sig)
    return Fo, Po, So, None, np.array(info.F_bar[:]).reshape(3, 3)


_PLANE_AXES = (0, 2)  # in-plane axes of the (nx, 1, ny) grid a 2D problem runs on


def plane_law_supported(model_type, plane_mode, capping=False):
    """Whether the 2D device step covers a law in a plane mode: svk,
    landau (capped or not), murnaghan and neo_hookean in both modes (the
    reference's 2D laws, finite_strain_simulator.py:591-1091)."""
    return model_type in ("svk", "landau", "murnaghan", "neo_hookean") and \
        plane_mode in ("plane_strain", "plane_stress")


def _embed2(A, fill=1.0):
    """(2, 2, nx, ny) [or (2, 2)] -> (3, 3, nx, 1, ny) [(3, 3)], out-of-plane diagonal = fill."""
    A = np.asarray(A, dtype=np.float64)
    grid = A.shape[2:]
    out = np.zeros((3, 3) + ((grid[0], 1, grid[1]) if grid else ()))
    for a in range(2):
        for b in range(2):
            out[_PLANE_AXES[a], _PLANE_AXES[b]] = A[a, b].reshape(out.shape[2:]) if grid else A[a, b]
    out[1, 1] = fill
    return out


def _extract2(A3):
    """(3, 3, nx, 1, ny) [or (3, 3)] -> (2, 2, nx, ny) [(2, 2)]."""
    g = A3.shape[2:]
    return np.stack([np.stack([A3[_PLANE_AXES[a], _PLANE_AXES[b]].reshape((g[0], g[2]) if g else ())
                               for b in range(2)]) for a in range(2)])


def finite_strain_solver_step_2d(F, F_bar, Ghat4, C4, I2, I4, I4rt, Fp=None, driving_component=(0, 0),
                                 P_target=None, P_mask=None, E_avg=100e9, nu_avg=0.3, tol_NW=1e-5, tol_CG=1e-6,
                                 max_NW=50, tol_macro=1e6, max_iter_macro=20, enable_console=True, model_type="svk",
                                 plane_mode="plane_strain", A_m=0.0, B_m=0.0, C_m=0.0, solver="al", pixel=1.0,
                                 v1=0.0, v2=0.0, v3=0.0, g1=0.0, g2=0.0, g3=0.0, g4=0.0,
                                 strain_capping_enabled=False, strain_capping_limit=None,
                                 strain_capping_tangent_ratio=0.1, strain_capping_type="piecewise",
                                 strain_capping_smooth_power=1.0):
    """One 2D finite-strain load step (finite_strain_simulator.py:2356-2572),
    solver "newton_cg", "dbfft" (:1381-1507) or "al" (:2007-2140) -> (F, P, Sig, None, F_bar).  The plane problem runs
    through the 3D device step on an (nx, 1, ny) grid: F_yy = 1, no
    out-of-plane shear, the projection's y wave number 0, so the in-plane
    fields, the Krylov iterates and the Newton / macro decisions are those of
    the 2D step (its stop test over the in-plane |F|, and in plane stress
    Cauchy's J = det F * F33).  Laws: svk, landau and murnaghan in both plane
    modes (plane-stress landau uncapped: the reference's capped E33 iteration
    stops on a grid-wide 1e-9 with linear convergence), neo_hookean in plane
    strain; lam = C4[0,0,1,1] and mu = C4[0,1,0,1] of the 2D stiffness are
    read."""
    if solver not in ("newton_cg", "dbfft", "al"):
        raise NotImplementedError(f"solver={solver!r}: the 2D device step implements 'newton_cg', 'al' and 'dbfft'")
    if not plane_law_supported(model_type, plane_mode, strain_capping_enabled):
        raise NotImplementedError(f"2D {model_type} in {plane_mode}: the device step covers svk, landau, "
                                  "murnaghan and neo_hookean")
    C4 = np.asarray(C4)
    nx, ny = C4.shape[4:]
    sh3 = (nx, 1, ny)
    lam = np.ascontiguousarray(C4[0, 0, 1, 1]).reshape(sh3)
    mu = np.ascontiguousarray(C4[0, 1, 0, 1]).reshape(sh3)
    s = FsSolver(lam, mu, precision=0, Fp=None if Fp is None else _embed2(Fp))
    s.F.copy_(dev.upload(np.ascontiguousarray(_embed2(F)), 0))
    mat = material(model_type, A_m, B_m, C_m, v1, v2, v3, g1, g2, g3, g4, strain_capping_enabled,
                   strain_capping_limit, strain_capping_tangent_ratio, strain_capping_type,
                   strain_capping_smooth_power, mu_mean=float(np.mean(mu)), ndim=2)
    mat.plane = 2 if plane_mode == "plane_stress" else 1
    Pt = None if P_target is None else _embed2(P_target, 0.0)
    Pm = None if P_mask is None else _embed2(np.asarray(P_mask, dtype=float), 0.0).astype(bool)
    drv = (_PLANE_AXES[driving_component[0]], _PLANE_AXES[driving_component[1]])
    info = s.step(_embed2(F_bar), mat, drv, Pt, Pm, E_avg, nu_avg, tol_NW, tol_CG, max_NW, tol_macro,
                  max_iter_macro, solver=solver, pixel=pixel)
    return (_extract2(dev.download(s.F)), _extract2(dev.download(s.P)), _extract2(dev.download(s.sig)), None,
            _extract2(np.array(info.F_bar[:]).reshape(3, 3)))


def constitutive_hyperelastic_3d(F, C4, I2, I4, I4rt, Fp=None, model_type="svk", A_m=0.0, B_m=0.0, C_m=0.0, v1=0.0,
                                 v2=0.0, v3=0.0, g1=0.0, g2=0.0, g3=0.0, g4=0.0, strain_capping_enabled=False,
                                 strain_capping_limit=None, strain_capping_tangent_ratio=0.1,
                                 strain_capping_type="piecewise", strain_capping_smooth_power=1.0, need_tangent=True,
                                 tangent_method="mandel"):
    """(P, K4) of the reference (:1647-1885).  K4 (81 comps per voxel) is
    assembled from nine tangent actions only when need_tangent is True."""
    torch = dev.torch
    dev.require_gpu()
    C4 = np.asarray(C4)
    lam = np.ascontiguousarray(C4[0, 0, 1, 1])
    mu = np.ascontiguousarray(C4[0, 1, 0, 1])
    sh = lam.shape
    nv = int(np.prod(sh))
    mat = material(model_type, A_m, B_m, C_m, v1, v2, v3, g1, g2, g3, g4, strain_capping_enabled,
                   strain_capping_limit, strain_capping_tangent_ratio, strain_capping_type,
                   strain_capping_smooth_power, mu_mean=float(np.mean(mu)))
    Fd = dev.upload(np.ascontiguousarray(F, dtype=np.float64), 0)
    Fpd = None if Fp is None else dev.upload(np.ascontiguousarray(Fp, dtype=np.float64), 0)
    ld, md = dev.upload(lam, 0), dev.upload(mu, 0)
    P = torch.empty((3, 3) + sh, dtype=torch.float64, device="cuda")
    lib = _lib.load()
    st = ctypes.c_void_p(dev.stream_ptr())
    _lib.check(lib.mgk_fs_constitutive(nv, 0, _lib.dptr(Fd), _lib.dptr(Fpd), _lib.dptr(ld), _lib.dptr(md),
                                       ctypes.byref(mat), _lib.dptr(P), None, None, st), "mgk_fs_constitutive")
    K4 = None
    if need_tangent:
        K4 = np.empty((3, 3, 3, 3) + sh)
        dP = torch.empty((3, 3) + sh, dtype=torch.float64, device="cuda")
        for k in range(3):
            for l in range(3):
                e = torch.zeros((3, 3) + sh, dtype=torch.float64, device="cuda")
                e[k, l] = 1.0
                _lib.check(lib.mgk_fs_constitutive(nv, 0, _lib.dptr(Fd), _lib.dptr(Fpd), _lib.dptr(ld),
                                                   _lib.dptr(md), ctypes.byref(mat), None, _lib.dptr(e),
                                                   _lib.dptr(dP), st), "mgk_fs_constitutive")
                # the matvec convention: (K4 : dF^T)^T_ij = sum_kl K4[j,i,k,l] dF[k,l]
                K4[:, :, k, l] = dev.download(dP).swapaxes(0, 1)
    return dev.download(P), K4


def constitutive_hyperelastic_2d(F, C4, I2, I4, I4rt, Fp=None, model_type="svk", plane_mode="plane_strain",
                                 A_m=0.0, B_m=0.0, C_m=0.0, v1=0.0, v2=0.0, v3=0.0, g1=0.0, g2=0.0, g3=0.0, g4=0.0,
                                 strain_capping_enabled=False, strain_capping_limit=None,
                                 strain_capping_tangent_ratio=0.1, strain_capping_type="piecewise",
                                 strain_capping_smooth_power=1.0, need_tangent=True, tangent_method="mandel"):
    """(P, K4, F33) of the reference's 2D law (:591-1091) on the device, through
    the (nx, 1, ny) embedding of the plane problem (mgk_fs_constitutive_2d).  K4
    (16 comps per voxel) is assembled from four tangent actions only when
    need_tangent is True."""
    torch = dev.torch
    dev.require_gpu()
    if not plane_law_supported(model_type, plane_mode):
        raise NotImplementedError(f"2D {model_type} in {plane_mode}")
    C4 = np.asarray(C4)
    nx, ny = C4.shape[4:]
    sh3 = (nx, 1, ny)
    lam = np.ascontiguousarray(C4[0, 0, 1, 1]).reshape(sh3)
    mu = np.ascontiguousarray(C4[0, 1, 0, 1]).reshape(sh3)
    nv = nx * ny
    mat = material(model_type, A_m, B_m, C_m, v1, v2, v3, g1, g2, g3, g4, strain_capping_enabled,
                   strain_capping_limit, strain_capping_tangent_ratio, strain_capping_type,
                   strain_capping_smooth_power, mu_mean=float(np.mean(mu)), ndim=2)
    mat.plane = 2 if plane_mode == "plane_stress" else 1
    Fd = dev.upload(np.ascontiguousarray(_embed2(F)), 0)
    Fpd = None if Fp is None else dev.upload(np.ascontiguousarray(_embed2(Fp)), 0)
    ld, md = dev.upload(lam, 0), dev.upload(mu, 0)
    P = torch.empty((3, 3) + sh3, dtype=torch.float64, device="cuda")
    F33 = torch.empty(sh3, dtype=torch.float64, device="cuda")
    lib = _lib.load()
    st = ctypes.c_void_p(dev.stream_ptr())

    def call(Pout, e, dP, f33):
        rc = lib.mgk_fs_constitutive_2d(nv, 0, _lib.dptr(Fd), _lib.dptr(Fpd), _lib.dptr(ld), _lib.dptr(md),
                                        ctypes.byref(mat), _lib.dptr(Pout), _lib.dptr(e), _lib.dptr(dP),
                                        _lib.dptr(f33), st)
        if rc == _lib.MGK_ERR_NONFINITE:
            raise FloatingPointError(_lib.last_error().split(" (/")[0])
        _lib.check(rc, "mgk_fs_constitutive_2d")

    call(P, None, None, F33)
    K4 = None
    if need_tangent:
        K4 = np.empty((2, 2, 2, 2, nx, ny))
        dP = torch.empty((3, 3) + sh3, dtype=torch.float64, device="cuda")
        for k in range(2):
            for l in range(2):
                e2 = np.zeros((2, 2, nx, ny))
                e2[k, l] = 1.0
                e = dev.upload(np.ascontiguousarray(_embed2(e2, 0.0)), 0)
                call(None, e, dP, None)
                # the matvec convention: (K4 : dF^T)^T_ij = sum_kl K4[j,i,k,l] dF[k,l]
                K4[:, :, k, l] = _extract2(dev.download(dP)).swapaxes(0, 1)
    return _extract2(dev.download(P)), K4, dev.download(F33).reshape(nx, ny)


def cauchy_from_P_3d(P, F):
    """sigma = P F^T / det F (:1892-1904)."""
    torch = dev.torch
    dev.require_gpu()
    Pd = dev.upload(np.ascontiguousarray(P, dtype=np.float64), 0)
    Fd = dev.upload(np.ascontiguousarray(F, dtype=np.float64), 0)
    S = torch.empty_like(Pd)
    nv = int(np.prod(Pd.shape[2:]))
    _lib.check(_lib.load().mgk_fs_cauchy(nv, 0, _lib.dptr(Pd), _lib.dptr(Fd), _lib.dptr(S),
                                         ctypes.c_void_p(dev.stream_ptr())), "mgk_fs_cauchy")
    return dev.download(S)


def build_finite_strain_bc_3d(driving_component, eps_target_step, mixed_targets, ndim=3, F_bar_initial=None):
    """Macro boundary conditions of one load step (:2852-2899), host logic."""
    F_bar = np.eye(ndim) if F_bar_initial is None else F_bar_initial.copy()
    F_mask = np.zeros((ndim, ndim), dtype=bool)
    P_target = np.zeros((ndim, ndim))
    P_mask = np.zeros((ndim, ndim), dtype=bool)
    i_drv, j_drv = driving_component
    if i_drv == j_drv:
        F_bar[i_drv, j_drv] = 1.0 + eps_target_step
    else:
        F_bar[i_drv, j_drv] = eps_target_step
    F_mask[i_drv, j_drv] = True
    for ii in range(ndim):
        for jj in range(ndim):
            if ii == jj or (ii, jj) == (i_drv, j_drv):
                continue
            F_bar[ii, jj] = 0.0
            F_mask[ii, jj] = True
    for k in range(ndim):
        if k == i_drv and i_drv == j_drv:
            continue
        if (k, k) in mixed_targets:
            P_target[k, k] = mixed_targets[(k, k)]
            P_mask[k, k] = True
        else:
            F_bar[k, k] = 1.0
            F_mask[k, k] = True
    return F_bar, F_mask, P_target, P_mask


def build_finite_strain_bc(driving_component, eps_target_step, mixed_targets, plane_mode, ndim=2,
                           F_bar_initial=None):
    """2D macro boundary conditions of one load step (:1912-1974): the same
    rules as the 3D builder (plane_mode does not enter them)."""
    return build_finite_strain_bc_3d(driving_component, eps_target_step, mixed_targets, ndim=ndim,
                                     F_bar_initial=F_bar_initial)


def finite_strain_simulation_2d(E, nu, driving_component, eps_target, n_steps, mixed_targets=None,
                                plane_mode="plane_strain", pixel=1.0, tol_NW=1e-5, tol_CG=1e-6, max_NW=20,
                                tol_macro=1e6, max_iter_macro=20, store=True, log_path=None, global_log_path=None,
                                enable_console=True, checkpoint_interval="none", checkpoint_path=None,
                                vtk_interval="none", vtk_path=None, model_type="svk", A_m=0.0, B_m=0.0, C_m=0.0,
                                solver="al", v1=0.0, v2=0.0, v3=0.0, g1=0.0, g2=0.0, g3=0.0, g4=0.0,
                                strain_capping_enabled=False, strain_capping_limit=None,
                                strain_capping_tangent_ratio=0.1, strain_capping_type="piecewise",
                                strain_capping_smooth_power=1.0):
    """2D incremental load stepping (:2580-2845) with F resident on the device
    (the plane problem on an (nx, 1, ny) grid, as finite_strain_solver_step_2d).
    Logs, checkpoints and VTK output are accepted and ignored.  Returns
    (F_macro, Sig_macro, P_macro, F_list, Sig_list) in 2D shapes."""
    if solver not in ("newton_cg", "dbfft", "al"):
        raise NotImplementedError(f"solver={solver!r}: the 2D device step implements 'newton_cg', 'al' and 'dbfft'")
    if not plane_law_supported(model_type, plane_mode, strain_capping_enabled):
        raise NotImplementedError(f"2D {model_type} in {plane_mode}")
    mixed_targets = {} if mixed_targets is None else mixed_targets
    E = np.asarray(E, dtype=np.float64)
    nu = np.asarray(nu, dtype=np.float64)
    nx, ny = E.shape
    sh3 = (nx, 1, ny)
    mu = E / (2.0 * (1.0 + nu))
    lam = E * nu / (1.0 - nu ** 2) if plane_mode == "plane_stress" else E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
    s = FsSolver(lam.reshape(sh3), mu.reshape(sh3), precision=0)
    mat = material(model_type, A_m, B_m, C_m, v1, v2, v3, g1, g2, g3, g4, strain_capping_enabled,
                   strain_capping_limit, strain_capping_tangent_ratio, strain_capping_type,
                   strain_capping_smooth_power, mu_mean=float(mu.mean()), ndim=2)
    mat.plane = 2 if plane_mode == "plane_stress" else 1
    drv = (_PLANE_AXES[driving_component[0]], _PLANE_AXES[driving_component[1]])
    Fm, Sm, Pm, Fl, Sl = [], [], [], [], []
    F_bar_init = None
    for eps_s in np.linspace(0.0, eps_target, n_steps + 1):
        F_bar, _, P_tgt, P_mask = build_finite_strain_bc(driving_component, eps_s, mixed_targets, plane_mode,
                                                         F_bar_initial=F_bar_init)
        info = s.step(_embed2(F_bar), mat, drv, _embed2(P_tgt, 0.0), _embed2(P_mask.astype(float), 0.0).astype(bool),
                      E.mean(), nu.mean(), tol_NW, tol_CG, max_NW, tol_macro, max_iter_macro, solver=solver,
                      pixel=pixel)
        F_bar_init = _extract2(np.array(info.F_bar[:]).reshape(3, 3))
        Fm.append(_extract2(np.array(info.F_mean[:]).reshape(3, 3)))
        Sm.append(_extract2(np.array(info.sig_mean[:]).reshape(3, 3)))
        Pm.append(_extract2(s.P.double().mean(dim=(2, 3, 4)).cpu().numpy()))
        if store:
            Fl.append(_extract2(dev.download(s.F)))
            Sl.append(_extract2(dev.download(s.sig)))
    return np.array(Fm), np.array(Sm), np.array(Pm), Fl, Sl


def finite_strain_simulation_3d(E, nu, driving_component, eps_target, n_steps, mixed_targets=None, pixel=1.0,
                                tol_NW=1e-5, tol_CG=1e-6, max_NW=20, tol_macro=1e6, max_iter_macro=20, store=True,
                                log_path=None, global_log_path=None, enable_console=True,
                                checkpoint_interval="none", checkpoint_path=None, vtk_interval="none", vtk_path=None,
                                model_type="svk", A_m=0.0, B_m=0.0, C_m=0.0, solver="al", v1=0.0, v2=0.0, v3=0.0,
                                g1=0.0, g2=0.0, g3=0.0, g4=0.0, strain_capping_enabled=False,
                                strain_capping_limit=None, strain_capping_tangent_ratio=0.1,
                                strain_capping_type="piecewise", strain_capping_smooth_power=1.0, precision=0,
                                step_callback=None):
    """Incremental load stepping (:3175-3413) with F resident on the device.
    Logs, checkpoints and VTK output of the reference are accepted and
    ignored.  Returns (F_macro_list, Sig_macro_list, P_macro_list, F_list,
    Sig_list); the field lists are filled only when store=True."""
    if solver not in ("newton_cg", "al", "dbfft"):
        raise NotImplementedError("the device path implements solver='newton_cg', 'al' and 'dbfft'")
    mixed_targets = {} if mixed_targets is None else mixed_targets
    E = np.asarray(E, dtype=np.float64)
    nu = np.asarray(nu, dtype=np.float64)
    mu = E / (2.0 * (1.0 + nu))
    lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
    s = FsSolver(lam, mu, precision=precision)
    mat = material(model_type, A_m, B_m, C_m, v1, v2, v3, g1, g2, g3, g4, strain_capping_enabled,
                   strain_capping_limit, strain_capping_tangent_ratio, strain_capping_type,
                   strain_capping_smooth_power, mu_mean=float(mu.mean()))
    eps_steps = np.linspace(0.0, eps_target, n_steps + 1)
    Fm, Sm, Pm, Fl, Sl = [], [], [], [], []
    F_bar_init = None
    for st, eps_s in enumerate(eps_steps):
        F_bar, F_mask, P_tgt, P_mask = build_finite_strain_bc_3d(driving_component, eps_s, mixed_targets,
                                                                 F_bar_initial=F_bar_init)
        info = s.step(F_bar, mat, driving_component, P_tgt, P_mask, E.mean(), nu.mean(), tol_NW, tol_CG, max_NW,
                      tol_macro, max_iter_macro, solver=solver, pixel=pixel)
        F_bar_init = np.array(info.F_bar[:]).reshape(3, 3)
        Fm.append(np.array(info.F_mean[:]).reshape(3, 3))
        Sm.append(np.array(info.sig_mean[:]).reshape(3, 3))
        Pm.append(s.P.double().mean(dim=(2, 3, 4)).cpu().numpy())
        if store:
            Fl.append(dev.download(s.F))
            Sl.append(dev.download(s.sig))
        if step_callback is not None:
            step_callback(st, info)
    return np.array(Fm), np.array(Sm), np.array(Pm), Fl, Sl
