"""Device-resident STZ/KMC state and the host-array drop-in helpers.

Layouts on the device (v = voxel index in the reference's x-major order):
  Q, Q0   (M, NV) FP64          catalog (CK, M, NV) FP64
  soft    (2, NV) = (g_p, g_t)  tlast (NV,)   prev (d*d, NV) (directional)
  eps_p   (C, NV) symmetric SoA in the LS solver's precision
CK = 2 for planar modes (gxx, gxy) -- the reference's default catalog
(mgkmc/stz/catalog.py:7-40, kmc_simulator_functions.py:78-98 pure shear) --
or d*d for general modes (stz_catalog_glass_3d, simple-shear 2D).
"""
import ctypes

import numpy as np

from . import _lib
from . import device as dev

SIG_SYM64, SIG_SYM32, SIG_GEN64 = 0, 1, 2
CAT_PLANAR, CAT_GENERAL = 0, 1


def params(volume=1.0, rate_volume=1.0, nu0=1e13, temperature=0.0, cap=0.0, tau=np.inf, time=0.0,
           scheme=0, cascade=False):
    p = _lib.MgkKmcParams()
    p.volume = float(volume)
    p.rate_volume = float(rate_volume)
    p.nu0 = float(nu0)
    p.temperature = float(temperature)
    p.cap = float(cap)
    p.tau = float(tau)
    p.time = float(time)
    p.scheme = int(scheme)
    p.cascade = 1 if cascade else 0
    return p


def is_planar(catalog, d):
    """True if every mode is [[a, b(,0)],[b, -a(,0)](,[0,0,0])] (the reference's
    default glass catalog), so two parameters per mode describe it exactly."""
    c = np.asarray(catalog)
    ok = np.array_equal(c[..., 1, 1], -c[..., 0, 0]) and np.array_equal(c[..., 0, 1], c[..., 1, 0])
    if d == 3:
        ok = ok and not np.any(c[..., 2, :]) and not np.any(c[..., :, 2])
    return bool(ok)


def catalog_to_device(catalog, d, planar):
    """(..., M, d, d) host catalog -> (CK, M, NV) device FP64."""
    c = np.asarray(catalog, dtype=np.float64)
    M = c.shape[-3]
    c = c.reshape(-1, M, d, d)  # (NV, M, d, d)
    if planar:
        arr = np.stack([c[..., 0, 0], c[..., 0, 1]], 0)  # (2, NV, M)
    else:
        arr = c.reshape(c.shape[0], M, d * d).transpose(2, 0, 1)  # (dd, NV, M)
    arr = np.ascontiguousarray(arr.transpose(0, 2, 1))  # (CK, M, NV)
    return dev.upload(arr, 0)


def modes_to_device(a):
    """(..., M) host -> (M, NV) device FP64."""
    a = np.asarray(a, dtype=np.float64)
    M = a.shape[-1]
    return dev.upload(np.ascontiguousarray(a.reshape(-1, M).T), 0)


def modes_from_device(t, shape):
    """(M, NV) device -> (*shape, M) host."""
    a = dev.download(t)
    return np.ascontiguousarray(a.T).reshape(tuple(shape) + (a.shape[0],))


class KmcWorkspace:
    def __init__(self, nvox):
        n = ctypes.c_size_t()
        _lib.check(_lib.load().mgk_kmc_workspace_bytes(int(nvox), ctypes.byref(n)), "mgk_kmc_workspace_bytes")
        self.buf = dev.torch.empty(int(n.value), dtype=dev.torch.uint8, device="cuda")

    @property
    def ptr(self):
        return _lib.dptr(self.buf)


class KmcState:
    """All per-voxel STZ/KMC arrays of one simulation, resident in HBM."""

    def __init__(self, shape, M, Q0, catalog, soft_prop=None, last_event_time=None, prev_strain_dir=None,
                 directional=False, eps_precision=0, eps_shadow32=False):
        dev.require_gpu()
        torch = dev.torch
        self.shape = tuple(int(s) for s in shape)
        self.d = len(self.shape)
        self.M = int(M)
        self.NV = int(np.prod(self.shape))
        d = self.d
        if Q0 is None and catalog is None:
            # zero planar modes, filled on the device afterwards
            self.planar = True
            self.cat_kind = CAT_PLANAR
            self.CK = 2
            self.Q0 = torch.zeros((self.M, self.NV), dtype=torch.float64, device="cuda")
            self.cat = torch.zeros((2, self.M, self.NV), dtype=torch.float64, device="cuda")
        else:
            self.planar = is_planar(catalog, d)
            self.cat_kind = CAT_PLANAR if self.planar else CAT_GENERAL
            self.CK = 2 if self.planar else d * d
            self.Q0 = modes_to_device(Q0)
            self.cat = catalog_to_device(catalog, d, self.planar)
        self.Q = None  # allocated on the first barrier pass that stores Q
        f64 = dict(dtype=torch.float64, device="cuda")
        if soft_prop is None:
            self.soft = torch.zeros((2, self.NV), **f64)
        else:
            soft = np.asarray(soft_prop)[..., :2]
            self.soft = dev.upload(np.ascontiguousarray(soft.reshape(-1, 2).T), 0)
        if last_event_time is None:
            self.tlast = torch.full((self.NV,), -np.inf, **f64)
        else:
            self.tlast = dev.upload(np.asarray(last_event_time, float).reshape(-1), 0)
        self.directional = bool(directional)
        self.prev = None
        if self.directional:
            if prev_strain_dir is None:
                self.prev = torch.zeros((d * d, self.NV), **f64)
            else:
                pv = np.asarray(prev_strain_dir, float)
                self.prev = dev.upload(np.ascontiguousarray(pv.reshape(-1, d * d).T), 0)
        C = 6 if d == 3 else 3
        self.eps_precision = eps_precision
        self.eps_p = torch.zeros((C,) + self.shape, dtype=dev.torch_dtype(eps_precision), device="cuda")
        # FP32 copy of an FP64 eps_p kept in step by the flip kernel (FP32 + FP64 refinement LS)
        self.eps_p32 = None
        if eps_shadow32:
            if eps_precision != 0:
                raise ValueError("eps_shadow32 needs an FP64 eps_p")
            self.eps_p32 = torch.zeros((C,) + self.shape, dtype=torch.float32, device="cuda")
        self.ws = KmcWorkspace(self.NV)
        self.sel = _lib.MgkKmcSel()
        self._shape_c = (ctypes.c_int * d)(*self.shape)
        self.gamma = torch.zeros(d * d, dtype=torch.float64, device="cuda")

    def barrier_rates(self, sig, sig_kind, p, write_q=True, Tfield=None):
        """Fused compute_barrier + compute_rates; returns the FP64 total rate.
        write_q: also store the barriers in self.Q ([m][v] FP64); Tfield: a
        per-voxel temperature (FP64 device field) instead of p.temperature."""
        if write_q and self.Q is None:
            self.Q = dev.torch.zeros_like(self.Q0)
        tot = ctypes.c_double()
        rc = _lib.load().mgk_kmc_barrier_rates(
            self.d, self.NV, self.M, sig_kind, self.cat_kind, _lib.dptr(sig), _lib.dptr(self.Q0),
            _lib.dptr(self.cat), _lib.dptr(self.soft), _lib.dptr(self.tlast), _lib.dptr(self.prev), _lib.dptr(Tfield),
            ctypes.byref(p), _lib.dptr(self.Q) if write_q else None, self.ws.ptr,
            ctypes.c_void_p(dev.stream_ptr()), ctypes.byref(tot))
        _lib.check(rc, "mgk_kmc_barrier_rates")
        return tot.value

    def unstable_modes(self, threshold):
        """(v, m, q) host arrays of the modes with stored Q < threshold, in the
        voxel-major scan order of find_unstable (mgkmc/stz/cascade.py:4-41):
        one compaction per mode row, then a sort by v * M + m."""
        torch = dev.torch
        if getattr(self, "_fb", None) is None:
            self._fb = torch.empty(self.NV, dtype=torch.int64, device="cuda")
        cnt = ctypes.c_longlong()
        vs, ms = [], []
        for m in range(self.M):
            rc = _lib.load().mgk_find_below(self.NV, _lib.dptr(self.Q[m]), float(threshold), _lib.dptr(self._fb),
                                            ctypes.c_void_p(dev.stream_ptr()), ctypes.byref(cnt))
            _lib.check(rc, "mgk_find_below")
            if cnt.value:
                vs.append(self._fb[: cnt.value].clone())
                ms.append(torch.full((cnt.value,), m, dtype=torch.int64, device="cuda"))
        if not vs:
            e = np.zeros(0, dtype=np.int64)
            return e, e, np.zeros(0)
        v = torch.cat(vs)
        m = torch.cat(ms)
        order = torch.argsort(v * self.M + m)
        v, m = v[order], m[order]
        q = self.Q[m, v]
        return v.cpu().numpy(), m.cpu().numpy(), q.cpu().numpy()

    def select(self, u, sig, sig_kind, p, Tfield=None):
        """(voxel, mode) of the event for draw u in [0, 1): r = total * u, after
        barrier_rates with the same sig, parameters and temperature field.
        self.sel.q holds the selected mode's barrier."""
        rc = _lib.load().mgk_kmc_select(self.d, self.NV, self.M, sig_kind, self.cat_kind, float(u),
                                        _lib.dptr(sig), _lib.dptr(self.Q0), _lib.dptr(self.cat),
                                        _lib.dptr(self.soft), _lib.dptr(self.tlast), _lib.dptr(self.prev),
                                        _lib.dptr(Tfield), ctypes.byref(p), self.ws.ptr,
                                        ctypes.c_void_p(dev.stream_ptr()), ctypes.byref(self.sel))
        _lib.check(rc, "mgk_kmc_select")
        return int(self.sel.item), int(self.sel.mode)

    def flip(self, v, m, time, jp, jt, gmax, jn, eps_div=1):
        """Flip mode m of voxel v; eps_div = N > 1 adds mode / N (last
        sub-step of a sub-stepped flip), -N adds mode / N and nothing else."""
        rc = _lib.load().mgk_kmc_flip(self.d, self._shape_c, self.M, self.cat_kind, int(v), int(m), float(time),
                                      float(jp), float(jt), float(gmax), float(jn), _lib.dptr(self.cat),
                                      _lib.dptr(self.eps_p), self.eps_precision, _lib.dptr(self.eps_p32),
                                      _lib.dptr(self.soft),
                                      _lib.dptr(self.tlast), _lib.dptr(self.prev), _lib.dptr(self.gamma),
                                      int(eps_div), ctypes.c_void_p(dev.stream_ptr()))
        _lib.check(rc, "mgk_kmc_flip")

    def set_voxel(self, v, newcat=None, newq0=None):
        """Replace the modes of voxel v: newcat (M, d, d) host, newq0 (M,) host."""
        torch = dev.torch
        nc = None
        if newcat is not None:
            c = np.asarray(newcat, dtype=np.float64).reshape(self.M, self.d, self.d)
            if self.planar:
                arr = np.stack([c[:, 0, 0], c[:, 0, 1]], 1)
            else:
                arr = c.reshape(self.M, self.d * self.d)
            nc = torch.from_numpy(np.ascontiguousarray(arr)).cuda()
        nq = None if newq0 is None else torch.from_numpy(np.asarray(newq0, dtype=np.float64).copy()).cuda()
        rc = _lib.load().mgk_kmc_set_voxel(self.NV, self.M, self.CK, int(v), _lib.dptr(nc), _lib.dptr(nq),
                                           _lib.dptr(self.cat), _lib.dptr(self.Q0),
                                           ctypes.c_void_p(dev.stream_ptr()))
        _lib.check(rc, "mgk_kmc_set_voxel")
        if nc is not None or nq is not None:
            dev.torch.cuda.current_stream().synchronize()

    def voxel_coords(self, v):
        return np.unravel_index(int(v), self.shape)


# ---------------------------------------------------------------------------
# host-array drop-in implementations (upload, run the kernels, download)
# ---------------------------------------------------------------------------
def barrier_host(d, Q_field, Q0_field, sigma_field, catalog, volume, soft_prop, last_event_time, current_time,
                 prev_strain_dir, softening_cap, scheme, tau):
    shape = Q_field.shape[:-1]
    M = Q_field.shape[-1]
    st = KmcState(shape, M, Q0_field, catalog, soft_prop=soft_prop, last_event_time=last_event_time,
                  prev_strain_dir=prev_strain_dir, directional=(scheme == 1))
    sig = np.asarray(sigma_field, dtype=np.float64).reshape(-1, d * d).T
    sig_d = dev.upload(np.ascontiguousarray(sig), 0)
    p = params(volume=volume, cap=softening_cap, tau=tau, time=current_time, scheme=scheme)
    st.barrier_rates(sig_d, SIG_GEN64, p)
    Q_field[...] = modes_from_device(st.Q, shape)


def rates_host(Q_field, volume, temperature, nu0, cascade_mode):
    torch = dev.torch
    dev.require_gpu()
    M = Q_field.shape[-1]
    nvox = int(np.prod(Q_field.shape[:-1]))
    q = dev.upload(np.ascontiguousarray(Q_field, dtype=np.float64).reshape(-1), 0)
    n = nvox * M
    full = torch.empty(n, dtype=torch.float64, device="cuda")
    idx = torch.empty(n, dtype=torch.int64, device="cuda")
    val = torch.empty(n, dtype=torch.float64, device="cuda")
    ws = KmcWorkspace(nvox)
    tf = None
    if not np.isscalar(temperature) and not isinstance(temperature, (int, float, np.number)):
        tf = dev.upload(np.asarray(temperature, dtype=np.float64).reshape(-1), 0)
        temperature = 0.0
    p = params(rate_volume=volume, nu0=nu0, temperature=float(temperature), cascade=cascade_mode)
    cnt = ctypes.c_longlong()
    tot = ctypes.c_double()
    rc = _lib.load().mgk_kmc_rates_from_q(nvox, M, _lib.dptr(q), _lib.dptr(tf), ctypes.byref(p), _lib.dptr(full),
                                          _lib.dptr(idx), _lib.dptr(val), ws.ptr,
                                          ctypes.c_void_p(dev.stream_ptr()), ctypes.byref(cnt), ctypes.byref(tot))
    _lib.check(rc, "mgk_kmc_rates_from_q")
    k = cnt.value
    return dev.download(val[:k]), idx[:k].cpu().numpy(), tot.value


def select_host(rates, r):
    dev.require_gpu()
    a = dev.upload(np.ascontiguousarray(rates, dtype=np.float64), 0)
    ws = KmcWorkspace(max(1, a.numel()))
    sel = _lib.MgkKmcSel()
    rc = _lib.load().mgk_kmc_select_array(a.numel(), _lib.dptr(a), float(r), ws.ptr,
                                          ctypes.c_void_p(dev.stream_ptr()), ctypes.byref(sel))
    _lib.check(rc, "mgk_kmc_select_array")
    return int(sel.item)


def find_below_host(Q_field, threshold):
    torch = dev.torch
    dev.require_gpu()
    q = dev.upload(np.ascontiguousarray(Q_field, dtype=np.float64).reshape(-1), 0)
    idx = torch.empty(q.numel(), dtype=torch.int64, device="cuda")
    cnt = ctypes.c_longlong()
    rc = _lib.load().mgk_find_below(q.numel(), _lib.dptr(q), float(threshold), _lib.dptr(idx),
                                    ctypes.c_void_p(dev.stream_ptr()), ctypes.byref(cnt))
    _lib.check(rc, "mgk_find_below")
    flat = idx[: cnt.value].cpu().numpy()
    return np.stack(np.unravel_index(flat, Q_field.shape), 1).astype(np.int64)


def flip_host(d, eps_plastic_field, soft_prop_field, last_event_time_field, catalog, coords, m, current_time,
              jp, jt, g_max, jn_frac):
    shape = eps_plastic_field.shape[:d]
    M = catalog.shape[d]
    v = int(np.ravel_multi_index(tuple(int(c) for c in coords), shape))
    st = KmcState(shape, M, np.zeros(shape + (M,)), catalog, soft_prop=soft_prop_field,
                  last_event_time=last_event_time_field)
    sym, anti = dev.to_sym_soa(eps_plastic_field, d)
    st.eps_p.copy_(dev.upload(sym, 0))
    st.flip(v, m, current_time, jp, jt, g_max, jn_frac)
    eps_new = dev.from_sym_soa(dev.download(st.eps_p), d) + anti
    if not st.planar:
        # general modes may carry an antisymmetric part (2D simple shear)
        c = np.asarray(catalog)[tuple(int(x) for x in coords) + (m,)]
        a = 0.5 * (c - c.T)
        eps_new[tuple(int(x) for x in coords)] += a
    eps_plastic_field[...] = eps_new
    soft = dev.download(st.soft)
    soft_prop_field[..., 0] = soft[0].reshape(shape)
    soft_prop_field[..., 1] = soft[1].reshape(shape)
    last_event_time_field[...] = dev.download(st.tlast).reshape(shape)
