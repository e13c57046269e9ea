"""mgk_ls_solve_noalloc (the LS solve whose workspace and twiddle tables live in
the library's static device scratch) against mgk_ls_solve with a caller
workspace, through the C ABI: identical fields, means and iteration counts on
two calls in a row, the second on a smaller grid, in both precisions; a grid
beyond the scratch is refused before any launch."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _inputs(shape, prec, seed):
    import torch
    from mgkmc_amd.elasticity import compute_lame_2d, compute_lame_3d
    rng = np.random.default_rng(seed)
    d = len(shape)
    E = 70e9 * rng.uniform(0.9, 1.1, shape)
    nu = np.full(shape, 0.3)
    lam, mu = compute_lame_3d(E, nu) if d == 3 else compute_lame_2d(E, nu)
    lam0, mu0 = (compute_lame_3d if d == 3 else compute_lame_2d)(E.mean(), nu.mean())
    C = 6 if d == 3 else 3
    ep = rng.standard_normal((C,) + shape) * 1e-3
    dt = torch.float64 if prec == 0 else torch.float32
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(device="cuda", dtype=dt)
    eb = np.zeros((d, d))
    eb[0, 0] = 0.01
    eb[0, 1] = eb[1, 0] = 0.002
    # FP32 stops at the FP32 noise floor: its tolerance is the refined form's FP32 stage's
    return dict(lam=t(lam), mu=t(mu), lam0=float(lam0), mu0=float(mu0), ep=t(ep), eb=eb, C=C,
                tol=1e-9 if prec == 0 else 1e-6)


def _solve(lib, shape, prec, x, noalloc):
    import torch
    from mgkmc_amd import _lib
    d = len(shape)
    nv = int(np.prod(shape))
    eps = torch.zeros((x["C"], nv), dtype=x["lam"].dtype, device="cuda")
    sig = torch.zeros_like(eps)
    info = _lib.MgkLsInfo()
    eb = np.ascontiguousarray(x["eb"].reshape(-1))
    ebp = eb.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    sh = _lib.shape_arg(shape)
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    if noalloc:
        rc = lib.mgk_ls_solve_noalloc(d, sh, prec, _lib.dptr(x["lam"]), _lib.dptr(x["mu"]), x["lam0"], x["mu0"],
                                      ebp, _lib.dptr(x["ep"]), _lib.dptr(eps), _lib.dptr(sig), 0, 200, x["tol"], stream,
                                      ctypes.byref(info))
    else:
        nb = ctypes.c_size_t()
        assert lib.mgk_ls_workspace_bytes(d, sh, prec, ctypes.byref(nb)) == 0
        ws = torch.empty(int(nb.value), dtype=torch.uint8, device="cuda")
        rc = lib.mgk_ls_solve(d, sh, prec, _lib.dptr(x["lam"]), _lib.dptr(x["mu"]), x["lam0"], x["mu0"], ebp,
                              _lib.dptr(x["ep"]), _lib.dptr(eps), _lib.dptr(sig), 0, 200, x["tol"], _lib.dptr(ws),
                              int(nb.value), stream, ctypes.byref(info))
    torch.cuda.synchronize()
    return rc, eps.cpu().numpy(), sig.cpu().numpy(), info


@pytest.mark.parametrize("prec", [0, 1])
def test_noalloc_solve_equals_workspace_solve(gpu, prec):
    from mgkmc_amd import _lib
    lib = _lib.load()
    for k, shape in enumerate([(40, 36, 32), (24, 20, 16), (48, 40)]):
        print(f"[noalloc] precision {prec} shape {shape}", flush=True)
        x = _inputs(shape, prec, 10 + k)
        rc0, e0, s0, i0 = _solve(lib, shape, prec, x, False)
        rc1, e1, s1, i1 = _solve(lib, shape, prec, x, True)
        assert rc0 == 0 and rc1 == 0, (rc0, rc1, _lib.last_error())
        assert i0.iters == i1.iters and i0.iters > 2
        # the same kernels on the same inputs and twiddles: the same bits
        assert np.array_equal(e0, e1) and np.array_equal(s0, s1)
        assert list(i0.sig_mean) == list(i1.sig_mean) and i0.diff == i1.diff


def test_noalloc_refuses_grid_beyond_scratch(gpu):
    from mgkmc_amd import _lib
    lib = _lib.load()
    info = _lib.MgkLsInfo()
    eb = np.zeros(9)
    dummy = ctypes.c_void_p(16)  # never dereferenced: the size check comes first
    rc = lib.mgk_ls_solve_noalloc(3, _lib.shape_arg((128, 128, 128)), 0, dummy, dummy, 1.0, 1.0,
                                  eb.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), None, dummy, None, 0, 10,
                                  1e-6, None, ctypes.byref(info))
    assert rc == _lib.MGK_ERR_WORKSPACE
    assert "scratch" in _lib.last_error()
