"""Device memory and stream plumbing (PyTorch-ROCm tensors as HBM buffers).

Symmetric tensor fields live on the device as SoA component-major arrays:
3D (6, nx, ny, nz) with components xx, yy, zz, yz, xz, xy; 2D (3, nx, ny) with
xx, yy, xy.  The reference keeps (nx, ny[, nz], d, d) AoS host arrays; the
conversions below are the only place the two layouts meet.
"""
import numpy as np

try:
    import torch
except Exception:  # pragma: no cover - torch is part of the image
    torch = None

SYM3 = ((0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1))
SYM2 = ((0, 0), (1, 1), (0, 1))


def sym_pairs(d):
    return SYM3 if d == 3 else SYM2


def require_gpu():
    if torch is None or not torch.cuda.is_available():
        raise RuntimeError("mgkmc_amd needs a ROCm GPU (torch.cuda.is_available() is False); "
                           "there is no CPU fallback")
    from . import _lib
    _lib.load()


def torch_dtype(precision):
    return torch.float64 if precision == 0 else torch.float32


def stream_ptr():
    return torch.cuda.current_stream().cuda_stream


def to_sym_soa(a, d):
    """(..., d, d) host array -> (C, ...) symmetric components (off-diagonals
    averaged) and the antisymmetric part (..., d, d)."""
    a = np.asarray(a, dtype=np.float64)
    comps = []
    for (i, j) in sym_pairs(d):
        if i == j:
            comps.append(a[..., i, i])
        else:
            comps.append(0.5 * (a[..., i, j] + a[..., j, i]))
    anti = 0.5 * (a - np.swapaxes(a, -1, -2))
    return np.ascontiguousarray(np.stack(comps, 0)), anti


def from_sym_soa(s, d):
    """(C, ...) symmetric components -> (..., d, d) host array."""
    s = np.asarray(s)
    shape = s.shape[1:]
    out = np.empty(shape + (d, d), dtype=np.float64)
    for c, (i, j) in enumerate(sym_pairs(d)):
        out[..., i, j] = s[c]
        out[..., j, i] = s[c]
    return out


def upload(a, precision):
    return torch.from_numpy(np.ascontiguousarray(a)).to(device="cuda", dtype=torch_dtype(precision))


def download(t):
    return t.detach().to(device="cpu", dtype=torch.float64).numpy()


class Workspace:
    """Grow-only device scratch buffer owned by a solver object."""

    def __init__(self):
        self.buf = None

    def get(self, nbytes):
        if self.buf is None or self.buf.numel() < nbytes:
            self.buf = torch.empty(int(nbytes), dtype=torch.uint8, device="cuda")
        return self.buf
