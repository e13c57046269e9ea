// Fast patching: the stress predictor of the reference's KMC loops
// (mgkmc/aqs.py:1021-1054, mgkmc/kmc_simulator.py:764-774, :1044-1066,
// :1094-1096).  After a flip the stress field is not re-solved; instead
//   sig += sig_unit * a                (macro-unit stress over the waiting time)
//   sig[box around v] += sum_b c_b K_b (truncated point solutions, per offset)
//   sig += sum_b c_b mm_b              (the box's missing mean)
// in that order per voxel, each step rounded as numpy rounds it (products
// and sums separately: contraction off).  One pass over the field, which
// also leaves FP64 partial sums of the new field for its mean.
//
// Device layouts: sig symmetric SoA [c][v] (FP64 or FP32; 3D c over
// xx,yy,zz,yz,xz,xy, 2D xx,yy,xy), sig_unit FP64 in the same layout, kernels
// K [b][c][wx][wy][wz] FP64 with wx = 2 Rx + 1 (a flat axis has R = 0).  The
// 2D problem uses the (nx, 1, ny) grid of the LS solver.
#include "mgk_common.h"
#include "mgk_host.h"
#include "../../include/mgkmc_hip.h"
#include <algorithm>

#pragma clang fp contract(off)

namespace {

constexpr int kPatchBlocks = 1024;

struct PatchArgs {
  int nx, ny, nz;
  long long NV;
  int cx, cy, cz;
  int Rx, Ry, Rz;
  int nb;           // basis count (0: no patch, only the macro-unit term and the shift)
  double coef[5];
  double shift[6];  // per component, the combined missing mean
  double a;         // scale of sig_unit (unused when sig_unit == NULL)
  const double* K;
  // coefficients read on the device from the planar catalog (gxx, gxy of mode m at
  // v, before its redraw) with the missing means mm: used when cat != NULL
  const double* cat;
  long long off_xx, off_xy;
  double mm[5][6];
};

// basis weights and combined missing mean of the planar mode (gxx, gxy)
// [[gxx, gxy, 0], [gxy, -gxx, 0], [0, 0, 0]] (2D: [[gxx, gxy], [gxy, -gxx]]), in
// the reference's order (mgkmc/aqs.py:1030-1037, kmc_simulator.py:1055-1057)
template <int C>
__device__ __forceinline__ void patch_coefs_from_cat(const PatchArgs& p, double* coef, double* shift) {
  const double gxx = p.cat[p.off_xx], gxy = p.cat[p.off_xy];
  if (C == 6) {
    coef[0] = gxx;
    coef[1] = -0.0;  // -C22 of a planar mode
    coef[2] = gxy;
    coef[3] = 0.0;
    coef[4] = 0.0;
  } else {
    coef[0] = gxx;
    coef[1] = gxy;
  }
#pragma unroll
  for (int c = 0; c < C; ++c) {
    double t = coef[0] * p.mm[0][c];
#pragma unroll
    for (int b = 1; b < 5; ++b)
      if (b < p.nb) t = t + coef[b] * p.mm[b][c];
    shift[c] = t;
  }
}

// smallest offset d in [-R, R] with (c0 + d) mod n == c, or R + 1 if none
__device__ __forceinline__ int first_off(int c, int c0, int n, int R) {
  int d0 = (c - c0) % n;
  if (d0 < 0) d0 += n;
  return d0 - n * ((d0 + R) / n);
}

template <typename T, int C>
__global__ __launch_bounds__(MGK_NT) void k_kmc_patch(PatchArgs p, T* __restrict__ sig,
                                                      const double* __restrict__ unit, double* __restrict__ parts) {
  __shared__ double red[C * (MGK_NT / 64)];
  double acc[C];
#pragma unroll
  for (int c = 0; c < C; ++c) acc[c] = 0.0;
  double coef[5], shift[C];
  if (p.cat) {
    patch_coefs_from_cat<C>(p, coef, shift);
  } else {
#pragma unroll
    for (int b = 0; b < 5; ++b) coef[b] = p.coef[b];
#pragma unroll
    for (int c = 0; c < C; ++c) shift[c] = p.shift[c];
  }
  const int wx = 2 * p.Rx + 1, wy = 2 * p.Ry + 1, wz = 2 * p.Rz + 1;
  const size_t cstride = (size_t)wx * wy * wz;
  for (long long v = blockIdx.x * (long long)MGK_NT + threadIdx.x; v < p.NV; v += (long long)gridDim.x * MGK_NT) {
    const int z = (int)(v % p.nz);
    const long long r = v / p.nz;
    const int y = (int)(r % p.ny), x = (int)(r / p.ny);
    double s[C];
#pragma unroll
    for (int c = 0; c < C; ++c) {
      s[c] = (double)sig[(size_t)c * p.NV + v];
      if (unit) s[c] = s[c] + unit[(size_t)c * p.NV + v] * p.a;
    }
    if (p.nb > 0) {
      const int dx0 = first_off(x, p.cx, p.nx, p.Rx), dy0 = first_off(y, p.cy, p.ny, p.Ry),
                dz0 = first_off(z, p.cz, p.nz, p.Rz);
      for (int dx = dx0; dx <= p.Rx; dx += p.nx)
        for (int dy = dy0; dy <= p.Ry; dy += p.ny)
          for (int dz = dz0; dz <= p.Rz; dz += p.nz) {
            const size_t o = ((size_t)(dx + p.Rx) * wy + (dy + p.Ry)) * wz + (dz + p.Rz);
#pragma unroll
            for (int c = 0; c < C; ++c) {
              const double* Kc = p.K + (size_t)c * cstride + o;
              double q = coef[0] * Kc[0];
#pragma unroll
              for (int b = 1; b < 5; ++b)
                if (b < p.nb) q = q + coef[b] * Kc[(size_t)b * C * cstride];
              s[c] = s[c] + q;
            }
          }
    }
#pragma unroll
    for (int c = 0; c < C; ++c) {
      s[c] = s[c] + shift[c];
      const T w = (T)s[c];
      sig[(size_t)c * p.NV + v] = w;
      acc[c] += (double)w;
    }
  }
  block_sum<C>(acc, red);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int c = 0; c < C; ++c) parts[(size_t)blockIdx.x * C + c] = acc[c];
  }
}

// means[c] = (sum over blocks in order) / NV
template <int C>
__global__ __launch_bounds__(MGK_NT) void k_kmc_patch_mean(const double* __restrict__ parts, int nblk, double inv_nv,
                                                           double* __restrict__ means) {
  __shared__ double red[C * (MGK_NT / 64)];
  double acc[C];
#pragma unroll
  for (int c = 0; c < C; ++c) acc[c] = 0.0;
  for (int b = threadIdx.x; b < nblk; b += MGK_NT) {
#pragma unroll
    for (int c = 0; c < C; ++c) acc[c] += parts[(size_t)b * C + c];
  }
  block_sum<C>(acc, red);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int c = 0; c < C; ++c) means[c] = acc[c] * inv_nv;
  }
}

}  // namespace

static int patch_launch(PatchArgs& p, int ndim, int sig_kind, void* sig, const double* sig_unit, void* workspace,
                        void* stream, double* sig_mean);

extern "C" int mgk_kmc_patch_workspace_bytes(size_t* bytes) {
  if (!bytes) return MGK_ERR_ARG;
  *bytes = (size_t)(kPatchBlocks * 6 + 8) * sizeof(double);
  return MGK_OK;
}

extern "C" int mgk_kmc_patch(int ndim, const int64_t* shape, int sig_kind, void* sig, const double* sig_unit,
                             double a, long long v, const int* radius, int nbasis, const double* kernels,
                             const double* coef, const double* shift, void* workspace, size_t workspace_bytes,
                             void* stream, double* sig_mean) {
  size_t need = 0;
  mgk_kmc_patch_workspace_bytes(&need);
  if ((ndim != 2 && ndim != 3) || !shape || !sig || (sig_kind != 0 && sig_kind != 1) || nbasis < 0 ||
      nbasis > (ndim == 3 ? 5 : 2) || (nbasis > 0 && (!kernels || !coef || !radius)) || !shift || !workspace ||
      workspace_bytes < need) {
    mgk_set_error("mgk_kmc_patch: bad argument", __FILE__, __LINE__);
    return MGK_ERR_ARG;
  }
  PatchArgs p{};
  // the LS layout: 3D (nx, ny, nz); 2D (nx, 1, ny)
  p.nx = (int)shape[0];
  p.ny = ndim == 3 ? (int)shape[1] : 1;
  p.nz = ndim == 3 ? (int)shape[2] : (int)shape[1];
  if (p.nx < 1 || p.ny < 1 || p.nz < 1) {
    mgk_set_error("mgk_kmc_patch: bad shape", __FILE__, __LINE__);
    return MGK_ERR_ARG;
  }
  p.NV = (long long)p.nx * p.ny * p.nz;
  if (nbasis > 0 && (v < 0 || v >= p.NV)) {
    mgk_set_error("mgk_kmc_patch: voxel out of range", __FILE__, __LINE__);
    return MGK_ERR_ARG;
  }
  p.cz = (int)(v % p.nz);
  p.cy = (int)((v / p.nz) % p.ny);
  p.cx = (int)(v / ((long long)p.nz * p.ny));
  if (nbasis > 0) {
    // radius[] follows the simulation's axes; the 2D y axis is the LS grid's last
    p.Rx = radius[0];
    p.Ry = ndim == 3 ? radius[1] : 0;
    p.Rz = ndim == 3 ? radius[2] : radius[1];
    if (p.Rx < 0 || p.Ry < 0 || p.Rz < 0 || p.Rx > 64 || p.Ry > 64 || p.Rz > 64) {
      mgk_set_error("mgk_kmc_patch: patch radius outside 0..64", __FILE__, __LINE__);
      return MGK_ERR_ARG;
    }
    for (int b = 0; b < nbasis; ++b) p.coef[b] = coef[b];
  }
  p.nb = nbasis;
  p.K = kernels;
  const int C = ndim == 3 ? 6 : 3;
  for (int c = 0; c < C; ++c) p.shift[c] = shift[c];
  p.a = a;
  p.cat = nullptr;
  return patch_launch(p, ndim, sig_kind, sig, sig_unit, workspace, stream, sig_mean);
}

static int patch_launch(PatchArgs& p, int ndim, int sig_kind, void* sig, const double* sig_unit, void* workspace,
                        void* stream, double* sig_mean) {
  const int C = ndim == 3 ? 6 : 3;
  hipStream_t s = (hipStream_t)stream;
  double* parts = (double*)workspace;
  double* means = parts + (size_t)kPatchBlocks * 6;
  const int nblk = (int)std::min<long long>((p.NV + MGK_NT - 1) / MGK_NT, kPatchBlocks);
#define MGK_PATCH(TT, CC) k_kmc_patch<TT, CC><<<nblk, MGK_NT, 0, s>>>(p, (TT*)sig, sig_unit, parts)
  if (C == 6) {
    if (sig_kind == 0) MGK_PATCH(double, 6);
    else MGK_PATCH(float, 6);
    k_kmc_patch_mean<6><<<1, MGK_NT, 0, s>>>(parts, nblk, 1.0 / (double)p.NV, means);
  } else {
    if (sig_kind == 0) MGK_PATCH(double, 3);
    else MGK_PATCH(float, 3);
    k_kmc_patch_mean<3><<<1, MGK_NT, 0, s>>>(parts, nblk, 1.0 / (double)p.NV, means);
  }
#undef MGK_PATCH
  if (hipGetLastError() != hipSuccess) {
    mgk_set_error("mgk_kmc_patch: launch failed", __FILE__, __LINE__);
    return MGK_ERR_HIP;
  }
  if (sig_mean) {
    double m[6];
    if (hipMemcpyAsync(m, means, sizeof(double) * C, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
      mgk_set_error("mgk_kmc_patch: mean read-back failed", __FILE__, __LINE__);
      return MGK_ERR_HIP;
    }
    // d*d row-major, as mgk_ls_info.sig_mean
    if (ndim == 3) {
      const int ij[6][2] = {{0, 0}, {1, 1}, {2, 2}, {1, 2}, {0, 2}, {0, 1}};
      for (int c = 0; c < 6; ++c) {
        sig_mean[ij[c][0] * 3 + ij[c][1]] = m[c];
        sig_mean[ij[c][1] * 3 + ij[c][0]] = m[c];
      }
    } else {
      sig_mean[0] = m[0];
      sig_mean[3] = m[1];
      sig_mean[1] = sig_mean[2] = m[2];
    }
  }
  return MGK_OK;
}

// Predictor step of a KMC event inside the native loop: the weights of the
// planar mode m of voxel v are read from the catalog on the device (before its
// redraw), the combined missing mean is formed on the device in the reference's
// order, then the one-pass stamp as mgk_kmc_patch.  sig_mean: host d*d mean.
int mgk_kmc_patch_mode(const mgk_kmc_loop* L, const mgk_kmc_patch_cfg* P, long long v, int m, double a,
                       int with_unit, double* sig_mean, hipStream_t stream) {
  const int ndim = L->ndim;
  PatchArgs p{};
  p.nx = (int)L->shape[0];
  p.ny = ndim == 3 ? (int)L->shape[1] : 1;
  p.nz = ndim == 3 ? (int)L->shape[2] : (int)L->shape[1];
  p.NV = (long long)p.nx * p.ny * p.nz;
  if (v < 0 || v >= p.NV || m < 0 || m >= L->M || P->nbasis != (ndim == 3 ? 5 : 2)) {
    mgk_set_error("mgk_kmc_patch_mode: bad voxel, mode or basis count", __FILE__, __LINE__);
    return MGK_ERR_ARG;
  }
  p.cz = (int)(v % p.nz);
  p.cy = (int)((v / p.nz) % p.ny);
  p.cx = (int)(v / ((long long)p.nz * p.ny));
  p.Rx = P->radius[0];
  p.Ry = ndim == 3 ? P->radius[1] : 0;
  p.Rz = ndim == 3 ? P->radius[2] : P->radius[1];
  p.nb = P->nbasis;
  p.K = P->kernels;
  p.a = a;
  p.cat = L->cat;
  p.off_xx = (long long)m * p.NV + v;
  p.off_xy = ((long long)L->M + m) * p.NV + v;
  for (int b = 0; b < 5; ++b)
    for (int c = 0; c < 6; ++c) p.mm[b][c] = P->missing_mean[b][c];
  return patch_launch(p, ndim, L->sig_kind, L->sig, with_unit ? P->sig_unit : nullptr, P->workspace, stream,
                      sig_mean);
}

// the macro-unit stress over the remaining time of an increment (2D loops,
// mgkmc/kmc_simulator.py:1094-1096): sig += sig_unit * a, nothing else
int mgk_kmc_patch_unit(const mgk_kmc_loop* L, const mgk_kmc_patch_cfg* P, double a, double* sig_mean,
                       hipStream_t stream) {
  const int ndim = L->ndim;
  PatchArgs p{};
  p.nx = (int)L->shape[0];
  p.ny = ndim == 3 ? (int)L->shape[1] : 1;
  p.nz = ndim == 3 ? (int)L->shape[2] : (int)L->shape[1];
  p.NV = (long long)p.nx * p.ny * p.nz;
  p.nb = 0;
  p.a = a;
  p.cat = nullptr;
  return patch_launch(p, ndim, L->sig_kind, L->sig, P->sig_unit, P->workspace, stream, sig_mean);
}
