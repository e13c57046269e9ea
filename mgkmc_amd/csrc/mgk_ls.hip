// Small-strain Lippmann-Schwinger fixed point on gfx950.
//
// One LS iteration (reference: mgkmc/linear_elastic_simulator.py:307-364 for
// 3D, :43-81 for 2D) is five sweeps over HBM:
//   1. k_ls_tau_xfwd   : tau = C(x):(eps-eps_p) - C0:eps per voxel, then a
//                        real-to-complex FFT along x (two components packed
//                        as one complex sequence), half spectrum out.
//   2. k_fft_mid (fwd) : complex FFT along y (3D only), strided tiles.
//   3. k_ls_pencil     : complex FFT along the contiguous axis for all
//                        components of a pencil, eps_hat = -Gamma0(k):tau_hat
//                        evaluated from k in registers (no materialised
//                        Gamma), inverse FFT along the same axis.
//   4. k_fft_mid (inv) : inverse along y (3D only).
//   5. k_ls_xinv_update: complex-to-real along x, eps_new = eps_bar + eps~,
//                        FP64 partial sums of |eps_new-eps|^2 and |eps|^2.
//   6. k_ls_finalize   : one workgroup, FP64 diff and the stop decision,
//                        kept in device memory so that later launches of the
//                        same solve exit at once (no host round trip per
//                        iteration).
#include "mgk_common.h"
#include "mgk_host.h"
#include "mgk_fftpass.h"
#include "mgk_landau.h"
#include "mgk_secant.h"
#include "mgk_krylov.h"

// Compile-time x sweeps (tau + r2c, c2r + update) run wide tiles: twice the
// points of the other compile-time tiles on 512-thread workgroups, so each
// contiguous row segment of a field is 64 B (FP64) / 128 B (FP32) instead of
// 32 / 64 B.  FP64 4096 points: tau_xfwd 0.635 -> 0.593 ms, xinv_update
// 0.597 -> 0.581 ms at 256^3.  FP32 8192 points (two workgroups per CU at
// 128 VGPRs with the radix-16 passes): tau_xfwd -1%, xinv_update -3%.  FP32
// grids with fewer than 8192 / nx rows along the last axis take narrow tiles
// (4096 points, 256 threads).
template <typename T, bool WIDE>
constexpr int x_tile() {
  return ct_tile<T>() * (WIDE ? 2 : 1);
}
template <bool WIDE>
constexpr int x_nt() {
  return WIDE ? 512 : MGK_NT;
}
#include <algorithm>
#include <cstdio>

struct LsState {
  int done;    // set once the solve has stopped
  int iters;   // iterations executed
  int status;  // MGK_OK (converged), MGK_NOT_CONVERGED (max_iter), MGK_ERR_NONFINITE
  int pad;
  double diff;
  double d2, e2;
  double mean[12];  // eps means [0..C), sig means [C..2C) (after the stress pass)
  // 2D Landau: plane-stress Newton count (device max), e33 buffer state,
  // Anderson history length / ring head / mixing fallback flag
  int nmax, e33_valid, aa_len, aa_head, aa_plain, pad2;
  double aa_gamma[8];
};

template <typename T>
struct LsGeom {
  int nx, P, Q, nxh, C;
  long long NV;
  int logBx;
  FftPlan plx;
  const cplx<T>* twx;
  int logBy;
  FftPlan ply;
  const cplx<T>* twy;
  int logPP;
  FftPlan plq;
  const cplx<T>* twq;
  int npencil;
};

struct Sym6 {
  double v[6];
};

// One simulation of a batched LS launch: its fields, spectrum, partial sums,
// solve state and constants.  The sweep kernels take a table of up to
// MGK_MAXB members of one grid shape by value and pick their member from the
// grid (x sweeps: blockIdx.z, pencils: blockIdx.y, mid sweeps: blockIdx.z /
// components); each member stops on its own state's done flag, so a member's
// arithmetic and iteration count are those of a solve alone.
template <typename T>
struct LsMem {
  const T* lam;
  const T* mu;
  const T* epsp;
  T* eps;
  cplx<T>* spec;
  double* part;
  LsState* st;
  Sym6 eb;
  double lam0, mu0, e2_extra;
  T* sig;  // final stress (batched solves), or NULL
  // The first iteration's x sweeps read the starting field from elsewhere,
  // so it is never written out first:
  // eps_old32: FP64 solves started from an FP32 iterate read it here (exact
  // widening); NULL otherwise.
  // eps_uniform: cold starts of the linear law, eps = eps_bar everywhere,
  // read as that constant; 0 otherwise.
  const float* eps_old32;
  int eps_uniform;
};

// eps reads of the x sweeps: component c at voxel v, from the FP32 starting
// field or the uniform starting value when the member has one
template <typename T>
struct EpsRead {
  const T* eps;
  const float* e32;
  Sym6 eb;  // the member's eps_bar components
  size_t NV;
  bool uni;
  __device__ __forceinline__ void vload(int c, size_t v, T (&o)[16 / sizeof(T)]) const {
    if (uni) {
#pragma unroll
      for (int i = 0; i < 16 / (int)sizeof(T); ++i) o[i] = T(eb.v[c]);
      return;
    }
    const size_t off = (size_t)c * NV + v;
    if constexpr (sizeof(T) == 8) {
      if (e32) {
        const float2 w = *reinterpret_cast<const float2*>(e32 + off);
        o[0] = (T)w.x;
        o[1] = (T)w.y;
        return;
      }
    }
    vload16(eps + off, o);
  }
  __device__ __forceinline__ T ld(int c, size_t v) const {
    if (uni) return T(eb.v[c]);
    const size_t off = (size_t)c * NV + v;
    if constexpr (sizeof(T) == 8) {
      if (e32) return (T)e32[off];
    }
    return eps[off];
  }
};

template <typename T>
struct LsBatch {
  LsMem<T> m[MGK_MAXB];
};

// what k_ls_finalize needs of each member
struct LsFin {
  const double* part[MGK_MAXB];
  LsState* st[MGK_MAXB];
  double e2_extra[MGK_MAXB];
};

// ---------------------------------------------------------------------------
// frequency helpers
// ---------------------------------------------------------------------------
// signed frequency index of FFT bin i on an axis of length n (numpy fftfreq order)
__device__ __forceinline__ int fbin(int i, int n) { return (i <= (n - 1) / 2) ? i : i - n; }
__device__ __forceinline__ bool fnyq(int i, int n) { return (n % 2 == 0) && (i == n / 2); }

// eps_hat_kh = -[ (1/(2 mu0 |q|^2)) (q_h (tau q)_k + q_k (tau q)_h) - B q_k q_h (q.tau.q)/|q|^4 ]
// (mgkmc/elasticity.py:70-105 contracted with a symmetric tau).  3D components
// xx,yy,zz,yz,xz,xy.  Works on real and imaginary parts independently.
template <typename T>
__device__ __forceinline__ void gamma3_apply(const T q0, const T q1, const T q2, const cplx<T>* t, cplx<T>* o,
                                             T A2, T Bc, T scale) {
  const T k2 = q0 * q0 + q1 * q1 + q2 * q2;
  const T ik2 = T(1) / k2;
  const T a = A2 * ik2;
  const T bb = Bc * ik2 * ik2;
#pragma unroll
  for (int ri = 0; ri < 2; ++ri) {
    const T xx = ri ? t[0].y : t[0].x, yy = ri ? t[1].y : t[1].x, zz = ri ? t[2].y : t[2].x;
    const T yz = ri ? t[3].y : t[3].x, xz = ri ? t[4].y : t[4].x, xy = ri ? t[5].y : t[5].x;
    const T u0 = xx * q0 + xy * q1 + xz * q2;
    const T u1 = xy * q0 + yy * q1 + yz * q2;
    const T u2 = xz * q0 + yz * q1 + zz * q2;
    const T qtq = (q0 * u0 + q1 * u1 + q2 * u2) * bb;
    T r[6];
    r[0] = -(a * (2 * q0 * u0) - q0 * q0 * qtq);
    r[1] = -(a * (2 * q1 * u1) - q1 * q1 * qtq);
    r[2] = -(a * (2 * q2 * u2) - q2 * q2 * qtq);
    r[3] = -(a * (q2 * u1 + q1 * u2) - q1 * q2 * qtq);
    r[4] = -(a * (q2 * u0 + q0 * u2) - q0 * q2 * qtq);
    r[5] = -(a * (q1 * u0 + q0 * u1) - q0 * q1 * qtq);
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      if (ri) o[c].y = r[c] * scale;
      else o[c].x = r[c] * scale;
    }
  }
}

// 2D: components xx, yy, xy (mgkmc/elasticity.py:20-53)
template <typename T>
__device__ __forceinline__ void gamma2_apply(const T q0, const T q1, const cplx<T>* t, cplx<T>* o, T A2, T Bc,
                                             T scale) {
  const T k2 = q0 * q0 + q1 * q1;
  const T ik2 = T(1) / k2;
  const T a = A2 * ik2;
  const T bb = Bc * ik2 * ik2;
#pragma unroll
  for (int ri = 0; ri < 2; ++ri) {
    const T xx = ri ? t[0].y : t[0].x, yy = ri ? t[1].y : t[1].x, xy = ri ? t[2].y : t[2].x;
    const T u0 = xx * q0 + xy * q1;
    const T u1 = xy * q0 + yy * q1;
    const T qtq = (q0 * u0 + q1 * u1) * bb;
    T r[3];
    r[0] = -(a * (2 * q0 * u0) - q0 * q0 * qtq);
    r[1] = -(a * (2 * q1 * u1) - q1 * q1 * qtq);
    r[2] = -(a * (q1 * u0 + q0 * u1) - q0 * q1 * qtq);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      if (ri) o[c].y = r[c] * scale;
      else o[c].x = r[c] * scale;
    }
  }
}

// ---------------------------------------------------------------------------
// sweep 1: polarisation + forward x FFT
// ---------------------------------------------------------------------------
template <typename T, int D, int MAXV, int NN = 0, int BB = 0, int NTH = MGK_NT>
__global__ __launch_bounds__(NTH) __attribute__((amdgpu_waves_per_eu(NTH == 512 ? (sizeof(T) == 4 ? 4 : 2) : (sizeof(T) == 4 ? 3 : 2)))) void k_ls_tau_xfwd(LsGeom<T> g, LsBatch<T> bt) {
  const LsMem<T>& M = bt.m[blockIdx.z];
  if (M.st->done) return;
  const T* __restrict__ epsp = M.epsp;
  const T* __restrict__ lam = M.lam;
  const T* __restrict__ mu = M.mu;
  const T lam0 = T(M.lam0), mu0 = T(M.mu0);
  cplx<T>* __restrict__ spec = M.spec;
  extern __shared__ __align__(16) unsigned char smem[];
  cplx<T>* tw = reinterpret_cast<cplx<T>*>(smem);
  cplx<T>* tile = tw + g.nx;
  constexpr bool CT = NN > 0;
  const int nx = CT ? NN : g.nx, P = g.P, Q = g.Q, logB = CT ? ct_log2(BB) : g.logBx, B = CT ? BB : 1 << logB;
  const int ldb = B;
  const int tid_lin = xcd_tile(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
  const int q0 = (tid_lin % gridDim.x) * B, p = tid_lin / gridDim.x;
  const size_t NV = (size_t)g.NV;
  const size_t cs = (size_t)g.nxh * P * Q;  // spectrum component stride
  const EpsRead<T> er{M.eps, sizeof(T) == 8 ? M.eps_old32 : nullptr, M.eb, NV, M.eps_uniform != 0};
  for (int t = threadIdx.x; t < nx; t += NTH) tw[t] = g.twx[t];
  const Tile tl{nx, logB, ldb, 1, 0};
  const T m02 = T(2) * mu0;

  T keep[MAXV], mus[MAXV];
  // 16-byte vector path: compile-time tile, every column inside the grid and
  // rows 16-byte aligned; each lane then handles VW consecutive voxels
  constexpr int VW = vw16<T>();
  constexpr bool VOK = CT && (BB % VW == 0) && (MAXV % VW == 0);
  const bool vec = VOK && q0 + B <= Q && Q % VW == 0;
  if (vec) {
    constexpr int GPR = VOK ? BB / VW : 1;  // vector groups per tile row
    constexpr int SV = MAXV / VW;
    // pass 1: diagonal components, traces; (xx, yy) into the tile
#pragma unroll
    for (int s = 0; s < SV; ++s) {
      const int e = threadIdx.x + s * NTH;
      const int bg = e % GPR, x = e / GPR;
      const size_t v = ((size_t)x * P + p) * Q + q0 + bg * VW;
      T exx[VW], eyy[VW], ezz[VW], pxx[VW], pyy[VW], pzz[VW], l[VW], m[VW];
      er.vload(0, v, exx);
      er.vload(1, v, eyy);
      vload16(lam + v, l);
      vload16(mu + v, m);
      if (D == 3) er.vload(2, v, ezz);
      if (epsp) {
        vload16(epsp + v, pxx);
        vload16(epsp + NV + v, pyy);
        if (D == 3) vload16(epsp + 2 * NV + v, pzz);
      } else {
#pragma unroll
        for (int i = 0; i < VW; ++i) pxx[i] = pyy[i] = pzz[i] = T(0);
      }
#pragma unroll
      for (int i = 0; i < VW; ++i) {
        const T dxx = exx[i] - pxx[i], dyy = eyy[i] - pyy[i];
        T tre, tr0, ez = T(0), dzz = T(0);
        if (D == 3) {
          ez = ezz[i];
          dzz = ez - pzz[i];
          tre = (dxx + dyy) + dzz;
          tr0 = (exx[i] + eyy[i]) + ez;
        } else {
          tre = dxx + dyy;
          tr0 = exx[i] + eyy[i];
        }
        const T m2 = T(2) * m[i];
        const T ltr = l[i] * tre, l0tr = lam0 * tr0;
        const T txx = (m2 * dxx + ltr) - (m02 * exx[i] + l0tr);
        const T tyy = (m2 * dyy + ltr) - (m02 * eyy[i] + l0tr);
        keep[s * VW + i] = (D == 3) ? (m2 * dzz + ltr) - (m02 * ez + l0tr) : T(0);
        mus[s * VW + i] = m[i];
        tile[x * ldb + bg * VW + i] = cmk(txx, tyy);
      }
    }
  } else {
    // pass 1: diagonal components, traces; (xx, yy) into the tile
#pragma unroll
    for (int s = 0; s < MAXV; ++s) {
      keep[s] = T(0);
      mus[s] = T(0);
      const int e = threadIdx.x + s * NTH;
      if (e < nx * B) {
        const int b = e & (B - 1), x = e >> logB, q = q0 + b;
        T txx = T(0), tyy = T(0);
        if (q < Q) {
          const size_t v = ((size_t)x * P + p) * Q + q;
          const T exx = er.ld(0, v), eyy = er.ld(1, v);
          const T pxx = epsp ? epsp[v] : T(0), pyy = epsp ? epsp[NV + v] : T(0);
          const T l = lam[v], m = mu[v];
          const T dxx = exx - pxx, dyy = eyy - pyy;
          T tre, tr0, ezz = T(0), dzz = T(0);
          if (D == 3) {
            ezz = er.ld(2, v);
            const T pzz = epsp ? epsp[2 * NV + v] : T(0);
            dzz = ezz - pzz;
            tre = (dxx + dyy) + dzz;
            tr0 = (exx + eyy) + ezz;
          } else {
            tre = dxx + dyy;
            tr0 = exx + eyy;
          }
          const T m2 = T(2) * m;
          const T ltr = l * tre, l0tr = lam0 * tr0;
          txx = (m2 * dxx + ltr) - (m02 * exx + l0tr);
          tyy = (m2 * dyy + ltr) - (m02 * eyy + l0tr);
          if (D == 3) keep[s] = (m2 * dzz + ltr) - (m02 * ezz + l0tr);
          mus[s] = m;
        }
        tile[x * ldb + b] = cmk(txx, tyy);
      }
    }
  }
  __syncthreads();
  if constexpr (CT) tile_fft_ct<NN, BB, 1, false, NTH>(tile, tw, false);
  else tile_fft<MAXV>(tile, tl, g.plx, tw, false);
  store_half<T, NN, BB, NTH>(tile, vec, ldb, nx, g.nxh, logB, q0, Q, P, p, spec, spec + cs);
  __syncthreads();

  // remaining pairs: 3D (zz, yz), (xz, xy); 2D (xy, -)
  constexpr int NPAIR = (D == 3) ? 2 : 1;
#pragma unroll
  for (int pr = 0; pr < NPAIR; ++pr) {
    const int ca = (D == 3) ? (pr == 0 ? 2 : 4) : 2;
    const int cb = (D == 3) ? (pr == 0 ? 3 : 5) : -1;
    if (vec) {
      constexpr int GPR = VOK ? BB / VW : 1;
      constexpr int SV = MAXV / VW;
#pragma unroll
      for (int s = 0; s < SV; ++s) {
        const int e = threadIdx.x + s * NTH;
        const int bg = e % GPR, x = e / GPR;
        const size_t v = ((size_t)x * P + p) * Q + q0 + bg * VW;
        T ea[VW], pa[VW], eb[VW], pb[VW];
        if (!(D == 3 && pr == 0)) {
          er.vload(ca, v, ea);
          if (epsp) vload16(epsp + (size_t)ca * NV + v, pa);
        }
        if (cb >= 0) {
          er.vload(cb, v, eb);
          if (epsp) vload16(epsp + (size_t)cb * NV + v, pb);
        }
#pragma unroll
        for (int i = 0; i < VW; ++i) {
          const T m2 = T(2) * mus[s * VW + i];
          T ta, tb = T(0);
          if (D == 3 && pr == 0) ta = keep[s * VW + i];
          else ta = m2 * (ea[i] - (epsp ? pa[i] : T(0))) - m02 * ea[i];
          if (cb >= 0) tb = m2 * (eb[i] - (epsp ? pb[i] : T(0))) - m02 * eb[i];
          tile[x * ldb + bg * VW + i] = cmk(ta, tb);
        }
      }
    } else {
#pragma unroll
      for (int s = 0; s < MAXV; ++s) {
        const int e = threadIdx.x + s * NTH;
        if (e < nx * B) {
          const int b = e & (B - 1), x = e >> logB, q = q0 + b;
          T ta = T(0), tb = T(0);
          if (q < Q) {
            const size_t v = ((size_t)x * P + p) * Q + q;
            const T m2 = T(2) * mus[s];
            if (D == 3 && pr == 0) {
              ta = keep[s];
            } else {
              const T ea = er.ld(ca, v);
              const T pa = epsp ? epsp[(size_t)ca * NV + v] : T(0);
              ta = m2 * (ea - pa) - m02 * ea;
            }
            if (cb >= 0) {
              const T eb = er.ld(cb, v);
              const T pb = epsp ? epsp[(size_t)cb * NV + v] : T(0);
              tb = m2 * (eb - pb) - m02 * eb;
            }
          }
          tile[x * ldb + b] = cmk(ta, tb);
        }
      }
    }
    __syncthreads();
    if constexpr (CT) tile_fft_ct<NN, BB, 1, false, NTH>(tile, tw, false);
    else tile_fft<MAXV>(tile, tl, g.plx, tw, false);
    store_half<T, NN, BB, NTH>(tile, vec, ldb, nx, g.nxh, logB, q0, Q, P, p, spec + (size_t)ca * cs,
                          cb >= 0 ? spec + (size_t)cb * cs : (cplx<T>*)nullptr);
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// sweep 3: pencil FFT along the contiguous axis + Green operator + inverse
// grid ceil(npencil/PP); pencil id = kx*P + p
// ---------------------------------------------------------------------------
// FP64 compile-time pencils up to 256 long are held to 128 VGPRs (4 waves/SIMD
// instead of 3 at 129): FP64 LS solve +3.3% at 256^3, measured on one device
template <typename T, int D, int MAXV, int NN = 0, int BB = 0, int NTH = MGK_NT>
__global__ __launch_bounds__(NTH) __attribute__((amdgpu_waves_per_eu(NN > 0 && NN <= 256 && sizeof(T) == 8 ? 4 : 1))) void k_ls_pencil(LsGeom<T> g, LsBatch<T> bt, T scale) {
  const LsMem<T>& M = bt.m[blockIdx.y];
  if (M.st->done) return;
  cplx<T>* __restrict__ spec = M.spec;
  const T A2 = T(1.0 / (2.0 * M.mu0));
  const T Bc = T((M.lam0 + M.mu0) / (M.mu0 * (M.lam0 + 2.0 * M.mu0)));
  constexpr int C = (D == 3) ? 6 : 3;
  extern __shared__ __align__(16) unsigned char smem[];
  cplx<T>* tw = reinterpret_cast<cplx<T>*>(smem);
  cplx<T>* tile = tw + g.Q;
  constexpr bool CT = NN > 0;
  const int Q = CT ? NN : g.Q, P = g.P, logPP = CT ? ct_log2(BB) : g.logPP, PP = CT ? BB : 1 << logPP;
  const int ldb = PP, cstride = Q * PP;
  const int pid0 = blockIdx.x * PP;
  const int npen = g.npencil;
  for (int t = threadIdx.x; t < Q; t += NTH) tw[t] = g.twq[t];
  const int chunk = PP * Q;
  // tile index (padded for the compile-time passes)
  auto ti = [&](int c, int q, int pp) { return tpad<CT>(c * cstride + q * ldb + pp); };
  // FP32 with compile-time pencils all inside the grid: two complex values
  // (16 bytes) per lane and global access
  constexpr bool VOK = CT && sizeof(T) == 4;
  const bool vec = VOK && pid0 + PP <= npen;
  if (vec) {
    constexpr int QH = VOK ? NN / 2 : 1;
    for (int e = threadIdx.x; e < C * PP * QH; e += NTH) {
      const int h = e % QH, r = e / QH, pp = r % PP, c = r / PP;
      const float4 v = *reinterpret_cast<const float4*>(spec + ((size_t)c * npen + pid0 + pp) * Q + 2 * h);
      tile[ti(c, 2 * h, pp)] = cmk(T(v.x), T(v.y));
      tile[ti(c, 2 * h + 1, pp)] = cmk(T(v.z), T(v.w));
    }
  } else {
    for (int e = threadIdx.x; e < C * chunk; e += NTH) {
      const int c = e / chunk;
      const int r = e - c * chunk;
      const int pp = r / Q;
      const int q = r - pp * Q;
      const int pid = pid0 + pp;
      tile[ti(c, q, pp)] =
          (pid < npen) ? spec[((size_t)c * npen + pid) * Q + q] : cmk(T(0), T(0));
    }
  }
  __syncthreads();
  const Tile tl{Q, logPP, ldb, C, cstride};
  if constexpr (CT) tile_fft_ct<NN, BB, C, true, NTH>(tile, tw, false);
  else tile_fft<MAXV>(tile, tl, g.plq, tw, false);

  for (int e = threadIdx.x; e < Q * PP; e += NTH) {
    const int pp = e & (PP - 1), q = e >> logPP;
    const int pid = pid0 + pp;
    if (pid >= npen) continue;
    const int kx = pid / P, p = pid - kx * P;
    cplx<T> t[C], o[C];
#pragma unroll
    for (int c = 0; c < C; ++c) t[c] = tile[ti(c, q, pp)];
    const bool nyx = fnyq(kx, g.nx);
    if (D == 3) {
      const bool nyy = fnyq(p, P), nyz = fnyq(q, Q);
      const T qx = T(fbin(kx, g.nx)) / T(g.nx), qy = T(fbin(p, P)) / T(P), qz = T(fbin(q, Q)) / T(Q);
      if (kx == 0 && p == 0 && q == 0) {
#pragma unroll
        for (int c = 0; c < C; ++c) o[c] = cmk(T(0), T(0));
      } else if (!(nyx || nyy || nyz)) {
        gamma3_apply(qx, qy, qz, t, o, A2, Bc, scale);
      } else {
        // Hermitian part of the reference's full-spectrum operator: average
        // Gamma over q and q with its Nyquist components negated.
        cplx<T> o2[C];
        gamma3_apply(qx, qy, qz, t, o, A2, Bc, scale);
        gamma3_apply(nyx ? -qx : qx, nyy ? -qy : qy, nyz ? -qz : qz, t, o2, A2, Bc, scale);
#pragma unroll
        for (int c = 0; c < C; ++c) o[c] = cmk(T(0.5) * (o[c].x + o2[c].x), T(0.5) * (o[c].y + o2[c].y));
      }
    } else {
      const bool nyy = fnyq(q, Q);
      const T qx = T(fbin(kx, g.nx)) / T(g.nx), qy = T(fbin(q, Q)) / T(Q);
      if (kx == 0 && q == 0) {
#pragma unroll
        for (int c = 0; c < C; ++c) o[c] = cmk(T(0), T(0));
      } else if (!(nyx || nyy)) {
        gamma2_apply(qx, qy, t, o, A2, Bc, scale);
      } else {
        cplx<T> o2[C];
        gamma2_apply(qx, qy, t, o, A2, Bc, scale);
        gamma2_apply(nyx ? -qx : qx, nyy ? -qy : qy, t, o2, A2, Bc, scale);
#pragma unroll
        for (int c = 0; c < C; ++c) o[c] = cmk(T(0.5) * (o[c].x + o2[c].x), T(0.5) * (o[c].y + o2[c].y));
      }
    }
#pragma unroll
    for (int c = 0; c < C; ++c) tile[ti(c, q, pp)] = o[c];
  }
  __syncthreads();
  if constexpr (CT) tile_fft_ct<NN, BB, C, true, NTH>(tile, tw, true);
  else tile_fft<MAXV>(tile, tl, g.plq, tw, true);
  if (vec) {
    constexpr int QH = VOK ? NN / 2 : 1;
    for (int e = threadIdx.x; e < C * PP * QH; e += NTH) {
      const int h = e % QH, r = e / QH, pp = r % PP, c = r / PP;
      const cplx<T> a = tile[ti(c, 2 * h, pp)];
      const cplx<T> b = tile[ti(c, 2 * h + 1, pp)];
      *reinterpret_cast<float4*>(spec + ((size_t)c * npen + pid0 + pp) * Q + 2 * h) =
          make_float4(float(a.x), float(a.y), float(b.x), float(b.y));
    }
  } else {
    for (int e = threadIdx.x; e < C * chunk; e += NTH) {
      const int c = e / chunk;
      const int r = e - c * chunk;
      const int pp = r / Q;
      const int q = r - pp * Q;
      const int pid = pid0 + pp;
      if (pid < npen) spec[((size_t)c * npen + pid) * Q + q] = tile[ti(c, q, pp)];
    }
  }
}

// ---------------------------------------------------------------------------
// sweep 5: inverse x FFT (complex-to-real) + update + FP64 partial norms
// ---------------------------------------------------------------------------
template <typename T, int D, int MAXV, int NN = 0, int BB = 0, int NTH = MGK_NT>
__global__ __launch_bounds__(NTH) __attribute__((amdgpu_waves_per_eu(NTH == 512 ? (sizeof(T) == 4 ? 4 : 2) : (sizeof(T) == 4 ? 4 : 2)))) void k_ls_xinv_update(LsGeom<T> g, LsBatch<T> bt) {
  const LsMem<T>& M = bt.m[blockIdx.z];
  if (M.st->done) return;
  const cplx<T>* __restrict__ spec = M.spec;
  T* __restrict__ eps = M.eps;
  const Sym6 ebar = M.eb;
  double* __restrict__ partials = M.part;
  extern __shared__ __align__(16) unsigned char smem[];
  cplx<T>* tw = reinterpret_cast<cplx<T>*>(smem);
  cplx<T>* tile = tw + g.nx;
  constexpr bool CT = NN > 0;
  const int nx = CT ? NN : g.nx, P = g.P, Q = g.Q, logB = CT ? ct_log2(BB) : g.logBx, B = CT ? BB : 1 << logB;
  const int ldb = B;
  const int tid_lin = xcd_tile(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
  const int q0 = (tid_lin % gridDim.x) * B, p = tid_lin / gridDim.x;
  const size_t NV = (size_t)g.NV;
  const size_t cs = (size_t)g.nxh * P * Q;
  const EpsRead<T> er{eps, sizeof(T) == 8 ? M.eps_old32 : nullptr, ebar, NV, M.eps_uniform != 0};
  for (int t = threadIdx.x; t < nx; t += NTH) tw[t] = g.twx[t];
  const Tile tl{nx, logB, ldb, 1, 0};
  double acc[2] = {0.0, 0.0};
  constexpr int VW = vw16<T>();
  constexpr bool VOK = CT && (BB % VW == 0) && (MAXV % VW == 0);
  const bool vec = VOK && q0 + B <= Q && Q % VW == 0;
  constexpr int NPAIR = (D == 3) ? 3 : 2;
#pragma unroll
  for (int pr = 0; pr < NPAIR; ++pr) {
    const int ca = 2 * pr;
    const int cb = (D == 3 || pr == 0) ? 2 * pr + 1 : -1;
    const cplx<T>* sa = spec + (size_t)ca * cs;
    const cplx<T>* sb = cb >= 0 ? spec + (size_t)cb * cs : (const cplx<T>*)nullptr;
    bool merged = false;
    if constexpr (sizeof(T) == 4 && CT && BB % 2 == 0) {
      if (vec) {
        xmerge_load_v<NN, BB, NTH>(tile, q0, Q, P, p, sa, sb);
        merged = true;
      }
    }
    if (!merged) xmerge_load<T, NTH>(tile, ldb, nx, g.nxh, logB, q0, Q, P, p, sa, sb);
    if constexpr (CT) tile_fft_ct<NN, BB, 1, false, NTH>(tile, tw, true);
    else tile_fft<MAXV>(tile, tl, g.plx, tw, true);
    const double wa = (D == 3) ? (ca >= 3 ? 2.0 : 1.0) : (ca >= 2 ? 2.0 : 1.0);
    const double wb = (D == 3) ? (cb >= 3 ? 2.0 : 1.0) : (cb >= 2 ? 2.0 : 1.0);
    const T eba = T(ebar.v[ca]);
    const T ebb = cb >= 0 ? T(ebar.v[cb]) : T(0);
    if (vec) {
      constexpr int GPR = VOK ? BB / VW : 1;
      constexpr int SV = MAXV / VW;
#pragma unroll
      for (int s = 0; s < SV; ++s) {
        const int e = threadIdx.x + s * NTH;
        const int bg = e % GPR, x = e / GPR;
        const size_t v = ((size_t)x * P + p) * Q + q0 + bg * VW;
        T oa[VW], na[VW], ob[VW], nb[VW];
        er.vload(ca, v, oa);
        if (cb >= 0) er.vload(cb, v, ob);
#pragma unroll
        for (int i = 0; i < VW; ++i) {
          const cplx<T> z = tile[x * ldb + bg * VW + i];
          na[i] = eba + z.x;
          const double dd = (double)na[i] - (double)oa[i];
          acc[0] += wa * dd * dd;
          acc[1] += wa * (double)oa[i] * (double)oa[i];
          if (cb >= 0) {
            nb[i] = ebb + z.y;
            const double db = (double)nb[i] - (double)ob[i];
            acc[0] += wb * db * db;
            acc[1] += wb * (double)ob[i] * (double)ob[i];
          }
        }
        vstore16(eps + (size_t)ca * NV + v, na);
        if (cb >= 0) vstore16(eps + (size_t)cb * NV + v, nb);
      }
      __syncthreads();
      continue;
    }
#pragma unroll
    for (int s = 0; s < MAXV; ++s) {
      const int e = threadIdx.x + s * NTH;
      if (e < nx * B) {
        const int b = e & (B - 1), x = e >> logB, q = q0 + b;
        if (q < Q) {
          const size_t v = ((size_t)x * P + p) * Q + q;
          const cplx<T> z = tile[x * ldb + b];
          {
            const T nw = eba + z.x;
            const T od = er.ld(ca, v);
            const double dd = (double)nw - (double)od;
            acc[0] += wa * dd * dd;
            acc[1] += wa * (double)od * (double)od;
            eps[(size_t)ca * NV + v] = nw;
          }
          if (cb >= 0) {
            const T nw = ebb + z.y;
            const T od = er.ld(cb, v);
            const double dd = (double)nw - (double)od;
            acc[0] += wb * dd * dd;
            acc[1] += wb * (double)od * (double)od;
            eps[(size_t)cb * NV + v] = nw;
          }
        }
      }
    }
    __syncthreads();
  }
  double* red = reinterpret_cast<double*>(tile);
  block_sum<2, NTH>(acc, red);
  if (threadIdx.x == 0) {
    const size_t bid = (size_t)blockIdx.y * gridDim.x + blockIdx.x;
    partials[2 * bid] = acc[0];
    partials[2 * bid + 1] = acc[1];
  }
}

__global__ __launch_bounds__(MGK_NT) void k_ls_finalize(LsFin f, int nparts, double tol, int max_iter) {
  LsState* __restrict__ st = f.st[blockIdx.x];
  if (st->done) return;
  const double* __restrict__ partials = f.part[blockIdx.x];
  const double e2_extra = f.e2_extra[blockIdx.x];
  __shared__ double red[2 * (MGK_NT / 64)];
  double acc[2] = {0.0, 0.0};
  for (int i = threadIdx.x; i < nparts; i += MGK_NT) {
    acc[0] += partials[2 * i];
    acc[1] += partials[2 * i + 1];
  }
  block_sum<2>(acc, red);
  if (threadIdx.x == 0) {
    const double diff = sqrt(acc[0]) / (sqrt(acc[1] + e2_extra) + 1e-20);
    st->d2 = acc[0];
    st->e2 = acc[1];
    st->diff = diff;
    st->iters += 1;
    if (!isfinite(diff)) {
      st->done = 1;
      st->status = MGK_ERR_NONFINITE;
    } else if (diff < tol) {
      st->done = 1;
      st->status = MGK_OK;
    } else if (st->iters >= max_iter) {
      st->done = 1;
      st->status = MGK_NOT_CONVERGED;
    }
  }
}

// the members' solve states side by side (one read for the whole batch)
__global__ void k_ls_gather_states(LsFin f, LsState* __restrict__ out) {
  const unsigned int* src = (const unsigned int*)f.st[blockIdx.x];
  unsigned int* dst = (unsigned int*)(out + blockIdx.x);
  for (int i = threadIdx.x; i < (int)(sizeof(LsState) / 4); i += blockDim.x) dst[i] = src[i];
}

// ---------------------------------------------------------------------------
// init / warm start / stress / means
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void ls_fill_body(T* __restrict__ eps, int C, size_t NV, const Sym6& ebar) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < NV; i += (size_t)gridDim.x * blockDim.x) {
    for (int c = 0; c < C; ++c) eps[(size_t)c * NV + i] = T(ebar.v[c]);
  }
}

template <typename T>
__global__ void k_ls_fill(T* __restrict__ eps, int C, size_t NV, Sym6 ebar) {
  ls_fill_body<T>(eps, C, NV, ebar);
}

// per-block FP64 sums of C components (fixed grid => deterministic)
template <typename T>
__device__ __forceinline__ void comp_sums_body(const T* __restrict__ f, int C, size_t NV,
                                               double* __restrict__ partials) {
  __shared__ double red[6 * (MGK_NT / 64)];
  double acc[6] = {0, 0, 0, 0, 0, 0};
  for (size_t i = blockIdx.x * (size_t)MGK_NT + threadIdx.x; i < NV; i += (size_t)gridDim.x * MGK_NT) {
#pragma unroll
    for (int c = 0; c < 6; ++c)
      if (c < C) acc[c] += (double)f[(size_t)c * NV + i];
  }
  block_sum<6>(acc, red);
  if (threadIdx.x == 0)
    for (int c = 0; c < C; ++c) partials[(size_t)blockIdx.x * C + c] = acc[c];
}

template <typename T>
__global__ __launch_bounds__(MGK_NT) void k_comp_sums(const T* __restrict__ f, int C, size_t NV,
                                                      double* __restrict__ partials) {
  comp_sums_body<T>(f, C, NV, partials);
}

// means[off + c] = sum(partials[:, c]) / NV
__device__ __forceinline__ void means_finalize_body(const double* __restrict__ partials, int nparts, int C,
                                                    double NV, double* __restrict__ out) {
  __shared__ double red[12 * (MGK_NT / 64)];
  double acc[12];
#pragma unroll
  for (int c = 0; c < 12; ++c) acc[c] = 0.0;
  for (int i = threadIdx.x; i < nparts; i += MGK_NT) {
#pragma unroll
    for (int c = 0; c < 12; ++c)
      if (c < C) acc[c] += partials[(size_t)i * C + c];
  }
  block_sum<12>(acc, red);
  if (threadIdx.x == 0)
    for (int c = 0; c < C; ++c) out[c] = acc[c] / NV;
}

__global__ __launch_bounds__(MGK_NT) void k_means_finalize(const double* __restrict__ partials, int nparts, int C,
                                                           double NV, double* __restrict__ out) {
  means_finalize_body(partials, nparts, C, NV, out);
}

template <typename T>
__device__ __forceinline__ void ls_warm_adjust_body(T* __restrict__ eps, int C, size_t NV, const Sym6& ebar,
                                                    const double* __restrict__ means) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < NV; i += (size_t)gridDim.x * blockDim.x) {
    for (int c = 0; c < C; ++c) {
      const T add = T(ebar.v[c] - means[c]);
      eps[(size_t)c * NV + i] += add;
    }
  }
}

template <typename T>
__global__ void k_ls_warm_adjust(T* __restrict__ eps, int C, size_t NV, Sym6 ebar,
                                 const double* __restrict__ means) {
  ls_warm_adjust_body<T>(eps, C, NV, ebar, means);
}

// sig = 2 mu (eps-eps_p) + lam tr(eps-eps_p) I  (elasticity.py:12-18, 62-67),
// plus per-block FP64 sums of eps and sig components.
template <typename T, int D>
__device__ __forceinline__ void ls_stress_body(const T* __restrict__ eps, const T* __restrict__ epsp,
                                               const T* __restrict__ lam, const T* __restrict__ mu,
                                               T* __restrict__ sig, size_t NV, double* __restrict__ partials) {
  constexpr int C = (D == 3) ? 6 : 3;
  __shared__ double red[2 * C * (MGK_NT / 64)];
  double acc[2 * C];
#pragma unroll
  for (int c = 0; c < 2 * C; ++c) acc[c] = 0.0;
  for (size_t i = blockIdx.x * (size_t)MGK_NT + threadIdx.x; i < NV; i += (size_t)gridDim.x * MGK_NT) {
    T e[C], d[C];
#pragma unroll
    for (int c = 0; c < C; ++c) {
      e[c] = eps[(size_t)c * NV + i];
      d[c] = e[c] - (epsp ? epsp[(size_t)c * NV + i] : T(0));
    }
    const T tr = (D == 3) ? ((d[0] + d[1]) + d[2]) : (d[0] + d[1]);
    const T m2 = T(2) * mu[i], ltr = lam[i] * tr;
#pragma unroll
    for (int c = 0; c < C; ++c) {
      const bool diag = c < D;
      const T s = diag ? (m2 * d[c] + ltr) : (m2 * d[c]);
      if (sig) sig[(size_t)c * NV + i] = s;
      acc[c] += (double)e[c];
      acc[C + c] += (double)s;
    }
  }
  block_sum<2 * C>(acc, red);
  if (threadIdx.x == 0)
    for (int c = 0; c < 2 * C; ++c) partials[(size_t)blockIdx.x * 2 * C + c] = acc[c];
}

template <typename T, int D>
__global__ __launch_bounds__(MGK_NT) void k_ls_stress(const T* __restrict__ eps, const T* __restrict__ epsp,
                                                      const T* __restrict__ lam, const T* __restrict__ mu,
                                                      T* __restrict__ sig, size_t NV, double* __restrict__ partials) {
  ls_stress_body<T, D>(eps, epsp, lam, mu, sig, NV, partials);
}

// ---------------------------------------------------------------------------
// Nonlinear small-strain laws (3D Landau, secant-modulus laws): polarisation
// tau = sigma(eps - eps_p) - C0:eps into a C-component real field, r2c of that
// field along x, and the final stress with FP64 sums of eps and sigma.  Law
// objects map C strain components to C stress components per voxel.
// ---------------------------------------------------------------------------
struct LandauLaw {
  LandauPar P;
  static constexpr int C = 6;
  template <typename T>
  __device__ __forceinline__ void operator()(const T (&e)[6], T lam, T mu, T (&s)[6]) const {
    landau_sigma_sym(e, lam, mu, P, s);
  }
};

template <typename T, class Law>
__global__ __launch_bounds__(MGK_NT) void k_ls_pol_tau(const T* __restrict__ eps, const T* __restrict__ epsp,
                                                       const T* __restrict__ lam, const T* __restrict__ mu,
                                                       T lam0, T mu0, Law law, T* __restrict__ tau, size_t NV,
                                                       const LsState* __restrict__ st) {
  constexpr int C = Law::C, ND = C == 6 ? 3 : 2;
  if (st->done) return;
  for (size_t i = blockIdx.x * (size_t)MGK_NT + threadIdx.x; i < NV; i += (size_t)gridDim.x * MGK_NT) {
    T e[C], d[C], sg[C];
#pragma unroll
    for (int c = 0; c < C; ++c) {
      e[c] = eps[(size_t)c * NV + i];
      d[c] = e[c] - (epsp ? epsp[(size_t)c * NV + i] : T(0));
    }
    law(d, lam[i], mu[i], sg);
    const T l0tr = lam0 * (C == 6 ? (e[0] + e[1]) + e[2] : e[0] + e[1]);
    const T m02 = T(2) * mu0;
#pragma unroll
    for (int c = 0; c < C; ++c) tau[(size_t)c * NV + i] = sg[c] - ((c < ND ? l0tr : T(0)) + m02 * e[c]);
  }
}

template <typename T, int MAXV, int NN = 0, int BB = 0, int NTH = MGK_NT>
__global__ __launch_bounds__(NTH) void k_ls_xfwd_real(LsGeom<T> g, const T* __restrict__ f,
                                                         cplx<T>* __restrict__ spec, const LsState* __restrict__ st) {
  if (st->done) return;
  extern __shared__ __align__(16) unsigned char smem[];
  cplx<T>* tw = reinterpret_cast<cplx<T>*>(smem);
  cplx<T>* tile = tw + g.nx;
  constexpr bool CT = NN > 0;
  const int nx = CT ? NN : g.nx, P = g.P, Q = g.Q, logB = CT ? ct_log2(BB) : g.logBx, B = CT ? BB : 1 << logB;
  const int ldb = B;
  const int tid_lin = xcd_tile(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
  const int q0 = (tid_lin % gridDim.x) * B, p = tid_lin / gridDim.x;
  const size_t NV = (size_t)g.NV;
  const size_t cs = (size_t)g.nxh * P * Q;
  for (int t = threadIdx.x; t < nx; t += NTH) tw[t] = g.twx[t];
  const Tile tl{nx, logB, ldb, 1, 0};
  constexpr int VW = vw16<T>();
  constexpr bool VOK = CT && (BB % VW == 0);
  const bool vec = VOK && q0 + B <= Q && Q % VW == 0;
  const int npair = (g.C + 1) / 2;
#pragma unroll 1
  for (int pr = 0; pr < npair; ++pr) {
    const bool hasb = 2 * pr + 1 < g.C;
    const T* fa = f + (size_t)(2 * pr) * NV;
    const T* fb = f + (size_t)(hasb ? 2 * pr + 1 : 2 * pr) * NV;
    if (vec) {
      constexpr int GPR = VOK ? BB / VW : 1;
      for (int e = threadIdx.x; e < nx * GPR; e += NTH) {
        const int bg = e % GPR, x = e / GPR;
        const size_t v = ((size_t)x * P + p) * Q + q0 + bg * VW;
        T a[VW], b[VW];
        vload16(fa + v, a);
        vload16(fb + v, b);
#pragma unroll
        for (int i = 0; i < VW; ++i) tile[x * ldb + bg * VW + i] = cmk(a[i], hasb ? b[i] : T(0));
      }
    } else {
      for (int e = threadIdx.x; e < nx * B; e += NTH) {
        const int b = e & (B - 1), x = e >> logB, q = q0 + b;
        cplx<T> z = cmk(T(0), T(0));
        if (q < Q) {
          const size_t v = ((size_t)x * P + p) * Q + q;
          z = cmk(fa[v], hasb ? fb[v] : T(0));
        }
        tile[x * ldb + b] = z;
      }
    }
    __syncthreads();
    if constexpr (CT) tile_fft_ct<NN, BB, 1, false, NTH>(tile, tw, false);
    else tile_fft<MAXV>(tile, tl, g.plx, tw, false);
    store_half<T, NN, BB, NTH>(tile, vec, ldb, nx, g.nxh, logB, q0, Q, P, p, spec + (size_t)(2 * pr) * cs,
                          hasb ? spec + (size_t)(2 * pr + 1) * cs : (cplx<T>*)nullptr);
    __syncthreads();
  }
}

template <typename T, class Law>
__global__ __launch_bounds__(MGK_NT) void k_ls_pol_stress(const T* __restrict__ eps, const T* __restrict__ epsp,
                                                          const T* __restrict__ lam, const T* __restrict__ mu,
                                                          Law law, T* __restrict__ sig, size_t NV,
                                                          double* __restrict__ partials) {
  constexpr int C = Law::C;
  __shared__ double red[2 * C * (MGK_NT / 64)];
  double acc[2 * C];
#pragma unroll
  for (int c = 0; c < 2 * C; ++c) acc[c] = 0.0;
  for (size_t i = blockIdx.x * (size_t)MGK_NT + threadIdx.x; i < NV; i += (size_t)gridDim.x * MGK_NT) {
    T e[C], d[C], sg[C];
#pragma unroll
    for (int c = 0; c < C; ++c) {
      e[c] = eps[(size_t)c * NV + i];
      d[c] = e[c] - (epsp ? epsp[(size_t)c * NV + i] : T(0));
    }
    law(d, lam[i], mu[i], sg);
#pragma unroll
    for (int c = 0; c < C; ++c) {
      if (sig) sig[(size_t)c * NV + i] = sg[c];
      acc[c] += (double)e[c];
      acc[C + c] += (double)sg[c];
    }
  }
  block_sum<2 * C>(acc, red);
  if (threadIdx.x == 0)
    for (int c = 0; c < 2 * C; ++c) partials[(size_t)blockIdx.x * 2 * C + c] = acc[c];
}

// ---------------------------------------------------------------------------
// 2D Landau (FP64): plane-stress e33 Newton as the reference runs it -- every
// voxel takes the same number of steps N, the first step at which all steps
// are below 1e-9 (at most 20) -- found by a first pass (per-voxel first
// convergence step, device max), then the stress pass takes exactly N steps
// and evaluates the stress at the state before the last update.
// e33 buffer: warm start when st->e33_valid (persistent mode), else the
// reference's initial guess -lam/(lam + 2 mu) (a11 + a22).
// ---------------------------------------------------------------------------
__device__ __forceinline__ double l2_e33_init(const double* e33, int valid, size_t i, double a11, double a22,
                                              double lam, double mu) {
  return valid ? e33[i] : -lam / (lam + 2.0 * mu) * (a11 + a22);
}

__global__ __launch_bounds__(MGK_NT) void k_l2_count(const double* __restrict__ eps,
                                                     const double* __restrict__ epsp, const double* __restrict__ lam,
                                                     const double* __restrict__ mu, LandauPar P,
                                                     const double* __restrict__ e33, size_t NV,
                                                     LsState* __restrict__ st, int final_pass) {
  if (st->done && !final_pass) return;
  const int valid = st->e33_valid;
  int kmax = 0;
  bool bad = false;
  for (size_t i = blockIdx.x * (size_t)MGK_NT + threadIdx.x; i < NV; i += (size_t)gridDim.x * MGK_NT) {
    const double a11 = eps[i] - (epsp ? epsp[i] : 0.0), a22 = eps[NV + i] - (epsp ? epsp[NV + i] : 0.0);
    const double a12 = eps[2 * NV + i] - (epsp ? epsp[2 * NV + i] : 0.0);
    const double l = lam[i], m = mu[i];
    double e = l2_e33_init(e33, valid, i, a11, a22, l, m);
    int k = 20;
    for (int it = 1; it <= 20; ++it) {
      const Landau2<double> sv = landau2_eval(a11, a22, a12, e, l, m, P);
      const double d = landau2_delta(sv, e, l, m, P);
      if (!isfinite(d)) {
        bad = true;
        break;
      }
      e = e + d;
      if (fabs(d) < 1e-9) {
        k = it;
        break;
      }
    }
    kmax = max(kmax, k);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) kmax = max(kmax, __shfl_xor(kmax, off, 64));
  if ((threadIdx.x & 63) == 0) atomicMax(&st->nmax, kmax);
  if (bad) {
    atomicExch(&st->status, MGK_ERR_NONFINITE);
    atomicExch(&st->done, 1);
  }
}

// tau = sigma_L(eps - eps_p) - C0:eps (tau != NULL) and/or the final stress
// with FP64 sums of eps and sigma (sig != NULL); updates e33
__global__ __launch_bounds__(MGK_NT) void k_l2_stress(const double* __restrict__ eps,
                                                      const double* __restrict__ epsp,
                                                      const double* __restrict__ lam,
                                                      const double* __restrict__ mu, double lam0, double mu0,
                                                      LandauPar P, int plane_stress, double* __restrict__ e33,
                                                      double* __restrict__ tau, double* __restrict__ sig,
                                                      double* __restrict__ partials, size_t NV,
                                                      const LsState* __restrict__ st, int final_pass) {
  if (st->done && !final_pass) return;
  __shared__ double red[6 * (MGK_NT / 64)];
  double acc[6] = {0, 0, 0, 0, 0, 0};
  const int valid = st->e33_valid, N = st->nmax;
  for (size_t i = blockIdx.x * (size_t)MGK_NT + threadIdx.x; i < NV; i += (size_t)gridDim.x * MGK_NT) {
    const double e11 = eps[i], e22 = eps[NV + i], e12 = eps[2 * NV + i];
    const double a11 = e11 - (epsp ? epsp[i] : 0.0), a22 = e22 - (epsp ? epsp[NV + i] : 0.0);
    const double a12 = e12 - (epsp ? epsp[2 * NV + i] : 0.0);
    const double l = lam[i], m = mu[i];
    Landau2<double> sv;
    if (plane_stress) {
      double e = l2_e33_init(e33, valid, i, a11, a22, l, m);
      sv = landau2_eval(a11, a22, a12, e, l, m, P);
      for (int it = 1; it <= N; ++it) {
        if (it > 1) sv = landau2_eval(a11, a22, a12, e, l, m, P);
        e = e + landau2_delta(sv, e, l, m, P);
      }
      e33[i] = e;
    } else {
      sv = landau2_eval(a11, a22, a12, 0.0, l, m, P);
    }
    double sg[3];
    landau2_sigma(sv, a11, a22, a12, m, P, sg);
    if (tau) {
      const double l0tr = lam0 * (e11 + e22), m02 = 2.0 * mu0;
      tau[i] = sg[0] - (l0tr + m02 * e11);
      tau[NV + i] = sg[1] - (l0tr + m02 * e22);
      tau[2 * NV + i] = sg[2] - m02 * e12;
    }
    if (sig) {
      sig[i] = sg[0];
      sig[NV + i] = sg[1];
      sig[2 * NV + i] = sg[2];
      acc[0] += e11;
      acc[1] += e22;
      acc[2] += e12;
      acc[3] += sg[0];
      acc[4] += sg[1];
      acc[5] += sg[2];
    }
  }
  if (partials) {
    block_sum<6>(acc, red);
    if (threadIdx.x == 0)
      for (int c = 0; c < 6; ++c) partials[(size_t)blockIdx.x * 6 + c] = acc[c];
  }
}

__global__ void k_l2_after_eval(LsState* __restrict__ st, int persist) {
  if (persist) st->e33_valid = 1;
  st->nmax = 0;
}

// ---------------------------------------------------------------------------
// Type-II Anderson mixing of the last m+1 iterates (the reference's dbfft path:
// numpy lstsq on residual differences, mgkmc/linear_elastic_simulator.py:
// 1178-1212), on the device: history ring of g = G(eps) and f = G(eps) - eps,
// Gram sums of the residual differences (off-diagonal component weight 2, as
// the flattened full tensor counts it), a one-thread pseudo-inverse solve, mix.
// ---------------------------------------------------------------------------
// Ring of nslot = m + 1 history entries, oldest first: entry a lives in slot
// (aa_head + a) % nslot; g and f slots are n = 3 NV values apart.
__device__ __forceinline__ size_t aa_slot(const LsState* st, int a, int nslot) {
  return (size_t)((st->aa_head + a) % nslot);
}

// new entry (g = eps, f = eps - eps_old) at the ring's end; skipped once the
// solve has stopped
__global__ void k_aa_store(const double* __restrict__ eps, const double* __restrict__ eps_old, double* __restrict__ gb,
                           double* __restrict__ fb, size_t n, int nslot, const LsState* __restrict__ st) {
  if (st->done) return;
  const size_t sl = st->aa_len < nslot ? aa_slot(st, st->aa_len, nslot) : (size_t)st->aa_head;
  double* gs = gb + sl * n;
  double* fs = fb + sl * n;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const double v = eps[i];
    gs[i] = v;
    fs[i] = v - eps_old[i];
  }
}

__global__ void k_aa_advance(int nslot, LsState* __restrict__ st) {
  if (st->done) return;
  if (st->aa_len < nslot) st->aa_len += 1;
  else st->aa_head = (st->aa_head + 1) % nslot;
}

// partials per block: G[a][b] (a <= b) of dF_a = f_{a+1} - f_a (a < L-1 <= 5),
// then r_a = (dF_a, f_{L-1}); the off-diagonal component counts twice
__global__ __launch_bounds__(MGK_NT) void k_aa_gram(const double* __restrict__ fb, size_t NV, int nslot,
                                                    double* __restrict__ partials, const LsState* __restrict__ st) {
  const int L = st->aa_len;
  if (st->done || L < 2) return;
  __shared__ double red[20 * (MGK_NT / 64)];
  double acc[20];
#pragma unroll
  for (int k = 0; k < 20; ++k) acc[k] = 0.0;
  const int K = L - 1;
  const size_t n = 3 * NV;
  const double* f[6];
#pragma unroll
  for (int a = 0; a < 6; ++a) f[a] = fb + aa_slot(st, a < L ? a : L - 1, nslot) * n;
  for (size_t i = blockIdx.x * (size_t)MGK_NT + threadIdx.x; i < n; i += (size_t)gridDim.x * MGK_NT) {
    const double wgt = i >= 2 * NV ? 2.0 : 1.0;
    double d[5];
    const double fl = f[K][i];
#pragma unroll
    for (int a = 0; a < 5; ++a) d[a] = a < K ? f[a + 1][i] - f[a][i] : 0.0;
    int k = 0;
#pragma unroll
    for (int a = 0; a < 5; ++a)
#pragma unroll
      for (int b = a; b < 5; ++b) acc[k++] += wgt * d[a] * d[b];
#pragma unroll
    for (int a = 0; a < 5; ++a) acc[15 + a] += wgt * d[a] * fl;
  }
  block_sum<20>(acc, red);
  if (threadIdx.x == 0)
    for (int k = 0; k < 20; ++k) partials[(size_t)blockIdx.x * 20 + k] = acc[k];
}

// one thread: reduce the partials, solve the K x K normal equations through a
// Jacobi eigen-decomposition with the least-squares singular-value cutoff
// (rcond = eps * rows), store gamma; aa_plain = 1 if it is not finite
__global__ void k_aa_solve(const double* __restrict__ partials, int nparts, double rows, LsState* __restrict__ st) {
  const int L = st->aa_len;
  if (st->done || L < 2 || threadIdx.x != 0) return;
  const int K = L - 1;
  double acc[20];
  for (int k = 0; k < 20; ++k) acc[k] = 0.0;
  for (int b = 0; b < nparts; ++b)
    for (int k = 0; k < 20; ++k) acc[k] += partials[(size_t)b * 20 + k];
  double G[5][5], V[5][5], rhs[5];
  int k = 0;
  for (int a = 0; a < 5; ++a)
    for (int b = a; b < 5; ++b) {
      G[a][b] = G[b][a] = acc[k++];
    }
  for (int a = 0; a < 5; ++a) rhs[a] = acc[15 + a];
  for (int a = 0; a < 5; ++a)
    for (int b = 0; b < 5; ++b) V[a][b] = (a == b) ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 60; ++sweep) {
    double off = 0.0, dia = 0.0;
    for (int a = 0; a < K; ++a) {
      dia += G[a][a] * G[a][a];
      for (int b = a + 1; b < K; ++b) off += G[a][b] * G[a][b];
    }
    if (off <= 1e-32 * dia) break;
    for (int p = 0; p < K; ++p)
      for (int q = p + 1; q < K; ++q) {
        if (G[p][q] == 0.0) continue;
        const double th = (G[q][q] - G[p][p]) / (2.0 * G[p][q]);
        const double t = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
        const double c = 1.0 / sqrt(t * t + 1.0), sn = t * c;
        for (int r = 0; r < K; ++r) {
          const double gp = G[r][p], gq = G[r][q];
          G[r][p] = c * gp - sn * gq;
          G[r][q] = sn * gp + c * gq;
        }
        for (int r = 0; r < K; ++r) {
          const double gp = G[p][r], gq = G[q][r];
          G[p][r] = c * gp - sn * gq;
          G[q][r] = sn * gp + c * gq;
        }
        for (int r = 0; r < K; ++r) {
          const double vp = V[r][p], vq = V[r][q];
          V[r][p] = c * vp - sn * vq;
          V[r][q] = sn * vp + c * vq;
        }
      }
  }
  double smax = 0.0;
  for (int a = 0; a < K; ++a) smax = fmax(smax, sqrt(fmax(G[a][a], 0.0)));
  const double cut = 2.220446049250313e-16 * rows * smax;
  double gam[5] = {0, 0, 0, 0, 0};
  for (int a = 0; a < K; ++a) {
    const double sv = sqrt(fmax(G[a][a], 0.0));
    if (!(sv > cut)) continue;
    double proj = 0.0;
    for (int r = 0; r < K; ++r) proj += V[r][a] * rhs[r];
    proj /= G[a][a];
    for (int r = 0; r < K; ++r) gam[r] += V[r][a] * proj;
  }
  bool ok = true;
  for (int a = 0; a < K; ++a) {
    st->aa_gamma[a] = gam[a];
    ok = ok && isfinite(gam[a]);
  }
  st->aa_plain = ok ? 0 : 1;
}

// eps = g_{L-1} - sum_a gamma_a (g_{a+1} - g_a), or the plain step g_{L-1}
__global__ void k_aa_mix(const double* __restrict__ gb, size_t n, int nslot, double* __restrict__ eps,
                         const LsState* __restrict__ st) {
  if (st->done) return;
  const int L = st->aa_len, K = L - 1;
  const bool plain = L < 2 || st->aa_plain;
  double gam[5] = {0, 0, 0, 0, 0};
  const double* g[6];
  for (int a = 0; a < 6; ++a) g[a] = gb + aa_slot(st, a < L ? a : L - 1, nslot) * n;
  for (int a = 0; a < K && a < 5; ++a) gam[a] = st->aa_gamma[a];
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    double v = g[K][i];
    if (!plain) {
      double dg = 0.0;
      for (int a = 0; a < K; ++a) dg += (g[a + 1][i] - g[a][i]) * gam[a];
      v = v - dg;
    }
    eps[i] = v;
  }
}

// after a plain fallback the history restarts from its newest entry
__global__ void k_aa_finish(int nslot, LsState* __restrict__ st) {
  if (st->done) return;
  if (st->aa_len >= 2 && st->aa_plain) {
    st->aa_head = (st->aa_head + st->aa_len - 1) % nslot;
    st->aa_len = 1;
  }
  st->aa_plain = 0;
}

__global__ void k_copy(const double* __restrict__ a, double* __restrict__ b, size_t n, const LsState* __restrict__ st) {
  if (st->done) return;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) b[i] = a[i];
}

// clears n 32-bit words (a fill kernel rather than hipMemsetAsync, so the
// enqueue sequence stays a plain kernel sequence)
__global__ void k_zero_words(unsigned int* __restrict__ p, size_t n) {
  for (size_t i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0u;
}

// ---------------------------------------------------------------------------
// Batched forms of the set-up, mean and stress passes (member
// blockIdx.y, or blockIdx.x for the one-block passes): each member runs the
// single pass's body on its own buffers.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_ls_init_b(LsBatch<T> bt) {
  const LsMem<T>& M = bt.m[blockIdx.y];
  for (int i = threadIdx.x; i < (int)(sizeof(LsState) / 4); i += blockDim.x) ((unsigned int*)M.st)[i] = 0u;
}

template <typename T>
__global__ __launch_bounds__(MGK_NT) void k_comp_sums_b(LsBatch<T> bt, int C, size_t NV) {
  const LsMem<T>& M = bt.m[blockIdx.y];
  comp_sums_body<T>(M.eps, C, NV, M.part);
}

template <typename T>
__global__ __launch_bounds__(MGK_NT) void k_means_finalize_b(LsBatch<T> bt, int nparts, int C, double NV) {
  const LsMem<T>& M = bt.m[blockIdx.x];
  means_finalize_body(M.part, nparts, C, NV, M.st->mean);
}

template <typename T>
__global__ void k_ls_warm_adjust_b(LsBatch<T> bt, int C, size_t NV) {
  const LsMem<T>& M = bt.m[blockIdx.y];
  ls_warm_adjust_body<T>(M.eps, C, NV, M.eb, M.st->mean);
}

template <typename T, int D>
__global__ __launch_bounds__(MGK_NT) void k_ls_stress_b(LsBatch<T> bt, size_t NV) {
  const LsMem<T>& M = bt.m[blockIdx.y];
  ls_stress_body<T, D>(M.eps, M.epsp, M.lam, M.mu, M.sig, NV, M.part);
}

// ---------------------------------------------------------------------------
// host driver
// ---------------------------------------------------------------------------
// the C ABI's structs (mgk_ls_info, mgk_landau_params, mgk_secant_params, mgk_ls_member)
#include "../../include/mgkmc_hip.h"

namespace {

template <typename T>
constexpr int maxv() {
  return sizeof(T) == 8 ? 8 : 16;
}

struct LsPlan {
  int D, C, nx, P, Q, nxh;
  long long NV;
  int logBx, logBy, logPP;
  int ctx, cty, ctq;  // compile-time specialised sweeps in use
  size_t lds_x, lds_y, lds_q;
  int nblk_upd, nblk_red;
  size_t off_spec, off_part, off_state, total;
};

template <typename T>
int ls_plan(int ndim, const int64_t* shape, LsPlan* pl, bool allow_ct = true) {
  if (ndim != 2 && ndim != 3) return MGK_ERR_ARG;
  pl->D = ndim;
  pl->C = ndim == 3 ? 6 : 3;
  pl->nx = (int)shape[0];
  pl->P = ndim == 3 ? (int)shape[1] : 1;
  pl->Q = ndim == 3 ? (int)shape[2] : (int)shape[1];
  if (pl->nx < 1 || pl->P < 1 || pl->Q < 1) return MGK_ERR_ARG;
  pl->nxh = pl->nx / 2 + 1;
  pl->NV = (long long)pl->nx * pl->P * pl->Q;
  FftPlan t;
  if (!mgk_make_plan(pl->nx, &t) || !mgk_make_plan(pl->P, &t) || !mgk_make_plan(pl->Q, &t)) return MGK_ERR_ARG;
  const int TILE = maxv<T>() * MGK_NT;
  const size_t lim = mgk_lds_limit();
  const size_t cb = sizeof(cplx<T>);
  // x pass: tile nx x B
  if (pl->nx > TILE) return MGK_ERR_ARG;
  int B = 1;
  while (2 * B * pl->nx <= TILE && 2 * B <= mgk_pow2_ceil(pl->Q)) B *= 2;
  while (B > 1 && (size_t)(pl->nx + pl->nx * B) * cb > lim) B /= 2;
  pl->logBx = mgk_ilog2(B);
  pl->lds_x = (size_t)(pl->nx + pl->nx * B) * cb;
  if (pl->lds_x > lim) return MGK_ERR_ARG;
  // mid pass
  B = 1;
  if (pl->P > TILE) return MGK_ERR_ARG;
  while (2 * B * pl->P <= TILE && 2 * B <= mgk_pow2_ceil(pl->Q)) B *= 2;
  while (B > 1 && (size_t)(pl->P + pl->P * B) * cb > lim) B /= 2;
  pl->logBy = mgk_ilog2(B);
  pl->lds_y = (size_t)(pl->P + pl->P * B) * cb;
  // pencil pass: C x Q x PP (the generic path stages at most TILE values per
  // workgroup; longer pencils need the compile-time path below)
  const bool gen_q = pl->C * pl->Q <= TILE;
  int PP = 1;
  const int npen = pl->nxh * pl->P;
  while (2 * PP * pl->C * pl->Q <= TILE && 2 * PP <= mgk_pow2_ceil(npen)) PP *= 2;
  while (PP > 1 && (size_t)(pl->Q + pl->C * pl->Q * PP) * cb > lim) PP /= 2;
  pl->logPP = mgk_ilog2(PP);
  pl->lds_q = (size_t)(pl->Q + pl->C * pl->Q * PP) * cb;
  // compile-time specialised sweeps (lengths of ct_len: 64..512 in powers of two
  // and 96, 192, 384; tiles of up to 4096 points, power-of-two column counts)
  const int CTT = ct_tile<T>();
  auto ct_ok = [](int n) { return ct_len(n); };
  pl->ctx = pl->cty = pl->ctq = 0;
  // x sweeps: wide tiles (ctx = 2), or narrow FP32 tiles (ctx = 1)
  for (int wide = 1; wide >= (sizeof(T) == 4 ? 0 : 1) && !pl->ctx; --wide) {
    const int XTT = wide ? x_tile<T, true>() : x_tile<T, false>();
    if (allow_ct && ct_ok(pl->nx) && pl->Q >= XTT / pl->nx && (size_t)(pl->nx + XTT) * cb <= lim) {
      pl->ctx = wide ? 2 : 1;
      pl->logBx = mgk_ilog2(XTT / pl->nx);
      pl->lds_x = (size_t)(pl->nx + XTT) * cb;
    }
  }
  if (allow_ct && ct_ok(pl->P) && pl->Q >= CTT / pl->P && (size_t)(pl->P + CTT) * cb <= lim) {
    pl->cty = 1;
    pl->logBy = mgk_ilog2(CTT / pl->P);
    pl->lds_y = (size_t)(pl->P + CTT) * cb;
  }
  if (allow_ct && ct_ok(pl->Q)) {
    int pp = 1;
    while (2 * pp * pl->C * pl->Q <= CTT) pp *= 2;
    if (npen >= pp && (size_t)(pl->Q + tpad_size(pl->C * pl->Q * pp)) * cb <= lim) {
      pl->ctq = 1;
      pl->logPP = mgk_ilog2(pp);
      pl->lds_q = (size_t)(pl->Q + tpad_size(pl->C * pl->Q * pp)) * cb;
    }
  }
  if (!pl->ctq && (!gen_q || pl->lds_q > lim)) return MGK_ERR_ARG;
  const int nqt = (pl->Q + (1 << pl->logBx) - 1) >> pl->logBx;
  pl->nblk_upd = nqt * pl->P;
  long long nr = (pl->NV + MGK_NT - 1) / MGK_NT;
  pl->nblk_red = (int)std::min<long long>(nr, 1024);
  pl->off_spec = 0;
  pl->off_part = mgk_align((size_t)pl->C * pl->nxh * pl->P * pl->Q * cb);
  // partial sums: update norms, stress means, Anderson Gram sums (20 per block)
  size_t npart = (size_t)std::max(pl->nblk_upd * 2, pl->nblk_red * std::max(2 * pl->C, 20));
  pl->off_state = pl->off_part + mgk_align(npart * sizeof(double));
  pl->total = pl->off_state + mgk_align(sizeof(LsState));
  return MGK_OK;
}

// threads per workgroup of the compile-time mid-axis sweep (its tiles are
// LDS-bound to four per CU, so wider workgroups give more waves per CU)
constexpr int MID_NT = 512;

template <typename T, int MV>
void launch_mid(const LsPlan& pl, dim3 gy, hipStream_t stream, const FftPlan& ply, const cplx<T>* twy,
                const MidPtrs<T>& mp, int inv) {
  if (pl.cty) {
    ct_dispatch(pl.P, [&](auto nn) {
      constexpr int NN = decltype(nn)::value, BB = ct_cols(ct_tile<T>(), NN);
      k_fft_mid<T, ct_tile<T>() / MGK_NT, NN, BB, MID_NT><<<gy, MID_NT, pl.lds_y, stream>>>(pl.P, pl.Q, pl.nxh,
                                                                                          pl.logBy, ply, twy, mp, inv);
    });
  } else {
    k_fft_mid<T, MV><<<gy, MGK_NT, pl.lds_y, stream>>>(pl.P, pl.Q, pl.nxh, pl.logBy, ply, twy, mp, inv);
  }
}

// Resources of a solve that allocates nothing: twiddle tables already in
// device memory and a host buffer for the state reads (two LsState).
struct NoAllocCtx {
  const void* tw[3];
  LsState* host_state;
};

// last iteration count per workspace (0 = unknown); set when iters > 0
int ls_iter_hint(const void* ws, int iters) {
  static std::mutex mu;
  static std::map<const void*, int> last;
  std::lock_guard<std::mutex> lk(mu);
  if (iters > 0) {
    last[ws] = iters;
    return iters;
  }
  auto it = last.find(ws);
  return it == last.end() ? 0 : it->second;
}

// ---------------------------------------------------------------------------
// Heat conduction of the local temperature field (the thermal extension of
// the KMC simulators, mgkmc/aqs.py:462-478, mgkmc/kmc_simulator.py:559-575):
// T <- F^-1[exp(-|k|^2 D dt) F[T]], then the thermostat and the cap.  The x
// pass is k_ls_xfwd_real with one component, the y pass k_fft_mid; this
// pencil kernel multiplies along z, k_heat_xinv transforms back.
// ---------------------------------------------------------------------------
struct HeatPar {
  double kx, ky, kz;  // wave number per frequency bin along each axis (2 pi / L)
  double Ddt;         // diffusivity * dt
  double scale;       // 1 / NV (the transforms are unnormalised)
};

template <typename T, int MAXV>
__global__ __launch_bounds__(MGK_NT) void k_heat_pencil(LsGeom<T> g, cplx<T>* __restrict__ spec, HeatPar hp) {
  extern __shared__ __align__(16) unsigned char smem[];
  cplx<T>* tw = reinterpret_cast<cplx<T>*>(smem);
  cplx<T>* tile = tw + g.Q;
  const int Q = g.Q, P = g.P, logPP = g.logPP, PP = 1 << logPP;
  const int pid0 = blockIdx.x * PP, npen = g.npencil;
  for (int t = threadIdx.x; t < Q; t += MGK_NT) tw[t] = g.twq[t];
  for (int e = threadIdx.x; e < Q * PP; e += MGK_NT) {
    const int pp = e / Q, q = e - pp * Q, pid = pid0 + pp;
    tile[q * PP + pp] = (pid < npen) ? spec[(size_t)pid * Q + q] : cmk(T(0), T(0));
  }
  __syncthreads();
  const Tile tl{Q, logPP, PP, 1, Q * PP};
  tile_fft<MAXV>(tile, tl, g.plq, tw, false);
  for (int e = threadIdx.x; e < Q * PP; e += MGK_NT) {
    const int pp = e & (PP - 1), q = e >> logPP, pid = pid0 + pp;
    if (pid >= npen) continue;
    const int kx = pid / P, p = pid - kx * P;
    const double ax = hp.kx * fbin(kx, g.nx), ay = hp.ky * fbin(p, P), az = hp.kz * fbin(q, Q);
    const double k2 = (ax * ax + ay * ay) + az * az;
    const T f = T(exp(-k2 * hp.Ddt) * hp.scale);
    const cplx<T> v = tile[q * PP + pp];
    tile[q * PP + pp] = cmk(v.x * f, v.y * f);
  }
  __syncthreads();
  tile_fft<MAXV>(tile, tl, g.plq, tw, true);
  for (int e = threadIdx.x; e < Q * PP; e += MGK_NT) {
    const int pp = e / Q, q = e - pp * Q, pid = pid0 + pp;
    if (pid < npen) spec[(size_t)pid * Q + q] = tile[q * PP + pp];
  }
}

// inverse x transform of one half-spectrum component into a real field
template <typename T, int MAXV>
__global__ __launch_bounds__(MGK_NT) void k_heat_xinv(LsGeom<T> g, const cplx<T>* __restrict__ spec,
                                                      T* __restrict__ out) {
  extern __shared__ __align__(16) unsigned char smem[];
  cplx<T>* tw = reinterpret_cast<cplx<T>*>(smem);
  cplx<T>* tile = tw + g.nx;
  const int nx = g.nx, P = g.P, Q = g.Q, logB = g.logBx, B = 1 << logB;
  const int q0 = blockIdx.x * B, p = blockIdx.y;
  for (int t = threadIdx.x; t < nx; t += MGK_NT) tw[t] = g.twx[t];
  xmerge_load<T, MGK_NT>(tile, B, nx, g.nxh, logB, q0, Q, P, p, spec, (const cplx<T>*)nullptr);
  const Tile tl{nx, logB, B, 1, 0};
  tile_fft<MAXV>(tile, tl, g.plx, tw, true);
  for (int e = threadIdx.x; e < nx * B; e += MGK_NT) {
    const int b = e & (B - 1), x = e >> logB, q = q0 + b;
    if (q < Q) out[((size_t)x * P + p) * Q + q] = tile[x * B + b].x;
  }
}

// thermostat (mode 1: T0 + (T - T0) f; mode 2: T - (<T> - T0)) and cap (cap > 0)
__global__ void k_heat_post(double* __restrict__ Tf, size_t NV, double T0, int mode, double f, double cap,
                            const double* __restrict__ mean) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < NV; i += (size_t)gridDim.x * blockDim.x) {
    double t = Tf[i];
    if (mode == 1) t = T0 + (t - T0) * f;
    else if (mode == 2) t = t - (mean[0] - T0);
    if (cap > 0) t = fmin(t, cap);
    Tf[i] = t;
  }
}

// 2D Landau mode of the LS loop (FP64)
struct L2Mode {
  LandauPar P;
  int plane_stress;
  int persist, e33_have;  // e33 buffer kept between evaluations / holds a warm start now
  double* e33;            // NV
  double* tau;            // 3 NV
  int aa_m;               // Anderson depth (0: plain fixed point)
  double *aa_g, *aa_f;    // (m + 1) x 3 NV each
  double* eps_old;        // 3 NV
};

__global__ void k_l2_init(LsState* __restrict__ st, int e33_valid) { st->e33_valid = e33_valid; }


// ---------------------------------------------------------------------------
// Launches of the linear sweeps over a table of nb members of one grid shape.
// The grid of each sweep gains the member dimension; a member's workgroups
// run the same code on its own buffers as in a solve of that member alone.
// ---------------------------------------------------------------------------
template <typename T>
LsGeom<T> ls_geom(const LsPlan& pl, int C) {
  LsGeom<T> g;
  g.nx = pl.nx;
  g.P = pl.P;
  g.Q = pl.Q;
  g.nxh = pl.nxh;
  g.C = C;
  g.NV = pl.NV;
  g.logBx = pl.logBx;
  g.logBy = pl.logBy;
  g.logPP = pl.logPP;
  g.npencil = pl.nxh * pl.P;
  mgk_make_plan(pl.nx, &g.plx);
  mgk_make_plan(pl.P, &g.ply);
  mgk_make_plan(pl.Q, &g.plq);
  g.twx = g.twy = g.twq = nullptr;
  return g;
}

// symmetric part of the macro strain in component order and the squared
// norm of its antisymmetric part (per voxel)
template <int D>
Sym6 ls_sym_bar(const double* e, double* anti2) {
  Sym6 eb{};
  if (D == 3) {
    eb.v[0] = e[0];
    eb.v[1] = e[4];
    eb.v[2] = e[8];
    eb.v[3] = 0.5 * (e[5] + e[7]);
    eb.v[4] = 0.5 * (e[2] + e[6]);
    eb.v[5] = 0.5 * (e[1] + e[3]);
    const double a1 = 0.5 * (e[5] - e[7]), a2 = 0.5 * (e[2] - e[6]), a3 = 0.5 * (e[1] - e[3]);
    *anti2 = 2.0 * (a1 * a1 + a2 * a2 + a3 * a3);
  } else {
    eb.v[0] = e[0];
    eb.v[1] = e[3];
    eb.v[2] = 0.5 * (e[1] + e[2]);
    const double a = 0.5 * (e[1] - e[2]);
    *anti2 = 2.0 * a * a;
  }
  return eb;
}

// macro means of a finished solve into the caller's info record
template <int D>
void ls_fill_info(const LsState& s, const double* eps_bar, mgk_ls_info* info) {
  constexpr int C = (D == 3) ? 6 : 3;
  info->iters = s.iters;
  info->status = s.status;
  info->diff = s.diff;
  const double* em = s.mean;
  const double* sm = s.mean + C;
  if (D == 3) {
    const double a1 = 0.5 * (eps_bar[5] - eps_bar[7]), a2 = 0.5 * (eps_bar[2] - eps_bar[6]),
                 a3 = 0.5 * (eps_bar[1] - eps_bar[3]);
    const double E[9] = {em[0], em[5] + a3, em[4] + a2, em[5] - a3, em[1], em[3] + a1, em[4] - a2, em[3] - a1, em[2]};
    const double S[9] = {sm[0], sm[5], sm[4], sm[5], sm[1], sm[3], sm[4], sm[3], sm[2]};
    for (int i = 0; i < 9; ++i) {
      info->eps_mean[i] = E[i];
      info->sig_mean[i] = S[i];
    }
  } else {
    const double a = 0.5 * (eps_bar[1] - eps_bar[2]);
    const double E[4] = {em[0], em[2] + a, em[2] - a, em[1]};
    const double S[4] = {sm[0], sm[2], sm[2], sm[1]};
    for (int i = 0; i < 9; ++i) {
      info->eps_mean[i] = i < 4 ? E[i] : 0.0;
      info->sig_mean[i] = i < 4 ? S[i] : 0.0;
    }
  }
}

// the compile-time x sweeps of length NN on wide or narrow tiles
template <typename T, int D, int NN, bool WIDE>
void launch_tau_ct(const LsPlan& pl, const LsGeom<T>& g, dim3 gx, hipStream_t stream, const LsBatch<T>& bt) {
  constexpr int NT = x_nt<WIDE>(), BB = ct_cols(x_tile<T, WIDE>(), NN);
  k_ls_tau_xfwd<T, D, NN * BB / NT, NN, BB, NT><<<gx, NT, pl.lds_x, stream>>>(g, bt);
}
template <typename T, int D, int NN, bool WIDE>
void launch_xinv_ct(const LsPlan& pl, const LsGeom<T>& g, dim3 gx, hipStream_t stream, const LsBatch<T>& bt) {
  constexpr int NT = x_nt<WIDE>(), BB = ct_cols(x_tile<T, WIDE>(), NN);
  k_ls_xinv_update<T, D, NN * BB / NT, NN, BB, NT><<<gx, NT, pl.lds_x, stream>>>(g, bt);
}
template <typename T, int NN, bool WIDE>
void launch_xfwd_real_ct(const LsPlan& pl, const LsGeom<T>& g, dim3 gx, hipStream_t stream, const T* tau,
                         cplx<T>* spec, const LsState* st) {
  constexpr int NT = x_nt<WIDE>(), BB = ct_cols(x_tile<T, WIDE>(), NN);
  k_ls_xfwd_real<T, NN * BB / NT, NN, BB, NT><<<gx, NT, pl.lds_x, stream>>>(g, tau, spec, st);
}

template <typename T, int D>
void launch_tau(const LsPlan& pl, const LsGeom<T>& g, dim3 gx, hipStream_t stream, const LsBatch<T>& bt) {
  constexpr int MV = maxv<T>();
  if (pl.ctx) {
    ct_dispatch(pl.nx, [&](auto nn) {
      constexpr int NN = decltype(nn)::value;
      if (pl.ctx == 2) launch_tau_ct<T, D, NN, true>(pl, g, gx, stream, bt);
      else if constexpr (sizeof(T) == 4) launch_tau_ct<T, D, NN, false>(pl, g, gx, stream, bt);
    });
  } else {
    k_ls_tau_xfwd<T, D, MV><<<gx, MGK_NT, pl.lds_x, stream>>>(g, bt);
  }
}

template <typename T, int D>
void launch_pencil(const LsPlan& pl, const LsGeom<T>& g, dim3 gp, hipStream_t stream, const LsBatch<T>& bt,
                   T scale) {
  constexpr int MV = maxv<T>();
  constexpr int TT = ct_tile<T>();
  constexpr int C = (D == 3) ? 6 : 3;
  if (pl.ctq) {
    ct_dispatch(pl.Q, [&](auto nn) {
      constexpr int NN = decltype(nn)::value;
      k_ls_pencil<T, D, TT / MGK_NT, NN, ct_pp(TT, C, NN)><<<gp, MGK_NT, pl.lds_q, stream>>>(g, bt, scale);
    });
  } else {
    k_ls_pencil<T, D, MV><<<gp, MGK_NT, pl.lds_q, stream>>>(g, bt, scale);
  }
}

template <typename T, int D>
void launch_xinv(const LsPlan& pl, const LsGeom<T>& g, dim3 gx, hipStream_t stream, const LsBatch<T>& bt) {
  constexpr int MV = maxv<T>();
  if (pl.ctx) {
    ct_dispatch(pl.nx, [&](auto nn) {
      constexpr int NN = decltype(nn)::value;
      if (pl.ctx == 2) launch_xinv_ct<T, D, NN, true>(pl, g, gx, stream, bt);
      else if constexpr (sizeof(T) == 4) launch_xinv_ct<T, D, NN, false>(pl, g, gx, stream, bt);
    });
  } else {
    k_ls_xinv_update<T, D, MV><<<gx, MGK_NT, pl.lds_x, stream>>>(g, bt);
  }
}

// r2c along x of a C-component real field (the nonlinear laws' polarisation)
template <typename T>
void launch_xfwd_real(const LsPlan& pl, const LsGeom<T>& g, dim3 gx, hipStream_t stream, const T* tau,
                      cplx<T>* spec, const LsState* st) {
  constexpr int MV = maxv<T>();
  if (pl.ctx) {
    ct_dispatch(pl.nx, [&](auto nn) {
      constexpr int NN = decltype(nn)::value;
      if (pl.ctx == 2) launch_xfwd_real_ct<T, NN, true>(pl, g, gx, stream, tau, spec, st);
      else if constexpr (sizeof(T) == 4) launch_xfwd_real_ct<T, NN, false>(pl, g, gx, stream, tau, spec, st);
    });
  } else {
    k_ls_xfwd_real<T, MV><<<gx, MGK_NT, pl.lds_x, stream>>>(g, tau, spec, st);
  }
}

template <typename T>
MidPtrs<T> mid_table(const LsBatch<T>& bt, int nb, int C) {
  MidPtrs<T> mp{};
  for (int b = 0; b < nb; ++b) {
    mp.spec[b] = bt.m[b].spec;
    mp.done[b] = &bt.m[b].st->done;
  }
  mp.ncomp = C;
  return mp;
}

template <typename T>
LsFin fin_table(const LsBatch<T>& bt, int nb) {
  LsFin f{};
  for (int b = 0; b < nb; ++b) {
    f.part[b] = bt.m[b].part;
    f.st[b] = bt.m[b].st;
    f.e2_extra[b] = bt.m[b].e2_extra;
  }
  return f;
}

template <typename T, int D>
int ls_solve_t(const LsPlan& pl, const T* lam, const T* mu, double lam0, double mu0, const double* eps_bar,
               const T* epsp, T* eps, T* sig, int warm, int max_iter, double tol, char* ws, hipStream_t stream,
               mgk_ls_info* info, bool final_stress = true, const LandauPar* lp = nullptr, T* taubuf = nullptr,
               const L2Mode* l2 = nullptr, const SecLaw* sl = nullptr, const NoAllocCtx* na = nullptr,
               const float* eps_from32 = nullptr) {
  constexpr int MV = maxv<T>();
  constexpr int C = (D == 3) ? 6 : 3;
  const int prec = sizeof(T) == 8 ? 0 : 1;
  LsGeom<T> g = ls_geom<T>(pl, C);
  if (na) {  // twiddle tables staged in the caller's scratch (nothing allocated here)
    g.twx = (const cplx<T>*)na->tw[0];
    g.twy = (const cplx<T>*)na->tw[1];
    g.twq = (const cplx<T>*)na->tw[2];
  } else {
    g.twx = (const cplx<T>*)mgk_twiddles(pl.nx, prec);
    g.twy = (const cplx<T>*)mgk_twiddles(pl.P, prec);
    g.twq = (const cplx<T>*)mgk_twiddles(pl.Q, prec);
  }
  if (!g.twx || !g.twy || !g.twq) {
    mgk_set_error("twiddle table allocation failed", __FILE__, __LINE__);
    return MGK_ERR_HIP;
  }
  cplx<T>* spec = (cplx<T>*)(ws + pl.off_spec);
  double* part = (double*)(ws + pl.off_part);
  LsState* st = (LsState*)(ws + pl.off_state);

  // symmetric part of eps_bar in component order, antisymmetric norm extra
  double anti2 = 0.0;
  const Sym6 eb = ls_sym_bar<D>(eps_bar, &anti2);
  const double e2_extra = anti2 * (double)pl.NV;
  const size_t NV = (size_t)pl.NV;

  k_zero_words<<<1, 64, 0, stream>>>((unsigned int*)st, sizeof(LsState) / 4);
  if (l2) k_l2_init<<<1, 1, 0, stream>>>(st, l2->persist && l2->e33_have ? 1 : 0);
  const int nfill = (int)std::min<long long>((pl.NV + MGK_NT - 1) / MGK_NT, 4096);
  // cold starts of the linear law read eps = eps_bar as a constant in the
  // first iteration (nothing else reads eps before that iteration's update
  // has written it); the nonlinear laws read eps in their own kernels
  const bool uniform_start = !warm && !lp && !sl && !l2;
  if (!warm && !uniform_start) {
    k_ls_fill<T><<<nfill, MGK_NT, 0, stream>>>(eps, C, NV, eb);
  } else if (warm == 1) {
    k_comp_sums<T><<<pl.nblk_red, MGK_NT, 0, stream>>>(eps, C, NV, part);
    k_means_finalize<<<1, MGK_NT, 0, stream>>>(part, pl.nblk_red, C, (double)NV, st->mean);
    k_ls_warm_adjust<T><<<nfill, MGK_NT, 0, stream>>>(eps, C, NV, eb, st->mean);
  }
  MGK_HIP(hipGetLastError());

  const T scale = T(1.0 / (double)pl.NV);
  const int nqx = (pl.Q + (1 << pl.logBx) - 1) >> pl.logBx;
  const int nqy = (pl.Q + (1 << pl.logBy) - 1) >> pl.logBy;
  const int npb = (g.npencil + (1 << pl.logPP) - 1) >> pl.logPP;
  const dim3 gx(nqx, pl.P), gy(nqy, pl.nxh, C), gp(npb);
  // this solve as a one-member table of the linear sweeps
  LsBatch<T> bt{};
  bt.m[0] = LsMem<T>{lam, mu, epsp, eps, spec, part, st, eb, lam0, mu0, e2_extra};
  // an FP32 starting field is read by the first iteration's x sweeps only
  // (the linear law's; the nonlinear laws read eps in their own kernels)
  if (eps_from32 && (lp || sl || l2 || warm != 2)) {
    mgk_set_error("an FP32 starting field needs the linear law and warm = 2", __FILE__, __LINE__);
    return MGK_ERR_ARG;
  }
  bt.m[0].eps_old32 = eps_from32;
  bt.m[0].eps_uniform = uniform_start ? 1 : 0;
  const MidPtrs<T> mp = mid_table(bt, 1, C);
  const LsFin fin = fin_table(bt, 1);

  // algorithmic HBM bytes per launch (what each sweep must read + write)
  const double sT = (double)sizeof(T);
  const double specB = (double)C * pl.nxh * pl.P * pl.Q * 2.0 * sT;
  const double fieldB = (double)pl.NV * sT;
  const double b_tau = fieldB * ((epsp ? 2.0 * C : (double)C) + 2.0) + specB;
  const double b_fft = 2.0 * specB;
  const double b_upd = specB + 2.0 * C * fieldB;
  // the first iteration reads its starting field as a constant (uniform
  // start) or in FP32 (start from an FP32 iterate) in both x sweeps
  const double eps_in = (double)C * fieldB;
  const double eps_in0 = uniform_start ? 0.0 : (eps_from32 ? (double)C * pl.NV * 4.0 : eps_in);
  const bool light0 = uniform_start || eps_from32;  // timed under the _START ids
  MgkProfRec prof(stream, na == nullptr);
  int iter_idx = 0;
  const int nel = (int)std::min<long long>((pl.NV + MGK_NT - 1) / MGK_NT, 8192);
  auto enqueue_iter = [&]() {
    int h = -1;
    if constexpr (sizeof(T) == 8 && D == 2) {
      if (l2) {
        if (l2->plane_stress) k_l2_count<<<nel, MGK_NT, 0, stream>>>(eps, epsp, lam, mu, l2->P, l2->e33, NV, st, 0);
        k_l2_stress<<<nel, MGK_NT, 0, stream>>>(eps, epsp, lam, mu, lam0, mu0, l2->P, l2->plane_stress, l2->e33,
                                                l2->tau, nullptr, nullptr, NV, st, 0);
        k_l2_after_eval<<<1, 1, 0, stream>>>(st, l2->persist);
        launch_xfwd_real<T>(pl, g, gx, stream, l2->tau, spec, st);
      }
    }
    if (l2) {
    } else if (lp || sl) {
      // nonlinear law: elementwise polarisation, then r2c of the C-component field
      if (sl) {
        k_ls_pol_tau<T, SecantLaw<D>><<<nel, MGK_NT, 0, stream>>>(eps, epsp, lam, mu, T(lam0), T(mu0),
                                                                  SecantLaw<D>{*sl}, taubuf, NV, st);
      } else if constexpr (D == 3) {
        k_ls_pol_tau<T, LandauLaw><<<nel, MGK_NT, 0, stream>>>(eps, epsp, lam, mu, T(lam0), T(mu0), LandauLaw{*lp},
                                                               taubuf, NV, st);
      }
      launch_xfwd_real<T>(pl, g, gx, stream, taubuf, spec, st);
    } else {
    if (iter_idx == 0 && light0) h = prof.begin(KID_LS_TAU_START, iter_idx, b_tau - eps_in + eps_in0);
    else h = prof.begin(KID_LS_TAU, iter_idx, b_tau);
    launch_tau<T, D>(pl, g, gx, stream, bt);
    prof.end(h);
    }
    if (D == 3 && pl.P > 1) {
      h = prof.begin(KID_LS_MID, iter_idx, b_fft);
      launch_mid<T, MV>(pl, gy, stream, g.ply, g.twy, mp, 0);
      prof.end(h);
    }
    h = prof.begin(KID_LS_PENCIL, iter_idx, b_fft);
    launch_pencil<T, D>(pl, g, gp, stream, bt, scale);
    prof.end(h);
    if (D == 3 && pl.P > 1) {
      h = prof.begin(KID_LS_MID, iter_idx, b_fft);
      launch_mid<T, MV>(pl, gy, stream, g.ply, g.twy, mp, 1);
      prof.end(h);
    }
    if constexpr (sizeof(T) == 8) {
      if (l2 && l2->aa_m > 0)
        k_copy<<<nel, MGK_NT, 0, stream>>>(eps, l2->eps_old, (size_t)C * NV, st);
    }
    if (iter_idx == 0 && light0) h = prof.begin(KID_LS_UPDATE_START, iter_idx, b_upd - eps_in + eps_in0);
    else h = prof.begin(KID_LS_UPDATE, iter_idx, b_upd);
    launch_xinv<T, D>(pl, g, gx, stream, bt);
    prof.end(h);
    bt.m[0].eps_old32 = nullptr;  // eps now holds the first iterate
    bt.m[0].eps_uniform = 0;
    k_ls_finalize<<<1, MGK_NT, 0, stream>>>(fin, pl.nblk_upd, tol, max_iter);
    if constexpr (sizeof(T) == 8) {
      if (l2 && l2->aa_m > 0) {
        const size_t n3 = (size_t)C * NV;
        const int ns = l2->aa_m + 1;
        k_aa_store<<<nel, MGK_NT, 0, stream>>>(eps, l2->eps_old, l2->aa_g, l2->aa_f, n3, ns, st);
        k_aa_advance<<<1, 1, 0, stream>>>(ns, st);
        k_aa_gram<<<pl.nblk_red, MGK_NT, 0, stream>>>(l2->aa_f, NV, ns, part, st);
        k_aa_solve<<<1, 64, 0, stream>>>(part, pl.nblk_red, 4.0 * (double)NV, st);
        k_aa_mix<<<nel, MGK_NT, 0, stream>>>(l2->aa_g, n3, ns, eps, st);
        k_aa_finish<<<1, 1, 0, stream>>>(ns, st);
      }
    }
    ++iter_idx;
  };

  LsState* hst = na ? na->host_state : (LsState*)mgk_pinned(2 * sizeof(LsState));
  if (!hst) {
    mgk_set_error("pinned host allocation failed", __FILE__, __LINE__);
    return MGK_ERR_HIP;
  }
  hipEvent_t ev[2] = {nullptr, nullptr};
  if (!na) {
    MGK_HIP(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
    MGK_HIP(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
  }
  // Iterations are enqueued in chunks; the stop flag of chunk i is read while
  // chunk i+1 runs.  The first chunk is the iteration count of the previous
  // solve on this workspace (solves in a KMC loop converge alike), so at most
  // one chunk of early-exit launches trails the converged iteration.  The
  // no-allocation form keeps no per-workspace history and reads the flag
  // after a stream synchronisation per chunk instead of events.
  const int hint = na ? 0 : ls_iter_hint(ws, 0);
  int enq = 0, chunk = hint > 0 ? hint : 2, slot = 0;
  bool pending = false, stop = false;
  bool first = true;
  while (!stop && enq < max_iter) {
    const int k = std::min(chunk, max_iter - enq);
    for (int i = 0; i < k; ++i) enqueue_iter();
    enq += k;
    MGK_HIP(hipGetLastError());
    MGK_HIP(hipMemcpyAsync(&hst[slot], st, sizeof(LsState), hipMemcpyDeviceToHost, stream));
    if (na) {
      MGK_HIP(hipStreamSynchronize(stream));
      if (hst[slot].done) stop = true;
    } else {
      MGK_HIP(hipEventRecord(ev[slot], stream));
      if (pending) {
        MGK_HIP(hipEventSynchronize(ev[slot ^ 1]));
        if (hst[slot ^ 1].done) stop = true;
      }
      pending = true;
      slot ^= 1;
    }
    chunk = first ? 1 : std::min(chunk * 2, 16);
    first = false;
  }
  // final stress and macro means
  bool l2_done = false;
  if constexpr (sizeof(T) == 8 && D == 2) {
    if (final_stress && l2) {
      if (l2->plane_stress) k_l2_count<<<nel, MGK_NT, 0, stream>>>(eps, epsp, lam, mu, l2->P, l2->e33, NV, st, 1);
      k_l2_stress<<<pl.nblk_red, MGK_NT, 0, stream>>>(eps, epsp, lam, mu, lam0, mu0, l2->P, l2->plane_stress,
                                                      l2->e33, nullptr, sig, part, NV, st, 1);
      k_l2_after_eval<<<1, 1, 0, stream>>>(st, l2->persist);
      k_means_finalize<<<1, MGK_NT, 0, stream>>>(part, pl.nblk_red, 2 * C, (double)NV, st->mean);
      MGK_HIP(hipGetLastError());
      l2_done = true;
    }
  }
  if (l2_done) {
  } else if (final_stress && (lp || sl)) {
    if (sl) {
      k_ls_pol_stress<T, SecantLaw<D>><<<pl.nblk_red, MGK_NT, 0, stream>>>(eps, epsp, lam, mu, SecantLaw<D>{*sl},
                                                                           sig, NV, part);
    } else if constexpr (D == 3) {
      k_ls_pol_stress<T, LandauLaw><<<pl.nblk_red, MGK_NT, 0, stream>>>(eps, epsp, lam, mu, LandauLaw{*lp}, sig, NV,
                                                                        part);
    }
    k_means_finalize<<<1, MGK_NT, 0, stream>>>(part, pl.nblk_red, 2 * C, (double)NV, st->mean);
    MGK_HIP(hipGetLastError());
  } else if (final_stress) {
    k_ls_stress<T, D><<<pl.nblk_red, MGK_NT, 0, stream>>>(eps, epsp, lam, mu, sig, NV, part);
    k_means_finalize<<<1, MGK_NT, 0, stream>>>(part, pl.nblk_red, 2 * C, (double)NV, st->mean);
    MGK_HIP(hipGetLastError());
  }
  MGK_HIP(hipMemcpyAsync(&hst[0], st, sizeof(LsState), hipMemcpyDeviceToHost, stream));
  MGK_HIP(hipStreamSynchronize(stream));
  prof.commit(hst[0].iters);
  if (!na) {
    ls_iter_hint(ws, hst[0].iters);
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
  }
  const LsState& s = hst[0];
  if (info) ls_fill_info<D>(s, eps_bar, info);
  return s.status;
}


// ---------------------------------------------------------------------------
// Batched linear solve: nb (<= MGK_MAXB) simulations of one grid shape and
// precision, each with its own fields, reference medium, macro strain and
// workspace, iterated together with one launch per sweep.  Every member runs
// the kernel sequence of a solve alone (same kernels, same partial-sum
// order) and stops on its own state, so its fields, iteration count and
// status are bit-identical to mgk_ls_solve on the same inputs; the batch
// polls until every member has stopped.
// ---------------------------------------------------------------------------
template <typename T>
struct LsMemArgs {
  const T* lam;
  const T* mu;
  double lam0, mu0;
  const double* eps_bar;
  const T* epsp;
  T* eps;
  T* sig;
  char* ws;
  const float* eps_from32;  // FP32 starting field of an FP64 stage (warm = 2), or NULL
};

// pinned host copies of the members' solve states, per thread (two chunks
// in flight); allocated once and kept until the process ends
LsState* batch_host_states() {
  thread_local LsState* p = nullptr;
  if (!p && hipHostMalloc((void**)&p, 2 * MGK_MAXB * sizeof(LsState), hipHostMallocDefault) != hipSuccess)
    p = nullptr;
  return p;
}

template <typename T, int D>
int ls_solve_batch_t(const LsPlan& pl, int nb, const LsMemArgs<T>* a, int warm, int max_iter, double tol,
                     hipStream_t stream, mgk_ls_info* info, bool final_stress) {
  constexpr int MV = maxv<T>();
  constexpr int C = (D == 3) ? 6 : 3;
  const int prec = sizeof(T) == 8 ? 0 : 1;
  LsGeom<T> g = ls_geom<T>(pl, C);
  g.twx = (const cplx<T>*)mgk_twiddles(pl.nx, prec);
  g.twy = (const cplx<T>*)mgk_twiddles(pl.P, prec);
  g.twq = (const cplx<T>*)mgk_twiddles(pl.Q, prec);
  if (!g.twx || !g.twy || !g.twq) {
    mgk_set_error("twiddle table allocation failed", __FILE__, __LINE__);
    return MGK_ERR_HIP;
  }
  // device staging: the members' states side by side, two polling slots + the final read
  LsState* dst = (LsState*)mgk_device_stage(0, 3 * MGK_MAXB * sizeof(LsState));
  LsState* hst = batch_host_states();
  if (!dst || !hst) {
    mgk_set_error("batch staging allocation failed", __FILE__, __LINE__);
    return MGK_ERR_HIP;
  }
  const size_t NV = (size_t)pl.NV;
  const int nfill = (int)std::min<long long>((pl.NV + MGK_NT - 1) / MGK_NT, 4096);
  LsBatch<T> bt{};
  int hint = 0;
  for (int b = 0; b < nb; ++b) {
    double anti2 = 0.0;
    const Sym6 eb = ls_sym_bar<D>(a[b].eps_bar, &anti2);
    bt.m[b] = LsMem<T>{a[b].lam, a[b].mu, a[b].epsp, a[b].eps, (cplx<T>*)(a[b].ws + pl.off_spec),
                       (double*)(a[b].ws + pl.off_part), (LsState*)(a[b].ws + pl.off_state), eb, a[b].lam0,
                       a[b].mu0, anti2 * (double)pl.NV, a[b].sig, a[b].eps_from32, warm == 0 ? 1 : 0};
    hint = std::max(hint, ls_iter_hint(a[b].ws, 0));
  }
  // set-up: state cleared, eps shifted to mean eps_bar (warm 1); a cold start
  // reads eps = eps_bar as a constant in the first iteration (eps_uniform)
  k_ls_init_b<T><<<dim3(1, nb), 64, 0, stream>>>(bt);
  if (warm == 1) {
    k_comp_sums_b<T><<<dim3(pl.nblk_red, nb), MGK_NT, 0, stream>>>(bt, C, NV);
    k_means_finalize_b<T><<<nb, MGK_NT, 0, stream>>>(bt, pl.nblk_red, C, (double)NV);
    k_ls_warm_adjust_b<T><<<dim3(nfill, nb), MGK_NT, 0, stream>>>(bt, C, NV);
  }
  MGK_HIP(hipGetLastError());

  const T scale = T(1.0 / (double)pl.NV);
  const int nqx = (pl.Q + (1 << pl.logBx) - 1) >> pl.logBx;
  const int nqy = (pl.Q + (1 << pl.logBy) - 1) >> pl.logBy;
  const int npb = (g.npencil + (1 << pl.logPP) - 1) >> pl.logPP;
  const dim3 gx(nqx, pl.P, nb), gy(nqy, pl.nxh, C * nb), gp(npb, nb);
  const MidPtrs<T> mp = mid_table(bt, nb, C);
  const LsFin fin = fin_table(bt, nb);

  const double sT = (double)sizeof(T);
  const double specB = (double)C * pl.nxh * pl.P * pl.Q * 2.0 * sT;
  const double fieldB = (double)pl.NV * sT;
  const double b_tau = nb * (fieldB * ((a[0].epsp ? 2.0 * C : (double)C) + 2.0) + specB);
  const double b_fft = nb * 2.0 * specB;
  const double b_upd = nb * (specB + 2.0 * C * fieldB);
  // first iteration: starting fields read as constants (cold) or in FP32
  double eps_in0 = 0.0;
  for (int b = 0; b < nb; ++b)
    eps_in0 += warm == 0 ? 0.0 : (a[b].eps_from32 ? (double)C * pl.NV * 4.0 : (double)C * fieldB);
  const double eps_in = nb * (double)C * fieldB;
  const bool light0 = eps_in0 < eps_in;  // timed under the _START ids
  MgkProfRec prof(stream);
  int iter_idx = 0;
  auto enqueue_iter = [&]() {
    int h = (iter_idx == 0 && light0) ? prof.begin(KID_LS_TAU_START, iter_idx, b_tau - eps_in + eps_in0)
                                      : prof.begin(KID_LS_TAU, iter_idx, b_tau);
    launch_tau<T, D>(pl, g, gx, stream, bt);
    prof.end(h);
    if (D == 3 && pl.P > 1) {
      h = prof.begin(KID_LS_MID, iter_idx, b_fft);
      launch_mid<T, MV>(pl, gy, stream, g.ply, g.twy, mp, 0);
      prof.end(h);
    }
    h = prof.begin(KID_LS_PENCIL, iter_idx, b_fft);
    launch_pencil<T, D>(pl, g, gp, stream, bt, scale);
    prof.end(h);
    if (D == 3 && pl.P > 1) {
      h = prof.begin(KID_LS_MID, iter_idx, b_fft);
      launch_mid<T, MV>(pl, gy, stream, g.ply, g.twy, mp, 1);
      prof.end(h);
    }
    h = (iter_idx == 0 && light0) ? prof.begin(KID_LS_UPDATE_START, iter_idx, b_upd - eps_in + eps_in0)
                                  : prof.begin(KID_LS_UPDATE, iter_idx, b_upd);
    launch_xinv<T, D>(pl, g, gx, stream, bt);
    prof.end(h);
    for (int b = 0; b < nb; ++b) {  // eps now holds the first iterate
      bt.m[b].eps_old32 = nullptr;
      bt.m[b].eps_uniform = 0;
    }
    k_ls_finalize<<<nb, MGK_NT, 0, stream>>>(fin, pl.nblk_upd, tol, max_iter);
    ++iter_idx;
  };

  hipEvent_t ev[2] = {nullptr, nullptr};
  MGK_HIP(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
  MGK_HIP(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
  // chunked enqueue as in the single solve; each chunk ends with one read of
  // every member's state, and the batch stops once all of them read done
  int enq = 0, chunk = hint > 0 ? hint : 2, slot = 0;
  bool pending = false, stop = false, first = true;
  while (!stop && enq < max_iter) {
    const int k = std::min(chunk, max_iter - enq);
    for (int i = 0; i < k; ++i) enqueue_iter();
    enq += k;
    k_ls_gather_states<<<nb, 64, 0, stream>>>(fin, dst + slot * MGK_MAXB);
    MGK_HIP(hipGetLastError());
    MGK_HIP(hipMemcpyAsync(&hst[slot * MGK_MAXB], dst + slot * MGK_MAXB, nb * sizeof(LsState),
                           hipMemcpyDeviceToHost, stream));
    MGK_HIP(hipEventRecord(ev[slot], stream));
    if (pending) {
      MGK_HIP(hipEventSynchronize(ev[slot ^ 1]));
      bool all = true;
      for (int b = 0; b < nb; ++b) all = all && hst[(slot ^ 1) * MGK_MAXB + b].done;
      if (all) stop = true;
    }
    pending = true;
    slot ^= 1;
    chunk = first ? 1 : std::min(chunk * 2, 16);
    first = false;
  }
  if (final_stress) {
    k_ls_stress_b<T, D><<<dim3(pl.nblk_red, nb), MGK_NT, 0, stream>>>(bt, NV);
    k_means_finalize_b<T><<<nb, MGK_NT, 0, stream>>>(bt, pl.nblk_red, 2 * C, (double)NV);
  }
  k_ls_gather_states<<<nb, 64, 0, stream>>>(fin, dst + 2 * MGK_MAXB);
  MGK_HIP(hipGetLastError());
  MGK_HIP(hipMemcpyAsync(hst, dst + 2 * MGK_MAXB, nb * sizeof(LsState), hipMemcpyDeviceToHost, stream));
  MGK_HIP(hipStreamSynchronize(stream));
  (void)hipEventDestroy(ev[0]);
  (void)hipEventDestroy(ev[1]);
  int iters_max = 0, rc = MGK_OK;
  for (int b = 0; b < nb; ++b) {
    const LsState& sb = hst[b];
    ls_iter_hint(a[b].ws, sb.iters);
    iters_max = std::max(iters_max, sb.iters);
    if (info) ls_fill_info<D>(sb, a[b].eps_bar, &info[b]);
    if (sb.status < 0 && rc >= 0) rc = sb.status;
    else if (sb.status > 0 && rc == MGK_OK) rc = sb.status;
  }
  prof.commit(iters_max);
  return rc;
}

// ---------------------------------------------------------------------------
// 2D Landau small strain, Newton-CG (the reference's solver="newton_cg",
// mgkmc/linear_elastic_simulator.py:1011-1117).  Newton steps on eps with
// BiCGSTAB (SciPy's recurrence, rtol = tol, maxiter 150) on A = G C_h G,
// G = the reference medium's -Gamma0 and C_h d = (sigma(eps_el + h d) -
// sigma(eps_el - h d)) / 2h with h = 1e-7 (central differences of the
// Landau stress, each evaluation with its own cold plane-stress e33 Newton).
// The Krylov vectors hold the reference's full 2x2 field per voxel (xx, yy,
// xy, yx), so their dot products are those of the reference's flattened
// vectors; G reads the symmetric part and writes xy into both slots.
// ---------------------------------------------------------------------------
constexpr double kNcgH = 1e-7;

__global__ void k_ncg_sym(size_t NV, const double* __restrict__ p4, double* __restrict__ q3,
                          const int* __restrict__ skip) {
  if (skip && *skip) return;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < NV; i += (size_t)gridDim.x * blockDim.x) {
    q3[i] = p4[i];
    q3[NV + i] = p4[NV + i];
    q3[2 * NV + i] = 0.5 * (p4[2 * NV + i] + p4[3 * NV + i]);
  }
}

__global__ void k_ncg_expand(size_t NV, const double* __restrict__ a3, double sgn, double* __restrict__ p4,
                             const int* __restrict__ skip) {
  if (skip && *skip) return;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < NV; i += (size_t)gridDim.x * blockDim.x) {
    const double xy = sgn * a3[2 * NV + i];
    p4[i] = sgn * a3[i];
    p4[NV + i] = sgn * a3[NV + i];
    p4[2 * NV + i] = xy;
    p4[3 * NV + i] = xy;
  }
}

// (eps - eps_p) + h d  (sign > 0) or (eps - eps_p) - h d, as the reference forms them
__global__ void k_ncg_pert(size_t n3, const double* __restrict__ eps, const double* __restrict__ epsp,
                           const double* __restrict__ d, double h, int sign, double* __restrict__ out,
                           const int* __restrict__ skip) {
  if (skip && *skip) return;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n3; i += (size_t)gridDim.x * blockDim.x) {
    const double el = eps[i] - (epsp ? epsp[i] : 0.0);
    const double hd = h * d[i];
    out[i] = sign > 0 ? el + hd : el - hd;
  }
}

__global__ void k_ncg_fd(size_t n3, const double* __restrict__ sp, const double* __restrict__ sm, double h2,
                         double* __restrict__ out, const int* __restrict__ skip) {
  if (skip && *skip) return;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n3; i += (size_t)gridDim.x * blockDim.x)
    out[i] = (sp[i] - sm[i]) / h2;
}

// constant field of the three symmetric components of a 2x2 matrix
__global__ void k_ncg_const(size_t NV, Sym6 v, double* __restrict__ out) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < NV; i += (size_t)gridDim.x * blockDim.x) {
    out[i] = v.v[0];
    out[NV + i] = v.v[1];
    out[2 * NV + i] = v.v[2];
  }
}

// eps <- (eps + dbar) + g  (first Newton step) or eps + g
__global__ void k_ncg_update(size_t NV, double* __restrict__ eps, const double* __restrict__ g, int with_dbar,
                             Sym6 dbar) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < NV; i += (size_t)gridDim.x * blockDim.x) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      double e = eps[(size_t)c * NV + i];
      if (with_dbar) e = e + dbar.v[c];
      eps[(size_t)c * NV + i] = e + g[(size_t)c * NV + i];
    }
  }
}

// per-block (a, b) and (c, d) dot products, stride 2 (the Krylov stop kernels' layout)
__global__ __launch_bounds__(MGK_NT) void k_ncg_dot2(size_t n, const double* __restrict__ a,
                                                     const double* __restrict__ b, const double* __restrict__ c,
                                                     const double* __restrict__ d, double* __restrict__ partials,
                                                     const int* __restrict__ skip) {
  if (skip && *skip) return;
  __shared__ double red[2 * (MGK_NT / 64)];
  double acc[2] = {0.0, 0.0};
  for (size_t i = blockIdx.x * (size_t)MGK_NT + threadIdx.x; i < n; i += (size_t)gridDim.x * MGK_NT) {
    acc[0] += a[i] * b[i];
    if (c) acc[1] += c[i] * d[i];
  }
  block_sum<2>(acc, red);
  if (threadIdx.x == 0) {
    partials[2 * blockIdx.x] = acc[0];
    partials[2 * blockIdx.x + 1] = acc[1];
  }
}

// sum over voxels of the full 2x2 squared norm (xy counted twice) of a 3-component field
__global__ __launch_bounds__(MGK_NT) void k_ncg_norm2w(size_t NV, const double* __restrict__ f,
                                                       double* __restrict__ partials) {
  __shared__ double red[MGK_NT / 64];
  double acc[1] = {0.0};
  for (size_t i = blockIdx.x * (size_t)MGK_NT + threadIdx.x; i < NV; i += (size_t)gridDim.x * MGK_NT) {
    const double a = f[i], b = f[NV + i], c = f[2 * NV + i];
    acc[0] += (a * a + b * b) + 2.0 * (c * c);
  }
  block_sum<1>(acc, red);
  if (threadIdx.x == 0) partials[blockIdx.x] = acc[0];
}

// the law and Gamma kernels of the Krylov matvec skip once the solve is done
__global__ void k_ncg_gate(const int* __restrict__ flag, LsState* __restrict__ st) { st->done = flag ? *flag : 0; }

struct NcgCtx {
  LsPlan pl;
  LsGeom<double> g;
  hipStream_t s;
  size_t NV;
  const double *lam, *mu, *epsp;
  double lam0, mu0;
  LandauPar P;
  int plane_stress;
  LsBatch<double> bt;  // member 0: the Gamma application's spectrum, partials, state
  LsState* lst;        // Gamma and law state (done gate, e33 count, status)
  double* e33;
  double *q3, *dc3, *pp3, *sp3, *sm3, *dt3, *o3;
  dim3 gx, gp;
  int nel;
};

// out3 = -Gamma0 : in3 (r2c along x, pencil along y with the Hermitian Nyquist part, c2r)
void ncg_gamma(NcgCtx& c, const double* in3, double* out3) {
  launch_xfwd_real<double>(c.pl, c.g, c.gx, c.s, in3, c.bt.m[0].spec, c.lst);
  launch_pencil<double, 2>(c.pl, c.g, c.gp, c.s, c.bt, 1.0 / (double)c.pl.NV);
  LsBatch<double> b = c.bt;
  b.m[0].eps = out3;
  launch_xinv<double, 2>(c.pl, c.g, c.gx, c.s, b);
}

// sig3 = sigma_Landau(strain3) with a cold plane-stress e33 Newton
void ncg_sigma(NcgCtx& c, const double* strain3, double* sig3) {
  if (c.plane_stress)
    k_l2_count<<<c.nel, MGK_NT, 0, c.s>>>(strain3, nullptr, c.lam, c.mu, c.P, c.e33, c.NV, c.lst, 0);
  k_l2_stress<<<c.nel, MGK_NT, 0, c.s>>>(strain3, nullptr, c.lam, c.mu, c.lam0, c.mu0, c.P, c.plane_stress, c.e33,
                                         nullptr, sig3, nullptr, c.NV, c.lst, 0);
  k_l2_after_eval<<<1, 1, 0, c.s>>>(c.lst, 0);
}

// out3 = C_h d3 at the current eps
void ncg_cfd(NcgCtx& c, const double* eps, const double* d3, double* out3) {
  const size_t n3 = 3 * c.NV;
  const int* sk = &c.lst->done;
  k_ncg_pert<<<c.nel, MGK_NT, 0, c.s>>>(n3, eps, c.epsp, d3, kNcgH, 1, c.pp3, sk);
  ncg_sigma(c, c.pp3, c.sp3);
  k_ncg_pert<<<c.nel, MGK_NT, 0, c.s>>>(n3, eps, c.epsp, d3, kNcgH, -1, c.pp3, sk);
  ncg_sigma(c, c.pp3, c.sm3);
  k_ncg_fd<<<c.nel, MGK_NT, 0, c.s>>>(n3, c.sp3, c.sm3, 2.0 * kNcgH, out3, sk);
}

// out4 = G C_h G p4
void ncg_matvec(NcgCtx& c, const double* eps, const double* p4, double* out4) {
  const int* sk = &c.lst->done;
  k_ncg_sym<<<c.nel, MGK_NT, 0, c.s>>>(c.NV, p4, c.q3, sk);
  ncg_gamma(c, c.q3, c.dc3);
  ncg_cfd(c, eps, c.dc3, c.dt3);
  ncg_gamma(c, c.dt3, c.o3);
  k_ncg_expand<<<c.nel, MGK_NT, 0, c.s>>>(c.NV, c.o3, 1.0, out4, sk);
}

int ncg_read_sum(double* part, int n, hipStream_t s, double* out) {
  double* h = (double*)mgk_pinned((size_t)n * sizeof(double));
  if (!h) return MGK_ERR_HIP;
  MGK_HIP(hipMemcpyAsync(h, part, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, s));
  MGK_HIP(hipStreamSynchronize(s));
  double t = 0.0;
  for (int i = 0; i < n; ++i) t += h[i];
  *out = t;
  return MGK_OK;
}

size_t ncg_extra_bytes(const LsPlan& pl) {
  const size_t f3 = mgk_align((size_t)3 * pl.NV * sizeof(double)), f4 = mgk_align((size_t)4 * pl.NV * sizeof(double));
  return 7 * f4 + 9 * f3 + mgk_align((size_t)pl.NV * sizeof(double)) + mgk_align((size_t)KRY_NBLK * 2 * sizeof(double)) +
         mgk_align(sizeof(KryState)) + mgk_align(sizeof(LsState));
}

int l2_newton_cg(const LsPlan& pl, const double* lam, const double* mu, double lam0, double mu0, const LandauPar& P,
                 int plane_stress, const double* eps_bar, const double* epsp, double* eps, double* sig, int warm,
                 int max_iter, double tol, char* ws, hipStream_t s, mgk_ls_info* info, int* krylov_total) {
  NcgCtx c;
  c.pl = pl;
  c.g = ls_geom<double>(pl, 3);
  c.g.twx = (const cplx<double>*)mgk_twiddles(pl.nx, 0);
  c.g.twy = (const cplx<double>*)mgk_twiddles(pl.P, 0);
  c.g.twq = (const cplx<double>*)mgk_twiddles(pl.Q, 0);
  if (!c.g.twx || !c.g.twy || !c.g.twq) {
    mgk_set_error("twiddle table allocation failed", __FILE__, __LINE__);
    return MGK_ERR_HIP;
  }
  c.s = s;
  c.NV = (size_t)pl.NV;
  c.lam = lam;
  c.mu = mu;
  c.epsp = epsp;
  c.lam0 = lam0;
  c.mu0 = mu0;
  c.P = P;
  c.plane_stress = plane_stress;
  const size_t NV = c.NV, n3 = 3 * NV, n4 = 4 * NV;
  cplx<double>* spec = (cplx<double>*)(ws + pl.off_spec);
  double* part = (double*)(ws + pl.off_part);
  LsState* gst_unused = (LsState*)(ws + pl.off_state);
  (void)gst_unused;
  char* w = ws + pl.total;
  const size_t f3 = mgk_align(n3 * sizeof(double)), f4 = mgk_align(n4 * sizeof(double));
  double* kv[7];
  for (int i = 0; i < 7; ++i) {
    kv[i] = (double*)w;
    w += f4;
  }
  double *kb = kv[0], *kx = kv[1], *kr = kv[2], *krt = kv[3], *kp = kv[4], *kvv = kv[5], *kt = kv[6];
  double* f[9];
  for (int i = 0; i < 9; ++i) {
    f[i] = (double*)w;
    w += f3;
  }
  c.q3 = f[0];
  c.dc3 = f[1];
  c.pp3 = f[2];
  c.sp3 = f[3];
  c.sm3 = f[4];
  c.dt3 = f[5];
  c.o3 = f[6];
  double* sigc = sig ? sig : f[7];  // the current stress
  double* dbar3 = f[8];
  c.e33 = (double*)w;
  w += mgk_align(NV * sizeof(double));
  double* kpart = (double*)w;
  w += mgk_align((size_t)KRY_NBLK * 2 * sizeof(double));
  KryState* kst = (KryState*)w;
  w += mgk_align(sizeof(KryState));
  c.lst = (LsState*)w;
  c.bt = LsBatch<double>{};
  c.bt.m[0] = LsMem<double>{lam, mu, nullptr, nullptr, spec, part, c.lst, Sym6{}, lam0, mu0, 0.0};
  const int nqx = (pl.Q + (1 << pl.logBx) - 1) >> pl.logBx;
  const int npb = (c.g.npencil + (1 << pl.logPP) - 1) >> pl.logPP;
  c.gx = dim3(nqx, pl.P);
  c.gp = dim3(npb);
  c.nel = (int)std::min<long long>((pl.NV + MGK_NT - 1) / MGK_NT, 8192);
  const int nb = (int)std::min<size_t>((n4 + MGK_NT - 1) / MGK_NT, KRY_NBLK);

  // fields start defined: the Gamma output buffers are read by the update sweep
  // before it overwrites them
  k_zero_words<<<1, 64, 0, s>>>((unsigned int*)c.lst, sizeof(LsState) / 4);
  for (int i = 0; i < 9; ++i) k_ncg_const<<<c.nel, MGK_NT, 0, s>>>(NV, Sym6{}, f[i]);
  double anti2 = 0.0;
  const Sym6 eb = ls_sym_bar<2>(eps_bar, &anti2);
  if (!warm) {
    k_ls_fill<double><<<c.nel, MGK_NT, 0, s>>>(eps, 3, NV, eb);
  } else {
    k_comp_sums<double><<<pl.nblk_red, MGK_NT, 0, s>>>(eps, 3, NV, part);
    k_means_finalize<<<1, MGK_NT, 0, s>>>(part, pl.nblk_red, 3, (double)NV, c.lst->mean);
    k_ls_warm_adjust<double><<<c.nel, MGK_NT, 0, s>>>(eps, 3, NV, eb, c.lst->mean);
  }
  // Dbar = eps_bar - <eps>
  k_comp_sums<double><<<pl.nblk_red, MGK_NT, 0, s>>>(eps, 3, NV, part);
  k_means_finalize<<<1, MGK_NT, 0, s>>>(part, pl.nblk_red, 3, (double)NV, c.lst->mean);
  LsState* hst = (LsState*)mgk_pinned(sizeof(LsState));
  if (!hst) return MGK_ERR_HIP;
  MGK_HIP(hipMemcpyAsync(hst, c.lst, sizeof(LsState), hipMemcpyDeviceToHost, s));
  MGK_HIP(hipStreamSynchronize(s));
  Sym6 dbar{};
  for (int k = 0; k < 3; ++k) dbar.v[k] = eb.v[k] - hst->mean[k];
  k_ncg_const<<<c.nel, MGK_NT, 0, s>>>(NV, dbar, dbar3);
  // sigma of the start field
  k_ncg_pert<<<c.nel, MGK_NT, 0, s>>>(n3, eps, epsp, dbar3, 0.0, 1, c.pp3, nullptr);
  ncg_sigma(c, c.pp3, sigc);
  MGK_HIP(hipGetLastError());

  int newton = 0, kry_total = 0, status = MGK_NOT_CONVERGED;
  double res = 0.0;
  for (int it = 0; it < max_iter; ++it) {
    // right-hand side: -G C_h Dbar (first step) or -G sigma
    k_ncg_gate<<<1, 1, 0, s>>>(nullptr, c.lst);
    if (it == 0) {
      ncg_cfd(c, eps, dbar3, c.dt3);
      ncg_gamma(c, c.dt3, c.o3);
    } else {
      ncg_gamma(c, sigc, c.o3);
    }
    k_ncg_expand<<<c.nel, MGK_NT, 0, s>>>(NV, c.o3, -1.0, kb, nullptr);
    // BiCGSTAB(A, b, rtol = tol, maxiter = 150) from x = 0
    KryState init{};
    init.rtol = tol;
    init.maxiter = 150;
    init.rhotol = 4.930380657631324e-32;  // eps(double)^2, as SciPy for float64
    init.omtol = init.rhotol;
    MGK_HIP(hipMemcpyAsync(kst, &init, sizeof(KryState), hipMemcpyHostToDevice, s));
    k_kry_init<double><<<nb, MGK_NT, 0, s>>>(n4, kb, kx, kr, krt, kpart);
    k_kry_fin_init<<<1, MGK_NT, 0, s>>>(kpart, nb, kst);
    auto enqueue = [&]() {
      k_kry_p<double><<<nb, MGK_NT, 0, s>>>(n4, kst, kp, kr, kvv);
      k_ncg_gate<<<1, 1, 0, s>>>(&kst->done, c.lst);
      ncg_matvec(c, eps, kp, kvv);
      k_ncg_dot2<<<nb, MGK_NT, 0, s>>>(n4, krt, kvv, nullptr, nullptr, kpart, &kst->done);
      k_kry_fin_rv<<<1, MGK_NT, 0, s>>>(kpart, nb, kst);
      k_kry_s<double><<<nb, MGK_NT, 0, s>>>(n4, kst, kr, kvv, kpart);
      k_kry_fin_s<<<1, MGK_NT, 0, s>>>(kpart, nb, kst);
      k_ncg_gate<<<1, 1, 0, s>>>(&kst->done2, c.lst);
      ncg_matvec(c, eps, kr, kt);
      k_ncg_dot2<<<nb, MGK_NT, 0, s>>>(n4, kt, kr, kt, kt, kpart, &kst->done2);
      k_kry_fin_t<<<1, MGK_NT, 0, s>>>(kpart, nb, kst);
      k_kry_xr<double><<<nb, MGK_NT, 0, s>>>(n4, kst, kx, kp, kr, kt, krt, kpart, (const double*)nullptr);
      k_kry_fin_xr<<<1, MGK_NT, 0, s>>>(kpart, nb, kst);
    };
    int kit = 0, kstat = 0, enq = 0;
    int rc = kry_drive(kst, s, 150, enqueue, &kit, &kstat, &enq);
    if (rc != MGK_OK) return rc;
    kry_total += kit;
    // eps update: (eps + Dbar) + G dx (first step) or eps + G dx
    k_ncg_gate<<<1, 1, 0, s>>>(nullptr, c.lst);
    k_ncg_sym<<<c.nel, MGK_NT, 0, s>>>(NV, kx, c.q3, nullptr);
    ncg_gamma(c, c.q3, c.o3);
    k_ncg_update<<<c.nel, MGK_NT, 0, s>>>(NV, eps, c.o3, it == 0 ? 1 : 0, dbar);
    k_ncg_pert<<<c.nel, MGK_NT, 0, s>>>(n3, eps, epsp, dbar3, 0.0, 1, c.pp3, nullptr);
    ncg_sigma(c, c.pp3, sigc);
    // |dx| / (|eps| + 1e-20) over the full 2x2 fields
    k_norm2<double><<<nb, MGK_NT, 0, s>>>(n4, kx, kpart);
    k_ncg_norm2w<<<nb, MGK_NT, 0, s>>>(NV, eps, kpart + KRY_NBLK);
    MGK_HIP(hipGetLastError());
    double nx2 = 0.0, ne2 = 0.0;
    rc = ncg_read_sum(kpart, nb, s, &nx2);
    if (rc == MGK_OK) rc = ncg_read_sum(kpart + KRY_NBLK, nb, s, &ne2);
    if (rc != MGK_OK) return rc;
    MGK_HIP(hipMemcpyAsync(hst, c.lst, sizeof(LsState), hipMemcpyDeviceToHost, s));
    MGK_HIP(hipStreamSynchronize(s));
    newton = it + 1;
    res = std::sqrt(nx2) / (std::sqrt(ne2) + 1e-20);
    if (hst->status == MGK_ERR_NONFINITE || !std::isfinite(res)) {
      status = MGK_ERR_NONFINITE;
      break;
    }
    if (res < tol && it > 0) {
      status = MGK_OK;
      break;
    }
  }
  // macro means of eps and sigma
  k_comp_sums<double><<<pl.nblk_red, MGK_NT, 0, s>>>(eps, 3, NV, part);
  k_means_finalize<<<1, MGK_NT, 0, s>>>(part, pl.nblk_red, 3, (double)NV, c.lst->mean);
  k_comp_sums<double><<<pl.nblk_red, MGK_NT, 0, s>>>(sigc, 3, NV, part);
  k_means_finalize<<<1, MGK_NT, 0, s>>>(part, pl.nblk_red, 3, (double)NV, c.lst->mean + 3);
  MGK_HIP(hipMemcpyAsync(hst, c.lst, sizeof(LsState), hipMemcpyDeviceToHost, s));
  MGK_HIP(hipStreamSynchronize(s));
  LsState out = *hst;
  out.iters = newton;
  out.status = status;
  out.diff = res;
  if (info) ls_fill_info<2>(out, eps_bar, info);
  if (krylov_total) *krylov_total = kry_total;
  return status;
}
}  // namespace

extern "C" int mgk_ls_workspace_bytes(int ndim, const int64_t* shape, int precision, size_t* bytes) {
  LsPlan pl;
  int rc = precision == 0 ? ls_plan<double>(ndim, shape, &pl) : ls_plan<float>(ndim, shape, &pl);
  if (rc != MGK_OK) {
    mgk_set_error("unsupported grid shape for the LS solver", __FILE__, __LINE__);
    return rc;
  }
  *bytes = pl.total;
  return MGK_OK;
}

extern "C" int mgk_ls_solve(int ndim, const int64_t* shape, int precision, const void* lam, const void* mu,
                            double lam0, double mu0, const double* eps_bar, const void* eps_plastic, void* eps,
                            void* sig, int warm, int max_iter, double tol, void* workspace, size_t workspace_bytes,
                            void* stream, mgk_ls_info* info) {
  if (!lam || !mu || !eps || !eps_bar || !workspace || max_iter < 1) {
    mgk_set_error("null argument or max_iter < 1", __FILE__, __LINE__);
    return MGK_ERR_ARG;
  }
  LsPlan pl;
  int rc = precision == 0 ? ls_plan<double>(ndim, shape, &pl) : ls_plan<float>(ndim, shape, &pl);
  if (rc != MGK_OK) {
    mgk_set_error("unsupported grid shape for the LS solver", __FILE__, __LINE__);
    return rc;
  }
  if (workspace_bytes < pl.total) {
    mgk_set_error("workspace too small", __FILE__, __LINE__);
    return MGK_ERR_WORKSPACE;
  }
  hipStream_t s = (hipStream_t)stream;
  char* ws = (char*)workspace;
  warm = warm ? 1 : 0;
  if (precision == 0) {
    if (ndim == 3)
      return ls_solve_t<double, 3>(pl, (const double*)lam, (const double*)mu, lam0, mu0, eps_bar,
                                   (const double*)eps_plastic, (double*)eps, (double*)sig, warm, max_iter, tol, ws, s,
                                   info);
    return ls_solve_t<double, 2>(pl, (const double*)lam, (const double*)mu, lam0, mu0, eps_bar,
                                 (const double*)eps_plastic, (double*)eps, (double*)sig, warm, max_iter, tol, ws, s,
                                 info);
  }
  if (ndim == 3)
    return ls_solve_t<float, 3>(pl, (const float*)lam, (const float*)mu, lam0, mu0, eps_bar,
                                (const float*)eps_plastic, (float*)eps, (float*)sig, warm, max_iter, tol, ws, s, info);
  return ls_solve_t<float, 2>(pl, (const float*)lam, (const float*)mu, lam0, mu0, eps_bar, (const float*)eps_plastic,
                              (float*)eps, (float*)sig, warm, max_iter, tol, ws, s, info);
}

// ---------------------------------------------------------------------------
// No-allocation form of mgk_ls_solve.  The workspace and the FFT twiddle
// tables live in `scratch`, a file-scope device array of a fixed size that
// each GPU holds for the life of the process.  Its device address is looked
// up once per device (hipGetSymbolAddress with that device current) and only
// that address is handed to kernels and copies: the array's name is a
// host-side stand-in, not a device pointer.  Every call on one GPU shares its
// array, so calls that overlap on one GPU (two streams, two threads) race;
// the solve state is cleared by a kernel at the start of every call.
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[MGK_LS_NOALLOC_BYTES];

namespace {
constexpr int kMaxDevices = 64;
constexpr int kMaxAxis = 4096;  // longest axis the compile-time and generic FFT passes take
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices];

// device address of this device's scratch array (looked up on first use)
int scratch_address(void** out) {
  int dev = 0;
  MGK_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) {
    mgk_set_error("device index out of range for the no-allocation scratch", __FILE__, __LINE__);
    return MGK_ERR_ARG;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    MGK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) {
      mgk_set_error("hipGetSymbolAddress returned no address for the scratch array", __FILE__, __LINE__);
      return MGK_ERR_HIP;
    }
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return MGK_OK;
}

// per-thread host staging: three twiddle tables (doubles cover both
// precisions) and the two solve-state reads
thread_local double g_na_tw[3][2 * kMaxAxis];
thread_local LsState g_na_state[2];
}  // namespace

extern "C" int mgk_ls_solve_noalloc(int ndim, const int64_t* shape, int precision, const void* lam,
                                    const void* mu, double lam0, double mu0, const double* eps_bar,
                                    const void* eps_plastic, void* eps, void* sig, int warm, int max_iter,
                                    double tol, void* stream, mgk_ls_info* info) {
  if (!lam || !mu || !eps || !eps_bar || !shape || max_iter < 1 || (precision != 0 && precision != 1)) {
    mgk_set_error("null argument, max_iter < 1 or precision not 0/1", __FILE__, __LINE__);
    return MGK_ERR_ARG;
  }
  LsPlan pl;
  int rc = precision == 0 ? ls_plan<double>(ndim, shape, &pl) : ls_plan<float>(ndim, shape, &pl);
  if (rc != MGK_OK) {
    mgk_set_error("unsupported grid shape for the LS solver", __FILE__, __LINE__);
    return rc;
  }
  const int axis[3] = {pl.nx, pl.P, pl.Q};
  const size_t cb = precision == 0 ? sizeof(cplx<double>) : sizeof(cplx<float>);
  size_t off = 0, tw_off[3];
  for (int a = 0; a < 3; ++a) {
    if (axis[a] > kMaxAxis) {
      mgk_set_error("axis longer than the no-allocation path takes (4096)", __FILE__, __LINE__);
      return MGK_ERR_ARG;
    }
    tw_off[a] = off;
    off += mgk_align((size_t)axis[a] * cb);
  }
  const size_t need = off + pl.total;
  if (need > MGK_LS_NOALLOC_BYTES) {
    char msg[160];
    snprintf(msg, sizeof(msg), "grid needs %zu bytes of scratch; the no-allocation path holds %zu", need,
             (size_t)MGK_LS_NOALLOC_BYTES);
    mgk_set_error(msg, __FILE__, __LINE__);
    return MGK_ERR_WORKSPACE;
  }
  char* base = nullptr;
  rc = scratch_address((void**)&base);
  if (rc != MGK_OK) return rc;
  hipStream_t s = (hipStream_t)stream;
  NoAllocCtx na;
  for (int a = 0; a < 3; ++a) {
    mgk_twiddle_values(axis[a], precision, g_na_tw[a]);
    MGK_HIP(hipMemcpyAsync(base + tw_off[a], g_na_tw[a], (size_t)axis[a] * cb, hipMemcpyHostToDevice, s));
    na.tw[a] = base + tw_off[a];
  }
  na.host_state = g_na_state;
  char* ws = base + off;
  warm = warm ? 1 : 0;
  if (precision == 0) {
    if (ndim == 3)
      return ls_solve_t<double, 3>(pl, (const double*)lam, (const double*)mu, lam0, mu0, eps_bar,
                                   (const double*)eps_plastic, (double*)eps, (double*)sig, warm, max_iter, tol, ws, s,
                                   info, true, nullptr, nullptr, nullptr, nullptr, &na);
    return ls_solve_t<double, 2>(pl, (const double*)lam, (const double*)mu, lam0, mu0, eps_bar,
                                 (const double*)eps_plastic, (double*)eps, (double*)sig, warm, max_iter, tol, ws, s,
                                 info, true, nullptr, nullptr, nullptr, nullptr, &na);
  }
  if (ndim == 3)
    return ls_solve_t<float, 3>(pl, (const float*)lam, (const float*)mu, lam0, mu0, eps_bar,
                                (const float*)eps_plastic, (float*)eps, (float*)sig, warm, max_iter, tol, ws, s, info,
                                true, nullptr, nullptr, nullptr, nullptr, &na);
  return ls_solve_t<float, 2>(pl, (const float*)lam, (const float*)mu, lam0, mu0, eps_bar, (const float*)eps_plastic,
                              (float*)eps, (float*)sig, warm, max_iter, tol, ws, s, info, true, nullptr, nullptr,
                              nullptr, nullptr, &na);
}

extern "C" int mgk_ls_solve_refined(int ndim, const int64_t* shape, const void* lam32, const void* mu32,
                                    const void* lam64, const void* mu64, double lam0, double mu0,
                                    const double* eps_bar, const void* epsp32, const void* epsp64, void* eps32,
                                    void* eps64, void* sig64, int warm, int max_iter, double tol, double tol32,
                                    void* ws32, size_t ws32_bytes, void* ws64, size_t ws64_bytes, void* stream,
                                    mgk_ls_info* info, int* iters32) {
  if (!lam32 || !mu32 || !lam64 || !mu64 || !eps32 || !eps64 || !eps_bar || !ws32 || !ws64 || max_iter < 1 ||
      (epsp32 == nullptr) != (epsp64 == nullptr)) {
    mgk_set_error("null argument, max_iter < 1, or only one of the eps_plastic copies given", __FILE__, __LINE__);
    return MGK_ERR_ARG;
  }
  LsPlan p32, p64;
  int rc = ls_plan<float>(ndim, shape, &p32);
  if (rc == MGK_OK) rc = ls_plan<double>(ndim, shape, &p64);
  if (rc != MGK_OK) {
    mgk_set_error("unsupported grid shape for the LS solver", __FILE__, __LINE__);
    return rc;
  }
  if (ws32_bytes < p32.total || ws64_bytes < p64.total) {
    mgk_set_error("workspace too small", __FILE__, __LINE__);
    return MGK_ERR_WORKSPACE;
  }
  hipStream_t s = (hipStream_t)stream;
  warm = warm ? 1 : 0;
  mgk_ls_info i32{};
  // stage 1: FP32 fixed point to tol32 (no stress / means)
  if (ndim == 3)
    rc = ls_solve_t<float, 3>(p32, (const float*)lam32, (const float*)mu32, lam0, mu0, eps_bar, (const float*)epsp32,
                              (float*)eps32, nullptr, warm, max_iter, tol32, (char*)ws32, s, &i32, false);
  else
    rc = ls_solve_t<float, 2>(p32, (const float*)lam32, (const float*)mu32, lam0, mu0, eps_bar, (const float*)epsp32,
                              (float*)eps32, nullptr, warm, max_iter, tol32, (char*)ws32, s, &i32, false);
  if (iters32) *iters32 = i32.iters;
  if (rc < 0) {
    if (info) *info = i32;
    return rc;
  }
  // stage 2: refine in FP64 with the reference's stop test at tol, starting
  // from the FP32 iterate widened exactly.  It already has mean eps_bar up to
  // FP32 rounding, and every update restores the exact mean (Gamma0(0) = 0),
  // so no mean adjustment pass runs (warm = 2); the first iteration's x
  // sweeps read the FP32 field directly and its update writes eps64 whole.
  const float* from32 = (const float*)eps32;
  if (ndim == 3)
    return ls_solve_t<double, 3>(p64, (const double*)lam64, (const double*)mu64, lam0, mu0, eps_bar,
                                 (const double*)epsp64, (double*)eps64, (double*)sig64, 2, max_iter, tol, (char*)ws64,
                                 s, info, true, nullptr, nullptr, nullptr, nullptr, nullptr, from32);
  return ls_solve_t<double, 2>(p64, (const double*)lam64, (const double*)mu64, lam0, mu0, eps_bar,
                               (const double*)epsp64, (double*)eps64, (double*)sig64, 2, max_iter, tol, (char*)ws64, s,
                               info, true, nullptr, nullptr, nullptr, nullptr, nullptr, from32);
}


// ---------------------------------------------------------------------------
// Batched entry points (no counterpart in the reference: the serving form's
// one-launch-per-sweep solve of several simulations of one grid shape).
// ---------------------------------------------------------------------------
static_assert(MGK_MAXB == MGK_LS_MAX_BATCH, "kernel member tables and the C ABI batch limit differ");

namespace {
template <typename T>
int ls_batch_args(const LsPlan& pl, int nb, const mgk_ls_member* m, int refined32, LsMemArgs<T>* a) {
  for (int b = 0; b < nb; ++b) {
    const mgk_ls_member& x = m[b];
    const void* lam = refined32 ? x.lam32 : x.lam;
    const void* mu = refined32 ? x.mu32 : x.mu;
    const void* epsp = refined32 ? x.eps_plastic32 : x.eps_plastic;
    void* eps = refined32 ? x.eps32 : x.eps;
    void* ws = refined32 ? x.workspace32 : x.workspace;
    const size_t wsb = refined32 ? x.workspace32_bytes : x.workspace_bytes;
    if (!lam || !mu || !eps || !x.eps_bar || !ws) {
      mgk_set_error("batch member with a null field, macro strain or workspace", __FILE__, __LINE__);
      return MGK_ERR_ARG;
    }
    if (wsb < pl.total) {
      mgk_set_error("batch member workspace too small", __FILE__, __LINE__);
      return MGK_ERR_WORKSPACE;
    }
    a[b] = LsMemArgs<T>{(const T*)lam, (const T*)mu, x.lam0, x.mu0, x.eps_bar, (const T*)epsp, (T*)eps,
                        refined32 ? nullptr : (T*)x.sig, (char*)ws};
  }
  return MGK_OK;
}
}  // namespace

extern "C" int mgk_ls_solve_batch(int ndim, const int64_t* shape, int precision, int nb, const mgk_ls_member* members,
                                  int warm, int max_iter, double tol, void* stream, mgk_ls_info* info) {
  if (!members || !shape || nb < 1 || nb > MGK_MAXB || max_iter < 1 || (precision != 0 && precision != 1)) {
    mgk_set_error("null argument, nb outside 1..MGK_LS_MAX_BATCH, max_iter < 1 or precision not 0/1", __FILE__,
                  __LINE__);
    return MGK_ERR_ARG;
  }
  LsPlan pl;
  int rc = precision == 0 ? ls_plan<double>(ndim, shape, &pl) : ls_plan<float>(ndim, shape, &pl);
  if (rc != MGK_OK) {
    mgk_set_error("unsupported grid shape for the LS solver", __FILE__, __LINE__);
    return rc;
  }
  hipStream_t s = (hipStream_t)stream;
  warm = warm ? 1 : 0;
  if (precision == 0) {
    LsMemArgs<double> a[MGK_MAXB];
    rc = ls_batch_args<double>(pl, nb, members, 0, a);
    if (rc != MGK_OK) return rc;
    return ndim == 3 ? ls_solve_batch_t<double, 3>(pl, nb, a, warm, max_iter, tol, s, info, true)
                     : ls_solve_batch_t<double, 2>(pl, nb, a, warm, max_iter, tol, s, info, true);
  }
  LsMemArgs<float> a[MGK_MAXB];
  rc = ls_batch_args<float>(pl, nb, members, 0, a);
  if (rc != MGK_OK) return rc;
  return ndim == 3 ? ls_solve_batch_t<float, 3>(pl, nb, a, warm, max_iter, tol, s, info, true)
                   : ls_solve_batch_t<float, 2>(pl, nb, a, warm, max_iter, tol, s, info, true);
}

extern "C" int mgk_ls_solve_refined_batch(int ndim, const int64_t* shape, int nb, const mgk_ls_member* members,
                                          int warm, int max_iter, double tol, double tol32, void* stream,
                                          mgk_ls_info* info, int* iters32) {
  if (!members || !shape || nb < 1 || nb > MGK_MAXB || max_iter < 1) {
    mgk_set_error("null argument, nb outside 1..MGK_LS_MAX_BATCH or max_iter < 1", __FILE__, __LINE__);
    return MGK_ERR_ARG;
  }
  for (int b = 0; b < nb; ++b) {
    if ((members[b].eps_plastic32 == nullptr) != (members[b].eps_plastic == nullptr) || !members[b].eps32 ||
        !members[b].lam32 || !members[b].mu32) {
      mgk_set_error("batch member without its FP32 fields, or with only one eps_plastic copy", __FILE__, __LINE__);
      return MGK_ERR_ARG;
    }
  }
  LsPlan p32, p64;
  int rc = ls_plan<float>(ndim, shape, &p32);
  if (rc == MGK_OK) rc = ls_plan<double>(ndim, shape, &p64);
  if (rc != MGK_OK) {
    mgk_set_error("unsupported grid shape for the LS solver", __FILE__, __LINE__);
    return rc;
  }
  LsMemArgs<float> a32[MGK_MAXB];
  LsMemArgs<double> a64[MGK_MAXB];
  rc = ls_batch_args<float>(p32, nb, members, 1, a32);
  if (rc == MGK_OK) rc = ls_batch_args<double>(p64, nb, members, 0, a64);
  if (rc != MGK_OK) return rc;
  hipStream_t s = (hipStream_t)stream;
  warm = warm ? 1 : 0;
  // stage 1 (FP32 to tol32, no stress), stage 2 (FP64 from the FP32 iterate,
  // warm = 2), as mgk_ls_solve_refined does for one member
  mgk_ls_info i32[MGK_MAXB] = {};
  rc = ndim == 3 ? ls_solve_batch_t<float, 3>(p32, nb, a32, warm, max_iter, tol32, s, i32, false)
                 : ls_solve_batch_t<float, 2>(p32, nb, a32, warm, max_iter, tol32, s, i32, false);
  for (int b = 0; b < nb; ++b)
    if (iters32) iters32[b] = i32[b].iters;
  if (rc < 0 && rc != MGK_ERR_NONFINITE) {  // the batch itself failed
    if (info)
      for (int b = 0; b < nb; ++b) info[b] = i32[b];
    return rc;
  }
  // members whose FP32 stage failed keep its record (as the single call
  // returns it); the others continue in FP64
  LsMemArgs<double> a2[MGK_MAXB];
  int idx[MGK_MAXB], n2 = 0, rc1 = MGK_OK;
  for (int b = 0; b < nb; ++b) {
    if (i32[b].status < 0) {
      if (info) info[b] = i32[b];
      if (rc1 == MGK_OK) rc1 = i32[b].status;
    } else {
      idx[n2] = b;
      a2[n2++] = a64[b];
    }
  }
  if (n2 == 0) return rc1;
  for (int k = 0; k < n2; ++k) a2[k].eps_from32 = a32[idx[k]].eps;
  mgk_ls_info i64[MGK_MAXB] = {};
  const int rc2 = ndim == 3 ? ls_solve_batch_t<double, 3>(p64, n2, a2, 2, max_iter, tol, s, i64, true)
                            : ls_solve_batch_t<double, 2>(p64, n2, a2, 2, max_iter, tol, s, i64, true);
  if (info)
    for (int k = 0; k < n2; ++k) info[idx[k]] = i64[k];
  if (rc2 < 0 && rc2 != MGK_ERR_NONFINITE) return rc2;
  if (rc1 < 0) return rc1;
  return rc2;
}

extern "C" int mgk_ls_solve_landau(const int64_t* shape, int precision, const void* lam, const void* mu, double lam0,
                                   double mu0, const mgk_landau_params* lpar, const double* eps_bar,
                                   const void* eps_plastic, void* eps, void* sig, void* tau_buf, int warm,
                                   int max_iter, double tol, void* workspace, size_t workspace_bytes, void* stream,
                                   mgk_ls_info* info) {
  if (!lam || !mu || !eps || !eps_bar || !workspace || !tau_buf || !lpar || max_iter < 1) {
    mgk_set_error("null argument or max_iter < 1", __FILE__, __LINE__);
    return MGK_ERR_ARG;
  }
  LsPlan pl;
  int rc = precision == 0 ? ls_plan<double>(3, shape, &pl) : ls_plan<float>(3, shape, &pl);
  if (rc != MGK_OK) {
    mgk_set_error("unsupported grid shape for the LS solver", __FILE__, __LINE__);
    return rc;
  }
  if (workspace_bytes < pl.total) {
    mgk_set_error("workspace too small", __FILE__, __LINE__);
    return MGK_ERR_WORKSPACE;
  }
  LandauPar P;
  P.v1 = lpar->v1;
  P.v2 = lpar->v2;
  P.v3 = lpar->v3;
  P.g1 = lpar->g1;
  P.g2 = lpar->g2;
  P.g3 = lpar->g3;
  P.g4 = lpar->g4;
  P.Ecap = lpar->cap_limit;
  P.cap_p = lpar->cap_power;
  P.cap_ratio = lpar->cap_ratio;
  P.capping = lpar->capping;
  P.cap_smooth = lpar->cap_smooth;
  hipStream_t s = (hipStream_t)stream;
  warm = warm ? 1 : 0;
  if (precision == 0)
    return ls_solve_t<double, 3>(pl, (const double*)lam, (const double*)mu, lam0, mu0, eps_bar,
                                 (const double*)eps_plastic, (double*)eps, (double*)sig, warm, max_iter, tol,
                                 (char*)workspace, s, info, true, &P, (double*)tau_buf);
  return ls_solve_t<float, 3>(pl, (const float*)lam, (const float*)mu, lam0, mu0, eps_bar, (const float*)eps_plastic,
                              (float*)eps, (float*)sig, warm, max_iter, tol, (char*)workspace, s, info, true, &P,
                              (float*)tau_buf);
}


extern "C" int mgk_ls_landau2d_ncg_workspace_bytes(const int64_t* shape, size_t* bytes) {
  LsPlan pl;
  int rc = ls_plan<double>(2, shape, &pl);
  if (rc != MGK_OK) {
    mgk_set_error("unsupported grid shape for the LS solver", __FILE__, __LINE__);
    return rc;
  }
  *bytes = pl.total + ncg_extra_bytes(pl);
  return MGK_OK;
}

extern "C" int mgk_ls_solve_landau2d_ncg(const int64_t* shape, const double* lam, const double* mu, double lam0,
                                         double mu0, const mgk_landau_params* lpar, int plane_stress,
                                         const double* eps_bar, const double* eps_plastic, double* eps, double* sig,
                                         int warm, int max_iter, double tol, void* workspace, size_t workspace_bytes,
                                         void* stream, mgk_ls_info* info, int* krylov_iters) {
  if (!lam || !mu || !eps || !eps_bar || !workspace || !lpar || max_iter < 1) {
    mgk_set_error("bad argument (null pointer or max_iter < 1)", __FILE__, __LINE__);
    return MGK_ERR_ARG;
  }
  LsPlan pl;
  int rc = ls_plan<double>(2, shape, &pl);
  if (rc != MGK_OK) {
    mgk_set_error("unsupported grid shape for the LS solver", __FILE__, __LINE__);
    return rc;
  }
  if (workspace_bytes < pl.total + ncg_extra_bytes(pl)) {
    mgk_set_error("workspace too small", __FILE__, __LINE__);
    return MGK_ERR_WORKSPACE;
  }
  LandauPar P;
  P.v1 = lpar->v1;
  P.v2 = lpar->v2;
  P.v3 = lpar->v3;
  P.g1 = lpar->g1;
  P.g2 = lpar->g2;
  P.g3 = lpar->g3;
  P.g4 = lpar->g4;
  P.Ecap = lpar->cap_limit;
  P.cap_p = lpar->cap_power;
  P.cap_ratio = lpar->cap_ratio;
  P.capping = lpar->capping;
  P.cap_smooth = lpar->cap_smooth;
  return l2_newton_cg(pl, lam, mu, lam0, mu0, P, plane_stress ? 1 : 0, eps_bar, eps_plastic, eps, sig, warm ? 1 : 0,
                      max_iter, tol, (char*)workspace, (hipStream_t)stream, info, krylov_iters);
}

extern "C" int mgk_ls_solve_secant(int ndim, const int64_t* shape, int precision, const void* lam, const void* mu,
                                   double lam0, double mu0, const mgk_secant_params* par, const double* eps_bar,
                                   const void* eps_plastic, void* eps, void* sig, void* tau_buf, int warm,
                                   int max_iter, double tol, void* workspace, size_t workspace_bytes, void* stream,
                                   mgk_ls_info* info) {
  if (!lam || !mu || !eps || !eps_bar || !workspace || !tau_buf || !par || max_iter < 1 || (ndim != 2 && ndim != 3) ||
      (precision != 0 && precision != 1) || par->law < 0 || par->law > 1 || (par->law == 1 && ndim != 2)) {
    mgk_set_error("bad argument (null pointer, max_iter < 1, ndim, precision, or a law not defined in 3D)", __FILE__,
                  __LINE__);
    return MGK_ERR_ARG;
  }
  LsPlan pl;
  int rc = precision == 0 ? ls_plan<double>(ndim, shape, &pl) : ls_plan<float>(ndim, shape, &pl);
  if (rc != MGK_OK) {
    mgk_set_error("unsupported grid shape for the LS solver", __FILE__, __LINE__);
    return rc;
  }
  if (workspace_bytes < pl.total) {
    mgk_set_error("workspace too small", __FILE__, __LINE__);
    return MGK_ERR_WORKSPACE;
  }
  SecLaw L{par->law, par->plane_stress ? 1 : 0, par->p1, par->p2};
  hipStream_t s = (hipStream_t)stream;
  warm = warm ? 1 : 0;
  if (precision == 0) {
    if (ndim == 3)
      return ls_solve_t<double, 3>(pl, (const double*)lam, (const double*)mu, lam0, mu0, eps_bar,
                                   (const double*)eps_plastic, (double*)eps, (double*)sig, warm, max_iter, tol,
                                   (char*)workspace, s, info, true, nullptr, (double*)tau_buf, nullptr, &L);
    return ls_solve_t<double, 2>(pl, (const double*)lam, (const double*)mu, lam0, mu0, eps_bar,
                                 (const double*)eps_plastic, (double*)eps, (double*)sig, warm, max_iter, tol,
                                 (char*)workspace, s, info, true, nullptr, (double*)tau_buf, nullptr, &L);
  }
  if (ndim == 3)
    return ls_solve_t<float, 3>(pl, (const float*)lam, (const float*)mu, lam0, mu0, eps_bar,
                                (const float*)eps_plastic, (float*)eps, (float*)sig, warm, max_iter, tol,
                                (char*)workspace, s, info, true, nullptr, (float*)tau_buf, nullptr, &L);
  return ls_solve_t<float, 2>(pl, (const float*)lam, (const float*)mu, lam0, mu0, eps_bar, (const float*)eps_plastic,
                              (float*)eps, (float*)sig, warm, max_iter, tol, (char*)workspace, s, info, true, nullptr,
                              (float*)tau_buf, nullptr, &L);
}

namespace {
struct HeatWs {
  size_t off_part, off_state, total;
};
HeatWs heat_ws(const LsPlan& pl) {
  HeatWs w;
  w.off_part = mgk_align((size_t)pl.nxh * pl.P * pl.Q * sizeof(cplx<double>));
  w.off_state = w.off_part + mgk_align((size_t)pl.nblk_red * sizeof(double));
  w.total = w.off_state + mgk_align(sizeof(LsState));
  return w;
}
}  // namespace

extern "C" int mgk_heat_workspace_bytes(int ndim, const int64_t* shape, size_t* bytes) {
  LsPlan pl;
  if (!shape || !bytes || ls_plan<double>(ndim, shape, &pl, false) != MGK_OK) {
    mgk_set_error("unsupported grid shape for the heat solver", __FILE__, __LINE__);
    return MGK_ERR_ARG;
  }
  *bytes = heat_ws(pl).total;
  return MGK_OK;
}

extern "C" int mgk_heat_step(int ndim, const int64_t* shape, double* Tf, const double* lengths, double Ddt, double T0,
                             int thermostat, double bath_factor, double cap, void* workspace, size_t workspace_bytes,
                             void* stream) {
  LsPlan pl;
  if (!shape || !Tf || !lengths || !workspace || thermostat < 0 || thermostat > 2 ||
      ls_plan<double>(ndim, shape, &pl, false) != MGK_OK) {
    mgk_set_error("mgk_heat_step: bad argument or unsupported grid shape", __FILE__, __LINE__);
    return MGK_ERR_ARG;
  }
  const HeatWs w = heat_ws(pl);
  if (workspace_bytes < w.total) {
    mgk_set_error("workspace too small", __FILE__, __LINE__);
    return MGK_ERR_WORKSPACE;
  }
  constexpr int MV = maxv<double>();
  hipStream_t s = (hipStream_t)stream;
  char* ws = (char*)workspace;
  cplx<double>* spec = (cplx<double>*)ws;
  double* part = (double*)(ws + w.off_part);
  LsState* st = (LsState*)(ws + w.off_state);
  LsGeom<double> g;
  g.nx = pl.nx;
  g.P = pl.P;
  g.Q = pl.Q;
  g.nxh = pl.nxh;
  g.C = 1;
  g.NV = pl.NV;
  g.logBx = pl.logBx;
  g.logBy = pl.logBy;
  g.logPP = pl.logPP;
  g.npencil = pl.nxh * pl.P;
  mgk_make_plan(pl.nx, &g.plx);
  mgk_make_plan(pl.P, &g.ply);
  mgk_make_plan(pl.Q, &g.plq);
  g.twx = (const cplx<double>*)mgk_twiddles(pl.nx, 0);
  g.twy = (const cplx<double>*)mgk_twiddles(pl.P, 0);
  g.twq = (const cplx<double>*)mgk_twiddles(pl.Q, 0);
  if (!g.twx || !g.twy || !g.twq) {
    mgk_set_error("twiddle table allocation failed", __FILE__, __LINE__);
    return MGK_ERR_HIP;
  }
  HeatPar hp;
  const double two_pi = 6.283185307179586;
  hp.kx = two_pi / lengths[0];
  hp.ky = ndim == 3 ? two_pi / lengths[1] : 0.0;
  hp.kz = two_pi / lengths[ndim - 1];
  hp.Ddt = Ddt;
  hp.scale = 1.0 / (double)pl.NV;
  MGK_HIP(hipMemsetAsync(st, 0, sizeof(LsState), s));
  const int nqx = (pl.Q + (1 << pl.logBx) - 1) >> pl.logBx;
  const int nqy = (pl.Q + (1 << pl.logBy) - 1) >> pl.logBy;
  const int npb = (g.npencil + (1 << pl.logPP) - 1) >> pl.logPP;
  const dim3 gx(nqx, pl.P), gy(nqy, pl.nxh, 1);
  k_ls_xfwd_real<double, MV><<<gx, MGK_NT, pl.lds_x, s>>>(g, Tf, spec, st);
  if (pl.P > 1) launch_mid<double, MV>(pl, gy, s, g.ply, g.twy, mid_one(spec, (const int*)nullptr, (int)gy.z), 0);
  k_heat_pencil<double, MV><<<npb, MGK_NT, pl.lds_q, s>>>(g, spec, hp);
  if (pl.P > 1) launch_mid<double, MV>(pl, gy, s, g.ply, g.twy, mid_one(spec, (const int*)nullptr, (int)gy.z), 1);
  k_heat_xinv<double, MV><<<gx, MGK_NT, pl.lds_x, s>>>(g, spec, Tf);
  if (thermostat == 2) {
    k_comp_sums<double><<<pl.nblk_red, MGK_NT, 0, s>>>(Tf, 1, (size_t)pl.NV, part);
    k_means_finalize<<<1, MGK_NT, 0, s>>>(part, pl.nblk_red, 1, (double)pl.NV, st->mean);
  }
  if (thermostat != 0 || cap > 0) {
    const int nb = (int)std::min<long long>((pl.NV + MGK_NT - 1) / MGK_NT, 4096);
    k_heat_post<<<nb, MGK_NT, 0, s>>>(Tf, (size_t)pl.NV, T0, thermostat, bath_factor, cap, st->mean);
  }
  MGK_HIP(hipGetLastError());
  return MGK_OK;
}

extern "C" int mgk_ls_solve_landau2d(const int64_t* shape, const double* lam, const double* mu, double lam0,
                                     double mu0, const mgk_landau_params* lpar, int plane_stress,
                                     const double* eps_bar, const double* eps_plastic, double* eps, double* sig,
                                     double* e33, int e33_persist, int e33_have, double* tau_buf, int anderson_m,
                                     double* aa_buf, int warm, int max_iter, double tol, void* workspace,
                                     size_t workspace_bytes, void* stream, mgk_ls_info* info) {
  if (!lam || !mu || !eps || !eps_bar || !workspace || !tau_buf || !lpar || max_iter < 1 || anderson_m < 0 ||
      anderson_m > 5 || (plane_stress && !e33) || (anderson_m > 0 && !aa_buf)) {
    mgk_set_error("bad argument (null pointer, max_iter < 1 or anderson_m outside 0..5)", __FILE__, __LINE__);
    return MGK_ERR_ARG;
  }
  LsPlan pl;
  int rc = ls_plan<double>(2, shape, &pl);
  if (rc != MGK_OK) {
    mgk_set_error("unsupported grid shape for the LS solver", __FILE__, __LINE__);
    return rc;
  }
  if (workspace_bytes < pl.total) {
    mgk_set_error("workspace too small", __FILE__, __LINE__);
    return MGK_ERR_WORKSPACE;
  }
  L2Mode m;
  m.P.v1 = lpar->v1;
  m.P.v2 = lpar->v2;
  m.P.v3 = lpar->v3;
  m.P.g1 = lpar->g1;
  m.P.g2 = lpar->g2;
  m.P.g3 = lpar->g3;
  m.P.g4 = lpar->g4;
  m.P.Ecap = lpar->cap_limit;
  m.P.cap_p = lpar->cap_power;
  m.P.cap_ratio = lpar->cap_ratio;
  m.P.capping = lpar->capping;
  m.P.cap_smooth = lpar->cap_smooth;
  m.plane_stress = plane_stress ? 1 : 0;
  m.persist = e33_persist ? 1 : 0;
  m.e33_have = e33_have ? 1 : 0;
  m.e33 = e33;
  m.tau = tau_buf;
  m.aa_m = anderson_m;
  const size_t n3 = (size_t)3 * pl.NV;
  m.aa_g = aa_buf;
  m.aa_f = aa_buf ? aa_buf + (size_t)(anderson_m + 1) * n3 : nullptr;
  m.eps_old = aa_buf ? aa_buf + (size_t)2 * (anderson_m + 1) * n3 : nullptr;
  return ls_solve_t<double, 2>(pl, lam, mu, lam0, mu0, eps_bar, eps_plastic, eps, sig, warm ? 1 : 0, max_iter, tol,
                               (char*)workspace, (hipStream_t)stream, info, true, nullptr, nullptr, &m);
}
