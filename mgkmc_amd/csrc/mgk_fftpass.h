// FFT sweep building blocks shared by the LS and finite-strain kernels:
// half-spectrum split/merge for the real-to-complex x pass and the strided
// complex FFT along the middle axis.  Layout conventions: mgk_common.h.
#pragma once
#include "mgk_common.h"

// ---------------------------------------------------------------------------
// x pass helpers
// ---------------------------------------------------------------------------
// After a forward FFT of z = a + i b along x (tile rows), split into the half
// spectra of a and b and store them.
template <typename T>
__device__ __forceinline__ void xsplit_store(const cplx<T>* tile, int ldb, int nx, int nxh, int logB, int q0, int Q,
                                             int P, int p, cplx<T>* __restrict__ sa, cplx<T>* __restrict__ sb) {
  const int B = 1 << logB;
  for (int e = threadIdx.x; e < nxh * B; e += MGK_NT) {
    const int b = e & (B - 1), k = e >> logB, q = q0 + b;
    if (q >= Q) continue;
    const cplx<T> z1 = tile[k * ldb + b];
    const int kr = (k == 0) ? 0 : nx - k;
    const cplx<T> z2 = tile[kr * ldb + b];
    const size_t idx = ((size_t)k * P + p) * Q + q;
    sa[idx] = cmk(T(0.5) * (z1.x + z2.x), T(0.5) * (z1.y - z2.y));
    if (sb) sb[idx] = cmk(T(0.5) * (z1.y + z2.y), T(-0.5) * (z1.x - z2.x));
  }
}

// Build the full x-spectrum of z = a + i b from half spectra of a and b.
template <typename T>
__device__ __forceinline__ void xmerge_load(cplx<T>* tile, int ldb, int nx, int nxh, int logB, int q0, int Q, int P,
                                            int p, const cplx<T>* __restrict__ sa, const cplx<T>* __restrict__ sb) {
  const int B = 1 << logB;
  for (int e = threadIdx.x; e < nx * B; e += MGK_NT) {
    const int b = e & (B - 1), k = e >> logB, q = q0 + b;
    cplx<T> z = cmk(T(0), T(0));
    if (q < Q) {
      if (k < nxh) {
        const size_t idx = ((size_t)k * P + p) * Q + q;
        const cplx<T> A = sa[idx];
        const cplx<T> Bv = sb ? sb[idx] : cmk(T(0), T(0));
        z = cmk(A.x - Bv.y, A.y + Bv.x);
      } else {
        const size_t idx = ((size_t)(nx - k) * P + p) * Q + q;
        const cplx<T> A = sa[idx];
        const cplx<T> Bv = sb ? sb[idx] : cmk(T(0), T(0));
        z = cmk(A.x + Bv.y, -A.y + Bv.x);
      }
    }
    tile[k * ldb + b] = z;
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------
// sweep 2/4: complex FFT along the middle axis (length P, stride Q)
// grid (ceil(Q/B), nxh, C)
// ---------------------------------------------------------------------------
template <typename T, int MAXV, int NN = 0, int BB = 0>
__global__ __launch_bounds__(MGK_NT) void k_fft_mid(int P, int Q, int nxh, int logB, FftPlan pl,
                                                    const cplx<T>* __restrict__ twg, cplx<T>* __restrict__ spec,
                                                    int inv, const int* __restrict__ done) {
  if (done && *done) return;
  extern __shared__ __align__(16) unsigned char smem[];
  cplx<T>* tw = reinterpret_cast<cplx<T>*>(smem);
  cplx<T>* tile = tw + P;
  constexpr bool CT = NN > 0;
  if (CT) {
    P = NN;
    logB = ct_log2(BB);
  }
  const int B = CT ? BB : 1 << logB, ldb = B;
  const int q0 = blockIdx.x * B, kx = blockIdx.y, c = blockIdx.z;
  cplx<T>* base = spec + ((size_t)c * nxh + kx) * (size_t)P * Q;
  for (int t = threadIdx.x; t < P; t += MGK_NT) tw[t] = twg[t];
  for (int e = threadIdx.x; e < P * B; e += MGK_NT) {
    const int b = e & (B - 1), y = e >> logB, q = q0 + b;
    tile[y * ldb + b] = (q < Q) ? base[(size_t)y * Q + q] : cmk(T(0), T(0));
  }
  __syncthreads();
  if constexpr (CT) tile_fft_ct<NN, BB, 1>(tile, tw, inv != 0);
  else tile_fft<MAXV>(tile, Tile{P, logB, ldb, 1, 0}, pl, tw, inv != 0);
  for (int e = threadIdx.x; e < P * B; e += MGK_NT) {
    const int b = e & (B - 1), y = e >> logB, q = q0 + b;
    if (q < Q) base[(size_t)y * Q + q] = tile[y * ldb + b];
  }
}

