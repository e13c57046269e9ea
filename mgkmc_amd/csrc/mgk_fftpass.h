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
template <typename T, int NT = MGK_NT>
__device__ __forceinline__ void xsplit_store(const cplx<T>* tile, int ldb, int nx, int nxh, int logB, int q0, int Q,
                                             int P, int p, cplx<T>* __restrict__ sa, cplx<T>* __restrict__ sb) {
  const int B = 1 << logB;
  for (int e = threadIdx.x; e < nxh * B; e += NT) {
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

// Same, FP32 with a compile-time tile (nx = NN, B = BB columns) whose columns
// are all inside the grid and whose rows start 16-byte aligned: two complex
// values (16 bytes) per lane and access.
template <int NN, int BB, int NT = MGK_NT>
__device__ __forceinline__ void xsplit_store_v(const cplx<float>* tile, int q0, int Q, int P, int p,
                                               cplx<float>* __restrict__ sa, cplx<float>* __restrict__ sb) {
  constexpr int H = BB / 2, NXH = NN / 2 + 1;
  for (int e = threadIdx.x; e < NXH * H; e += NT) {
    const int h = e % H, k = e / H;
    const int kr = (k == 0) ? 0 : NN - k;
    const float4 z1 = *reinterpret_cast<const float4*>(tile + k * BB + 2 * h);
    const float4 z2 = *reinterpret_cast<const float4*>(tile + kr * BB + 2 * h);
    const size_t idx = ((size_t)k * P + p) * Q + q0 + 2 * h;
    *reinterpret_cast<float4*>(sa + idx) = make_float4(0.5f * (z1.x + z2.x), 0.5f * (z1.y - z2.y),
                                                       0.5f * (z1.z + z2.z), 0.5f * (z1.w - z2.w));
    if (sb)
      *reinterpret_cast<float4*>(sb + idx) = make_float4(0.5f * (z1.y + z2.y), -0.5f * (z1.x - z2.x),
                                                         0.5f * (z1.w + z2.w), -0.5f * (z1.z - z2.z));
  }
}

template <int NN, int BB, int NT = MGK_NT>
__device__ __forceinline__ void xmerge_load_v(cplx<float>* tile, int q0, int Q, int P, int p,
                                              const cplx<float>* __restrict__ sa,
                                              const cplx<float>* __restrict__ sb) {
  constexpr int H = BB / 2, NXH = NN / 2 + 1;
  // each stored half-spectrum row k is read once and gives tile rows k and
  // NN - k (conjugate-symmetric halves)
  for (int e = threadIdx.x; e < NXH * H; e += NT) {
    const int h = e % H, k = e / H;
    const size_t idx = ((size_t)k * P + p) * Q + q0 + 2 * h;
    const float4 A = *reinterpret_cast<const float4*>(sa + idx);
    const float4 Bv = sb ? *reinterpret_cast<const float4*>(sb + idx) : make_float4(0.f, 0.f, 0.f, 0.f);
    *reinterpret_cast<float4*>(tile + k * BB + 2 * h) =
        make_float4(A.x - Bv.y, A.y + Bv.x, A.z - Bv.w, A.w + Bv.z);
    if (k > 0 && k < NN - k)
      *reinterpret_cast<float4*>(tile + (NN - k) * BB + 2 * h) =
          make_float4(A.x + Bv.y, -A.y + Bv.x, A.z + Bv.w, -A.w + Bv.z);
  }
  __syncthreads();
}

// xsplit_store, through the vector form when the tile allows it (vec: every
// column inside the grid, rows 16-byte aligned).
template <typename T, int NN, int BB, int NT = MGK_NT>
__device__ __forceinline__ void store_half(const cplx<T>* tile, bool vec, int ldb, int nx, int nxh, int logB, int q0,
                                           int Q, int P, int p, cplx<T>* __restrict__ sa,
                                           cplx<T>* __restrict__ sb) {
  if constexpr (sizeof(T) == 4 && NN > 0 && BB % 2 == 0) {
    if (vec) {
      xsplit_store_v<NN, BB, NT>(tile, q0, Q, P, p, sa, sb);
      return;
    }
  }
  xsplit_store<T, NT>(tile, ldb, nx, nxh, logB, q0, Q, P, p, sa, sb);
}

// Build the full x-spectrum of z = a + i b from half spectra of a and b.
template <typename T, int NT = MGK_NT>
__device__ __forceinline__ void xmerge_load(cplx<T>* tile, int ldb, int nx, int nxh, int logB, int q0, int Q, int P,
                                            int p, const cplx<T>* __restrict__ sa, const cplx<T>* __restrict__ sb) {
  const int B = 1 << logB;
  // each stored row k < nxh is read once and gives tile rows k and nx - k
  for (int e = threadIdx.x; e < nxh * B; e += NT) {
    const int b = e & (B - 1), k = e >> logB, q = q0 + b;
    cplx<T> A = cmk(T(0), T(0)), Bv = cmk(T(0), T(0));
    if (q < Q) {
      const size_t idx = ((size_t)k * P + p) * Q + q;
      A = sa[idx];
      if (sb) Bv = sb[idx];
    }
    tile[k * ldb + b] = cmk(A.x - Bv.y, A.y + Bv.x);
    if (k > 0 && k < nx - k) tile[(nx - k) * ldb + b] = cmk(A.x + Bv.y, -A.y + Bv.x);
  }
  __syncthreads();
}

// Spectra of up to MGK_MAXB simulations of one grid shape for one mid-axis
// launch: member m's component c is grid slice z = m * ncomp + c.
template <typename T>
struct MidPtrs {
  cplx<T>* spec[MGK_MAXB];
  const int* done[MGK_MAXB];  // per-member stop flag (NULL: always run)
  int ncomp;
};

template <typename T>
__host__ __device__ inline MidPtrs<T> mid_one(cplx<T>* spec, const int* done, int ncomp) {
  MidPtrs<T> m{};
  m.spec[0] = spec;
  m.done[0] = done;
  m.ncomp = ncomp;
  return m;
}

// ---------------------------------------------------------------------------
// sweep 2/4: complex FFT along the middle axis (length P, stride Q)
// grid (ceil(Q/B), nxh, C * members)
// ---------------------------------------------------------------------------
template <typename T, int MAXV, int NN = 0, int BB = 0, int NTH = MGK_NT>
__global__ __launch_bounds__(NTH) void k_fft_mid(int P, int Q, int nxh, int logB, FftPlan pl,
                                                    const cplx<T>* __restrict__ twg, MidPtrs<T> mp, int inv) {
  const int mem = blockIdx.z / mp.ncomp;
  const int* __restrict__ done = mp.done[mem];
  if (done && *done) return;
  cplx<T>* __restrict__ spec = mp.spec[mem];
  extern __shared__ __align__(16) unsigned char smem[];
  cplx<T>* tw = reinterpret_cast<cplx<T>*>(smem);
  cplx<T>* tile = tw + P;
  constexpr bool CT = NN > 0;
  if (CT) {
    P = NN;
    logB = ct_log2(BB);
  }
  const int B = CT ? BB : 1 << logB, ldb = B;
  const int q0 = blockIdx.x * B, kx = blockIdx.y, c = blockIdx.z - mem * mp.ncomp;
  cplx<T>* base = spec + ((size_t)c * nxh + kx) * (size_t)P * Q;
  for (int t = threadIdx.x; t < P; t += NTH) tw[t] = twg[t];
  // FP32 with compile-time tiles: two complex values (16 bytes) per lane and
  // access, so one wave instruction moves 1 KB instead of 512 B
  constexpr bool VEC = CT && sizeof(T) == 4 && (BB % 2 == 0);
  if constexpr (VEC) {
    constexpr int H = BB / 2;
    const bool full = q0 + BB <= Q && (Q & 1) == 0;  // 16-byte aligned rows
    for (int e = threadIdx.x; e < NN * H; e += NTH) {
      const int h = e % H, y = e / H, q = q0 + 2 * h;
      float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
      if (full) v = *reinterpret_cast<const float4*>(base + (size_t)y * Q + q);
      else if (q < Q) {
        const cplx<T> a = base[(size_t)y * Q + q];
        v.x = a.x;
        v.y = a.y;
        if (q + 1 < Q) {
          const cplx<T> b = base[(size_t)y * Q + q + 1];
          v.z = b.x;
          v.w = b.y;
        }
      }
      *reinterpret_cast<float4*>(tile + y * ldb + 2 * h) = v;
    }
  } else {
    for (int e = threadIdx.x; e < P * B; e += NTH) {
      const int b = e & (B - 1), y = e >> logB, q = q0 + b;
      tile[y * ldb + b] = (q < Q) ? base[(size_t)y * Q + q] : cmk(T(0), T(0));
    }
  }
  __syncthreads();
  if constexpr (CT) tile_fft_ct<NN, BB, 1, false, NTH>(tile, tw, inv != 0);
  else tile_fft<MAXV>(tile, Tile{P, logB, ldb, 1, 0}, pl, tw, inv != 0);
  if constexpr (VEC) {
    constexpr int H = BB / 2;
    const bool full = q0 + BB <= Q && (Q & 1) == 0;  // 16-byte aligned rows
    for (int e = threadIdx.x; e < NN * H; e += NTH) {
      const int h = e % H, y = e / H, q = q0 + 2 * h;
      const float4 v = *reinterpret_cast<const float4*>(tile + y * ldb + 2 * h);
      if (full) *reinterpret_cast<float4*>(base + (size_t)y * Q + q) = v;
      else if (q < Q) {
        base[(size_t)y * Q + q] = cmk(T(v.x), T(v.y));
        if (q + 1 < Q) base[(size_t)y * Q + q + 1] = cmk(T(v.z), T(v.w));
      }
    }
  } else {
    for (int e = threadIdx.x; e < P * B; e += NTH) {
      const int b = e & (B - 1), y = e >> logB, q = q0 + b;
      if (q < Q) base[(size_t)y * Q + q] = tile[y * ldb + b];
    }
  }
}

