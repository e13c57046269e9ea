// Common device/host utilities for the mgkmc_amd HIP library (gfx950).
//
// Layout conventions used by every kernel in this library
// ---------------------------------------------------------
//  * A grid is (nx, P, Q) voxels, Q contiguous.  3D: P = ny, Q = nz.
//    2D: P = 1, Q = ny (so the x axis is always the outermost one).
//  * Symmetric tensor fields are SoA, component-major:  f[c][x][p][q].
//      3D components: xx, yy, zz, yz, xz, xy   (C = 6)
//      2D components: xx, yy, xy               (C = 3)
//    9-component (finite strain) fields: f[3*i+j][x][p][q], row-major ij.
//  * Half spectra (r2c along x) are complex, f[c][kx][p][q] with
//    kx in [0, nx/2].  y and z are full complex axes.
#pragma once
#include <type_traits>
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

#define MGK_NT 256  // threads per workgroup for every tiled kernel

template <typename T>
struct cplx {
  T x, y;
};

template <typename T>
__device__ __forceinline__ cplx<T> cmk(T a, T b) {
  cplx<T> r;
  r.x = a;
  r.y = b;
  return r;
}
template <typename T>
__device__ __forceinline__ cplx<T> cadd(cplx<T> a, cplx<T> b) { return cmk(a.x + b.x, a.y + b.y); }
template <typename T>
__device__ __forceinline__ cplx<T> csub(cplx<T> a, cplx<T> b) { return cmk(a.x - b.x, a.y - b.y); }
template <typename T>
__device__ __forceinline__ cplx<T> cmul(cplx<T> a, cplx<T> b) {
  return cmk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename T>
__device__ __forceinline__ cplx<T> cscale(cplx<T> a, T s) { return cmk(a.x * s, a.y * s); }
template <typename T>
__device__ __forceinline__ cplx<T> cconj(cplx<T> a) { return cmk(a.x, -a.y); }
// multiply by -i (forward) or +i (inverse)
template <typename T>
__device__ __forceinline__ cplx<T> mul_mi(cplx<T> a) { return cmk(a.y, -a.x); }
template <typename T>
__device__ __forceinline__ cplx<T> mul_pi(cplx<T> a) { return cmk(-a.y, a.x); }

// ---------------------------------------------------------------------------
// 16-byte vector access (one global_load/store_dwordx4 per lane): VW(T)
// consecutive reals.  The address must be 16-byte aligned.
// ---------------------------------------------------------------------------
template <typename T>
constexpr int vw16() { return 16 / (int)sizeof(T); }

template <typename T>
__device__ __forceinline__ void vload16(const T* __restrict__ p, T (&o)[16 / sizeof(T)]) {
  if constexpr (sizeof(T) == 4) {
    const float4 v = *reinterpret_cast<const float4*>(p);
    o[0] = v.x;
    o[1] = v.y;
    o[2] = v.z;
    o[3] = v.w;
  } else {
    const double2 v = *reinterpret_cast<const double2*>(p);
    o[0] = v.x;
    o[1] = v.y;
  }
}

template <typename T>
__device__ __forceinline__ void vstore16(T* __restrict__ p, const T (&a)[16 / sizeof(T)]) {
  if constexpr (sizeof(T) == 4) {
    *reinterpret_cast<float4*>(p) = make_float4(a[0], a[1], a[2], a[3]);
  } else {
    *reinterpret_cast<double2*>(p) = make_double2(a[0], a[1]);
  }
}

// ---------------------------------------------------------------------------
// FFT plan: length n factored into radices 8,5,4,3,2 (applied in order).
// ---------------------------------------------------------------------------
struct FftPlan {
  int n;
  int npass;
  int radix[12];
};

// Small fixed-size DFTs.  inv=false: exp(-2 pi i pq/R); inv=true: exp(+...).
template <typename T>
__device__ __forceinline__ void dft2(cplx<T>* v) {
  cplx<T> a = v[0], b = v[1];
  v[0] = cadd(a, b);
  v[1] = csub(a, b);
}

template <typename T>
__device__ __forceinline__ void dft4(cplx<T>* v, bool inv) {
  cplx<T> s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
  cplx<T> s13 = cadd(v[1], v[3]), d13 = csub(v[1], v[3]);
  cplx<T> r = inv ? mul_pi(d13) : mul_mi(d13);
  v[0] = cadd(s02, s13);
  v[2] = csub(s02, s13);
  v[1] = cadd(d02, r);
  v[3] = csub(d02, r);
}

template <typename T>
__device__ __forceinline__ void dft3(cplx<T>* v, bool inv) {
  const T c = T(-0.5);
  const T s = inv ? T(0.86602540378443864676372317075294) : T(-0.86602540378443864676372317075294);
  cplx<T> a = v[0], b = v[1], d = v[2];
  cplx<T> sbd = cadd(b, d), dbd = csub(b, d);
  cplx<T> t = cmk(a.x + c * sbd.x, a.y + c * sbd.y);
  // i*s*(b-d)
  cplx<T> u = cmk(-s * dbd.y, s * dbd.x);
  v[0] = cadd(a, sbd);
  v[1] = cadd(t, u);
  v[2] = csub(t, u);
}

template <typename T>
__device__ __forceinline__ void dft5(cplx<T>* v, bool inv) {
  // w = exp(-+2 pi i/5); c1 = cos(2pi/5), c2 = cos(4pi/5), s1 = sin(2pi/5), s2 = sin(4pi/5)
  const T c1 = T(0.30901699437494742410229341718282);
  const T c2 = T(-0.80901699437494742410229341718282);
  const T s1 = inv ? T(0.95105651629515357211643933337938) : T(-0.95105651629515357211643933337938);
  const T s2 = inv ? T(0.58778525229247312916870595463907) : T(-0.58778525229247312916870595463907);
  cplx<T> x0 = v[0];
  cplx<T> a1 = cadd(v[1], v[4]), b1 = csub(v[1], v[4]);
  cplx<T> a2 = cadd(v[2], v[3]), b2 = csub(v[2], v[3]);
  cplx<T> t1 = cmk(x0.x + c1 * a1.x + c2 * a2.x, x0.y + c1 * a1.y + c2 * a2.y);
  cplx<T> t2 = cmk(x0.x + c2 * a1.x + c1 * a2.x, x0.y + c2 * a1.y + c1 * a2.y);
  // i*(s1*b1 + s2*b2), i*(s2*b1 - s1*b2)
  cplx<T> u1 = cmk(s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y);
  cplx<T> u2 = cmk(s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y);
  u1 = mul_pi(u1);
  u2 = mul_pi(u2);
  v[0] = cadd(x0, cadd(a1, a2));
  v[1] = cadd(t1, u1);
  v[4] = csub(t1, u1);
  v[2] = cadd(t2, u2);
  v[3] = csub(t2, u2);
}

template <typename T>
__device__ __forceinline__ void dft8(cplx<T>* v, bool inv) {
  const T r = T(0.70710678118654752440084436210485);
  // radix-2 on (q, q+4)
  cplx<T> a[4], b[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    a[q] = cadd(v[q], v[q + 4]);
    b[q] = csub(v[q], v[q + 4]);
  }
  // twiddle b[q] by w8^q
  if (!inv) {
    b[1] = cmk(r * (b[1].x + b[1].y), r * (b[1].y - b[1].x));  // * (1-i)/sqrt2
    b[2] = mul_mi(b[2]);
    b[3] = cmk(r * (b[3].y - b[3].x), -r * (b[3].x + b[3].y));  // * (-1-i)/sqrt2
  } else {
    b[1] = cmk(r * (b[1].x - b[1].y), r * (b[1].y + b[1].x));  // * (1+i)/sqrt2
    b[2] = mul_pi(b[2]);
    b[3] = cmk(-r * (b[3].x + b[3].y), r * (b[3].x - b[3].y));  // * (-1+i)/sqrt2
  }
  dft4(a, inv);
  dft4(b, inv);
  v[0] = a[0];
  v[2] = a[1];
  v[4] = a[2];
  v[6] = a[3];
  v[1] = b[0];
  v[3] = b[1];
  v[5] = b[2];
  v[7] = b[3];
}

// w16^j = exp(-+2 pi i j / 16) times a (forward: -, inverse: +), j a compile-time
// multiple of pi/8 (cos(pi/8) = 0.9238795325..., sin(pi/8) = 0.3826834323...)
template <int J, typename T>
__device__ __forceinline__ cplx<T> w16mul(cplx<T> a, bool inv) {
  constexpr int j = J % 16;
  const T c8 = T(0.92387953251128675612818318939679), s8 = T(0.38268343236508977172845998403040);
  const T r = T(0.70710678118654752440084436210485);
  // cos and sin of j pi / 8 for j = 0..15
  const T cs[16] = {T(1), c8, r, s8, T(0), -s8, -r, -c8, T(-1), -c8, -r, -s8, T(0), s8, r, c8};
  const T sn[16] = {T(0), s8, r, c8, T(1), c8, r, s8, T(0), -s8, -r, -c8, T(-1), -c8, -r, -s8};
  const T c = cs[j], sg = inv ? sn[j] : -sn[j];
  return cmk(a.x * c - a.y * sg, a.x * sg + a.y * c);
}

// 16-point DFT as 4 x 4: DFT4 over the four stride-4 subsequences, twiddles
// w16^(b k1), DFT4 across them; natural order in and out
template <typename T>
__device__ __forceinline__ void dft16(cplx<T>* v, bool inv) {
  cplx<T> y[4][4];
#pragma unroll
  for (int b = 0; b < 4; ++b) {
#pragma unroll
    for (int q = 0; q < 4; ++q) y[b][q] = v[4 * q + b];
    dft4(y[b], inv);
  }
  y[1][1] = w16mul<1>(y[1][1], inv);
  y[1][2] = w16mul<2>(y[1][2], inv);
  y[1][3] = w16mul<3>(y[1][3], inv);
  y[2][1] = w16mul<2>(y[2][1], inv);
  y[2][2] = w16mul<4>(y[2][2], inv);
  y[2][3] = w16mul<6>(y[2][3], inv);
  y[3][1] = w16mul<3>(y[3][1], inv);
  y[3][2] = w16mul<6>(y[3][2], inv);
  y[3][3] = w16mul<9>(y[3][3], inv);
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1) {
    cplx<T> z[4] = {y[0][k1], y[1][k1], y[2][k1], y[3][k1]};
    dft4(z, inv);
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = z[k2];
  }
}

template <int R, typename T>
__device__ __forceinline__ void dftR(cplx<T>* v, bool inv) {
  if constexpr (R == 2) dft2(v);
  else if constexpr (R == 3) dft3(v, inv);
  else if constexpr (R == 4) dft4(v, inv);
  else if constexpr (R == 5) dft5(v, inv);
  else if constexpr (R == 16) dft16(v, inv);
  else dft8(v, inv);
}

// ---------------------------------------------------------------------------
// LDS tile descriptor and Stockham passes.
//   element (c, j, b) lives at buf[c*cstride + j*ldb + b]
//   j in [0,n) is the transform axis, b in [0,B) the column (B = 2^logB),
//   c in [0,C) a batch of independent tiles.
// tw: n twiddles exp(-2 pi i t/n) in LDS.
// Every thread of the workgroup must call these (they contain barriers).
// ---------------------------------------------------------------------------
struct Tile {
  int n, logB, ldb, C, cstride;
};

template <int R, int MAXV, typename T>
__device__ __forceinline__ void stockham_pass(cplx<T>* __restrict__ buf, const Tile& tl, int L,
                                              const cplx<T>* __restrict__ tw, bool inv) {
  constexpr int S = (MAXV + R - 1) / R;
  const int n = tl.n;
  const int m = n / R;
  const int B = 1 << tl.logB;
  const int nb = tl.C * m * B;
  const int tstride = n / (L * R);
  cplx<T> v[S][R];
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const int t = threadIdx.x + s * MGK_NT;
    if (t < nb) {
      const int b = t & (B - 1);
      const int u = t >> tl.logB;
      const int j = u % m;
      const int c = u / m;
      cplx<T>* base = buf + c * tl.cstride + b;
#pragma unroll
      for (int q = 0; q < R; ++q) v[s][q] = base[(j + q * m) * tl.ldb];
    }
  }
  __syncthreads();
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const int t = threadIdx.x + s * MGK_NT;
    if (t < nb) {
      const int b = t & (B - 1);
      const int u = t >> tl.logB;
      const int j = u % m;
      const int c = u / m;
      const int k = j % L;
      if (L > 1) {
#pragma unroll
        for (int q = 1; q < R; ++q) {
          cplx<T> w = tw[q * k * tstride];
          if (inv) w.y = -w.y;
          v[s][q] = cmul(v[s][q], w);
        }
      }
      dftR<R>(v[s], inv);
      cplx<T>* base = buf + c * tl.cstride + b;
      const int o = (j - k) * R + k;
#pragma unroll
      for (int q = 0; q < R; ++q) base[(o + q * L) * tl.ldb] = v[s][q];
    }
  }
  __syncthreads();
}

// Full in-LDS FFT of every column of the tile (forward: exp(-), inverse: exp(+),
// both unnormalised).  MAXV bounds (C*n*B)/MGK_NT, the values a thread stages.
template <int MAXV, typename T>
__device__ void tile_fft(cplx<T>* __restrict__ buf, const Tile& tl, const FftPlan& pl,
                         const cplx<T>* __restrict__ tw, bool inv) {
  int L = 1;
  for (int p = 0; p < pl.npass; ++p) {
    const int r = pl.radix[p];
    switch (r) {
      case 8: stockham_pass<8, MAXV>(buf, tl, L, tw, inv); break;
      case 4: stockham_pass<4, MAXV>(buf, tl, L, tw, inv); break;
      case 2: stockham_pass<2, MAXV>(buf, tl, L, tw, inv); break;
      case 3: stockham_pass<3, MAXV>(buf, tl, L, tw, inv); break;
      default: stockham_pass<5, MAXV>(buf, tl, L, tw, inv); break;
    }
    L *= r;
  }
}

// ---------------------------------------------------------------------------
// Compile-time specialised tile FFT for lengths 2^k and 3 * 2^k.  Same
// Stockham passes and data layout as above with row stride B and batch
// stride N*B, but every extent is a template constant, so the index
// arithmetic folds into shifts, immediates and constant divisions and the
// staged values stay in few registers.
// ---------------------------------------------------------------------------
constexpr int ct_log2(int v) { return v <= 1 ? 0 : 1 + ct_log2(v / 2); }

// radix of the pass that starts at sub-transform length L: a radix-3 pass
// first for lengths 3 * 2^k, then 8s, then 4/2.  With `wide` (FP32 tiles)
// 256 = 16 * 16 and 128 = 16 * 8 take two passes instead of three; a
// 16-point butterfly holds 64 VGPRs of FP64 values, which costs the FP64
// sweeps more occupancy than the saved LDS round trip gains.
constexpr int ct_radix(int N, int L, bool wide) {
  if ((N / L) % 3 == 0) return 3;
  const int k = ct_log2(N / L);
  if (wide && (k == 8 || k == 7 || k == 4)) return 16;
  return (k % 3 == 0) ? 8 : ((k % 3 == 2) ? 4 : (k >= 4 ? 4 : 2));
}

// Axis lengths with compile-time sweeps: 64..512 in powers of two and the
// mixed lengths 96, 192, 384 (3 * 2^k; one radix-3 pass).
constexpr bool ct_len(int n) {
  return n == 64 || n == 96 || n == 128 || n == 192 || n == 256 || n == 384 || n == 512;
}

// columns of a compile-time tile of at most `tile` points along an axis of
// length n: the largest power of two c with n * c <= tile
constexpr int ct_cols(int tile, int n) { return (2 * n <= tile) ? 2 * ct_cols(tile / 2, n) : 1; }

// Calls f(std::integral_constant<int, n>{}) for a length with compile-time
// sweeps; false (and no call) for any other length.
template <typename F>
bool ct_dispatch(int n, F&& f) {
  switch (n) {
    case 64: f(std::integral_constant<int, 64>{}); return true;
    case 96: f(std::integral_constant<int, 96>{}); return true;
    case 128: f(std::integral_constant<int, 128>{}); return true;
    case 192: f(std::integral_constant<int, 192>{}); return true;
    case 256: f(std::integral_constant<int, 256>{}); return true;
    case 384: f(std::integral_constant<int, 384>{}); return true;
    case 512: f(std::integral_constant<int, 512>{}); return true;
    default: return false;
  }
}

// points per compile-time tile: 32 KB of LDS in FP32, 32 KB in FP64 (so four
// tiles fit the 160 KB of a CU in either precision)
template <typename T>
constexpr int ct_tile() {
  return sizeof(T) == 8 ? 2048 : 4096;
}

// pencils per pencil-pass tile: largest power of two pp with pp*C*Q <= tile
constexpr int ct_pp(int tile, int C, int Q) { return (2 * C * Q <= tile) ? 2 * ct_pp(tile / 2, C, Q) : 1; }

// Optional LDS padding of the compile-time tiles: one spare complex slot per
// 16 (index i lives at i + i/16).  Stockham passes write with stride R*B; for
// narrow tiles (the pencil pass has B = 1 or 2 columns) that stride maps most
// lanes of a wave onto the same few LDS banks, and the padding spreads them.
template <bool PAD>
__device__ __forceinline__ int tpad(int i) {
  return PAD ? i + (i >> 4) : i;
}
constexpr int tpad_size(int n) { return n + n / 16 + 1; }

template <int N, int B, int C, int R, int L, bool PAD, int NT, typename T>
__device__ __forceinline__ void ct_pass(cplx<T>* __restrict__ buf, const cplx<T>* __restrict__ tw, bool inv) {
  constexpr int M = N / R;
  constexpr int NB = C * M * B;
  constexpr int S = (NB + NT - 1) / NT;
  constexpr int TS = N / (L * R);
  cplx<T> v[S][R];
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const int t = threadIdx.x + s * NT;
    if (NB % NT == 0 || t < NB) {
      const int b = t % B, u = t / B, j = u % M, c = u / M;
      const int base = c * (N * B) + b;
#pragma unroll
      for (int q = 0; q < R; ++q) v[s][q] = buf[tpad<PAD>(base + (j + q * M) * B)];
    }
  }
  __syncthreads();
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const int t = threadIdx.x + s * NT;
    if (NB % NT == 0 || t < NB) {
      const int b = t % B, u = t / B, j = u % M, c = u / M;
      const int k = j % L;
      if (L > 1) {
#pragma unroll
        for (int q = 1; q < R; ++q) {
          cplx<T> w = tw[q * k * TS];
          if (inv) w.y = -w.y;
          v[s][q] = cmul(v[s][q], w);
        }
      }
      dftR<R>(v[s], inv);
      const int base = c * (N * B) + b;
      const int o = (j - k) * R + k;
#pragma unroll
      for (int q = 0; q < R; ++q) buf[tpad<PAD>(base + (o + q * L) * B)] = v[s][q];
    }
  }
  __syncthreads();
}

template <int N, int B, int C, int L, bool PAD, int NT, typename T>
__device__ __forceinline__ void ct_fft_rec(cplx<T>* __restrict__ buf, const cplx<T>* __restrict__ tw, bool inv) {
  if constexpr (L < N) {
    constexpr int R = ct_radix(N, L, sizeof(T) == 4);
    ct_pass<N, B, C, R, L, PAD, NT>(buf, tw, inv);
    ct_fft_rec<N, B, C, L * R, PAD, NT>(buf, tw, inv);
  }
}

// element (c, j, b) at buf[tpad<PAD>(c*N*B + j*B + b)]; NT threads in the workgroup
template <int N, int B, int C, bool PAD = false, int NT = MGK_NT, typename T>
__device__ __forceinline__ void tile_fft_ct(cplx<T>* __restrict__ buf, const cplx<T>* __restrict__ tw, bool inv) {
  ct_fft_rec<N, B, C, 1, PAD, NT>(buf, tw, inv);
}

// ---------------------------------------------------------------------------
// XCD-aware tile order.  Workgroups are dispatched round-robin over the 8 XCDs
// of the MI355X, each with its own L2.  Mapping the linear workgroup id w to
// tile (w % 8) * per + w / 8 hands every XCD a contiguous range of tiles, so
// neighbouring tiles that share 128-byte lines are served by one L2 instead of
// being fetched from HBM once per XCD.
// ---------------------------------------------------------------------------
#define MGK_NXCD 8

// most simulations one batched LS launch serves (kernel-argument tables)
#define MGK_MAXB 8
__device__ __forceinline__ int xcd_tile(int w, int n) {
  const int per = n / MGK_NXCD;
  if (w >= per * MGK_NXCD) return w;
  return (w % MGK_NXCD) * per + w / MGK_NXCD;
}

// ---------------------------------------------------------------------------
// Block reductions in FP64 (wavefront = 64 lanes on CDNA).
// Deterministic: fixed shuffle tree, then fixed order over waves.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Sums NV values across the block; result valid in thread 0.  red: LDS of
// NV * (MGK_NT/64) doubles.
template <int NV, int NT = MGK_NT>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* red) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int i = 0; i < NV; ++i) v[i] = wave_sum(v[i]);
  __syncthreads();
  if (lane == 0) {
#pragma unroll
    for (int i = 0; i < NV; ++i) red[i * (NT / 64) + w] = v[i];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      double s = 0.0;
      for (int k = 0; k < NT / 64; ++k) s += red[i * (NT / 64) + k];
      v[i] = s;
    }
  }
  __syncthreads();
}

// Host-side helpers ---------------------------------------------------------
#define MGK_OK 0
#define MGK_NOT_CONVERGED 1
#define MGK_ERR_ARG -1
#define MGK_ERR_HIP -2
#define MGK_ERR_WORKSPACE -3
#define MGK_ERR_NONFINITE -4
#define MGK_ERR_SOLVER -5
