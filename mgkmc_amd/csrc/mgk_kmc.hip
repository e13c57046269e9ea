// KMC hot path on gfx950: STZ barriers, Arrhenius rates, FP64 rate sums and
// event selection, stream compaction, flips.
//
// Reference: mgkmc/stz/barriers.py:4-100, mgkmc/stz/kmc.py:4-103,
// mgkmc/stz/cascade.py:4-113 and the 2D twins in
// mgkmc/kmc_simulator_functions.py:5-218.
//
// Device layout (v = voxel index in reference order, x-major):
//   Q, Q0        [m][v]            FP64
//   catalog      [k][m][v]         FP64, k over CK parameters per mode:
//                 CK = 2 planar (gxx, gxy):  [[gxx,gxy,0],[gxy,-gxx,0],[0,0,0]]
//                 CK = D*D general, row-major ij
//   sigma        sym SoA [c][v] (FP32 or FP64, from the LS solver) or
//                general [ij][v] FP64
//   soft         [2][v] (g_p, g_t); tlast [v]; prev [ij][v] (D*D, directional)
//   Rv           [v] per-voxel FP64 rate sum (modes summed in order m = 0..M-1)
//   blocksums    [b] FP64 sum of Rv over voxels [b*VB, (b+1)*VB)
// Every floating-point step follows the reference's operation order with
// contraction off, so Q and rates agree with the reference to the last bits
// of exp().
#include "mgk_common.h"
#include "mgk_host.h"
#include <algorithm>
#include <math.h>
#include <string.h>

#define KMC_VPT 4                       // voxels per thread in the fused kernel
#define KMC_VB (MGK_NT * KMC_VPT)       // voxels per block-sum segment

struct KmcParams {
  double volume;       // barrier work volume (pixel^3, nm^3)
  double rate_volume;  // prefactor volume of the rates (1 unless scale_rate_by_volume)
  double nu0;
  double temperature;  // K (scalar); ignored if Tfield != NULL
  double cap;          // softening cap
  double tau;          // transient softening decay time (inf allowed)
  double time;         // current simulation time
  int scheme;          // 0 isotropic, 1 directional
  int cascade;         // 1: only q > 0 modes are events
};

#define KMC_SIG_SYM 0
#define KMC_SIG_GEN 1

template <int D>
__device__ __forceinline__ int sym_index(int i, int j) {
  // 3D: xx,yy,zz,yz,xz,xy; 2D: xx,yy,xy
  if (D == 3) {
    if (i == j) return i;
    const int s = i + j;  // (1,2)->3 (0,2)->2 (0,1)->1
    return s == 3 ? 3 : (s == 2 ? 4 : 5);
  } else {
    if (i == j) return i;
    return 2;
  }
}

#pragma clang fp contract(off)

// Per-voxel barrier state: sigma in (i, j) order, softening, previous
// direction, beta.  mode_q() is the barrier of one mode; the fused pass and the
// selection both call it, so their Q values are bit-identical.
template <int D, int CK, int SIGK, typename TS>
struct KmcVox {
  double sg[D * D];
  double pv[D * D];
  double g_base, beta;
  bool dir;

  __device__ __forceinline__ void load(size_t v, size_t NV, const TS* __restrict__ sig,
                                       const double* __restrict__ soft, const double* __restrict__ tlast,
                                       const double* __restrict__ prev, const double* __restrict__ Tfield,
                                       const KmcParams& P) {
    const double kB = 8.617e-5;
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
      for (int j = 0; j < D; ++j) {
        if (SIGK == KMC_SIG_SYM) sg[i * D + j] = (double)sig[(size_t)sym_index<D>(i, j) * NV + v];
        else sg[i * D + j] = (double)sig[(size_t)(i * D + j) * NV + v];
      }
    // transient softening (barriers.py:37-49)
    const double g_t = soft[NV + v];
    const double t_last = tlast[v];
    double g_tc = 0.0;
    if (g_t > 0) {
      if (isinf(P.tau) && P.tau > 0) {
        g_tc = g_t;
      } else if (isinf(t_last) && t_last < 0) {
        g_tc = 0.0;
      } else {
        double dt = P.time - t_last;
        if (dt < 0) dt = 0;
        g_tc = g_t * exp(-dt / P.tau);
      }
    }
    g_base = soft[v] + g_tc;
    if (P.cap > 0 && g_base > P.cap) g_base = P.cap;
    dir = P.scheme == 1;
    if (dir) {
#pragma unroll
      for (int k = 0; k < D * D; ++k) pv[k] = prev[(size_t)k * NV + v];
    }
    const double T = Tfield ? Tfield[v] : P.temperature;
    beta = (T > 0.0) ? 1.0 / (kB * T) : (double)INFINITY;
  }

  __device__ __forceinline__ double mode_q(int m, size_t v, size_t NV, size_t MNV, const double* __restrict__ Q0,
                                           const double* __restrict__ cat, const KmcParams& P) const {
    constexpr int DD = D * D;
    const size_t mo = (size_t)m * NV + v;
    double c[DD];
    if (CK == 2) {
      const double gxx = cat[mo], gxy = cat[MNV + mo];
#pragma unroll
      for (int k = 0; k < DD; ++k) c[k] = 0.0;
      c[0] = gxx;
      c[1] = gxy;
      c[D] = gxy;
      c[D + 1] = -gxx;
    } else {
#pragma unroll
      for (int k = 0; k < DD; ++k) c[k] = cat[(size_t)k * MNV + mo];
    }
    double modifier = 1.0;
    if (dir) {
      double dp = 0.0, nm = 0.0, npv2 = 0.0;
#pragma unroll
      for (int k = 0; k < DD; ++k) {
        dp = dp + c[k] * pv[k];
        nm = nm + c[k] * c[k];
        npv2 = npv2 + pv[k] * pv[k];
      }
      const double n_prev = sqrt(npv2), n_mode = sqrt(nm);
      if (!(n_prev < 1e-12 || n_mode < 1e-12)) {
        const double ct = dp / (n_mode * n_prev);
        const double a = 1.0 + ct;
        modifier = (a * a) / 4.0;
      }
    }
    const double g_eff = modifier * g_base;
    double w = 0.0;
#pragma unroll
    for (int k = 0; k < DD; ++k) w = w + sg[k] * c[k];
    const double w_val = ((0.5 * P.volume) * (w / 1e9)) * 6.241509;
    return Q0[mo] * exp(-g_eff) - w_val;
  }
};

// Arrhenius rate of a mode (kmc.py:52-60); 0 and *inc = false for a non-event
__device__ __forceinline__ double kmc_rate(double q, double beta, const KmcParams& P, bool* inc) {
  *inc = !P.cascade || q > 0;
  if (!*inc) return 0.0;
  return (q <= 0) ? (P.rate_volume * P.nu0) : ((P.rate_volume * P.nu0) * exp(-q * beta));
}

// Fused barrier + rate kernel: one thread per voxel (KMC_VPT voxels/thread),
// per-voxel FP64 rate sums (modes in order) and per-segment sums.  Q is
// written only when requested (the drop-in compute_barrier needs it; the
// device KMC loop does not).
template <int D, int CK, int SIGK, typename TS>
__device__ __forceinline__ void kmc_rates_block(
    size_t NV, int M, const TS* __restrict__ sig, const double* __restrict__ Q0, const double* __restrict__ cat,
    const double* __restrict__ soft, const double* __restrict__ tlast, const double* __restrict__ prev,
    const double* __restrict__ Tfield, const KmcParams& P, double* __restrict__ Q, double* __restrict__ Rv,
    double* __restrict__ blocksums) {
  const size_t MNV = (size_t)M * NV;
  __shared__ double red[MGK_NT / 64];
  double bsum = 0.0;
  const size_t base = (size_t)blockIdx.x * KMC_VB;
  for (int s = 0; s < KMC_VPT; ++s) {
    const size_t v = base + (size_t)s * MGK_NT + threadIdx.x;
    if (v >= NV) continue;
    KmcVox<D, CK, SIGK, TS> vx;
    vx.load(v, NV, sig, soft, tlast, prev, Tfield, P);
    double rsum = 0.0;
    for (int m = 0; m < M; ++m) {
      const double q = vx.mode_q(m, v, NV, MNV, Q0, cat, P);
      if (Q) Q[(size_t)m * NV + v] = q;
      bool inc;
      const double r = kmc_rate(q, vx.beta, P, &inc);
      if (inc) rsum = rsum + r;
    }
    Rv[v] = rsum;
    bsum += rsum;
  }
  double acc[1] = {bsum};
  block_sum<1>(acc, red);
  if (threadIdx.x == 0) blocksums[blockIdx.x] = acc[0];
}

template <int D, int CK, int SIGK, typename TS>
__global__ __launch_bounds__(MGK_NT) void k_kmc_barrier_rates(
    size_t NV, int M, const TS* __restrict__ sig, const double* __restrict__ Q0, const double* __restrict__ cat,
    const double* __restrict__ soft, const double* __restrict__ tlast, const double* __restrict__ prev,
    const double* __restrict__ Tfield, KmcParams P, double* __restrict__ Q, double* __restrict__ Rv,
    double* __restrict__ blocksums) {
  kmc_rates_block<D, CK, SIGK, TS>(NV, M, sig, Q0, cat, soft, tlast, prev, Tfield, P, Q, Rv, blocksums);
}

// Rates of an existing Q field (drop-in compute_rates): per-mode rate array in
// reference flat order v*M + m, included flags, per-voxel sums, block sums.
__global__ __launch_bounds__(MGK_NT) void k_kmc_rates_from_q(size_t NV, int M, const double* __restrict__ Qvm,
                                                              const double* __restrict__ Tfield, KmcParams P,
                                                              double* __restrict__ rates, double* __restrict__ Rv,
                                                              double* __restrict__ blocksums) {
  const double kB = 8.617e-5;
  __shared__ double red[MGK_NT / 64];
  double bsum = 0.0;
  const size_t base = (size_t)blockIdx.x * KMC_VB;
  for (int s = 0; s < KMC_VPT; ++s) {
    const size_t v = base + (size_t)s * MGK_NT + threadIdx.x;
    if (v >= NV) continue;
    const double T = Tfield ? Tfield[v] : P.temperature;
    const double beta = (T > 0.0) ? 1.0 / (kB * T) : (double)INFINITY;
    double rsum = 0.0;
    for (int m = 0; m < M; ++m) {
      const double q = Qvm[v * M + m];  // reference (x,y,z,m) layout
      const bool inc = !P.cascade || q > 0;
      double r = 0.0;
      if (inc) {
        r = (q <= 0) ? (P.rate_volume * P.nu0) : ((P.rate_volume * P.nu0) * exp(-q * beta));
        rsum = rsum + r;
      }
      rates[v * M + m] = inc ? r : -1.0;  // -1 marks "not an event"
    }
    Rv[v] = rsum;
    bsum += rsum;
  }
  double acc[1] = {bsum};
  block_sum<1>(acc, red);
  if (threadIdx.x == 0) blocksums[blockIdx.x] = acc[0];
}

// per-block sums of an arbitrary FP64 array (segments of KMC_VB items)
__global__ __launch_bounds__(MGK_NT) void k_kmc_block_sums(size_t n, const double* __restrict__ a,
                                                            double* __restrict__ blocksums) {
  __shared__ double red[MGK_NT / 64];
  double bsum = 0.0;
  const size_t base = (size_t)blockIdx.x * KMC_VB;
  for (int s = 0; s < KMC_VPT; ++s) {
    const size_t v = base + (size_t)s * MGK_NT + threadIdx.x;
    if (v < n) bsum += a[v];
  }
  double acc[1] = {bsum};
  block_sum<1>(acc, red);
  if (threadIdx.x == 0) blocksums[blockIdx.x] = acc[0];
}

__global__ __launch_bounds__(MGK_NT) void k_kmc_total(const double* __restrict__ blocksums, int nb,
                                                       double* __restrict__ out) {
  __shared__ double red[MGK_NT / 64];
  double acc[1] = {0.0};
  for (int i = threadIdx.x; i < nb; i += MGK_NT) acc[0] += blocksums[i];
  block_sum<1>(acc, red);
  if (threadIdx.x == 0) out[0] = acc[0];
}

struct KmcSel {
  long long item;    // selected voxel (or array item)
  int mode;          // selected mode within the voxel (0 for plain arrays)
  int fallthrough;   // 1 if no running sum reached r (last event chosen)
  double r;
  double cum_before;  // running sum before the selected event
  double cum_at;      // running sum including it
  double q;           // barrier of the selected mode
};

// Find the first index i with cum[i] >= r over `n` items with per-item sums
// `a` (grouped in segments of KMC_VB with sums `bs`).  Single workgroup.
// Returns item and the running sum before it (through LDS).
__device__ void first_reaching(const double* __restrict__ a, size_t n, const double* __restrict__ bs, int nb,
                               double r, long long* item_out, double* before_out) {
  __shared__ double tsum[MGK_NT];
  __shared__ int tmin;
  __shared__ long long seg;
  __shared__ double segbase;
  // level 1: segments
  const int cb = (nb + MGK_NT - 1) / MGK_NT;
  double s = 0.0;
  for (int k = 0; k < cb; ++k) {
    const int i = threadIdx.x * cb + k;
    if (i < nb) s += bs[i];
  }
  tsum[threadIdx.x] = s;
  if (threadIdx.x == 0) tmin = MGK_NT;
  __syncthreads();
  if (threadIdx.x == 0) {
    double c = 0.0;
    for (int t = 0; t < MGK_NT; ++t) {
      const double nc = c + tsum[t];
      tsum[t] = c;  // exclusive prefix
      c = nc;
    }
  }
  __syncthreads();
  // the first thread whose chunk reaches r scans it
  {
    double c = tsum[threadIdx.x];
    long long found = -1;
    double fb = 0.0;
    for (int k = 0; k < cb; ++k) {
      const int i = threadIdx.x * cb + k;
      if (i >= nb) break;
      const double nc = c + bs[i];
      if (nc >= r) {
        found = i;
        fb = c;
        break;
      }
      c = nc;
    }
    if (found >= 0) atomicMin(&tmin, (int)threadIdx.x);
    __syncthreads();
    if (found >= 0 && (int)threadIdx.x == tmin) {
      seg = found;
      segbase = fb;
    }
    if (threadIdx.x == 0 && tmin == MGK_NT) {
      seg = -1;
      segbase = 0.0;
    }
    __syncthreads();
  }
  if (seg < 0) {
    if (threadIdx.x == 0) {
      *item_out = -1;
      *before_out = 0.0;
    }
    __syncthreads();
    return;
  }
  // level 2: items of the segment, sequential running sum
  const size_t lo = (size_t)seg * KMC_VB;
  const size_t hi = std::min(n, lo + (size_t)KMC_VB);
  const int per = KMC_VB / MGK_NT;
  double ps = 0.0;
  for (int k = 0; k < per; ++k) {
    const size_t i = lo + (size_t)threadIdx.x * per + k;
    if (i < hi) ps += a[i];
  }
  tsum[threadIdx.x] = ps;
  if (threadIdx.x == 0) tmin = MGK_NT;
  __syncthreads();
  if (threadIdx.x == 0) {
    double c = segbase;
    for (int t = 0; t < MGK_NT; ++t) {
      const double nc = c + tsum[t];
      tsum[t] = c;
      c = nc;
    }
  }
  __syncthreads();
  {
    double c = tsum[threadIdx.x];
    long long found = -1;
    double fb = 0.0;
    for (int k = 0; k < per; ++k) {
      const size_t i = lo + (size_t)threadIdx.x * per + k;
      if (i >= hi) break;
      const double nc = c + a[i];
      if (nc >= r) {
        found = (long long)i;
        fb = c;
        break;
      }
      c = nc;
    }
    if (found >= 0) atomicMin(&tmin, (int)threadIdx.x);
    __syncthreads();
    if (found >= 0 && (int)threadIdx.x == tmin) {
      *item_out = found;
      *before_out = fb;
    }
    if (threadIdx.x == 0 && tmin == MGK_NT) {
      *item_out = -1;
      *before_out = 0.0;
    }
    __syncthreads();
  }
}

// Selection for the fused path: voxel via the two-level search over Rv, then
// the mode by recomputing that voxel's barriers and rates in order with the
// same code as the fused pass.  r = total * u with total read from `total`
// (device) so the host and device use one value.
template <int D, int CK, int SIGK, typename TS>
__device__ __forceinline__ void kmc_select_block(
    size_t NV, int M, const double* __restrict__ Rv, const double* __restrict__ bs, int nb,
    const double* __restrict__ total, double u, const TS* __restrict__ sig, const double* __restrict__ Q0,
    const double* __restrict__ cat, const double* __restrict__ soft, const double* __restrict__ tlast,
    const double* __restrict__ prev, const double* __restrict__ Tfield, const KmcParams& P,
    KmcSel* __restrict__ out, KmcSel* __restrict__ out2) {
  __shared__ long long item;
  __shared__ double before;
  double r = (*total) * u;
  if (!(r > 0.0)) r = 4.9406564584124654e-324;  // r == 0: first positive-rate event
  first_reaching(Rv, NV, bs, nb, r, &item, &before);
  __syncthreads();
  if (threadIdx.x != 0) return;
  const size_t MNV = (size_t)M * NV;
  long long v = item;
  double c = before;
  int fall = 0;
  if (v < 0) {
    // running sum never reached r (rounding at the very end): the reference
    // returns its last event; take the last voxel with a positive sum
    fall = 1;
    v = (long long)NV - 1;
    while (v > 0 && !(Rv[v] > 0.0)) --v;
    c = 0.0;
  }
  KmcVox<D, CK, SIGK, TS> vx;
  vx.load((size_t)v, NV, sig, soft, tlast, prev, Tfield, P);
  int mode = -1, last_inc = -1;
  double cb = c, ca = c, qsel = 0.0, qlast = 0.0;
  for (int m = 0; m < M; ++m) {
    const double q = vx.mode_q(m, (size_t)v, NV, MNV, Q0, cat, P);
    bool inc;
    const double rr = kmc_rate(q, vx.beta, P, &inc);
    if (!inc) continue;
    last_inc = m;
    qlast = q;
    const double nc = c + rr;
    if (!fall && nc >= r) {
      mode = m;
      cb = c;
      ca = nc;
      qsel = q;
      break;
    }
    c = nc;
  }
  if (mode < 0) {
    mode = last_inc < 0 ? 0 : last_inc;
    qsel = qlast;
    ca = c;
    fall = fall | 2;
  }
  KmcSel o;
  o.item = v;
  o.mode = mode;
  o.fallthrough = fall;
  o.r = r;
  o.cum_before = cb;
  o.cum_at = ca;
  o.q = qsel;
  *out = o;
  if (out2) *out2 = o;
}

template <int D, int CK, int SIGK, typename TS>
__global__ __launch_bounds__(MGK_NT) void k_kmc_select_voxel(
    size_t NV, int M, const double* __restrict__ Rv, const double* __restrict__ bs, int nb,
    const double* __restrict__ total, double u, const TS* __restrict__ sig, const double* __restrict__ Q0,
    const double* __restrict__ cat, const double* __restrict__ soft, const double* __restrict__ tlast,
    const double* __restrict__ prev, const double* __restrict__ Tfield, KmcParams P, KmcSel* __restrict__ out) {
  kmc_select_block<D, CK, SIGK, TS>(NV, M, Rv, bs, nb, total, u, sig, Q0, cat, soft, tlast, prev, Tfield, P, out,
                                    nullptr);
}

// Selection over a plain rates array (drop-in select_event): r given.
__global__ __launch_bounds__(MGK_NT) void k_kmc_select_array(size_t n, const double* __restrict__ a,
                                                              const double* __restrict__ bs, int nb, double r,
                                                              KmcSel* __restrict__ out) {
  __shared__ long long item;
  __shared__ double before;
  first_reaching(a, n, bs, nb, r, &item, &before);
  __syncthreads();
  if (threadIdx.x == 0) {
    out->item = item < 0 ? (long long)n - 1 : item;
    out->mode = 0;
    out->fallthrough = item < 0;
    out->r = r;
    out->cum_before = before;
    out->cum_at = item < 0 ? before : before + a[item];
    out->q = 0.0;
  }
}

// ---------------------------------------------------------------------------
// stream compaction of a predicate over the reference flat order v*M + m
//   kind 0: included rate (rates[i] >= 0) -> (rates, indices)
//   kind 1: Q[i] < thr                   -> indices
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool pred(int kind, const double* a, size_t i, double thr) {
  return kind == 0 ? (a[i] >= 0.0) : (a[i] < thr);
}

__global__ __launch_bounds__(MGK_NT) void k_compact_count(size_t n, int kind, const double* __restrict__ a,
                                                           double thr, long long* __restrict__ counts) {
  __shared__ int cnt;
  if (threadIdx.x == 0) cnt = 0;
  __syncthreads();
  const size_t base = (size_t)blockIdx.x * KMC_VB;
  int c = 0;
  for (int s = 0; s < KMC_VPT; ++s) {
    const size_t i = base + (size_t)s * MGK_NT + threadIdx.x;
    if (i < n && pred(kind, a, i, thr)) ++c;
  }
  atomicAdd(&cnt, c);
  __syncthreads();
  if (threadIdx.x == 0) counts[blockIdx.x] = cnt;
}

__global__ __launch_bounds__(MGK_NT) void k_compact_scan(long long* __restrict__ counts, int nb,
                                                          long long* __restrict__ total) {
  // exclusive scan in place, one workgroup (sequential over per-thread chunks)
  __shared__ long long ts[MGK_NT];
  const int cb = (nb + MGK_NT - 1) / MGK_NT;
  long long s = 0;
  for (int k = 0; k < cb; ++k) {
    const int i = threadIdx.x * cb + k;
    if (i < nb) s += counts[i];
  }
  ts[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    long long c = 0;
    for (int t = 0; t < MGK_NT; ++t) {
      const long long nc = c + ts[t];
      ts[t] = c;
      c = nc;
    }
    *total = c;
  }
  __syncthreads();
  long long c = ts[threadIdx.x];
  for (int k = 0; k < cb; ++k) {
    const int i = threadIdx.x * cb + k;
    if (i < nb) {
      const long long x = counts[i];
      counts[i] = c;
      c += x;
    }
  }
}

// order-preserving scatter: each block ranks its items in index order
__global__ __launch_bounds__(MGK_NT) void k_compact_scatter(size_t n, int kind, const double* __restrict__ a,
                                                             double thr, const long long* __restrict__ offs,
                                                             long long* __restrict__ idx_out,
                                                             double* __restrict__ val_out) {
  __shared__ int wsum[KMC_VPT][MGK_NT / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const size_t base = (size_t)blockIdx.x * KMC_VB;
  long long off = offs[blockIdx.x];
  for (int s = 0; s < KMC_VPT; ++s) {
    const size_t i = base + (size_t)s * MGK_NT + threadIdx.x;
    const bool p = i < n && pred(kind, a, i, thr);
    const unsigned long long bal = __ballot(p);
    const int below = __popcll(bal & ((lane == 0) ? 0ull : (~0ull >> (64 - lane))));
    if (lane == 0) wsum[s][w] = __popcll(bal);
    __syncthreads();
    int pre = 0;
    for (int k = 0; k < w; ++k) pre += wsum[s][k];
    int tot = 0;
    for (int k = 0; k < MGK_NT / 64; ++k) tot += wsum[s][k];
    if (p) {
      const long long o = off + pre + below;
      idx_out[o] = (long long)i;
      if (val_out) val_out[o] = a[i];
    }
    off += tot;
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// flip of mode m at voxel v (apply_flip_soa): eps_p += gamma, softening,
// event time, neighbour softening, previous strain direction.
// One workgroup; thread 0 does the voxel, threads 1..(3^D-1) the neighbours.
// ---------------------------------------------------------------------------
struct KmcFlip {
  int nx, ny, nz;  // 2D: nz = 1
  int M;
  long long v;
  int m;
  double time, jp, jt, gmax, jn;
  int eps_div;  // eps_p += mode / |eps_div|; eps_div < 0: that increment only; 0: no eps_p update
};

// the flip by the calling thread when `lead` (every update of one flip is done by
// one thread, in order)
template <int D, int CK, typename TE>
__device__ __forceinline__ void kmc_flip_apply(const KmcFlip& F, const double* __restrict__ cat, TE* __restrict__ epsp,
                                               float* __restrict__ shadow, double* __restrict__ soft,
                                               double* __restrict__ tlast, double* __restrict__ prev,
                                               double* __restrict__ gamma_out, bool lead) {
  constexpr int DD = D * D;
  const size_t NV = (size_t)F.nx * F.ny * F.nz;
  const size_t MNV = (size_t)F.M * NV;
  const size_t mo = (size_t)F.m * NV + (size_t)F.v;
  double c[DD];
  if (CK == 2) {
#pragma unroll
    for (int k = 0; k < DD; ++k) c[k] = 0.0;
    const double gxx = cat[mo], gxy = cat[MNV + mo];
    c[0] = gxx;
    c[1] = gxy;
    c[D] = gxy;
    c[D + 1] = -gxx;
  } else {
#pragma unroll
    for (int k = 0; k < DD; ++k) c[k] = cat[(size_t)k * MNV + mo];
  }
  double s;
  if (D == 3) {
    const double e11 = c[0], e22 = c[4], e33 = c[8], e12 = c[1], e13 = c[2], e23 = c[5];
    s = (e12 * e12 + e23 * e23 + e13 * e13) +
        ((e22 - e33) * (e22 - e33) + (e33 - e11) * (e33 - e11) + (e11 - e22) * (e11 - e22)) / 6.0;
  } else {
    const double e11 = c[0], e22 = c[3], e12 = c[1];
    s = (e12 * e12) + (e22 * e22 + e11 * e11 + (e11 - e22) * (e11 - e22)) / 6.0;
  }
  const int x = (int)(F.v / ((long long)F.ny * F.nz));
  const int y = (int)((F.v / F.nz) % F.ny);
  const int z = (int)(F.v % F.nz);
  if (lead) {
    const size_t v = (size_t)F.v;
    // eps_p += gamma (symmetric SoA storage)
    // symmetric part of the mode per stored component (exact for symmetric
    // modes: 0.5*(a+a) == a); the FP32 shadow, if any, gets the rounded sum
    double add[6];
    if (D == 3) {
      add[0] = c[0];
      add[1] = c[4];
      add[2] = c[8];
      add[3] = 0.5 * (c[5] + c[7]);
      add[4] = 0.5 * (c[2] + c[6]);
      add[5] = 0.5 * (c[1] + c[3]);
    } else {
      add[0] = c[0];
      add[1] = c[3];
      add[2] = 0.5 * (c[1] + c[2]);
    }
    constexpr int CS = (D == 3) ? 6 : 3;
    const double div = (double)(F.eps_div < 0 ? -F.eps_div : F.eps_div);
    if (F.eps_div != 0) {
#pragma unroll
      for (int k = 0; k < CS; ++k) {
        const TE nv = (TE)((double)epsp[k * NV + v] + add[k] / div);
        epsp[k * NV + v] = nv;
        if (shadow) shadow[k * NV + v] = (float)nv;
      }
    }
    if (F.eps_div < 0) return;
    double gp = soft[v] + F.jp * s;
    if (F.gmax > 0 && gp > F.gmax) gp = F.gmax;
    soft[v] = gp;
    soft[NV + v] = F.jt * s;
    tlast[v] = F.time;
    if (prev) {
#pragma unroll
      for (int k = 0; k < DD; ++k) prev[(size_t)k * NV + v] = c[k];
    }
    if (gamma_out) {
#pragma unroll
      for (int k = 0; k < DD; ++k) gamma_out[k] = c[k];
    }
  }
  if (lead && F.jn > 0.0 && F.eps_div >= 0) {
    // neighbours in the reference's (dx, dy[, dz]) loop order, sequentially
    // (on small periodic grids several offsets hit the same voxel)
    for (int dx = -1; dx <= 1; ++dx)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dz = (D == 3 ? -1 : 0); dz <= (D == 3 ? 1 : 0); ++dz) {
          if (dx == 0 && dy == 0 && dz == 0) continue;
          const int a = (x + dx + F.nx) % F.nx, b = (y + dy + F.ny) % F.ny;
          const int cc = (D == 3) ? (z + dz + F.nz) % F.nz : 0;
          const size_t w = ((size_t)a * F.ny + b) * F.nz + cc;
          double g = soft[w] + F.jn * F.jp * s;
          if (F.gmax > 0 && g > F.gmax) g = F.gmax;
          soft[w] = g;
          soft[NV + w] = soft[NV + w] + F.jn * F.jt * s;
        }
  }
}

template <int D, int CK, typename TE>
__device__ __forceinline__ void kmc_flip_body(const KmcFlip& F, const double* __restrict__ cat, TE* __restrict__ epsp,
                                              float* __restrict__ shadow, double* __restrict__ soft,
                                              double* __restrict__ tlast, double* __restrict__ prev,
                                              double* __restrict__ gamma_out) {
  kmc_flip_apply<D, CK, TE>(F, cat, epsp, shadow, soft, tlast, prev, gamma_out, threadIdx.x == 0);
}

template <int D, int CK, typename TE>
__global__ void k_kmc_flip(KmcFlip F, const double* __restrict__ cat, TE* __restrict__ epsp,
                           float* __restrict__ shadow, double* __restrict__ soft, double* __restrict__ tlast,
                           double* __restrict__ prev, double* __restrict__ gamma_out) {
  kmc_flip_body<D, CK, TE>(F, cat, epsp, shadow, soft, tlast, prev, gamma_out);
}

#pragma clang fp contract(on)

// write new catalog parameters and Q0 for all modes of voxel v
__device__ __forceinline__ void kmc_set_voxel_body(size_t NV, int M, int CK, long long v,
                                                   const double* __restrict__ newcat,
                                                   const double* __restrict__ newq0, double* __restrict__ cat,
                                                   double* __restrict__ Q0) {
  const size_t MNV = (size_t)M * NV;
  for (int i = threadIdx.x; i < M * CK; i += blockDim.x) {
    const int k = i / M, m = i % M;
    if (newcat) cat[(size_t)k * MNV + (size_t)m * NV + v] = newcat[(size_t)m * CK + k];
  }
  for (int m = threadIdx.x; m < M; m += blockDim.x)
    if (newq0) Q0[(size_t)m * NV + v] = newq0[m];
}

__global__ void k_kmc_set_voxel(size_t NV, int M, int CK, long long v, const double* __restrict__ newcat,
                                const double* __restrict__ newq0, double* __restrict__ cat,
                                double* __restrict__ Q0) {
  kmc_set_voxel_body(NV, M, CK, v, newcat, newq0, cat, Q0);
}

// ---------------------------------------------------------------------------
// Batched event-step passes: up to MGK_MAXB simulations of one grid, one
// launch per pass, member from the grid (rates: blockIdx.y, the others:
// blockIdx.x).  Each member runs the body of the single kernel on its own
// buffers, so its results equal those of the single launches bit for bit;
// small results also go to contiguous staging arrays (one copy for all).
// ---------------------------------------------------------------------------
struct KmcMem {
  const void* sig;
  const double *Q0, *cat, *soft, *tlast, *prev;
  KmcParams P;
  double *Rv, *bs, *total;
  KmcSel* sel;
  double u;
  // flip and redraw of the selected mode
  KmcFlip F;
  void* epsp;
  float* shadow;
  double* softw;
  double* tlastw;
  double* prevw;
  const double *newcat, *newq0;
  double *catw, *Q0w;
};

struct KmcBatch {
  KmcMem m[MGK_MAXB];
};

template <int D, int CK, int SIGK, typename TS>
__global__ __launch_bounds__(MGK_NT) void k_kmc_barrier_rates_b(size_t NV, int M, KmcBatch b) {
  const KmcMem& x = b.m[blockIdx.y];
  kmc_rates_block<D, CK, SIGK, TS>(NV, M, (const TS*)x.sig, x.Q0, x.cat, x.soft, x.tlast, x.prev, nullptr, x.P,
                                   nullptr, x.Rv, x.bs);
}

__global__ __launch_bounds__(MGK_NT) void k_kmc_total_b(int nb, KmcBatch b, double* __restrict__ tot_out) {
  const KmcMem& x = b.m[blockIdx.x];
  __shared__ double red[MGK_NT / 64];
  double acc[1] = {0.0};
  for (int i = threadIdx.x; i < nb; i += MGK_NT) acc[0] += x.bs[i];
  block_sum<1>(acc, red);
  if (threadIdx.x == 0) {
    x.total[0] = acc[0];
    tot_out[blockIdx.x] = acc[0];
  }
}

template <int D, int CK, int SIGK, typename TS>
__global__ __launch_bounds__(MGK_NT) void k_kmc_select_voxel_b(size_t NV, int M, int nb, KmcBatch b,
                                                                KmcSel* __restrict__ sel_out) {
  const KmcMem& x = b.m[blockIdx.x];
  kmc_select_block<D, CK, SIGK, TS>(NV, M, x.Rv, x.bs, nb, x.total, x.u, (const TS*)x.sig, x.Q0, x.cat, x.soft,
                                    x.tlast, x.prev, nullptr, x.P, x.sel, sel_out + blockIdx.x);
}

template <int D, int CK, typename TE>
__global__ void k_kmc_flip_b(KmcBatch b) {
  const KmcMem& x = b.m[blockIdx.x];
  kmc_flip_body<D, CK, TE>(x.F, x.catw, (TE*)x.epsp, x.shadow, x.softw, x.tlastw, x.prevw, nullptr);
}

__global__ void k_kmc_set_voxel_b(size_t NV, int M, int CK, KmcBatch b) {
  const KmcMem& x = b.m[blockIdx.x];
  kmc_set_voxel_body(NV, M, CK, x.F.v, x.newcat, x.newq0, x.catw, x.Q0w);
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
// the C ABI's structs (mgk_kmc_params, mgk_kmc_sel, mgk_kmc_loop)
#include "../../include/mgkmc_hip.h"

static KmcParams to_params(const mgk_kmc_params* p) {
  KmcParams P;
  P.volume = p->volume;
  P.rate_volume = p->rate_volume;
  P.nu0 = p->nu0;
  P.temperature = p->temperature;
  P.cap = p->cap;
  P.tau = p->tau;
  P.time = p->time;
  P.scheme = p->scheme;
  P.cascade = p->cascade;
  return P;
}

static int kmc_nblocks(size_t n) { return (int)((n + KMC_VB - 1) / KMC_VB); }

extern "C" int mgk_kmc_workspace_bytes(long long nvox, size_t* bytes) {
  const int nb = kmc_nblocks((size_t)nvox);
  // Rv, blocksums, total, selection record, compaction counts
  *bytes = mgk_align((size_t)nvox * 8) + mgk_align((size_t)nb * 8) + 256 + 256 + mgk_align((size_t)nb * 8) + 256;
  return MGK_OK;
}

struct KmcWs {
  double* Rv;
  double* bs;
  double* total;
  KmcSel* sel;
  long long* counts;
  long long* ctotal;
};

static KmcWs kmc_ws(void* ws, size_t nvox) {
  const int nb = kmc_nblocks(nvox);
  char* p = (char*)ws;
  KmcWs w;
  w.Rv = (double*)p;
  p += mgk_align(nvox * 8);
  w.bs = (double*)p;
  p += mgk_align((size_t)nb * 8);
  w.total = (double*)p;
  p += 256;
  w.sel = (KmcSel*)p;
  p += 256;
  w.counts = (long long*)p;
  p += mgk_align((size_t)nb * 8);
  w.ctotal = (long long*)p;
  return w;
}

template <int D, int CK, int SIGK, typename TS>
static void launch_rates(size_t NV, int M, const void* sig, const double* Q0, const double* cat, const double* soft,
                         const double* tlast, const double* prev, const double* Tf, const KmcParams& P, double* Q,
                         const KmcWs& w, hipStream_t s) {
  const int nb = kmc_nblocks(NV);
  k_kmc_barrier_rates<D, CK, SIGK, TS><<<nb, MGK_NT, 0, s>>>(NV, M, (const TS*)sig, Q0, cat, soft, tlast, prev, Tf,
                                                             P, Q, w.Rv, w.bs);
  k_kmc_total<<<1, MGK_NT, 0, s>>>(w.bs, nb, w.total);
}

template <int D, int CK, int SIGK, typename TS>
static void launch_select(size_t NV, int M, double u, const void* sig, const double* Q0, const double* cat,
                          const double* soft, const double* tlast, const double* prev, const double* Tf,
                          const KmcParams& P, const KmcWs& w, hipStream_t s) {
  const int nb = kmc_nblocks(NV);
  k_kmc_select_voxel<D, CK, SIGK, TS><<<1, MGK_NT, 0, s>>>(NV, M, w.Rv, w.bs, nb, w.total, u, (const TS*)sig, Q0,
                                                           cat, soft, tlast, prev, Tf, P, w.sel);
}

// sig_kind: 0 sym FP64, 1 sym FP32, 2 general FP64.  cat_kind: 0 planar, 1 general.
#define KMC_DISPATCH(FN, ...)                                                                      \
  do {                                                                                             \
    if (ndim == 3) {                                                                               \
      if (cat_kind == 0) {                                                                         \
        if (sig_kind == 0) FN<3, 2, KMC_SIG_SYM, double>(__VA_ARGS__);                             \
        else if (sig_kind == 1) FN<3, 2, KMC_SIG_SYM, float>(__VA_ARGS__);                         \
        else FN<3, 2, KMC_SIG_GEN, double>(__VA_ARGS__);                                           \
      } else {                                                                                     \
        if (sig_kind == 0) FN<3, 9, KMC_SIG_SYM, double>(__VA_ARGS__);                             \
        else if (sig_kind == 1) FN<3, 9, KMC_SIG_SYM, float>(__VA_ARGS__);                         \
        else FN<3, 9, KMC_SIG_GEN, double>(__VA_ARGS__);                                           \
      }                                                                                            \
    } else {                                                                                       \
      if (cat_kind == 0) {                                                                         \
        if (sig_kind == 0) FN<2, 2, KMC_SIG_SYM, double>(__VA_ARGS__);                             \
        else if (sig_kind == 1) FN<2, 2, KMC_SIG_SYM, float>(__VA_ARGS__);                         \
        else FN<2, 2, KMC_SIG_GEN, double>(__VA_ARGS__);                                           \
      } else {                                                                                     \
        if (sig_kind == 0) FN<2, 4, KMC_SIG_SYM, double>(__VA_ARGS__);                             \
        else if (sig_kind == 1) FN<2, 4, KMC_SIG_SYM, float>(__VA_ARGS__);                         \
        else FN<2, 4, KMC_SIG_GEN, double>(__VA_ARGS__);                                           \
      }                                                                                            \
    }                                                                                              \
  } while (0)

extern "C" int mgk_kmc_barrier_rates(int ndim, long long nvox, int M, int sig_kind, int cat_kind, const void* sig,
                                     const double* Q0, const double* cat, const double* soft, const double* tlast,
                                     const double* prev, const double* Tfield, const mgk_kmc_params* params,
                                     double* Q, void* workspace, void* stream, double* total_out) {
  if ((ndim != 2 && ndim != 3) || nvox < 1 || M < 1 || !sig || !Q0 || !cat || !soft || !tlast || !params ||
      !workspace || sig_kind < 0 || sig_kind > 2 || cat_kind < 0 || cat_kind > 1) {
    mgk_set_error("mgk_kmc_barrier_rates: bad argument", __FILE__, __LINE__);
    return MGK_ERR_ARG;
  }
  if (params->scheme == 1 && !prev) {
    mgk_set_error("directional softening needs prev_strain_dir", __FILE__, __LINE__);
    return MGK_ERR_ARG;
  }
  hipStream_t s = (hipStream_t)stream;
  const KmcParams P = to_params(params);
  const KmcWs w = kmc_ws(workspace, (size_t)nvox);
  MgkProfRec prof(s);
  const int nb = kmc_nblocks((size_t)nvox);
  const int CKn = cat_kind == 0 ? 2 : ndim * ndim;
  const double sigB = (sig_kind == 1 ? 4.0 : 8.0) * (sig_kind == 2 ? ndim * ndim : (ndim == 3 ? 6 : 3));
  const double bytes = (double)nvox * (sigB + 8.0 * M * (1 + CKn) + 8.0 * 3 + (params->scheme == 1 ? 8.0 * ndim * ndim : 0.0) +
                                       (Q ? 8.0 * M : 0.0) + 8.0);
  const int h = prof.begin(KID_KMC_RATES, 0, bytes);
  KMC_DISPATCH(launch_rates, (size_t)nvox, M, sig, Q0, cat, soft, tlast, prev, Tfield, P, Q, w, s);
  prof.end(h);
  MGK_HIP(hipGetLastError());
  if (total_out) {
    double* hp = (double*)mgk_pinned(sizeof(double));
    MGK_HIP(hipMemcpyAsync(hp, w.total, sizeof(double), hipMemcpyDeviceToHost, s));
    MGK_HIP(hipStreamSynchronize(s));
    *total_out = *hp;
    prof.commit(1);
  }
  (void)nb;
  return MGK_OK;
}

extern "C" int mgk_kmc_select(int ndim, long long nvox, int M, int sig_kind, int cat_kind, double u, const void* sig,
                              const double* Q0, const double* cat, const double* soft, const double* tlast,
                              const double* prev, const double* Tfield, const mgk_kmc_params* params,
                              void* workspace, void* stream, mgk_kmc_sel* out) {
  if ((ndim != 2 && ndim != 3) || nvox < 1 || M < 1 || !sig || !Q0 || !cat || !soft || !tlast || !params ||
      !workspace || !out || (params->scheme == 1 && !prev)) {
    mgk_set_error("mgk_kmc_select: bad argument", __FILE__, __LINE__);
    return MGK_ERR_ARG;
  }
  hipStream_t s = (hipStream_t)stream;
  const KmcParams P = to_params(params);
  const KmcWs w = kmc_ws(workspace, (size_t)nvox);
  MgkProfRec prof(s);
  const int h = prof.begin(KID_KMC_SELECT, 0, 8.0 * (KMC_VB + kmc_nblocks((size_t)nvox) + 2 * M));
  KMC_DISPATCH(launch_select, (size_t)nvox, M, u, sig, Q0, cat, soft, tlast, prev, Tfield, P, w, s);
  prof.end(h);
  MGK_HIP(hipGetLastError());
  KmcSel* hp = (KmcSel*)mgk_pinned(sizeof(KmcSel));
  MGK_HIP(hipMemcpyAsync(hp, w.sel, sizeof(KmcSel), hipMemcpyDeviceToHost, s));
  MGK_HIP(hipStreamSynchronize(s));
  prof.commit(1);
  out->item = hp->item;
  out->mode = hp->mode;
  out->fallthrough = hp->fallthrough;
  out->r = hp->r;
  out->cum_before = hp->cum_before;
  out->cum_at = hp->cum_at;
  out->q = hp->q;
  return MGK_OK;
}


// ---------------------------------------------------------------------------
// Cascade batch support (the native cascade loop): every mode with stored
// Q < thr, in the voxel-major scan order of find_unstable (v * M + m), in one
// count pass, one scan and one write pass; the batch's flips in levels of
// flips with disjoint neighbourhoods (levels assigned on the host); the
// redraws of the batch's voxels.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(MGK_NT) void k_below_count(size_t NV, int M, const double* __restrict__ Q, double thr,
                                                        long long* __restrict__ counts) {
  __shared__ int cnt;
  if (threadIdx.x == 0) cnt = 0;
  __syncthreads();
  const size_t v = (size_t)blockIdx.x * MGK_NT + threadIdx.x;
  int c = 0;
  if (v < NV)
    for (int m = 0; m < M; ++m) c += Q[(size_t)m * NV + v] < thr ? 1 : 0;
  if (c) atomicAdd(&cnt, c);
  __syncthreads();
  if (threadIdx.x == 0) counts[blockIdx.x] = cnt;
}

// inclusive prefix sum over the workgroup (Hillis-Steele in LDS; NT threads)
template <int NT, typename T>
__device__ __forceinline__ T block_incl_scan(T x, T* buf) {
  const int t = threadIdx.x;
  buf[t] = x;
  __syncthreads();
  for (int off = 1; off < NT; off <<= 1) {
    const T y = t >= off ? buf[t - off] : (T)0;
    __syncthreads();
    buf[t] += y;
    __syncthreads();
  }
  return buf[t];
}

// exclusive offsets of the block counts, and the total at offsets[nb]: each of
// the 1024 threads sums a contiguous run of counts, one scan over the runs
constexpr int kScanNT = 1024;
__global__ __launch_bounds__(kScanNT) void k_below_scan(const long long* __restrict__ counts, int nb,
                                                        long long* __restrict__ offsets) {
  __shared__ long long buf[kScanNT];
  const int per = (nb + kScanNT - 1) / kScanNT;
  const int b0 = min((int)threadIdx.x * per, nb), b1 = min(b0 + per, nb);
  long long c = 0;
  for (int b = b0; b < b1; ++b) c += counts[b];
  const long long inc = block_incl_scan<kScanNT>(c, buf);
  long long o = inc - c;
  for (int b = b0; b < b1; ++b) {
    offsets[b] = o;
    o += counts[b];
  }
  if (threadIdx.x == kScanNT - 1) offsets[nb] = inc;
}

__global__ __launch_bounds__(MGK_NT) void k_below_write(size_t NV, int M, const double* __restrict__ Q, double thr,
                                                        const long long* __restrict__ offsets, long long cap,
                                                        long long* __restrict__ vv, int* __restrict__ mm,
                                                        double* __restrict__ qq) {
  __shared__ int pre[MGK_NT];
  const size_t v = (size_t)blockIdx.x * MGK_NT + threadIdx.x;
  int c = 0;
  if (v < NV)
    for (int m = 0; m < M; ++m) c += Q[(size_t)m * NV + v] < thr ? 1 : 0;
  const int inc = block_incl_scan<MGK_NT>(c, pre);
  if (v >= NV || !c) return;
  long long o = offsets[blockIdx.x] + (inc - c);
  for (int m = 0; m < M; ++m) {
    const double q = Q[(size_t)m * NV + v];
    if (q < thr) {
      if (o < cap) {
        vv[o] = (long long)v;
        mm[o] = m;
        qq[o] = q;
      }
      ++o;
    }
  }
}

// the batch's flips level by level: flips of one level touch disjoint voxels (the
// flipped voxel and the neighbours it softens), and every voxel takes the updates of
// the flips that touch it in pair order, since a later flip touching a voxel that an
// earlier one touched sits on a higher level; each flip is done by one thread
constexpr int kLevelNT = 1024;
template <int D, int CK, typename TE>
__global__ __launch_bounds__(kLevelNT) void k_kmc_flip_levels(KmcFlip F, int nlev, const int* __restrict__ loff,
                                                              const long long* __restrict__ vs,
                                                              const int* __restrict__ ms,
                                                              const double* __restrict__ cat, TE* __restrict__ epsp,
                                                              float* __restrict__ shadow, double* __restrict__ soft,
                                                              double* __restrict__ tlast, double* __restrict__ prev) {
  for (int l = 0; l < nlev; ++l) {
    const int e = loff[l + 1];
    for (int i = loff[l] + (int)threadIdx.x; i < e; i += kLevelNT) {
      KmcFlip f = F;
      f.v = vs[i];
      f.m = ms[i];
      kmc_flip_apply<D, CK, TE>(f, cat, epsp, shadow, soft, tlast, prev, nullptr, true);
    }
    __syncthreads();
  }
}

__global__ void k_kmc_set_voxels(size_t NV, int M, int CK, const long long* __restrict__ vs,
                                 const double* __restrict__ red, size_t stride, int dirs, int bars,
                                 double* __restrict__ cat, double* __restrict__ Q0) {
  const double* r = red + (size_t)blockIdx.x * stride;
  kmc_set_voxel_body(NV, M, CK, vs[blockIdx.x], dirs ? r : nullptr, bars ? r + 2 * (size_t)M : nullptr, cat, Q0);
}

int mgk_kmc_below_all(const mgk_kmc_loop* L, const double* Q, double thr, long long* counts, long long* offsets,
                      hipStream_t s) {
  long long nvox = 1;
  for (int i = 0; i < L->ndim; ++i) nvox *= L->shape[i];
  const int nb = (int)((nvox + MGK_NT - 1) / MGK_NT);
  k_below_count<<<nb, MGK_NT, 0, s>>>((size_t)nvox, L->M, Q, thr, counts);
  k_below_scan<<<1, kScanNT, 0, s>>>(counts, nb, offsets);
  MGK_HIP(hipGetLastError());
  return MGK_OK;
}

int mgk_kmc_below_write(const mgk_kmc_loop* L, const double* Q, double thr, const long long* offsets, long long cap,
                        long long* vv, int* mm, double* qq, hipStream_t s) {
  long long nvox = 1;
  for (int i = 0; i < L->ndim; ++i) nvox *= L->shape[i];
  const int nb = (int)((nvox + MGK_NT - 1) / MGK_NT);
  k_below_write<<<nb, MGK_NT, 0, s>>>((size_t)nvox, L->M, Q, thr, offsets, cap, vv, mm, qq);
  MGK_HIP(hipGetLastError());
  return MGK_OK;
}

int mgk_kmc_flip_redraw_levels(const mgk_kmc_loop* L, int n, int nlev, const int* dloff, const long long* dlv,
                               const int* dlm, const long long* dv, double time, const double* dred, size_t stride,
                               hipStream_t s) {
  const int ndim = L->ndim;
  long long nvox = 1;
  for (int i = 0; i < ndim; ++i) nvox *= L->shape[i];
  KmcFlip F;
  F.nx = (int)L->shape[0];
  F.ny = (int)L->shape[1];
  F.nz = ndim == 3 ? (int)L->shape[2] : 1;
  F.M = L->M;
  F.v = 0;
  F.m = 0;
  F.time = time;
  F.jp = L->jp;
  F.jt = L->jt;
  F.gmax = L->kp.cap;
  F.jn = L->jn;
  F.eps_div = 1;
  if (n < 1 || nlev < 1 || nlev > n || (L->eps_precision != 0 && L->epsp_shadow)) {
    mgk_set_error("cascade flips: bad batch or an FP32 eps_p with an FP32 shadow", __FILE__, __LINE__);
    return MGK_ERR_ARG;
  }
#define MGK_LEVELS(DD, TE)                                                                                         \
  k_kmc_flip_levels<DD, 2, TE><<<1, kLevelNT, 0, s>>>(F, nlev, dloff, dlv, dlm, L->cat, (TE*)L->epsp, L->epsp_shadow, \
                                                      L->soft, L->tlast, L->prev)
  if (ndim == 3) {
    if (L->eps_precision == 0) MGK_LEVELS(3, double);
    else MGK_LEVELS(3, float);
  } else {
    if (L->eps_precision == 0) MGK_LEVELS(2, double);
    else MGK_LEVELS(2, float);
  }
#undef MGK_LEVELS
  k_kmc_set_voxels<<<n, 64, 0, s>>>((size_t)nvox, L->M, 2, dv, dred, stride, L->redraw_directions,
                                    L->redraw_barriers, L->cat, L->Q0);
  MGK_HIP(hipGetLastError());
  return MGK_OK;
}

// ---------------------------------------------------------------------------
// Batched passes of the event step for mgk_kmc_run_batch: every member shares
// ndim, shape, M and sig_kind (the caller checks); planar catalogs, no
// temperature field, no Q output, as the single loop calls them.
// ---------------------------------------------------------------------------
static KmcMem kmc_member(const mgk_kmc_loop* L, double time) {
  long long nvox = 1;
  for (int i = 0; i < L->ndim; ++i) nvox *= L->shape[i];
  mgk_kmc_params p = L->kp;
  p.time = time;
  const KmcWs w = kmc_ws(L->kmc_ws, (size_t)nvox);
  KmcMem x{};
  x.sig = L->sig;
  x.Q0 = L->Q0;
  x.cat = L->cat;
  x.soft = L->soft;
  x.tlast = L->tlast;
  x.prev = L->prev;
  x.P = to_params(&p);
  x.Rv = w.Rv;
  x.bs = w.bs;
  x.total = w.total;
  x.sel = w.sel;
  return x;
}

#define KMC_BATCH_DISPATCH(FN, ...)                                                   \
  do {                                                                                \
    if (ndim == 3) {                                                                  \
      if (sig_kind == 0) FN<3, 2, KMC_SIG_SYM, double>(__VA_ARGS__);                  \
      else if (sig_kind == 1) FN<3, 2, KMC_SIG_SYM, float>(__VA_ARGS__);              \
      else FN<3, 2, KMC_SIG_GEN, double>(__VA_ARGS__);                                \
    } else {                                                                          \
      if (sig_kind == 0) FN<2, 2, KMC_SIG_SYM, double>(__VA_ARGS__);                  \
      else if (sig_kind == 1) FN<2, 2, KMC_SIG_SYM, float>(__VA_ARGS__);              \
      else FN<2, 2, KMC_SIG_GEN, double>(__VA_ARGS__);                                \
    }                                                                                 \
  } while (0)

template <int D, int CK, int SIGK, typename TS>
static void launch_rates_b(size_t NV, int M, int nmem, const KmcBatch& b, double* tot, hipStream_t s) {
  const int nb = kmc_nblocks(NV);
  k_kmc_barrier_rates_b<D, CK, SIGK, TS><<<dim3(nb, nmem), MGK_NT, 0, s>>>(NV, M, b);
  k_kmc_total_b<<<nmem, MGK_NT, 0, s>>>(nb, b, tot);
}

template <int D, int CK, int SIGK, typename TS>
static void launch_select_b(size_t NV, int M, int nmem, const KmcBatch& b, KmcSel* out, hipStream_t s) {
  k_kmc_select_voxel_b<D, CK, SIGK, TS><<<nmem, MGK_NT, 0, s>>>(NV, M, kmc_nblocks(NV), b, out);
}

int mgk_kmc_rates_batch(int nmem, const mgk_kmc_loop* const* L, const double* times, double* dev_tot,
                        hipStream_t s) {
  const int ndim = L[0]->ndim, sig_kind = L[0]->sig_kind;
  long long nvox = 1;
  for (int i = 0; i < ndim; ++i) nvox *= L[0]->shape[i];
  KmcBatch b{};
  for (int k = 0; k < nmem; ++k) b.m[k] = kmc_member(L[k], times[k]);
  KMC_BATCH_DISPATCH(launch_rates_b, (size_t)nvox, L[0]->M, nmem, b, dev_tot, s);
  MGK_HIP(hipGetLastError());
  return MGK_OK;
}

int mgk_kmc_select_batch(int nmem, const mgk_kmc_loop* const* L, const double* times, const double* u,
                         void* dev_rec, hipStream_t s) {
  const int ndim = L[0]->ndim, sig_kind = L[0]->sig_kind;
  long long nvox = 1;
  for (int i = 0; i < ndim; ++i) nvox *= L[0]->shape[i];
  KmcBatch b{};
  for (int k = 0; k < nmem; ++k) {
    b.m[k] = kmc_member(L[k], times[k]);
    b.m[k].u = u[k];
  }
  KMC_BATCH_DISPATCH(launch_select_b, (size_t)nvox, L[0]->M, nmem, b, (KmcSel*)dev_rec, s);
  MGK_HIP(hipGetLastError());
  return MGK_OK;
}

int mgk_kmc_flip_redraw_batch(int nmem, const mgk_kmc_loop* const* L, const long long* v, const int* m,
                              const double* times, const double* dev_redraw, size_t stride, hipStream_t s) {
  const int ndim = L[0]->ndim, M = L[0]->M, prec = L[0]->eps_precision;
  long long nvox = 1;
  for (int i = 0; i < ndim; ++i) nvox *= L[0]->shape[i];
  KmcBatch b{};
  for (int k = 0; k < nmem; ++k) {
    const mgk_kmc_loop& x = *L[k];
    if (x.eps_precision != prec || (x.eps_precision != 0 && x.epsp_shadow) || v[k] < 0 || v[k] >= nvox ||
        m[k] < 0 || m[k] >= M) {
      mgk_set_error("mgk_kmc_run_batch: flip of a member out of range or with mixed eps_p precisions", __FILE__,
                    __LINE__);
      return MGK_ERR_ARG;
    }
    KmcMem& y = b.m[k];
    KmcFlip& F = y.F;
    F.nx = (int)x.shape[0];
    F.ny = (int)x.shape[1];
    F.nz = ndim == 3 ? (int)x.shape[2] : 1;
    F.M = M;
    F.v = v[k];
    F.m = m[k];
    F.time = times[k];
    F.jp = x.jp;
    F.jt = x.jt;
    F.gmax = x.kp.cap;
    F.jn = x.jn;
    F.eps_div = 1;
    y.epsp = x.epsp;
    y.shadow = x.epsp_shadow;
    y.softw = x.soft;
    y.tlastw = x.tlast;
    y.prevw = x.prev;
    y.catw = x.cat;
    y.Q0w = x.Q0;
    y.newcat = x.redraw_directions ? dev_redraw + (size_t)k * stride : nullptr;
    y.newq0 = x.redraw_barriers ? dev_redraw + (size_t)k * stride + 2 * (size_t)M : nullptr;
  }
  if (ndim == 3) {
    if (prec == 0) k_kmc_flip_b<3, 2, double><<<nmem, 64, 0, s>>>(b);
    else k_kmc_flip_b<3, 2, float><<<nmem, 64, 0, s>>>(b);
  } else {
    if (prec == 0) k_kmc_flip_b<2, 2, double><<<nmem, 64, 0, s>>>(b);
    else k_kmc_flip_b<2, 2, float><<<nmem, 64, 0, s>>>(b);
  }
  k_kmc_set_voxel_b<<<nmem, 64, 0, s>>>((size_t)nvox, M, 2, b);
  MGK_HIP(hipGetLastError());
  return MGK_OK;
}

size_t mgk_kmc_sel_rec_bytes() { return sizeof(KmcSel); }

void mgk_kmc_sel_unpack(const void* host_rec, mgk_kmc_sel* out) {
  const KmcSel* h = (const KmcSel*)host_rec;
  out->item = h->item;
  out->mode = h->mode;
  out->fallthrough = h->fallthrough;
  out->r = h->r;
  out->cum_before = h->cum_before;
  out->cum_at = h->cum_at;
  out->q = h->q;
}

// Rates of a Q field in the reference (v, m) layout: full per-mode rates
// (-1 for non-events), compacted (rates, indices) in reference order, total.
// rates_full: nvox*M doubles (device).  idx_out/val_out: capacity nvox*M.
extern "C" int mgk_kmc_rates_from_q(long long nvox, int M, const double* Qvm, const double* Tfield,
                                    const mgk_kmc_params* params, double* rates_full, long long* idx_out,
                                    double* val_out, void* workspace, void* stream, long long* count_out,
                                    double* total_out) {
  if (nvox < 1 || M < 1 || !Qvm || !params || !rates_full || !workspace) {
    mgk_set_error("mgk_kmc_rates_from_q: bad argument", __FILE__, __LINE__);
    return MGK_ERR_ARG;
  }
  hipStream_t s = (hipStream_t)stream;
  const KmcParams P = to_params(params);
  const KmcWs w = kmc_ws(workspace, (size_t)nvox);
  const int nbv = kmc_nblocks((size_t)nvox);
  k_kmc_rates_from_q<<<nbv, MGK_NT, 0, s>>>((size_t)nvox, M, Qvm, Tfield, P, rates_full, w.Rv, w.bs);
  k_kmc_total<<<1, MGK_NT, 0, s>>>(w.bs, nbv, w.total);
  const size_t n = (size_t)nvox * M;
  const int nbi = kmc_nblocks(n);
  long long* counts = nullptr;
  MGK_HIP(hipMallocAsync((void**)&counts, (size_t)(nbi + 1) * sizeof(long long), s));
  k_compact_count<<<nbi, MGK_NT, 0, s>>>(n, 0, rates_full, 0.0, counts);
  k_compact_scan<<<1, MGK_NT, 0, s>>>(counts, nbi, counts + nbi);
  k_compact_scatter<<<nbi, MGK_NT, 0, s>>>(n, 0, rates_full, 0.0, counts, idx_out, val_out);
  MGK_HIP(hipGetLastError());
  long long* hc = (long long*)mgk_pinned(2 * sizeof(double));
  MGK_HIP(hipMemcpyAsync(hc, counts + nbi, sizeof(long long), hipMemcpyDeviceToHost, s));
  MGK_HIP(hipMemcpyAsync(hc + 1, w.total, sizeof(double), hipMemcpyDeviceToHost, s));
  MGK_HIP(hipStreamSynchronize(s));
  MGK_HIP(hipFreeAsync(counts, s));
  if (count_out) *count_out = hc[0];
  if (total_out) {
    double t;
    memcpy(&t, hc + 1, sizeof(double));
    *total_out = t;
  }
  return MGK_OK;
}

// First index whose sequential running sum over `rates` reaches r.
extern "C" int mgk_kmc_select_array(long long n, const double* rates, double r, void* workspace, void* stream,
                                    mgk_kmc_sel* out) {
  if (n < 1 || !rates || !workspace || !out) {
    mgk_set_error("mgk_kmc_select_array: bad argument", __FILE__, __LINE__);
    return MGK_ERR_ARG;
  }
  hipStream_t s = (hipStream_t)stream;
  const KmcWs w = kmc_ws(workspace, (size_t)n);
  const int nb = kmc_nblocks((size_t)n);
  k_kmc_block_sums<<<nb, MGK_NT, 0, s>>>((size_t)n, rates, w.bs);
  k_kmc_select_array<<<1, MGK_NT, 0, s>>>((size_t)n, rates, w.bs, nb, r, w.sel);
  MGK_HIP(hipGetLastError());
  KmcSel* hp = (KmcSel*)mgk_pinned(sizeof(KmcSel));
  MGK_HIP(hipMemcpyAsync(hp, w.sel, sizeof(KmcSel), hipMemcpyDeviceToHost, s));
  MGK_HIP(hipStreamSynchronize(s));
  out->item = hp->item;
  out->mode = hp->mode;
  out->fallthrough = hp->fallthrough;
  out->r = hp->r;
  out->cum_before = hp->cum_before;
  out->cum_at = hp->cum_at;
  out->q = hp->q;
  return MGK_OK;
}

// Indices (reference flat order) of entries of `a` below thr (find_unstable).
extern "C" int mgk_find_below(long long n, const double* a, double thr, long long* idx_out, void* stream,
                              long long* count_out) {
  if (n < 1 || !a || !idx_out || !count_out) {
    mgk_set_error("mgk_find_below: bad argument", __FILE__, __LINE__);
    return MGK_ERR_ARG;
  }
  hipStream_t s = (hipStream_t)stream;
  const int nb = kmc_nblocks((size_t)n);
  long long* counts = nullptr;
  MGK_HIP(hipMallocAsync((void**)&counts, (size_t)(nb + 1) * sizeof(long long), s));
  k_compact_count<<<nb, MGK_NT, 0, s>>>((size_t)n, 1, a, thr, counts);
  k_compact_scan<<<1, MGK_NT, 0, s>>>(counts, nb, counts + nb);
  k_compact_scatter<<<nb, MGK_NT, 0, s>>>((size_t)n, 1, a, thr, counts, idx_out, nullptr);
  MGK_HIP(hipGetLastError());
  long long* hc = (long long*)mgk_pinned(sizeof(long long));
  MGK_HIP(hipMemcpyAsync(hc, counts + nb, sizeof(long long), hipMemcpyDeviceToHost, s));
  MGK_HIP(hipStreamSynchronize(s));
  MGK_HIP(hipFreeAsync(counts, s));
  *count_out = hc[0];
  return MGK_OK;
}

// Flip mode m of voxel v.  epsp: symmetric SoA eigenstrain (precision 0 FP64,
// 1 FP32).  prev (nullable): D*D previous strain direction, set to the mode.
// gamma_out (nullable, device): the D*D mode tensor.
extern "C" int mgk_kmc_flip(int ndim, const int* shape, int M, int cat_kind, long long v, int m, double time,
                            double jp, double jt, double gmax, double jn, const double* cat, void* epsp,
                            int precision, float* epsp_shadow, double* soft, double* tlast, double* prev, double* gamma_out,
                            int eps_div, void* stream) {
  if ((ndim != 2 && ndim != 3) || !shape || !cat || (!epsp && eps_div != 0) || !soft || !tlast ||
      (precision != 0 && epsp_shadow)) {
    mgk_set_error("mgk_kmc_flip: bad argument", __FILE__, __LINE__);
    return MGK_ERR_ARG;
  }
  KmcFlip F;
  F.nx = shape[0];
  F.ny = shape[1];
  F.nz = ndim == 3 ? shape[2] : 1;
  F.M = M;
  F.v = v;
  F.m = m;
  F.time = time;
  F.jp = jp;
  F.jt = jt;
  F.gmax = gmax;
  F.jn = jn;
  F.eps_div = eps_div;
  const long long NV = (long long)F.nx * F.ny * F.nz;
  if (v < 0 || v >= NV || m < 0 || m >= M) {
    mgk_set_error("mgk_kmc_flip: voxel/mode out of range", __FILE__, __LINE__);
    return MGK_ERR_ARG;
  }
  hipStream_t s = (hipStream_t)stream;
  if (ndim == 3) {
    if (cat_kind == 0) {
      if (precision == 0) k_kmc_flip<3, 2, double><<<1, 64, 0, s>>>(F, cat, (double*)epsp, epsp_shadow, soft, tlast, prev, gamma_out);
      else k_kmc_flip<3, 2, float><<<1, 64, 0, s>>>(F, cat, (float*)epsp, epsp_shadow, soft, tlast, prev, gamma_out);
    } else {
      if (precision == 0) k_kmc_flip<3, 9, double><<<1, 64, 0, s>>>(F, cat, (double*)epsp, epsp_shadow, soft, tlast, prev, gamma_out);
      else k_kmc_flip<3, 9, float><<<1, 64, 0, s>>>(F, cat, (float*)epsp, epsp_shadow, soft, tlast, prev, gamma_out);
    }
  } else {
    if (cat_kind == 0) {
      if (precision == 0) k_kmc_flip<2, 2, double><<<1, 64, 0, s>>>(F, cat, (double*)epsp, epsp_shadow, soft, tlast, prev, gamma_out);
      else k_kmc_flip<2, 2, float><<<1, 64, 0, s>>>(F, cat, (float*)epsp, epsp_shadow, soft, tlast, prev, gamma_out);
    } else {
      if (precision == 0) k_kmc_flip<2, 4, double><<<1, 64, 0, s>>>(F, cat, (double*)epsp, epsp_shadow, soft, tlast, prev, gamma_out);
      else k_kmc_flip<2, 4, float><<<1, 64, 0, s>>>(F, cat, (float*)epsp, epsp_shadow, soft, tlast, prev, gamma_out);
    }
  }
  MGK_HIP(hipGetLastError());
  return MGK_OK;
}

// Replace all modes (catalog parameters, mode-major (m, k) host-side order,
// and Q0) of voxel v.  newcat/newq0: device pointers (nullable).
extern "C" int mgk_kmc_set_voxel(long long nvox, int M, int CK, long long v, const double* newcat,
                                 const double* newq0, double* cat, double* Q0, void* stream) {
  if (nvox < 1 || v < 0 || v >= nvox || M < 1 || CK < 1) {
    mgk_set_error("mgk_kmc_set_voxel: bad argument", __FILE__, __LINE__);
    return MGK_ERR_ARG;
  }
  k_kmc_set_voxel<<<1, 64, 0, (hipStream_t)stream>>>((size_t)nvox, M, CK, v, newcat, newq0, cat, Q0);
  MGK_HIP(hipGetLastError());
  return MGK_OK;
}

// ---------------------------------------------------------------------------
// Device initial modes: Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel
// random numbers: as easy as 1, 2, 3", SC'11) keyed by the seed, counter
// (mode, voxel lo, voxel hi, draw), Box-Muller normals.  No reference
// counterpart: the reference draws these on the host from the legacy stream.
// ---------------------------------------------------------------------------
struct Philox4 {
  uint32_t v[4];
};

__device__ __forceinline__ Philox4 philox4x32_10(Philox4 c, uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)M0 * c.v[0];
    const uint64_t p1 = (uint64_t)M1 * c.v[2];
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    Philox4 n;
    n.v[0] = hi1 ^ c.v[1] ^ k0;
    n.v[1] = lo1;
    n.v[2] = hi0 ^ c.v[3] ^ k1;
    n.v[3] = lo0;
    c = n;
    k0 += W0;
    k1 += W1;
  }
  return c;
}

// 53-bit uniform in (0, 1) from two 32-bit words
__device__ __forceinline__ double u53(uint32_t a, uint32_t b) {
  const uint64_t x = ((uint64_t)a << 21) ^ (uint64_t)(b >> 11);
  return ((double)x + 0.5) * (1.0 / 9007199254740992.0);
}

__device__ __forceinline__ void box_muller(double u1, double u2, double* z0, double* z1) {
  const double r = sqrt(-2.0 * log(u1));
  double s, c;
  sincospi(2.0 * u2, &s, &c);
  *z0 = r * c;
  *z1 = r * s;
}

__global__ void k_kmc_init_modes(size_t NV, int M, uint32_t k0, uint32_t k1, double mean, double sd, int has_lo,
                                 double lo, int has_hi, double hi, double half_g0, double* __restrict__ Q0,
                                 double* __restrict__ cat) {
  const size_t MNV = (size_t)M * NV;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < MNV; i += (size_t)gridDim.x * blockDim.x) {
    const size_t m = i / NV, v = i - m * NV;
    Philox4 c;
    c.v[0] = (uint32_t)m;
    c.v[1] = (uint32_t)v;
    c.v[2] = (uint32_t)(v >> 32);
    c.v[3] = 0;
    const Philox4 a = philox4x32_10(c, k0, k1);
    c.v[3] = 1;
    const Philox4 b = philox4x32_10(c, k0, k1);
    double gxx, gxy, zq, unused;
    box_muller(u53(a.v[0], a.v[1]), u53(a.v[2], a.v[3]), &gxx, &gxy);
    box_muller(u53(b.v[0], b.v[1]), u53(b.v[2], b.v[3]), &zq, &unused);
    double q = mean + sd * zq;
    if (has_lo && q < lo) q = lo;
    if (has_hi && q > hi) q = hi;
    Q0[i] = q;
    cat[i] = half_g0 * gxx;
    cat[MNV + i] = half_g0 * gxy;
  }
}

extern "C" int mgk_kmc_init_modes(long long nvox, int M, uint64_t seed, double mean, double std_dev, int has_min,
                                  double min_cutoff, int has_max, double max_cutoff, double gamma0, double* Q0,
                                  double* cat, void* stream) {
  if (nvox < 1 || M < 1 || !Q0 || !cat) {
    mgk_set_error("mgk_kmc_init_modes: bad argument", __FILE__, __LINE__);
    return MGK_ERR_ARG;
  }
  const size_t n = (size_t)nvox * (size_t)M;
  const int nb = (int)std::min<size_t>((n + MGK_NT - 1) / MGK_NT, 16384);
  k_kmc_init_modes<<<nb, MGK_NT, 0, (hipStream_t)stream>>>((size_t)nvox, M, (uint32_t)seed, (uint32_t)(seed >> 32),
                                                             mean, std_dev, has_min, min_cutoff, has_max, max_cutoff,
                                                             0.5 * gamma0, Q0, cat);
  MGK_HIP(hipGetLastError());
  return MGK_OK;
}
