// BiCGSTAB building blocks shared by the finite-strain solvers and the 2D
// Landau Newton-CG solve: SciPy 1.15.3's recurrence (scipy/sparse/linalg/
// _isolve/iterative.py:157-260) with every scalar kept on the device in a
// KryState, vector kernels that exit once the state says done, and the host
// driver that enqueues iterations in chunks and polls one chunk behind.
#pragma once
#include "mgk_common.h"
#include "mgk_host.h"
#include <algorithm>
#include <map>
#include <mutex>

// ---------------------------------------------------------------------------
// BiCGSTAB (SciPy 1.15.3 recurrence) with device-resident scalars
// ---------------------------------------------------------------------------
struct KryState {
  int done;    // stop: every kernel of later iterations exits
  int done2;   // done || s-converged (skips the second matvec)
  int sconv;
  int status;  // 0 converged, maxiter, -10 rho breakdown, -11 omega/rv breakdown
  int iters;
  int maxiter;
  int pad0, pad1;
  double rtol, atol, bnrm2, rho, rho_prev, alpha, omega, beta, rv, ts, tt, rnorm, snorm, rhotol, omtol;
};

#define KRY_NBLK 1024

template <typename T>
__global__ __launch_bounds__(MGK_NT) void k_kry_init(size_t n, const T* __restrict__ b, T* __restrict__ x,
                                                     T* __restrict__ r, T* __restrict__ rt,
                                                     double* __restrict__ partials) {
  __shared__ double red[MGK_NT / 64];
  double acc[1] = {0.0};
  // 16-byte vectors of VW values (the Krylov vectors are 256-byte aligned),
  // then the tail element by element
  constexpr int VW = vw16<T>();
  const size_t nvec = n / VW, gs = (size_t)gridDim.x * MGK_NT, g0 = blockIdx.x * (size_t)MGK_NT + threadIdx.x;
  for (size_t j = g0; j < nvec; j += gs) {
    T v[VW], z[VW];
    vload16(b + j * VW, v);
#pragma unroll
    for (int k = 0; k < VW; ++k) {
      z[k] = T(0);
      acc[0] += (double)v[k] * (double)v[k];
    }
    vstore16(x + j * VW, z);
    vstore16(r + j * VW, v);
    vstore16(rt + j * VW, v);
  }
  for (size_t i = nvec * VW + g0; i < n; i += gs) {
    const T v = b[i];
    x[i] = T(0);
    r[i] = v;
    rt[i] = v;
    acc[0] += (double)v * (double)v;
  }
  block_sum<1>(acc, red);
  if (threadIdx.x == 0) partials[blockIdx.x] = acc[0];
}

static __device__ __forceinline__ void kry_sum(const double* __restrict__ partials, int n, int stride, int nv, double* out,
                                        double* red) {
  double acc[2] = {0.0, 0.0};
  for (int i = threadIdx.x; i < n; i += MGK_NT) {
    acc[0] += partials[(size_t)i * stride];
    if (nv > 1) acc[1] += partials[(size_t)i * stride + 1];
  }
  block_sum<2>(acc, red);
  if (threadIdx.x == 0) {
    out[0] = acc[0];
    out[1] = acc[1];
  }
}

static __global__ __launch_bounds__(MGK_NT) void k_kry_fin_init(const double* __restrict__ partials, int n,
                                                          KryState* __restrict__ st) {
  __shared__ double red[2 * (MGK_NT / 64)];
  double s[2];
  kry_sum(partials, n, 1, 1, s, red);
  if (threadIdx.x == 0) {
    st->bnrm2 = sqrt(s[0]);
    st->atol = fmax(0.0, st->rtol * st->bnrm2);
    st->rnorm = st->bnrm2;
    st->iters = 0;
    st->sconv = 0;
    st->done = 0;
    st->status = 0;
    if (st->bnrm2 == 0.0) {
      st->done = 1;  // x = 0 (= b)
    } else if (st->rnorm < st->atol) {
      st->done = 1;
    } else {
      st->rho = s[0];  // (rtilde, r) with rtilde = r = b
      if (fabs(st->rho) < st->rhotol) {
        st->done = 1;
        st->status = -10;
      }
    }
    st->done2 = st->done;
  }
}

// p = r (first spin) or p = ((p - omega v) * beta) + r
template <typename T>
__global__ __launch_bounds__(MGK_NT) void k_kry_p(size_t n, const KryState* __restrict__ st, T* __restrict__ p,
                                                  const T* __restrict__ r, const T* __restrict__ v) {
  if (st->done) return;
  const int first = st->iters == 0;
  const T om = (T)st->omega, be = (T)st->beta;
  constexpr int VW = vw16<T>();
  const size_t nvec = n / VW, gs = (size_t)gridDim.x * MGK_NT, g0 = blockIdx.x * (size_t)MGK_NT + threadIdx.x;
  for (size_t j = g0; j < nvec; j += gs) {
    T pv[VW], rv[VW], vv[VW];
    vload16(r + j * VW, rv);
    if (!first) {
      vload16(p + j * VW, pv);
      vload16(v + j * VW, vv);
#pragma unroll
      for (int k = 0; k < VW; ++k) {
        T t = pv[k] - om * vv[k];
        t = t * be;
        rv[k] = t + rv[k];
      }
    }
    vstore16(p + j * VW, rv);
  }
  for (size_t i = nvec * VW + g0; i < n; i += gs) {
    if (first) p[i] = r[i];
    else {
      T t = p[i] - om * v[i];
      t = t * be;
      p[i] = t + r[i];
    }
  }
}

// after v = A p: rv = (rtilde, v), alpha = rho / rv
static __global__ __launch_bounds__(MGK_NT) void k_kry_fin_rv(const double* __restrict__ partials, int n,
                                                        KryState* __restrict__ st) {
  if (st->done) return;
  __shared__ double red[2 * (MGK_NT / 64)];
  double s[2];
  kry_sum(partials, n, 2, 1, s, red);
  if (threadIdx.x == 0) {
    st->rv = s[0];
    if (s[0] == 0.0) {
      st->done = 1;
      st->done2 = 1;
      st->status = -11;
    } else {
      st->alpha = st->rho / s[0];
    }
  }
}

// s = r - alpha v (in place in r), |s|^2
template <typename T>
__global__ __launch_bounds__(MGK_NT) void k_kry_s(size_t n, const KryState* __restrict__ st, T* __restrict__ r,
                                                  const T* __restrict__ v, double* __restrict__ partials) {
  if (st->done) return;
  __shared__ double red[MGK_NT / 64];
  const T al = (T)st->alpha;
  double acc[1] = {0.0};
  constexpr int VW = vw16<T>();
  const size_t nvec = n / VW, gs = (size_t)gridDim.x * MGK_NT, g0 = blockIdx.x * (size_t)MGK_NT + threadIdx.x;
  for (size_t j = g0; j < nvec; j += gs) {
    T rv[VW], vv[VW];
    vload16(r + j * VW, rv);
    vload16(v + j * VW, vv);
#pragma unroll
    for (int k = 0; k < VW; ++k) {
      rv[k] = rv[k] - al * vv[k];
      acc[0] += (double)rv[k] * (double)rv[k];
    }
    vstore16(r + j * VW, rv);
  }
  for (size_t i = nvec * VW + g0; i < n; i += gs) {
    const T s = r[i] - al * v[i];
    r[i] = s;
    acc[0] += (double)s * (double)s;
  }
  block_sum<1>(acc, red);
  if (threadIdx.x == 0) partials[blockIdx.x] = acc[0];
}

static __global__ __launch_bounds__(MGK_NT) void k_kry_fin_s(const double* __restrict__ partials, int n,
                                                       KryState* __restrict__ st) {
  if (st->done) return;
  __shared__ double red[2 * (MGK_NT / 64)];
  double s[2];
  kry_sum(partials, n, 1, 1, s, red);
  if (threadIdx.x == 0) {
    st->snorm = sqrt(s[0]);
    if (st->snorm < st->atol) {
      st->sconv = 1;
      st->done2 = 1;
    }
  }
}

// after t = A s: omega = (t, s) / (t, t)
static __global__ __launch_bounds__(MGK_NT) void k_kry_fin_t(const double* __restrict__ partials, int n,
                                                       KryState* __restrict__ st) {
  if (st->done2) return;
  __shared__ double red[2 * (MGK_NT / 64)];
  double s[2];
  kry_sum(partials, n, 2, 2, s, red);
  if (threadIdx.x == 0) {
    st->ts = s[0];
    st->tt = s[1];
    st->omega = s[0] / s[1];
  }
}

// x += alpha p; x += omega s; r = s - omega t; partials |r|^2, (rtilde, r).
// Preconditioned: p is M p and sh = M s replaces s in the x update.
template <typename T>
__global__ __launch_bounds__(MGK_NT) void k_kry_xr(size_t n, const KryState* __restrict__ st, T* __restrict__ x,
                                                   const T* __restrict__ p, T* __restrict__ r, const T* __restrict__ t,
                                                   const T* __restrict__ rt, double* __restrict__ partials,
                                                   const T* __restrict__ sh) {
  if (st->done) return;
  __shared__ double red[2 * (MGK_NT / 64)];
  const T al = (T)st->alpha, om = (T)st->omega;
  const int sc = st->sconv;
  double acc[2] = {0.0, 0.0};
  constexpr int VW = vw16<T>();
  const size_t nvec = n / VW, gs = (size_t)gridDim.x * MGK_NT, g0 = blockIdx.x * (size_t)MGK_NT + threadIdx.x;
  for (size_t j = g0; j < nvec; j += gs) {
    T xv[VW], pv[VW];
    vload16(x + j * VW, xv);
    vload16(p + j * VW, pv);
    if (sc) {
#pragma unroll
      for (int k = 0; k < VW; ++k) xv[k] = xv[k] + al * pv[k];
    } else {
      T sv[VW], tv[VW], rtv[VW], hv[VW];
      vload16(r + j * VW, sv);
      vload16(t + j * VW, tv);
      vload16(rt + j * VW, rtv);
      if (sh) vload16(sh + j * VW, hv);
#pragma unroll
      for (int k = 0; k < VW; ++k) {
        T q = xv[k] + al * pv[k];
        xv[k] = q + om * (sh ? hv[k] : sv[k]);
        const T rn = sv[k] - om * tv[k];
        sv[k] = rn;
        acc[0] += (double)rn * (double)rn;
        acc[1] += (double)rtv[k] * (double)rn;
      }
      vstore16(r + j * VW, sv);
    }
    vstore16(x + j * VW, xv);
  }
  for (size_t i = nvec * VW + g0; i < n; i += gs) {
    if (sc) {
      x[i] = x[i] + al * p[i];
    } else {
      const T s = r[i];
      T xv = x[i] + al * p[i];
      xv = xv + om * (sh ? sh[i] : s);
      x[i] = xv;
      const T rn = s - om * t[i];
      r[i] = rn;
      acc[0] += (double)rn * (double)rn;
      acc[1] += (double)rt[i] * (double)rn;
    }
  }
  block_sum<2>(acc, red);
  if (threadIdx.x == 0) {
    partials[2 * blockIdx.x] = acc[0];
    partials[2 * blockIdx.x + 1] = acc[1];
  }
}

static __global__ __launch_bounds__(MGK_NT) void k_kry_fin_xr(const double* __restrict__ partials, int n,
                                                        KryState* __restrict__ st) {
  if (st->done) return;
  __shared__ double red[2 * (MGK_NT / 64)];
  double s[2];
  kry_sum(partials, n, 2, 2, s, red);
  if (threadIdx.x == 0) {
    if (st->sconv) {
      st->done = 1;
      st->done2 = 1;
      st->status = 0;
      return;
    }
    st->iters += 1;
    st->rnorm = sqrt(s[0]);
    const double rho_new = s[1];
    st->rho_prev = st->rho;
    if (st->iters >= st->maxiter) {
      st->done = 1;
      st->status = st->maxiter;
    } else if (st->rnorm < st->atol) {
      st->done = 1;
      st->status = 0;
    } else if (fabs(rho_new) < st->rhotol) {
      st->done = 1;
      st->status = -10;
    } else if (fabs(st->omega) < st->omtol) {
      st->done = 1;
      st->status = -11;
    } else {
      st->beta = (rho_new / st->rho_prev) * (st->alpha / st->omega);
      st->rho = rho_new;
    }
    st->done2 = st->done;
  }
}

// |x|^2 (FP64) into out[0]
template <typename T>
__global__ __launch_bounds__(MGK_NT) void k_norm2(size_t n, const T* __restrict__ x, double* __restrict__ partials) {
  __shared__ double red[MGK_NT / 64];
  double acc[1] = {0.0};
  constexpr int VW = vw16<T>();
  const size_t nvec = n / VW, gs = (size_t)gridDim.x * MGK_NT, g0 = blockIdx.x * (size_t)MGK_NT + threadIdx.x;
  for (size_t j = g0; j < nvec; j += gs) {
    T v[VW];
    vload16(x + j * VW, v);
#pragma unroll
    for (int k = 0; k < VW; ++k) acc[0] += (double)v[k] * (double)v[k];
  }
  for (size_t i = nvec * VW + g0; i < n; i += gs) acc[0] += (double)x[i] * (double)x[i];
  block_sum<1>(acc, red);
  if (threadIdx.x == 0) partials[blockIdx.x] = acc[0];
}


// last Krylov iteration count per workspace (-1 query)
static int kry_hint(const void* key, int iters) {
  static std::mutex mu;
  static std::map<const void*, int> last;
  std::lock_guard<std::mutex> lk(mu);
  if (iters >= 0) {
    last[key] = iters;
    return iters;
  }
  auto it = last.find(key);
  return it == last.end() ? 0 : it->second;
}

// Enqueues BiCGSTAB iterations in growing chunks and polls the device state
// one chunk behind, so the queue never drains while the host waits
template <typename Enq>
int kry_drive(KryState* kst, hipStream_t s, int maxiter, Enq&& enqueue, int* iters, int* status, int* enq_out) {
  KryState* hs = (KryState*)mgk_pinned(2 * sizeof(KryState));
  hipEvent_t ev[2];
  MGK_HIP(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
  MGK_HIP(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
  // first chunk: the iteration count of the previous solve on this workspace
  // (Newton steps of one load path need similar Krylov counts)
  const int hint = kry_hint(kst, -1);
  int enq = 0, chunk = hint > 0 ? hint + 1 : 4, slot = 0;
  bool pending = false, stop = false, first = true;
  while (!stop && enq < maxiter) {
    const int k = std::min(chunk, maxiter - enq);
    for (int i = 0; i < k; ++i) enqueue();
    enq += k;
    MGK_HIP(hipGetLastError());
    MGK_HIP(hipMemcpyAsync(&hs[slot], kst, sizeof(KryState), hipMemcpyDeviceToHost, s));
    MGK_HIP(hipEventRecord(ev[slot], s));
    if (pending) {
      MGK_HIP(hipEventSynchronize(ev[slot ^ 1]));
      if (hs[slot ^ 1].done) stop = true;
    }
    pending = true;
    slot ^= 1;
    chunk = first ? 2 : std::min(chunk * 2, 32);
    first = false;
  }
  MGK_HIP(hipMemcpyAsync(&hs[0], kst, sizeof(KryState), hipMemcpyDeviceToHost, s));
  MGK_HIP(hipStreamSynchronize(s));
  (void)hipEventDestroy(ev[0]);
  (void)hipEventDestroy(ev[1]);
  *iters = hs[0].iters;
  *status = hs[0].status;
  kry_hint(kst, hs[0].iters);
  *enq_out = enq;
  return MGK_OK;
}

