// Secant-modulus small-strain laws for the Lippmann-Schwinger sweeps, per voxel.
//   Exponential degradation (2D, 3D), mgkmc/elasticity.py:641-810:
//     sigma = lam tr(e) I + 2 mu_s e,  mu_s = mu (1 - d (1 - exp(-k e_eq)))
//   Rose (2D), mgkmc/elasticity.py:868-943, at constant K = lam + 2/3 mu:
//     sigma = (K - 2/3 G) tr I + 2 G e,  G = max(mu exp(-eta e_eq), floor mu);
//     plane stress: e33 from the linear guess -lam/(lam + 2 mu)(e11 + e22)
//     and five updates e33 = -(3K - 2G)/(3K + 4G)(e11 + e22), then G again.
// e_eq = sqrt(2/3 e':e'); a 2D strain is taken with e33 = 0 (exponential law)
// or the plane-stress e33 (Rose), which still enters the deviator.
// Sums are formed in the reference's left-to-right order.
#pragma once
#include "mgk_common.h"

struct SecLaw {
  int kind;          // 0: exponential degradation, 1: Rose
  int plane_stress;  // Rose only
  double p1, p2;     // (d, k) or (eta, floor fraction)
};

// e, s: xx, yy, zz, yz, xz, xy
template <typename T>
__device__ __forceinline__ void secant3_sigma(const T (&e)[6], T lam, T mu, const SecLaw& L, T (&s)[6]) {
  const T tr = (e[0] + e[1]) + e[2];
  const T tr3 = tr / T(3);
  const T a = e[0] - tr3, b = e[1] - tr3, c = e[2] - tr3;
  const T g12 = e[5] + e[5], g13 = e[4] + e[4], g23 = e[3] + e[3];
  const T eq = sqrt((T(2) / T(3)) * (((a * a + b * b) + c * c) + T(0.5) * ((g12 * g12 + g13 * g13) + g23 * g23)));
  const T ms = mu * (T(1) - T(L.p1) * (T(1) - exp(-T(L.p2) * eq)));
  const T ltr = lam * tr, m2 = T(2) * ms;
#pragma unroll
  for (int c3 = 0; c3 < 6; ++c3) s[c3] = (c3 < 3 ? ltr : T(0)) + m2 * e[c3];
}

// e, s: xx, yy, xy
template <typename T>
__device__ __forceinline__ void secant2_sigma(const T (&e)[3], T lam, T mu, const SecLaw& L, T (&s)[3]) {
  const T tr3 = (e[0] + e[1]) / T(3);
  const T a = e[0] - tr3, b = e[1] - tr3, c = -(e[0] + e[1]) / T(3);
  const T g = e[2] + e[2];
  const T eq = sqrt((T(2) / T(3)) * (((a * a + b * b) + c * c) + T(0.5) * (g * g)));
  const T ms = mu * (T(1) - T(L.p1) * (T(1) - exp(-T(L.p2) * eq)));
  const T ltr = lam * (e[0] + e[1]), m2 = T(2) * ms;
  s[0] = ltr + m2 * e[0];
  s[1] = ltr + m2 * e[1];
  s[2] = m2 * e[2];
}

template <typename T>
__device__ __forceinline__ T rose_shear(T exx, T eyy, T exy, T ezz, T mu, const SecLaw& L, T& tr) {
  tr = (exx + eyy) + ezz;
  const T tr3 = tr / T(3);
  const T a = exx - tr3, b = eyy - tr3, c = ezz - tr3;
  const T eq = sqrt((T(2) / T(3)) * (((a * a + b * b) + c * c) + T(2) * (exy * exy)));
  return fmax(mu * exp(-T(L.p1) * eq), T(L.p2) * mu);
}

template <typename T>
__device__ __forceinline__ void rose2_sigma(const T (&e)[3], T lam, T mu, const SecLaw& L, T (&s)[3]) {
  const T K = lam + (T(2) / T(3)) * mu;
  T ezz = T(0), tr, G;
  if (L.plane_stress) {
    ezz = -lam / (lam + T(2) * mu) * (e[0] + e[1]);
    for (int it = 0; it < 5; ++it) {
      G = rose_shear(e[0], e[1], e[2], ezz, mu, L, tr);
      ezz = -((T(3) * K - T(2) * G) / (T(3) * K + T(4) * G)) * (e[0] + e[1]);
    }
  }
  G = rose_shear(e[0], e[1], e[2], ezz, mu, L, tr);
  const T le = K - (T(2) / T(3)) * G;
  s[0] = le * tr + T(2) * G * e[0];
  s[1] = le * tr + T(2) * G * e[1];
  s[2] = T(2) * G * e[2];
}

// Per-voxel law objects for the generic polarisation kernels: C components in
// the solver's symmetric order.
template <int D>
struct SecantLaw {
  SecLaw L;
  static constexpr int C = D == 3 ? 6 : 3;
  template <typename T>
  __device__ __forceinline__ void operator()(const T (&e)[C], T lam, T mu, T (&s)[C]) const {
    if constexpr (D == 3) {
      secant3_sigma(e, lam, mu, L, s);
    } else {
      if (L.kind == 1) rose2_sigma(e, lam, mu, L, s);
      else secant2_sigma(e, lam, mu, L, s);
    }
  }
};
