// Landau small-strain stress (3D) for the Lippmann-Schwinger sweeps:
//   sigma = A I + B e + C e^2 (+ capping tangent term), with
//   A = lam I1 + v1/2 I1^2 + v2 I2 + g1/6 I1^3 + g2 I1 I2 + 4/3 g3 I3
//   B = 2 (mu + v2 I1 + g2/2 I1^2 + g4 I2),  C = 4 (v3 + g3 I1)
// and the optional deviatoric strain capping at E_cap (piecewise or smooth
// tanh), as mgkmc/elasticity.py:297-335 and :811-861 define them.
#pragma once
#include "mgk_common.h"

struct LandauPar {
  double v1, v2, v3, g1, g2, g3, g4;
  double Ecap, cap_p, cap_ratio;  // resolved capping strain, smooth power, tangent ratio
  int capping, cap_smooth;
};

// e, s: symmetric components xx, yy, zz, yz, xz, xy
template <typename T>
__device__ __forceinline__ void landau_sigma_sym(const T (&e)[6], T lam, T mu, const LandauPar& P, T (&s)[6]) {
  T E[9] = {e[0], e[5], e[4], e[5], e[1], e[3], e[4], e[3], e[2]};
  T Eo[9];
  T w = T(0);
  if (P.capping) {
#pragma unroll
    for (int k = 0; k < 9; ++k) Eo[k] = E[k];
    const T t3 = (E[0] + E[4] + E[8]) / T(3);
    T Ed[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) Ed[k] = E[k] - ((k % 4 == 0) ? t3 : T(0));
    T dd = T(0);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) dd += Ed[3 * i + j] * Ed[3 * j + i];
    const T Eeq = sqrt(fmax(T(0), (T(2) / T(3)) * dd));
    const T Ecap = T(P.Ecap);
    T f = T(1);
    if (P.cap_smooth) {
      const T p = T(P.cap_p);
      if (Eeq > T(1e-12)) f = (Ecap / Eeq) * pow(tanh(fmin(T(100), pow(Eeq / Ecap, p))), T(1) / p);
      const T tw = tanh(fmin(T(100), pow(Eeq / Ecap, p)));
      w = tw * tw;
    } else if (Eeq > Ecap) {
      f = Ecap / Eeq;
      w = T(1);
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) E[k] = Ed[k] * f + ((k % 4 == 0) ? t3 : T(0));
  }
  T E2[9];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int k = 0; k < 3; ++k) E2[3 * i + k] = E[3 * i] * E[k] + E[3 * i + 1] * E[3 + k] + E[3 * i + 2] * E[6 + k];
  const T I1 = E[0] + E[4] + E[8];
  T I2 = T(0), I3 = T(0);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      I2 += E[3 * i + j] * E[3 * j + i];
      I3 += E2[3 * i + j] * E[3 * j + i];
    }
  const T v1 = T(P.v1), v2 = T(P.v2), v3 = T(P.v3), g1 = T(P.g1), g2 = T(P.g2), g3 = T(P.g3), g4 = T(P.g4);
  const T A = lam * I1 + T(0.5) * v1 * I1 * I1 + v2 * I2 + (T(1) / T(6)) * g1 * I1 * I1 * I1 + g2 * I1 * I2 +
              (T(4) / T(3)) * g3 * I3;
  const T B = T(2) * (mu + v2 * I1 + T(0.5) * g2 * I1 * I1 + g4 * I2);
  const T C = T(4) * (v3 + g3 * I1);
  T S[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) S[k] = ((k % 4 == 0) ? A : T(0)) + B * E[k] + C * E2[k];
  if (P.capping) {
    const T Gt = T(P.cap_ratio) * mu;
#pragma unroll
    for (int k = 0; k < 9; ++k) S[k] += T(2) * Gt * (Eo[k] - E[k]) * w;
  }
  s[0] = S[0];
  s[1] = S[4];
  s[2] = S[8];
  s[3] = S[5];
  s[4] = S[2];
  s[5] = S[1];
}

// ---------------------------------------------------------------------------
// 2D (block-diagonal) Landau law: in-plane strain a11, a22, a12 (= a21) and
// the out-of-plane e33 (0 in plane strain, the root of sigma33 = 0 in plane
// stress), as mgkmc/elasticity.py:487-638 evaluates it.
// ---------------------------------------------------------------------------
template <typename T>
struct Landau2 {
  T I1, I2, c11, c22, c12, c33, w, A, B, C;
};

template <typename T>
__device__ __forceinline__ Landau2<T> landau2_eval(T a11, T a22, T a12, T e33, T lam, T mu, const LandauPar& P) {
  Landau2<T> s;
  s.w = T(0);
  if (P.capping) {
    const T m3 = (a11 + a22 + e33) / T(3);
    const T d11 = a11 - m3, d22 = a22 - m3, d33 = e33 - m3;
    const T Eeq = sqrt(fmax(T(0), (T(2) / T(3)) * (d11 * d11 + d22 * d22 + d33 * d33 + T(2) * a12 * a12)));
    const T Ecap = T(P.Ecap);
    T f = T(1);
    if (P.cap_smooth) {
      const T p = T(P.cap_p);
      if (Eeq > T(1e-12)) f = (Ecap / Eeq) * pow(tanh(fmin(T(100), pow(Eeq / Ecap, p))), T(1) / p);
      const T tw = tanh(fmin(T(100), pow(Eeq / Ecap, p)));
      s.w = tw * tw;
    } else if (Eeq > Ecap) {
      f = Ecap / Eeq;
      s.w = T(1);
    }
    s.c11 = f * d11 + m3;
    s.c22 = f * d22 + m3;
    s.c12 = f * a12;
    s.c33 = f * d33 + m3;
  } else {
    s.c11 = a11;
    s.c22 = a22;
    s.c12 = a12;
    s.c33 = e33;
  }
  s.I1 = s.c11 + s.c22 + s.c33;
  const T cross = s.c12 * s.c12;
  const T tr2 = s.c11 * s.c11 + s.c22 * s.c22 + T(2) * cross;
  s.I2 = tr2 + s.c33 * s.c33;
  const T I3 = (s.c11 + s.c22) * (tr2 - (s.c11 * s.c22 - cross)) + s.c33 * s.c33 * s.c33;
  const T v1 = T(P.v1), v2 = T(P.v2), v3 = T(P.v3), g1 = T(P.g1), g2 = T(P.g2), g3 = T(P.g3), g4 = T(P.g4);
  const T I1 = s.I1, I2 = s.I2;
  s.A = lam * I1 + T(0.5) * v1 * I1 * I1 + v2 * I2 + (T(1) / T(6)) * g1 * I1 * I1 * I1 + g2 * I1 * I2 +
        (T(4) / T(3)) * g3 * I3;
  s.B = T(2) * (mu + v2 * I1 + T(0.5) * g2 * I1 * I1 + g4 * I2);
  s.C = T(4) * (v3 + g3 * I1);
  return s;
}

// Newton step for sigma33(e33) = 0 at state s (mgkmc/elasticity.py:592-608)
template <typename T>
__device__ __forceinline__ T landau2_delta(const Landau2<T>& s, T e33, T lam, T mu, const LandauPar& P) {
  const T v1 = T(P.v1), v2 = T(P.v2), g1 = T(P.g1), g2 = T(P.g2), g3 = T(P.g3), g4 = T(P.g4);
  const T c33 = s.c33, I1 = s.I1, I2 = s.I2;
  T s33 = s.A + s.B * c33 + s.C * c33 * c33;
  const T dA = lam + v1 * I1 + T(0.5) * g1 * I1 * I1 + g2 * I2 + T(2) * (v2 + g2 * I1) * c33 + T(4) * g3 * c33 * c33;
  const T dB = T(2) * (v2 + g2 * I1) + T(4) * g4 * c33;
  T ds33 = dA + dB * c33 + s.B + T(4) * g3 * c33 * c33 + T(2) * s.C * c33;
  if (P.capping) {
    const T Gt = T(P.cap_ratio) * mu;
    s33 = s33 + T(2) * Gt * (e33 - c33) * s.w;
    ds33 = ds33 + T(2) * Gt * s.w;
  }
  return -s33 / (ds33 + T(1e-12));
}

// in-plane stress xx, yy, xy from state s (mgkmc/elasticity.py:622-636)
template <typename T>
__device__ __forceinline__ void landau2_sigma(const Landau2<T>& s, T a11, T a22, T a12, T mu, const LandauPar& P,
                                              T (&o)[3]) {
  const T cross = s.c12 * s.c12, t = s.c11 + s.c22;
  o[0] = s.A + s.B * s.c11 + s.C * (s.c11 * s.c11 + cross);
  o[1] = s.A + s.B * s.c22 + s.C * (s.c22 * s.c22 + cross);
  o[2] = s.B * s.c12 + s.C * (s.c12 * t);
  if (P.capping) {
    const T r = T(2) * (T(P.cap_ratio) * mu) * s.w;
    o[0] += r * (a11 - s.c11);
    o[1] += r * (a22 - s.c22);
    o[2] += r * (a12 - s.c12);
  }
}
