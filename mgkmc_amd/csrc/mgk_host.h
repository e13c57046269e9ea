// Host-side helpers shared by the library's translation units: FFT plans,
// twiddle-table cache, pinned-host scratch, launch geometry.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>
#include <mutex>
#include <map>
#include <vector>
#include <cmath>
#include "mgk_common.h"

#define MGK_HIP(call)                                  \
  do {                                                 \
    hipError_t e_ = (call);                            \
    if (e_ != hipSuccess) {                            \
      mgk_set_error(hipGetErrorString(e_), __FILE__, __LINE__); \
      return MGK_ERR_HIP;                              \
    }                                                  \
  } while (0)

void mgk_set_error(const char* msg, const char* file, int line);

// Factor n into radices (8,4,2,3,5 preference).  Returns false if n has a
// prime factor other than 2, 3, 5 or needs more than 12 passes.
bool mgk_make_plan(int n, FftPlan* p);

// Device twiddle table exp(-2 pi i t/n), t in [0,n), for the current device.
// prec 0 = double, 1 = float.  Owned by the library until process exit.
const void* mgk_twiddles(int n, int prec);

// Max dynamic LDS bytes per workgroup on the current device.
size_t mgk_lds_limit();

// Pinned host scratch (per thread, grown on demand) for small D2H reads.
void* mgk_pinned(size_t bytes);

inline int mgk_ilog2(int v) {
  int l = 0;
  while ((1 << (l + 1)) <= v) ++l;
  return l;
}
inline int mgk_pow2_ceil(int v) {
  int p = 1;
  while (p < v) p <<= 1;
  return p;
}
inline size_t mgk_align(size_t v) { return (v + 255) & ~size_t(255); }
