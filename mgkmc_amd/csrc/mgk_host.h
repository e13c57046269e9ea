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
#include "../../include/mgkmc_hip.h"

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

// Host values of the twiddle table exp(-2 pi i t/n), t in [0,n): 2n doubles
// (prec 0) or 2n floats (prec 1) written to out.
void mgk_twiddle_values(int n, int prec, void* out);

// Device twiddle table exp(-2 pi i t/n), t in [0,n), for the current device.
// prec 0 = double, 1 = float.  Owned by the library until process exit.
const void* mgk_twiddles(int n, int prec);

// Max dynamic LDS bytes per workgroup on the current device.
size_t mgk_lds_limit();

// Pinned host scratch (per thread, grown on demand) for small D2H reads.
void* mgk_pinned(size_t bytes);

// Rate pass and selection of one simulation of the event loop, enqueued on s
// with their small results copied into pinned host memory (host_total: one
// double; host_rec: mgk_kmc_sel_rec_bytes() bytes, read with
// mgk_kmc_sel_unpack after the stream has been synchronised).
int mgk_kmc_rates_async(const mgk_kmc_loop* L, double time, double* host_total, hipStream_t s);
size_t mgk_kmc_sel_rec_bytes();
int mgk_kmc_select_async(const mgk_kmc_loop* L, double time, double u, void* host_rec, hipStream_t s);
void mgk_kmc_sel_unpack(const void* host_rec, mgk_kmc_sel* out);

// ---------------------------------------------------------------------------
// Live per-kernel timing (HIP events on the launching stream).  Off by default;
// bench.py turns it on around its timed region.
// ---------------------------------------------------------------------------
enum MgkKid {
  KID_LS_TAU = 0,
  KID_LS_MID,
  KID_LS_PENCIL,
  KID_LS_UPDATE,
  KID_KMC_RATES,
  KID_KMC_SELECT,
  KID_FS_CONST,
  KID_FS_XFWD,
  KID_FS_MID,
  KID_FS_PENCIL,
  KID_FS_XINV,
  KID_FS_VEC,
  KID_COUNT = 16
};

bool mgk_prof_on();

// Records (kernel id, iteration, bytes, event pair) for one solve; commit()
// adds the elapsed times of records with iter < iters_done to the global
// counters after the stream has been synchronised.
class MgkProfRec {
 public:
  explicit MgkProfRec(hipStream_t s, bool allow = true) : s_(s), on_(allow && mgk_prof_on()) {}
  ~MgkProfRec();
  int begin(int kid, int iter, double bytes);
  void end(int h);
  void commit(int iters_done);
  // drop records with lo <= iter < hi (launches that exited early)
  void invalidate(int lo, int hi);
  bool on() const { return on_; }

 private:
  struct Rec {
    int kid, iter;
    double bytes;
    hipEvent_t e0, e1;
  };
  hipStream_t s_;
  bool on_;
  std::vector<Rec> r_;
};

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
