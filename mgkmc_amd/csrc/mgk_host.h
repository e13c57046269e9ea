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

// Device staging for the batched paths' small arrays (totals, selections,
// stop flags, solve states, redraws): per thread and slot (0..3), grown on
// demand; a smaller earlier block is dropped rather than freed, since hipFree
// would synchronise the device under other streams' work.
void* mgk_device_stage(int slot, size_t bytes);

// Batched passes of the event step (members share ndim, shape, M, sig_kind):
// the rate pass with every member's total into dev_tot[k]; the selection with
// every member's record (mgk_kmc_sel_rec_bytes() each, read back with
// mgk_kmc_sel_unpack) into dev_rec; the flips of (v[k], m[k]) at times[k] and
// the redraw of those voxels from dev_redraw + k * stride (2M directions, then
// M barriers).  Enqueued on s; nothing is synchronised.
int mgk_kmc_rates_batch(int nmem, const mgk_kmc_loop* const* L, const double* times, double* dev_tot,
                        hipStream_t s);
size_t mgk_kmc_sel_rec_bytes();
int mgk_kmc_select_batch(int nmem, const mgk_kmc_loop* const* L, const double* times, const double* u,
                         void* dev_rec, hipStream_t s);
void mgk_kmc_sel_unpack(const void* host_rec, mgk_kmc_sel* out);
int mgk_kmc_flip_redraw_batch(int nmem, const mgk_kmc_loop* const* L, const long long* v, const int* m,
                              const double* times, const double* dev_redraw, size_t stride, hipStream_t s);

// Cascade batches (native cascade loop): count of the modes with stored
// Q < thr (counts: one per 256 voxels; offsets: their exclusive scan with the
// total at offsets[nblocks]), then their (v, m, q) in voxel-major order (at
// most cap); the flips of a batch of n pairs at one time in nlev levels
// (level l: the pairs dlv/dlm[dloff[l] .. dloff[l+1]), disjoint neighbourhoods
// within a level), then the redraws of the voxels dv[i] (pair order) from
// dred + i * stride.  Enqueued on s.
int mgk_kmc_below_all(const mgk_kmc_loop* L, const double* Q, double thr, long long* counts, long long* offsets,
                      hipStream_t s);
int mgk_kmc_below_write(const mgk_kmc_loop* L, const double* Q, double thr, const long long* offsets, long long cap,
                        long long* vv, int* mm, double* qq, hipStream_t s);
int mgk_kmc_patch_mode(const mgk_kmc_loop* L, const mgk_kmc_patch_cfg* P, long long v, int m, double a,
                       int with_unit, double* sig_mean, hipStream_t stream);
int mgk_kmc_patch_unit(const mgk_kmc_loop* L, const mgk_kmc_patch_cfg* P, double a, double* sig_mean,
                       hipStream_t stream);

// Native finite-strain flip pieces (mgk_fsflip.hip): F_p <- dF F_p at voxel v
// (d = 3 all nine components, d = 2 the in-plane block of the (nx, 1, ny)
// grid; dF d x d row-major), and the NaN test of two n-long FP64 fields
// (*host = 1 on a NaN; synchronises the stream).
int mgk_fp_left_mul(double* Fp, long long NV, long long v, int d, const double* dF, hipStream_t s);
int mgk_any_nan2(const double* a, const double* b, size_t n, int* flag_dev, int* host, hipStream_t s);
int mgk_kmc_flip_redraw_levels(const mgk_kmc_loop* L, int n, int nlev, const int* dloff, const long long* dlv,
                               const int* dlm, const long long* dv, double time, const double* dred, size_t stride,
                               hipStream_t s);

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
  // first LS iteration when it reads its starting field as a constant or in
  // FP32 (fewer bytes than the steady-state launches above)
  KID_LS_TAU_START,
  KID_LS_UPDATE_START,
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
