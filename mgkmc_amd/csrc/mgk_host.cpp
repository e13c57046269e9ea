// Host-side helpers: error string, FFT plans, twiddle cache, pinned scratch.
#include "mgk_host.h"
#include <string.h>
#include <stdio.h>
#include <string>

static thread_local std::string g_err;

void mgk_set_error(const char* msg, const char* file, int line) {
  char buf[512];
  snprintf(buf, sizeof(buf), "%s (%s:%d)", msg, file, line);
  g_err = buf;
}

extern "C" const char* mgk_last_error(void) { return g_err.c_str(); }

bool mgk_make_plan(int n, FftPlan* p) {
  if (n < 1) return false;
  p->n = n;
  p->npass = 0;
  int m = n;
  int twos = 0;
  while (m % 2 == 0) {
    m /= 2;
    ++twos;
  }
  int r[64];  // radices (a fixed array: planning allocates nothing)
  int nr = 0;
  while (twos >= 3) {
    r[nr++] = 8;
    twos -= 3;
  }
  if (twos == 2) r[nr++] = 4;
  if (twos == 1) {
    // prefer 8*4*4 style splits over a trailing radix-2 when possible
    if (nr > 0 && r[nr - 1] == 8) {
      r[nr - 1] = 4;
      r[nr++] = 4;
    } else {
      r[nr++] = 2;
    }
  }
  while (m % 3 == 0 && nr < 64) {
    r[nr++] = 3;
    m /= 3;
  }
  while (m % 5 == 0 && nr < 64) {
    r[nr++] = 5;
    m /= 5;
  }
  if (m != 1 || nr > 12) return false;
  for (int i = 0; i < nr; ++i) p->radix[i] = r[i];
  p->npass = nr;
  return true;
}

namespace {
struct TwKey {
  int dev, n, prec;
  bool operator<(const TwKey& o) const {
    if (dev != o.dev) return dev < o.dev;
    if (n != o.n) return n < o.n;
    return prec < o.prec;
  }
};
std::mutex g_tw_mu;
std::map<TwKey, void*> g_tw;
constexpr int kLdsDevices = 64;
size_t g_lds[kLdsDevices];  // per device, 0 = not read yet (a fixed table: no allocation)
}  // namespace

void mgk_twiddle_values(int n, int prec, void* out) {
  // exp(-2 pi i t/n) from long double angles, rounded once to the target type
  const long double pi = 3.141592653589793238462643383279502884L;
  for (int t = 0; t < n; ++t) {
    const long double a = 2.0L * pi * (long double)t / (long double)n;
    if (prec == 0) {
      ((double*)out)[2 * t] = (double)cosl(a);
      ((double*)out)[2 * t + 1] = (double)(-sinl(a));
    } else {
      ((float*)out)[2 * t] = (float)cosl(a);
      ((float*)out)[2 * t + 1] = (float)(-sinl(a));
    }
  }
}

const void* mgk_twiddles(int n, int prec) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  std::lock_guard<std::mutex> lk(g_tw_mu);
  TwKey k{dev, n, prec};
  auto it = g_tw.find(k);
  if (it != g_tw.end()) return it->second;
  const size_t bytes = (size_t)2 * n * (prec == 0 ? sizeof(double) : sizeof(float));
  std::vector<double> h((bytes + sizeof(double) - 1) / sizeof(double));
  mgk_twiddle_values(n, prec, h.data());
  void* d = nullptr;
  if (hipMalloc(&d, bytes) != hipSuccess) return nullptr;
  if (hipMemcpy(d, h.data(), bytes, hipMemcpyHostToDevice) != hipSuccess) return nullptr;
  g_tw[k] = d;
  return d;
}

size_t mgk_lds_limit() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 65536;
  std::lock_guard<std::mutex> lk(g_tw_mu);
  if (dev >= 0 && dev < kLdsDevices && g_lds[dev]) return g_lds[dev];
  int v = 65536;
  if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) v = 65536;
  if (dev >= 0 && dev < kLdsDevices) g_lds[dev] = (size_t)v;
  return (size_t)v;
}

namespace {
struct Pinned {
  void* p = nullptr;
  size_t n = 0;
  ~Pinned() {
    // Left to process teardown: freeing pinned memory after the runtime has
    // shut down is unsafe, and the buffer is small.
  }
};
thread_local Pinned g_pin;
}  // namespace

namespace {
struct DevStage {
  void* p[4] = {nullptr, nullptr, nullptr, nullptr};
  size_t n[4] = {0, 0, 0, 0};
};
thread_local DevStage g_dstage;
}  // namespace

void* mgk_device_stage(int slot, size_t bytes) {
  if (slot < 0 || slot > 3) return nullptr;
  if (g_dstage.n[slot] >= bytes) return g_dstage.p[slot];
  // geometric growth: the blocks dropped on the way (not freed: hipFree would
  // synchronise the device under other threads' work) add up to less than the last
  void* np_ = nullptr;
  size_t want = bytes < 4096 ? 4096 : bytes;
  if (want < 2 * g_dstage.n[slot]) want = 2 * g_dstage.n[slot];
  if (hipMalloc(&np_, want) != hipSuccess) return nullptr;
  g_dstage.p[slot] = np_;
  g_dstage.n[slot] = want;
  return np_;
}

void* mgk_pinned(size_t bytes) {
  if (g_pin.n >= bytes) return g_pin.p;
  if (g_pin.p) (void)hipHostFree(g_pin.p);
  g_pin.p = nullptr;
  g_pin.n = 0;
  size_t want = bytes < 4096 ? 4096 : bytes;
  if (hipHostMalloc(&g_pin.p, want, hipHostMallocDefault) != hipSuccess) return nullptr;
  g_pin.n = want;
  return g_pin.p;
}

extern "C" int mgk_version(void) { return 1; }

extern "C" int mgk_device_info(int* lds_per_block, int* cu_count) {
  int dev = 0;
  MGK_HIP(hipGetDevice(&dev));
  MGK_HIP(hipDeviceGetAttribute(lds_per_block, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
  MGK_HIP(hipDeviceGetAttribute(cu_count, hipDeviceAttributeMultiprocessorCount, dev));
  return MGK_OK;
}

// ---------------------------------------------------------------------------
// live kernel timing
// ---------------------------------------------------------------------------
namespace {
std::mutex g_prof_mu;
bool g_prof_on = false;
double g_prof_ms[KID_COUNT];
double g_prof_bytes[KID_COUNT];
long long g_prof_n[KID_COUNT];
}  // namespace

bool mgk_prof_on() { return g_prof_on; }

int MgkProfRec::begin(int kid, int iter, double bytes) {
  if (!on_) return -1;
  Rec r{kid, iter, bytes, nullptr, nullptr};
  if (hipEventCreate(&r.e0) != hipSuccess || hipEventCreate(&r.e1) != hipSuccess) return -1;
  (void)hipEventRecord(r.e0, s_);
  r_.push_back(r);
  return (int)r_.size() - 1;
}

void MgkProfRec::end(int h) {
  if (!on_ || h < 0) return;
  (void)hipEventRecord(r_[h].e1, s_);
}

void MgkProfRec::invalidate(int lo, int hi) {
  for (auto& r : r_)
    if (r.iter >= lo && r.iter < hi) r.iter = -1 - r.iter;
}

void MgkProfRec::commit(int iters_done) {
  if (!on_) return;
  std::lock_guard<std::mutex> lk(g_prof_mu);
  for (auto& r : r_) {
    if (r.iter < 0 || r.iter >= iters_done) continue;
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, r.e0, r.e1) == hipSuccess) {
      g_prof_ms[r.kid] += ms;
      g_prof_bytes[r.kid] += r.bytes;
      g_prof_n[r.kid] += 1;
    }
  }
}

MgkProfRec::~MgkProfRec() {
  for (auto& r : r_) {
    if (r.e0) (void)hipEventDestroy(r.e0);
    if (r.e1) (void)hipEventDestroy(r.e1);
  }
}

extern "C" void mgk_profile_enable(int on) {
  std::lock_guard<std::mutex> lk(g_prof_mu);
  g_prof_on = on != 0;
}

extern "C" void mgk_profile_reset(void) {
  std::lock_guard<std::mutex> lk(g_prof_mu);
  for (int i = 0; i < KID_COUNT; ++i) {
    g_prof_ms[i] = 0.0;
    g_prof_bytes[i] = 0.0;
    g_prof_n[i] = 0;
  }
}

extern "C" int mgk_profile_read(int kid, double* total_ms, double* total_bytes, long long* launches) {
  if (kid < 0 || kid >= KID_COUNT) return MGK_ERR_ARG;
  std::lock_guard<std::mutex> lk(g_prof_mu);
  *total_ms = g_prof_ms[kid];
  *total_bytes = g_prof_bytes[kid];
  *launches = g_prof_n[kid];
  return MGK_OK;
}
