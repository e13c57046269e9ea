// Native KMC event loop: the per-event sequence of the reference's KMC loops
// (mgkmc/aqs.py:855-1087, mgkmc/kmc_simulator.py:871-1101; small strain,
// linear elasticity, KMC mode) driven from C++ over the library's own entry
// points, with numpy's legacy RandomState streams (MT19937 words, 53-bit
// doubles, polar-method Gaussians with one cached value) continued in C++
// from the caller's state and handed back, so an event sequence does not
// depend on which side of the boundary ran it.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <unordered_set>
#include <vector>
#include "mgk_host.h"
#include "../../include/mgkmc_hip.h"

namespace {

// MT19937 (Matsumoto & Nishimura 1998) in numpy's legacy RandomState usage.
struct Mt {
  mgk_mt_state* s;
  void regen() {
    uint32_t* k = s->key;
    const uint32_t up = 0x80000000u, lo = 0x7fffffffu, mag = 0x9908b0dfu;
    int i = 0;
    for (; i < 624 - 397; ++i) {
      const uint32_t y = (k[i] & up) | (k[i + 1] & lo);
      k[i] = k[i + 397] ^ (y >> 1) ^ ((y & 1u) ? mag : 0u);
    }
    for (; i < 623; ++i) {
      const uint32_t y = (k[i] & up) | (k[i + 1] & lo);
      k[i] = k[i + 397 - 624] ^ (y >> 1) ^ ((y & 1u) ? mag : 0u);
    }
    const uint32_t y = (k[623] & up) | (k[0] & lo);
    k[623] = k[396] ^ (y >> 1) ^ ((y & 1u) ? mag : 0u);
    s->pos = 0;
  }
  uint32_t word() {
    if (s->pos >= 624) regen();
    uint32_t y = s->key[s->pos++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
  }
  // random_sample(): 27 + 26 bits
  double uniform() {
    const uint32_t a = word() >> 5, b = word() >> 6;
    return (a * 67108864.0 + b) / 9007199254740992.0;
  }
  // standard normal (polar method; the second value is cached)
  double gauss() {
    if (s->has_gauss) {
      const double t = s->gauss;
      s->gauss = 0.0;
      s->has_gauss = 0;
      return t;
    }
    double x1, x2, r2;
    do {
      x1 = 2.0 * uniform() - 1.0;
      x2 = 2.0 * uniform() - 1.0;
      r2 = x1 * x1 + x2 * x2;
    } while (r2 >= 1.0 || r2 == 0.0);
    const double f = std::sqrt(-2.0 * std::log(r2) / r2);
    s->gauss = f * x1;
    s->has_gauss = 1;
    return f * x2;
  }
};

// Pinned host staging for the redraw uploads, per thread, grown on demand and
// kept until the process ends (a hipHostFree per call would synchronise the
// whole device and stall the other streams of a serving process).
struct RedrawStage {
  double* p = nullptr;
  size_t n = 0;
};
thread_local RedrawStage g_stage;

double* redraw_stage(size_t count) {
  if (g_stage.n >= count) return g_stage.p;
  double* np_ = nullptr;
  if (hipHostMalloc((void**)&np_, sizeof(double) * count, hipHostMallocDefault) != hipSuccess) return nullptr;
  // a smaller earlier buffer is dropped rather than freed (hipHostFree would
  // synchronise the device); growth happens only when M grows
  g_stage.p = np_;
  g_stage.n = count;
  return np_;
}

// the voxel's redraw from the host stream: 2M normals (gxx, gxy per mode),
// then M clipped Gaussian barriers (3M doubles into host)
void redraw_draws(const mgk_kmc_loop* L, Mt& R, double* host) {
  const int M = L->M;
  const double hg = 0.5 * L->gamma0;
  if (L->redraw_directions)
    for (int i = 0; i < 2 * M; ++i) host[i] = hg * R.gauss();
  if (L->redraw_barriers)
    for (int i = 0; i < M; ++i) {
      double q = L->q_mean + L->q_std * R.gauss();
      if (L->q_has_min) q = std::fmax(q, L->q_min);
      if (L->q_has_max) q = std::fmin(q, L->q_max);
      host[2 * M + i] = q;
    }
}

// the sub-stepped finite-strain flip (defined with mgk_kmc_fs_flip below)
int fs_flip_core(const mgk_kmc_loop* L, const mgk_kmc_fs_substep* c, int nflips, const long long* voxels,
                 const int* modes, double time, Mt& R, double* F_macro, double* eps_macro, int* substeps,
                 mgk_fs_info* info, hipStream_t s);

bool fs_ok(const mgk_kmc_loop* L, const mgk_kmc_fs_substep* c) {
  if (!c) return false;
  const mgk_fs_stepper* st = c->step;
  return st && st->F && st->Fp && st->sig && st->precision == 0 && c->flag && c->backup && c->nlive >= 0 &&
         c->nlive <= 12 && (L->ndim == 2) == (c->sig2 != nullptr) && c->drv_i >= 0 && c->drv_i < L->ndim &&
         c->drv_j >= 0 && c->drv_j < L->ndim && c->n_targets >= 0 && c->n_targets <= 3 &&
         (L->redraw_directions || L->redraw_barriers) &&
         st->shape[0] * st->shape[1] * st->shape[2] ==
             L->shape[0] * L->shape[1] * (L->ndim == 3 ? L->shape[2] : 1);
}

// the 2D nonlinear-law sub-stepped flip (defined with mgk_kmc_law_flip below)
int law_flip_core(const mgk_kmc_loop* L, const mgk_kmc_law_substep* c, int nflips, const long long* voxels,
                  const int* modes, double time, Mt& R, const double* eps_macro, int* substeps, mgk_ls_info* info,
                  hipStream_t s);

bool law_ok(const mgk_kmc_loop* L, const mgk_kmc_law_substep* c) {
  if (!L || !c || (L->ndim != 2 && L->ndim != 3) || L->M < 1 || !L->cat || !L->Q0 || !L->soft || !L->tlast || !L->epsp ||
      L->eps_precision != 0 || !L->redraw_dev || !(L->redraw_directions || L->redraw_barriers) || !c->backup ||
      c->nlive < 0 || c->nlive > 12)
    return false;
  const mgk_law_stepper* st = c->step;
  if (!st || !st->lam || !st->mu || !st->eps || !st->sig || !st->tau || !st->ws) return false;
  if (L->ndim == 3) return st->kind == 2 && st->landau;
  return (st->kind == 0 ? (st->landau && st->e33 && (st->anderson_m == 0 || st->aa))
                        : (st->kind == 1 && st->secant)) &&
         st->shape[0] == L->shape[0] && st->shape[1] == L->shape[1];
}

bool patch_ok(const mgk_kmc_loop* L) {
  const mgk_kmc_patch_cfg* P = L->patch;
  return !P || (P->kernels && P->sig_unit && P->workspace && P->flips_since_sync && P->unit_live && P->sig_mean &&
                P->mean_from_patch && P->sync_interval >= 1 && (L->sig_kind == 0 || L->sig_kind == 1));
}

bool loop_ok(const mgk_kmc_loop* L, const mgk_mt_state* rng, const mgk_mt_state* sel_rng) {
  return L && rng && sel_rng && (L->ndim == 2 || L->ndim == 3) && L->M >= 1 && L->Q0 && L->cat && L->soft &&
         L->tlast && L->epsp && L->kmc_ws && L->redraw_dev && L->sig &&
         (L->fs ? (fs_ok(L, L->fs) && L->fs->F_macro && !L->patch && !L->law)
                : L->law ? (law_ok(L, L->law) && !L->patch)
                         : (L->eps && L->ws &&
                            (L->ls_precision != 2 || (L->lam32 && L->mu32 && L->eps32 && L->ws32)))) &&
         rng->pos >= 0 && rng->pos <= 624 && sel_rng->pos >= 0 && sel_rng->pos <= 624 && patch_ok(L);
}

void fill_event(mgk_kmc_event& ev, long long v, int m, int instab, double time, const mgk_ls_info& inf, int it32,
                const double* eps_macro, int d) {
  ev.v = v;
  ev.m = m;
  ev.is_instab = instab;
  ev.time = time;
  ev.ls_iters = inf.iters;
  ev.ls_iters32 = it32;
  ev.substeps = 0;
  for (int k = 0; k < 9; ++k) {
    ev.eps_macro[k] = k < d * d ? eps_macro[k] : 0.0;
    ev.sig_mean[k] = inf.sig_mean[k];
  }
}

// the loop's LS re-solve at eps_macro (FP64, FP32 or FP32 + FP64 refinement)
int loop_solve(const mgk_kmc_loop* L, const double* eps_macro, mgk_ls_info* info, int* it32, void* stream) {
  const int d = L->ndim;
  *it32 = 0;
  if (L->ls_precision == 2)
    return mgk_ls_solve_refined(d, L->shape, L->lam32, L->mu32, L->lam, L->mu, L->lam0, L->mu0, eps_macro,
                                L->epsp_shadow, L->epsp, L->eps32, L->eps, L->sig, L->warm, L->max_iter, L->tol,
                                L->tol32, L->ws32, L->ws32_bytes, L->ws, L->ws_bytes, stream, info, it32);
  return mgk_ls_solve(d, L->shape, L->ls_precision, L->lam, L->mu, L->lam0, L->mu0, eps_macro, L->epsp, L->eps,
                      L->sig, L->warm, L->max_iter, L->tol, L->ws, L->ws_bytes, stream, info);
}

bool unstable_mean(const mgk_ls_info& inf, int d) {
  double mx = 0.0;
  for (int i = 0; i < d * d; ++i) {
    const double v = inf.sig_mean[i];
    if (!std::isfinite(v)) return true;
    mx = std::fmax(mx, std::fabs(v));
  }
  return inf.status == MGK_ERR_NONFINITE || mx / 1e9 > 20.0;
}

}  // namespace

extern "C" int mgk_mt_fill(mgk_mt_state* state, int kind, double loc, double scale, long long n, double* out) {
  if (!state || !out || n < 0 || state->pos < 0 || state->pos > 624) {
    mgk_set_error("mgk_mt_fill: bad argument", __FILE__, __LINE__);
    return MGK_ERR_ARG;
  }
  Mt r{state};
  for (long long i = 0; i < n; ++i) out[i] = kind == 0 ? r.uniform() : loc + scale * r.gauss();
  return MGK_OK;
}

extern "C" int mgk_kmc_run(const mgk_kmc_loop* L, mgk_mt_state* rng, mgk_mt_state* sel_rng, double* time,
                           double* eps_macro, double* remaining, int max_events, int max_seq, int* seq,
                           mgk_kmc_event* events, int* n_events, mgk_ls_info* info, int* iters32, void* stream) {
  if (!loop_ok(L, rng, sel_rng) || !time || !eps_macro || !remaining || !seq || !events || !n_events || !info ||
      max_events < 1) {
    mgk_set_error("mgk_kmc_run: bad argument", __FILE__, __LINE__);
    return MGK_ERR_ARG;
  }
  hipStream_t s = (hipStream_t)stream;
  const int d = L->ndim, M = L->M;
  long long NV = 1;
  int shape_i[3] = {1, 1, 1};
  for (int i = 0; i < d; ++i) {
    NV *= L->shape[i];
    shape_i[i] = (int)L->shape[i];
  }
  Mt R{rng}, S{sel_rng};
  double* host = redraw_stage(3 * (size_t)M);
  if (!host) {
    mgk_set_error("mgk_kmc_run: pinned host allocation failed", __FILE__, __LINE__);
    return MGK_ERR_HIP;
  }
  int n = 0, rc = MGK_KMC_RUN_TIME;
  *n_events = 0;
  while (true) {
    mgk_kmc_params p = L->kp;
    p.time = *time;
    double total = 0.0;
    int e = mgk_kmc_barrier_rates(d, NV, M, L->sig_kind, 0, L->sig, L->Q0, L->cat, L->soft, L->tlast, L->prev,
                                  nullptr, &p, nullptr, L->kmc_ws, stream, &total);
    if (e != MGK_OK) {
      rc = e;
      break;
    }
    bool event = false;
    if (total > 0) {
      const double u = R.uniform();
      const double tw = -std::log(u) / total;
      if (tw < *remaining) {
        event = true;
        *time += tw;
        for (int k = 0; k < d * d; ++k) eps_macro[k] += L->unit[k] * (L->strain_rate * tw);
        *remaining -= tw;
        const double us = S.uniform();
        mgk_kmc_sel so;
        e = mgk_kmc_select(d, NV, M, L->sig_kind, 0, us, L->sig, L->Q0, L->cat, L->soft, L->tlast, L->prev,
                           nullptr, &p, L->kmc_ws, stream, &so);
        if (e != MGK_OK) {
          rc = e;
          break;
        }
        const long long v = so.item;
        const int m = so.mode;
        const int instab = so.q <= L->stability_threshold ? 1 : 0;
        if (L->fs) {
          // finite strain: the sub-stepped flip (flip, redraw and the steps) replaces
          // the flip, redraw and LS re-solve
          int nsub = 0;
          mgk_fs_info fsi{};
          e = fs_flip_core(L, L->fs, 1, &v, &m, *time, R, L->fs->F_macro, eps_macro, &nsub, &fsi, s);
          if (L->fs->info) *L->fs->info = fsi;
          if (e != MGK_OK) {
            rc = e;
            break;
          }
          mgk_ls_info pin{};
          for (int a = 0; a < d; ++a)
            for (int b = 0; b < d; ++b)
              pin.sig_mean[a * d + b] = fsi.sig_mean[d == 3 ? 3 * a + b : 3 * (2 * a) + 2 * b];
          fill_event(events[n], v, m, instab, *time, pin, 0, eps_macro, d);
          events[n++].substeps = nsub;
          *n_events = n;
          *seq = instab ? *seq + 1 : 0;
          if (*seq > max_seq) {
            rc = MGK_KMC_RUN_SEQ;
            break;
          }
          if (n >= max_events) {
            rc = MGK_KMC_RUN_COUNT;
            break;
          }
          continue;
        }
        if (L->law) {
          // nonlinear law: the 2D sub-stepped flip (flip, redraw and the law's solves),
          // or the 3D Landau event (flip, redraw, cold solve)
          int nsub = 0;
          if (L->law->step->kind == 2) {
            e = mgk_kmc_flip(d, shape_i, M, 0, v, m, *time, L->jp, L->jt, L->kp.cap, L->jn, L->cat, L->epsp,
                             L->eps_precision, L->epsp_shadow, L->soft, L->tlast, L->prev, nullptr, 1, stream);
            if (e == MGK_OK) {
              redraw_draws(L, R, host);
              if (hipMemcpyAsync(L->redraw_dev, host, sizeof(double) * 3 * (size_t)M, hipMemcpyHostToDevice, s) !=
                  hipSuccess) {
                mgk_set_error("mgk_kmc_run: redraw upload failed", __FILE__, __LINE__);
                e = MGK_ERR_HIP;
              }
            }
            if (e == MGK_OK)
              e = mgk_kmc_set_voxel(NV, M, 2, v, L->redraw_directions ? L->redraw_dev : nullptr,
                                    L->redraw_barriers ? L->redraw_dev + 2 * M : nullptr, L->cat, L->Q0, stream);
            if (e == MGK_OK) {
              const mgk_law_stepper* st = L->law->step;
              e = mgk_ls_solve_landau(L->shape, 0, st->lam, st->mu, st->lam0, st->mu0, st->landau, eps_macro,
                                      L->epsp, st->eps, st->sig, st->tau, 0, st->max_iter, st->tol, st->ws,
                                      st->ws_bytes, stream, info);
              if (e == MGK_ERR_NONFINITE)
                mgk_set_error("spectral_solver_landau_3d: non-finite strain update in LS loop (Landau energy likely "
                              "unstable at this strain; consider strain capping or smaller load steps).",
                              __FILE__, __LINE__);
              else if (e == MGK_NOT_CONVERGED)
                e = MGK_OK;
            }
          } else {
            e = law_flip_core(L, L->law, 1, &v, &m, *time, R, eps_macro, &nsub, info, s);
          }
          if (e != MGK_OK) {
            rc = e;
            break;
          }
          fill_event(events[n], v, m, instab, *time, *info, 0, eps_macro, d);
          events[n++].substeps = nsub;
          *n_events = n;
          *seq = instab ? *seq + 1 : 0;
          if (*seq > max_seq) {
            rc = MGK_KMC_RUN_SEQ;
            break;
          }
          if (n >= max_events) {
            rc = MGK_KMC_RUN_COUNT;
            break;
          }
          continue;
        }
        e = mgk_kmc_flip(d, shape_i, M, 0, v, m, *time, L->jp, L->jt, L->kp.cap, L->jn, L->cat, L->epsp,
                         L->eps_precision, L->epsp_shadow, L->soft, L->tlast, L->prev, nullptr, 1, stream);
        if (e != MGK_OK) {
          rc = e;
          break;
        }
        // fast patching: the predictor from the flipped mode (read before its redraw;
        // the reference patches after the redraw with the mode it saved)
        const mgk_kmc_patch_cfg* P = L->patch;
        bool resolve = true;
        if (P) {
          e = mgk_kmc_patch_mode(L, P, v, m, L->strain_rate * tw, tw > 0 ? 1 : 0, P->sig_mean, s);
          if (e != MGK_OK) {
            rc = e;
            break;
          }
          *P->unit_live = 1;
          *P->mean_from_patch = 1;
          *P->flips_since_sync += 1;
          resolve = *P->flips_since_sync >= P->sync_interval;
          if (resolve) *P->flips_since_sync = 0;
        }
        redraw_draws(L, R, host);
        if (hipMemcpyAsync(L->redraw_dev, host, sizeof(double) * 3 * (size_t)M, hipMemcpyHostToDevice, s) !=
            hipSuccess) {
          mgk_set_error("mgk_kmc_run: redraw upload failed", __FILE__, __LINE__);
          rc = MGK_ERR_HIP;
          break;
        }
        e = mgk_kmc_set_voxel(NV, M, 2, v, L->redraw_directions ? L->redraw_dev : nullptr,
                              L->redraw_barriers ? L->redraw_dev + 2 * M : nullptr, L->cat, L->Q0, stream);
        if (e != MGK_OK) {
          rc = e;
          break;
        }
        if (!resolve) {
          // a predicted event: no solve, the patched field's mean in the record
          mgk_ls_info pin = *info;
          pin.iters = 0;
          for (int k = 0; k < 9; ++k) pin.sig_mean[k] = k < d * d ? P->sig_mean[k] : 0.0;
          fill_event(events[n++], v, m, instab, *time, pin, 0, eps_macro, d);
          *n_events = n;
          *seq = instab ? *seq + 1 : 0;
          if (*seq > max_seq) {
            rc = MGK_KMC_RUN_SEQ;
            break;
          }
          if (n >= max_events) {
            rc = MGK_KMC_RUN_COUNT;
            break;
          }
          continue;
        }
        if (P) *P->mean_from_patch = 0;
        int it32 = 0;
        if (L->ls_precision == 2) {
          e = mgk_ls_solve_refined(d, L->shape, L->lam32, L->mu32, L->lam, L->mu, L->lam0, L->mu0, eps_macro,
                                   L->epsp_shadow, L->epsp, L->eps32, L->eps, L->sig, L->warm, L->max_iter,
                                   L->tol, L->tol32, L->ws32, L->ws32_bytes, L->ws, L->ws_bytes, stream, info,
                                   &it32);
        } else {
          e = mgk_ls_solve(d, L->shape, L->ls_precision, L->lam, L->mu, L->lam0, L->mu0, eps_macro, L->epsp,
                           L->eps, L->sig, L->warm, L->max_iter, L->tol, L->ws, L->ws_bytes, stream, info);
        }
        if (iters32) *iters32 = it32;
        if (e < 0 && e != MGK_ERR_NONFINITE) {
          rc = e;
          break;
        }
        if (L->check_unstable && unstable_mean(*info, d)) {
          mgk_set_error("Unstable stress in small-strain elastic_run", __FILE__, __LINE__);
          rc = MGK_KMC_RUN_UNSTABLE;
          break;
        }
        fill_event(events[n++], v, m, instab, *time, *info, it32, eps_macro, d);
        *n_events = n;
        *seq = instab ? *seq + 1 : 0;
        if (*seq > max_seq) {
          rc = MGK_KMC_RUN_SEQ;
          break;
        }
        if (n >= max_events) {
          rc = MGK_KMC_RUN_COUNT;
          break;
        }
      }
    }
    if (!event) {
      const mgk_kmc_patch_cfg* P = L->patch;
      if (P && d == 2 && *P->unit_live) {
        // kmc_simulator.py:1094-1096 (the 3D loop re-solves at the increment's end without it)
        const int e2 = mgk_kmc_patch_unit(L, P, L->strain_rate * *remaining, P->sig_mean, s);
        if (e2 != MGK_OK) {
          rc = e2;
          break;
        }
        *P->mean_from_patch = 1;
      }
      for (int k = 0; k < d * d; ++k) eps_macro[k] += L->unit[k] * (L->strain_rate * *remaining);
      *time += *remaining;
      *remaining = 0.0;
      rc = MGK_KMC_RUN_TIME;
      break;
    }
  }
  (void)hipStreamSynchronize(s);
  return rc;
}

// ---------------------------------------------------------------------------
// Batched event loop: nb simulations of one grid shape and LS configuration
// stepped in lockstep on one stream.  Per step: every running member's rate
// pass is enqueued and one synchronisation reads all totals; the members'
// waiting times and selection draws come from their own streams; their
// selections are enqueued and read with one synchronisation; flips and
// redraws are enqueued; one batched LS solve re-equilibrates every member
// that had an event.  Each member runs the kernels and draws of mgk_kmc_run
// on its own buffers, so its events, fields and streams are bit-identical to
// mgk_kmc_run's; a member stops on its own condition while the others go on.
// ---------------------------------------------------------------------------
namespace {
struct BatchStage {
  void* p = nullptr;
  size_t n = 0;
};
thread_local BatchStage g_bstage;

// pinned host staging of the batch loop (grown on demand; a smaller earlier
// block is dropped, never freed, as a hipHostFree would synchronise the device)
void* batch_stage(size_t bytes) {
  if (g_bstage.n >= bytes) return g_bstage.p;
  void* np_ = nullptr;
  if (hipHostMalloc(&np_, bytes, hipHostMallocDefault) != hipSuccess) return nullptr;
  g_bstage.p = np_;
  g_bstage.n = bytes;
  return np_;
}
}  // namespace

extern "C" int mgk_kmc_run_batch(int nb, const mgk_kmc_loop* L, mgk_mt_state* rng, mgk_mt_state* sel_rng,
                                 double* time, double* eps_macro, double* remaining, const int* max_events,
                                 int max_seq, int* seq, mgk_kmc_event* events, int cap, int* n_events, int* status,
                                 mgk_ls_info* info, int* iters32, void* stream) {
  if (nb < 1 || nb > MGK_LS_MAX_BATCH || !L || !rng || !sel_rng || !time || !eps_macro || !remaining ||
      !max_events || !seq || !events || cap < 1 || !n_events || !status || !info) {
    mgk_set_error("mgk_kmc_run_batch: bad argument", __FILE__, __LINE__);
    return MGK_ERR_ARG;
  }
  for (int b = 0; b < nb; ++b)
    if (L[b].patch || L[b].fs || L[b].law) {
      mgk_set_error("mgk_kmc_run_batch: fast patching and finite strain run in mgk_kmc_run only", __FILE__,
                    __LINE__);
      return MGK_ERR_ARG;
    }
  const mgk_kmc_loop& L0 = L[0];
  const int d = L0.ndim;
  int Mmax = 0;
  for (int b = 0; b < nb; ++b) {
    const mgk_kmc_loop& x = L[b];
    bool same = loop_ok(&x, &rng[b], &sel_rng[b]) && x.ndim == d && x.ls_precision == L0.ls_precision &&
                (x.warm != 0) == (L0.warm != 0) && x.max_iter == L0.max_iter && x.tol == L0.tol &&
                x.tol32 == L0.tol32 && max_events[b] >= 1 && max_events[b] <= cap;
    for (int i = 0; i < d; ++i) same = same && x.shape[i] == L0.shape[i];
    if (!same) {
      mgk_set_error("mgk_kmc_run_batch: a member is invalid or differs from the first in grid shape, LS precision "
                    "or LS settings, or its max_events is outside 1..cap",
                    __FILE__, __LINE__);
      return MGK_ERR_ARG;
    }
    Mmax = std::max(Mmax, x.M);
  }
  for (int b = 1; b < nb; ++b) {
    if (L[b].M != L0.M || L[b].sig_kind != L0.sig_kind || L[b].eps_precision != L0.eps_precision) {
      mgk_set_error("mgk_kmc_run_batch: members differ in M, sig_kind or eps_p precision", __FILE__, __LINE__);
      return MGK_ERR_ARG;
    }
  }
  hipStream_t s = (hipStream_t)stream;
  long long NV = 1;
  int shape_i[3] = {1, 1, 1};
  for (int i = 0; i < d; ++i) {
    NV *= L0.shape[i];
    shape_i[i] = (int)L0.shape[i];
  }
  // staging, host (pinned) and device with one layout: totals, selection
  // records (side by side), redraws (3 Mmax per member)
  const size_t rec = mgk_kmc_sel_rec_bytes();
  const size_t off_rec = 64 * (size_t)MGK_LS_MAX_BATCH, off_red = off_rec + ((rec * MGK_LS_MAX_BATCH + 255) & ~(size_t)255);
  const size_t red_stride = 3 * (size_t)Mmax;
  const size_t stage_bytes = off_red + sizeof(double) * red_stride * MGK_LS_MAX_BATCH;
  char* stage = (char*)batch_stage(stage_bytes);
  char* dstage = (char*)mgk_device_stage(1, stage_bytes);
  if (!stage || !dstage) {
    mgk_set_error("mgk_kmc_run_batch: staging allocation failed", __FILE__, __LINE__);
    return MGK_ERR_HIP;
  }
  double* tot = (double*)stage;
  const mgk_kmc_loop* lp[MGK_LS_MAX_BATCH];
  int who[MGK_LS_MAX_BATCH];
  bool run[MGK_LS_MAX_BATCH], ev[MGK_LS_MAX_BATCH];
  double t0[MGK_LS_MAX_BATCH], us[MGK_LS_MAX_BATCH];
  for (int b = 0; b < nb; ++b) {
    run[b] = true;
    n_events[b] = 0;
    status[b] = MGK_KMC_RUN_TIME;
    if (iters32) iters32[b] = 0;
  }
  int rc_all = MGK_OK;
  auto stop = [&](int b, int code) {
    run[b] = false;
    status[b] = code;
  };
  auto fail_running = [&](int code) {
    for (int b = 0; b < nb; ++b)
      if (run[b]) stop(b, code);
    rc_all = code;
  };
  while (true) {
    int nrun = 0;
    for (int b = 0; b < nb; ++b) nrun += run[b];
    if (!nrun) break;
    // one rate pass for every running member, one read of all totals
    bool ok = true;
    double tm[MGK_LS_MAX_BATCH];
    int nr = 0;
    for (int b = 0; b < nb; ++b)
      if (run[b]) {
        lp[nr] = &L[b];
        tm[nr] = time[b];
        who[nr++] = b;
      }
    ok = mgk_kmc_rates_batch(nr, lp, tm, (double*)dstage, s) == MGK_OK &&
         hipMemcpyAsync(tot, dstage, nr * sizeof(double), hipMemcpyDeviceToHost, s) == hipSuccess;
    if (!ok || hipStreamSynchronize(s) != hipSuccess) {
      if (ok) mgk_set_error("mgk_kmc_run_batch: stream synchronisation failed", __FILE__, __LINE__);
      fail_running(MGK_ERR_HIP);
      break;
    }
    double totb[MGK_LS_MAX_BATCH];
    for (int k = 0; k < nr; ++k) totb[who[k]] = tot[k];
    int nev = 0;
    for (int b = 0; b < nb; ++b) {
      ev[b] = false;
      if (!run[b]) continue;
      Mt R{&rng[b]}, S{&sel_rng[b]};
      const mgk_kmc_loop& x = L[b];
      double* em = eps_macro + 9 * b;
      t0[b] = time[b];
      if (totb[b] > 0) {
        const double u = R.uniform();
        const double tw = -std::log(u) / totb[b];
        if (tw < remaining[b]) {
          ev[b] = true;
          ++nev;
          time[b] += tw;
          for (int k = 0; k < d * d; ++k) em[k] += x.unit[k] * (x.strain_rate * tw);
          remaining[b] -= tw;
          us[b] = S.uniform();
        }
      }
      if (!ev[b]) {
        for (int k = 0; k < d * d; ++k) em[k] += x.unit[k] * (x.strain_rate * remaining[b]);
        time[b] += remaining[b];
        remaining[b] = 0.0;
        stop(b, MGK_KMC_RUN_TIME);
      }
    }
    if (!nev) continue;
    // one selection pass and one read of all records
    int ne = 0;
    double usel[MGK_LS_MAX_BATCH];
    for (int b = 0; b < nb; ++b)
      if (ev[b]) {
        lp[ne] = &L[b];
        tm[ne] = t0[b];
        usel[ne] = us[b];
        who[ne++] = b;
      }
    ok = mgk_kmc_select_batch(ne, lp, tm, usel, dstage + off_rec, s) == MGK_OK &&
         hipMemcpyAsync(stage + off_rec, dstage + off_rec, ne * rec, hipMemcpyDeviceToHost, s) == hipSuccess;
    if (!ok || hipStreamSynchronize(s) != hipSuccess) {
      if (ok) mgk_set_error("mgk_kmc_run_batch: stream synchronisation failed", __FILE__, __LINE__);
      fail_running(MGK_ERR_HIP);
      break;
    }
    // flips and redraws: the draws on the host, one upload, one flip pass and
    // one redraw pass
    long long vv[MGK_LS_MAX_BATCH], vk[MGK_LS_MAX_BATCH];
    int mm[MGK_LS_MAX_BATCH], mk[MGK_LS_MAX_BATCH], ins[MGK_LS_MAX_BATCH];
    double* hred = (double*)(stage + off_red);
    for (int k = 0; k < ne; ++k) {
      const int b = who[k];
      const mgk_kmc_loop& x = L[b];
      mgk_kmc_sel so;
      mgk_kmc_sel_unpack(stage + off_rec + rec * k, &so);
      vv[b] = vk[k] = so.item;
      mm[b] = mk[k] = so.mode;
      ins[b] = so.q <= x.stability_threshold ? 1 : 0;
      tm[k] = time[b];
      Mt R{&rng[b]};
      redraw_draws(&x, R, hred + red_stride * k);
    }
    ok = hipMemcpyAsync(dstage + off_red, hred, sizeof(double) * red_stride * ne, hipMemcpyHostToDevice, s) ==
         hipSuccess;
    if (!ok) mgk_set_error("mgk_kmc_run_batch: redraw upload failed", __FILE__, __LINE__);
    int e = ok ? mgk_kmc_flip_redraw_batch(ne, lp, vk, mk, tm, (const double*)(dstage + off_red), red_stride, s)
               : MGK_ERR_HIP;
    if (e != MGK_OK) {
      fail_running(e);
      ok = false;
    }
    if (!ok) break;
    // one batched re-solve of every member with an event
    mgk_ls_member mem[MGK_LS_MAX_BATCH] = {};
    int idx[MGK_LS_MAX_BATCH], k = 0;
    for (int b = 0; b < nb; ++b) {
      if (!ev[b]) continue;
      const mgk_kmc_loop& x = L[b];
      mgk_ls_member& m = mem[k];
      m.lam = x.lam;
      m.mu = x.mu;
      m.lam0 = x.lam0;
      m.mu0 = x.mu0;
      m.eps_bar = eps_macro + 9 * b;
      m.eps_plastic = x.epsp;
      m.eps = x.eps;
      m.sig = x.sig;
      m.workspace = x.ws;
      m.workspace_bytes = x.ws_bytes;
      m.lam32 = x.lam32;
      m.mu32 = x.mu32;
      m.eps_plastic32 = x.epsp_shadow;
      m.eps32 = x.eps32;
      m.workspace32 = x.ws32;
      m.workspace32_bytes = x.ws32_bytes;
      idx[k++] = b;
    }
    mgk_ls_info inf[MGK_LS_MAX_BATCH] = {};
    int it32[MGK_LS_MAX_BATCH] = {};
    if (L0.ls_precision == 2)
      e = mgk_ls_solve_refined_batch(d, L0.shape, k, mem, L0.warm, L0.max_iter, L0.tol, L0.tol32, stream, inf, it32);
    else
      e = mgk_ls_solve_batch(d, L0.shape, L0.ls_precision, k, mem, L0.warm, L0.max_iter, L0.tol, stream, inf);
    if (e < 0 && e != MGK_ERR_NONFINITE) {  // the batch itself failed
      fail_running(e);
      break;
    }
    for (int j = 0; j < k; ++j) {
      const int b = idx[j];
      info[b] = inf[j];
      if (iters32) iters32[b] = it32[j];
      const int eb = inf[j].status;
      if (eb < 0 && eb != MGK_ERR_NONFINITE) {
        stop(b, eb);
        continue;
      }
      if (L[b].check_unstable && unstable_mean(inf[j], d)) {
        mgk_set_error("Unstable stress in small-strain elastic_run", __FILE__, __LINE__);
        stop(b, MGK_KMC_RUN_UNSTABLE);
        continue;
      }
      fill_event(events[(size_t)b * cap + n_events[b]], vv[b], mm[b], ins[b], time[b], inf[j], it32[j],
                 eps_macro + 9 * b, d);
      n_events[b] += 1;
      seq[b] = ins[b] ? seq[b] + 1 : 0;
      if (seq[b] > max_seq) stop(b, MGK_KMC_RUN_SEQ);
      else if (n_events[b] >= max_events[b]) stop(b, MGK_KMC_RUN_COUNT);
    }
  }
  (void)hipStreamSynchronize(s);
  return rc_all < 0 ? rc_all : MGK_OK;
}

// ---------------------------------------------------------------------------
// Native cascade loop (mgkmc/aqs.py:480-680, mgkmc/kmc_simulator.py:577-794;
// small strain, linear elasticity, planar catalogs): per batch one barrier
// pass storing Q, one scan for the modes with Q < stability_threshold in
// voxel-major order, a host sort by Q (ties by scan position) keeping the
// first mode per voxel, the batch's flips (levels of flips with disjoint
// neighbourhoods, each voxel's updates in pair order) and their redraws from
// the host stream in that order, one LS re-solve.  Each batch costs three host
// synchronisations (the count, the list, the solve), where the per-mode
// scans of the Python loop took one per mode.
// ---------------------------------------------------------------------------
namespace {
struct CascadeStage {
  void* h = nullptr;
  size_t hn = 0;
};
thread_local CascadeStage g_cstage;

void* cascade_host(size_t bytes) {
  if (g_cstage.hn >= bytes) return g_cstage.h;
  void* np_ = nullptr;
  const size_t want = std::max(bytes, 2 * g_cstage.hn);
  if (hipHostMalloc(&np_, want, hipHostMallocDefault) != hipSuccess) return nullptr;
  // the smaller earlier block is dropped (hipHostFree would synchronise the device);
  // growth doubles, so the dropped blocks add up to less than the current one
  g_cstage.h = np_;
  g_cstage.hn = want;
  return np_;
}

// levels of a batch's flips for k_kmc_flip_levels: a flip touches its voxel and,
// with neighbour softening on, the voxels one periodic step away along every
// axis; its level is one more than the highest level of the earlier flips that
// touch any voxel it touches, so flips of one level touch disjoint voxels and each
// voxel sees its flips in pair order.  touch (NV entries, -1 between batches)
// holds the highest level touching each voxel.
thread_local std::vector<int> g_touch;

template <typename Fn>
void for_hood(const mgk_kmc_loop* L, long long v, bool nbrs, Fn&& fn) {
  if (!nbrs) {
    fn(v);
    return;
  }
  const long long nx = L->shape[0], ny = L->shape[1], nz = L->ndim == 3 ? L->shape[2] : 1;
  const long long x = v / (ny * nz), y = (v / nz) % ny, z = v % nz;
  const int r = L->ndim == 3 ? 1 : 0;
  for (int dx = -1; dx <= 1; ++dx)
    for (int dy = -1; dy <= 1; ++dy)
      for (int dz = -r; dz <= r; ++dz)
        fn((((x + dx + nx) % nx) * ny + (y + dy + ny) % ny) * nz + (z + dz + nz) % nz);
}

int flip_levels(const mgk_kmc_loop* L, long long NV, const std::vector<long long>& pv, std::vector<int>& lev) {
  if ((long long)g_touch.size() < NV) g_touch.assign((size_t)NV, -1);
  const bool nbrs = L->jn > 0.0;
  int nlev = 0;
  lev.resize(pv.size());
  for (size_t k = 0; k < pv.size(); ++k) {
    int hi = -1;
    for_hood(L, pv[k], nbrs, [&](long long w) { hi = std::max(hi, g_touch[(size_t)w]); });
    lev[k] = hi + 1;
    for_hood(L, pv[k], nbrs, [&](long long w) { g_touch[(size_t)w] = hi + 1; });
    nlev = std::max(nlev, hi + 2);
  }
  for (long long v : pv) for_hood(L, v, nbrs, [&](long long w) { g_touch[(size_t)w] = -1; });
  return nlev;
}
}  // namespace

extern "C" int mgk_kmc_cascade(const mgk_kmc_loop* L, mgk_mt_state* rng, double* Q, double time,
                               const double* eps_macro, int max_batches, int truncate, int eps_comp,
                               double eps_target, mgk_cascade_flip* flips, int cap, int* n_flips, int* batch_sizes,
                               double* batch_sig, int batch_cap, int* n_batches, int* reason, mgk_ls_info* info,
                               int* iters32, void* stream) {
  mgk_mt_state sel_dummy{};
  sel_dummy.pos = 624;
  if (!loop_ok(L, rng, &sel_dummy) || L->fs || L->law || !Q || !eps_macro || !flips || cap < 1 || !n_flips || !batch_sizes ||
      batch_cap < 1 || !n_batches || !reason || !info || max_batches < 0 ||
      (!L->redraw_directions && !L->redraw_barriers)) {
    mgk_set_error("mgk_kmc_cascade: bad argument", __FILE__, __LINE__);
    return MGK_ERR_ARG;
  }
  hipStream_t s = (hipStream_t)stream;
  const int d = L->ndim, M = L->M;
  long long NV = 1;
  for (int i = 0; i < d; ++i) NV *= L->shape[i];
  const int nblk = (int)((NV + MGK_NT - 1) / MGK_NT);
  Mt R{rng};
  *n_flips = 0;
  *n_batches = 0;
  *reason = MGK_CASCADE_DONE;
  if (iters32) *iters32 = 0;
  // device: block counts, offsets (+ total), then the list (grown on demand)
  char* dcnt = (char*)mgk_device_stage(2, sizeof(long long) * (2 * (size_t)nblk + 1));
  if (!dcnt) {
    mgk_set_error("mgk_kmc_cascade: staging allocation failed", __FILE__, __LINE__);
    return MGK_ERR_HIP;
  }
  long long* counts = (long long*)dcnt;
  long long* offsets = counts + nblk;
  const size_t stride = 3 * (size_t)M;
  int local = 0, nf = 0;
  while (true) {
    mgk_kmc_params p = L->kp;
    p.time = time;
    int e = mgk_kmc_barrier_rates(d, NV, M, L->sig_kind, 0, L->sig, L->Q0, L->cat, L->soft, L->tlast, L->prev,
                                  nullptr, &p, Q, L->kmc_ws, stream, nullptr);
    if (e == MGK_OK) e = mgk_kmc_below_all(L, Q, L->stability_threshold, counts, offsets, s);
    if (e != MGK_OK) return e;
    long long* htot = (long long*)cascade_host(4096);
    if (!htot) {
      mgk_set_error("mgk_kmc_cascade: staging allocation failed", __FILE__, __LINE__);
      return MGK_ERR_HIP;
    }
    if (hipMemcpyAsync(htot, offsets + nblk, sizeof(long long), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
      mgk_set_error("mgk_kmc_cascade: count read failed", __FILE__, __LINE__);
      return MGK_ERR_HIP;
    }
    const long long total = *htot;
    if (total == 0) break;
    // the list: (v, m, q) per unstable mode in voxel-major order
    const size_t lb = ((size_t)total * (sizeof(long long) + sizeof(int) + sizeof(double)) + 255) & ~(size_t)255;
    const size_t npmax = (size_t)std::min<long long>(total, NV);
    const size_t sbytes = lb + npmax * (stride * sizeof(double) + 2 * sizeof(long long) + 2 * sizeof(int)) + 512;
    char* dl = (char*)mgk_device_stage(3, sbytes);
    char* hl = (char*)cascade_host(sbytes);
    if (!dl || !hl) {
      mgk_set_error("mgk_kmc_cascade: list allocation failed", __FILE__, __LINE__);
      return MGK_ERR_HIP;
    }
    long long* dvv = (long long*)dl;
    double* dqq = (double*)(dvv + total);
    int* dmm = (int*)(dqq + total);
    e = mgk_kmc_below_write(L, Q, L->stability_threshold, offsets, total, dvv, dmm, dqq, s);
    if (e != MGK_OK) return e;
    if (hipMemcpyAsync(hl, dl, (size_t)total * (sizeof(long long) + sizeof(double) + sizeof(int)),
                       hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
      mgk_set_error("mgk_kmc_cascade: list read failed", __FILE__, __LINE__);
      return MGK_ERR_HIP;
    }
    const long long* hvv = (const long long*)hl;
    const double* hqq = (const double*)(hvv + total);
    const int* hmm = (const int*)(hqq + total);
    // by barrier, ties by scan position; the first (most unstable) mode per voxel
    std::vector<long long> ord((size_t)total);
    for (long long i = 0; i < total; ++i) ord[(size_t)i] = i;
    std::stable_sort(ord.begin(), ord.end(), [&](long long a, long long b) { return hqq[a] < hqq[b]; });
    std::unordered_set<long long> seen;
    seen.reserve((size_t)total * 2);
    std::vector<long long> pv;
    std::vector<int> pm;
    for (long long i : ord) {
      const long long v = hvv[i];
      if (!seen.insert(v).second) continue;
      pv.push_back(v);
      pm.push_back(hmm[i]);
    }
    const int np = (int)pv.size();
    if (nf + np > cap || local >= batch_cap) {
      *reason = MGK_CASCADE_CAPACITY;  // records full: the caller continues with a new call
      break;
    }
    std::vector<int> lev;
    const int nlev = flip_levels(L, NV, pv, lev);
    // staged after the list, one upload: the voxels in pair order, the pairs
    // grouped by level, the level offsets, the redraws in pair order
    char* hp = hl + lb;
    long long* hpv = (long long*)hp;
    long long* hlv = hpv + np;
    int* hlm = (int*)(hlv + np);
    int* hloff = hlm + np;
    double* hred = (double*)(((uintptr_t)(hloff + nlev + 1) + 15) & ~(uintptr_t)15);
    for (int l = 0; l <= nlev; ++l) hloff[l] = 0;
    for (int k = 0; k < np; ++k) ++hloff[lev[k] + 1];
    for (int l = 0; l < nlev; ++l) hloff[l + 1] += hloff[l];
    {
      std::vector<int> fill(hloff, hloff + nlev);
      for (int k = 0; k < np; ++k) {
        const int at = fill[lev[k]]++;
        hlv[at] = pv[k];
        hlm[at] = pm[k];
      }
    }
    for (int k = 0; k < np; ++k) {
      hpv[k] = pv[k];
      redraw_draws(L, R, hred + stride * k);
    }
    const size_t pbytes = (size_t)((char*)(hred + stride * np) - hp);
    char* dp = dl + lb;
    if (hipMemcpyAsync(dp, hp, pbytes, hipMemcpyHostToDevice, s) != hipSuccess) {
      mgk_set_error("mgk_kmc_cascade: pair upload failed", __FILE__, __LINE__);
      return MGK_ERR_HIP;
    }
    auto dev = [&](const void* h) { return dp + ((const char*)h - hp); };
    e = mgk_kmc_flip_redraw_levels(L, np, nlev, (const int*)dev(hloff), (const long long*)dev(hlv),
                                   (const int*)dev(hlm), (const long long*)dev(hpv), time,
                                   (const double*)dev(hred), stride, s);
    if (e != MGK_OK) return e;
    int it32 = 0;
    if (L->ls_precision == 2) {
      e = mgk_ls_solve_refined(d, L->shape, L->lam32, L->mu32, L->lam, L->mu, L->lam0, L->mu0, eps_macro,
                               L->epsp_shadow, L->epsp, L->eps32, L->eps, L->sig, L->warm, L->max_iter, L->tol,
                               L->tol32, L->ws32, L->ws32_bytes, L->ws, L->ws_bytes, stream, info, &it32);
    } else {
      e = mgk_ls_solve(d, L->shape, L->ls_precision, L->lam, L->mu, L->lam0, L->mu0, eps_macro, L->epsp, L->eps,
                       L->sig, L->warm, L->max_iter, L->tol, L->ws, L->ws_bytes, stream, info);
    }
    if (iters32) *iters32 = it32;
    if (e < 0 && e != MGK_ERR_NONFINITE) return e;
    for (int k = 0; k < np; ++k) {
      flips[nf + k].batch = local;
      flips[nf + k].v = pv[k];
      flips[nf + k].m = pm[k];
    }
    nf += np;
    batch_sizes[local] = np;
    if (batch_sig)
      for (int k = 0; k < 9; ++k) batch_sig[9 * (size_t)local + k] = info->sig_mean[k];
    *n_flips = nf;
    if (L->check_unstable && unstable_mean(*info, d)) {
      mgk_set_error("Unstable stress in small-strain elastic_run", __FILE__, __LINE__);
      *n_batches = local + 1;
      *reason = MGK_CASCADE_UNSTABLE;
      break;
    }
    if (eps_comp >= 0 && eps_macro[eps_comp] >= eps_target) {
      *n_batches = local + 1;
      *reason = MGK_CASCADE_EPS_TARGET;
      break;
    }
    ++local;
    *n_batches = local;
    if (local > max_batches) {
      *reason = truncate ? MGK_CASCADE_TRUNCATED : MGK_CASCADE_DONE;
      break;
    }
  }
  (void)hipStreamSynchronize(s);
  return MGK_OK;
}

// ---------------------------------------------------------------------------
// Macro relaxation after an elastic increment (linear law, small strain).
// 3D (mgkmc/aqs.py:1105-1142): per iteration the stress error of the targeted
// components against the current mean; below tol: converged; else the
// Poisson-corrected compliance step (err - nu tr(err) I) / E with the driving
// component zeroed, then a re-solve.  2D (mgkmc/kmc_simulator.py:1106-1143):
// per iteration a re-solve first (with the unstable-stress check), then the
// error test, then per targeted in-plane component (in the given order) the
// compliance step, replaced by err / slope once two iterates exist and the
// secant slope ds / de is finite and above 0.01 E.
// ---------------------------------------------------------------------------
extern "C" int mgk_kmc_mixed_bc(const mgk_kmc_loop* L, int n_targets, const int* target_idx,
                                const double* target_val, int drv_i, int drv_j, double E_avg, double nu_avg,
                                double tol, int max_iter, double* eps_macro, int* converged, double* max_err,
                                int* iters, mgk_ls_info* info, int* iters32, void* stream) {
  if (!L || (L->ndim != 2 && L->ndim != 3) || !L->sig || !L->eps || !L->ws || n_targets < 0 || n_targets > 9 ||
      (n_targets > 0 && (!target_idx || !target_val)) || !eps_macro || !converged || !max_err || !iters || !info ||
      (L->ls_precision == 2 && !(L->lam32 && L->mu32 && L->eps32 && L->ws32))) {
    mgk_set_error("mgk_kmc_mixed_bc: bad argument", __FILE__, __LINE__);
    return MGK_ERR_ARG;
  }
  const int d = L->ndim;
  for (int t = 0; t < n_targets; ++t)
    if (target_idx[2 * t] < 0 || target_idx[2 * t] > 2 || target_idx[2 * t + 1] < 0 || target_idx[2 * t + 1] > 2) {
      mgk_set_error("mgk_kmc_mixed_bc: target index outside 0..2", __FILE__, __LINE__);
      return MGK_ERR_ARG;
    }
  *converged = 0;
  *max_err = 0.0;
  *iters = 0;
  int it32 = 0, e = MGK_OK;
  if (d == 3) {
    for (int it = 0; it < max_iter; ++it) {
      double err[9] = {0};
      double mx = 0.0;
      for (int t = 0; t < n_targets; ++t) {
        const int k = 3 * target_idx[2 * t] + target_idx[2 * t + 1];
        err[k] = target_val[t] - info->sig_mean[k];
        mx = std::fmax(mx, std::fabs(err[k]));
      }
      *max_err = mx;
      if (mx < tol) {
        *converged = 1;
        break;
      }
      const double tr = err[0] + err[4] + err[8];
      double corr[9];
      for (int k = 0; k < 9; ++k) corr[k] = (err[k] - nu_avg * tr * ((k % 4 == 0) ? 1.0 : 0.0)) / E_avg;
      corr[3 * drv_i + drv_j] = 0.0;
      for (int k = 0; k < 9; ++k) eps_macro[k] += corr[k];
      e = loop_solve(L, eps_macro, info, &it32, stream);
      ++*iters;
      if (e < 0 && e != MGK_ERR_NONFINITE) return e;
    }
  } else {
    bool have_prev[4] = {false, false, false, false};
    double prev_e[4] = {0}, prev_s[4] = {0};
    for (int it = 0; it < max_iter; ++it) {
      e = loop_solve(L, eps_macro, info, &it32, stream);
      ++*iters;
      if (e < 0 && e != MGK_ERR_NONFINITE) return e;
      if (L->check_unstable && unstable_mean(*info, 2)) {
        mgk_set_error("Unstable stress in small-strain elastic_run", __FILE__, __LINE__);
        if (iters32) *iters32 = it32;
        return MGK_KMC_RUN_UNSTABLE;
      }
      const double* sm = info->sig_mean;
      double err[4] = {0};
      double mx = 0.0;
      for (int t = 0; t < n_targets; ++t) {
        const int a = target_idx[2 * t], b = target_idx[2 * t + 1];
        if (a < 2 && b < 2) {
          err[2 * a + b] = target_val[t] - sm[2 * a + b];
          mx = std::fmax(mx, std::fabs(err[2 * a + b]));
        }
      }
      *max_err = mx;
      if (mx < tol) {
        *converged = 1;
        break;
      }
      const double tr = err[0] + err[3];
      double de[4];
      for (int k = 0; k < 4; ++k) de[k] = (err[k] - nu_avg * tr * ((k == 0 || k == 3) ? 1.0 : 0.0)) / E_avg;
      for (int t = 0; t < n_targets; ++t) {
        const int a = target_idx[2 * t], b = target_idx[2 * t + 1];
        if (a >= 2 || b >= 2) continue;
        const int k = 2 * a + b;
        double step = de[k];
        if (have_prev[k]) {
          const double dE = eps_macro[k] - prev_e[k];
          const double dS = sm[k] - prev_s[k];
          if (std::fabs(dE) > 1e-14 && std::isfinite(dS)) {
            const double slope = dS / dE;
            if (std::isfinite(slope) && slope > 0.01 * E_avg) step = err[k] / slope;
          }
        }
        have_prev[k] = true;
        prev_e[k] = eps_macro[k];
        prev_s[k] = sm[k];
        eps_macro[k] += step;
      }
    }
  }
  if (iters32) *iters32 = it32;
  return MGK_OK;
}

// ---------------------------------------------------------------------------
// Native finite-strain flip: the sub-stepped plastic flip of one event or one
// cascade batch (mgkmc/aqs.py:897-996, mgkmc/kmc_simulator.py:905-1000).
// ---------------------------------------------------------------------------
namespace {

// expm of a d x d matrix (d <= 3): scaling by 2^-s until the 1-norm is at most
// 1/4, the Taylor polynomial of degree 18 (Horner), then s squarings.  The
// truncation error is below 1e-25 of the norm, so the result is exact to
// rounding; the flips' increments C/N have norms of about 0.005.
void expm_small(const double* A, int d, double* E) {
  double nrm = 0.0;
  for (int j = 0; j < d; ++j) {
    double c = 0.0;
    for (int i = 0; i < d; ++i) c += std::fabs(A[i * d + j]);
    nrm = std::fmax(nrm, c);
  }
  int s = 0;
  while (nrm > 0.25 && s < 60) {
    nrm *= 0.5;
    ++s;
  }
  const double sc = std::ldexp(1.0, -s);
  double B[9], P[9], T[9];
  for (int k = 0; k < d * d; ++k) B[k] = A[k] * sc;
  // Horner: P = I + B/1 (I + B/2 (I + ... (I + B/18)))
  for (int k = 0; k < d * d; ++k) P[k] = (k % (d + 1) == 0) ? 1.0 : 0.0;
  for (int n = 18; n >= 1; --n) {
    for (int i = 0; i < d; ++i)
      for (int j = 0; j < d; ++j) {
        double acc = 0.0;
        for (int k = 0; k < d; ++k) acc += B[i * d + k] * P[k * d + j];
        T[i * d + j] = acc / n + (i == j ? 1.0 : 0.0);
      }
    std::memcpy(P, T, sizeof(double) * d * d);
  }
  for (int q = 0; q < s; ++q) {
    for (int i = 0; i < d; ++i)
      for (int j = 0; j < d; ++j) {
        double acc = 0.0;
        for (int k = 0; k < d; ++k) acc += P[i * d + k] * P[k * d + j];
        T[i * d + j] = acc;
      }
    std::memcpy(P, T, sizeof(double) * d * d);
  }
  std::memcpy(E, P, sizeof(double) * d * d);
}

// one finite-strain step with the stepper's solver
int fs_step_once(const mgk_fs_stepper* st, const double* Fb, const double* Pt, const int* Pm, int i3, int j3,
                 void* stream, mgk_fs_info* info) {
  if (st->solver == 2)
    return mgk_fs_solver_step_dbfft(st->shape, st->precision, st->F, st->Fp, st->lam, st->mu, st->mat, Fb, Pt, Pm,
                                    i3, j3, st->E_avg, st->nu_avg, st->tol_NW, st->tol_CG, st->max_NW,
                                    st->tol_macro, st->max_iter_macro, st->pixel, st->P, st->sig, st->ws,
                                    st->ws_bytes, stream, info);
  if (st->solver == 1)
    return mgk_fs_solver_step_al(st->shape, st->precision, st->F, st->Fp, st->lam, st->mu, st->mat, Fb, Pt, Pm, i3,
                                 j3, st->E_avg, st->nu_avg, st->tol_NW, st->max_NW, st->tol_macro,
                                 st->max_iter_macro, st->L0, st->P, st->sig, st->ws, st->ws_bytes, stream, info);
  if (st->mixed)
    return mgk_fs_solver_step_refined(st->shape, st->F, st->Fp, st->lam, st->mu, st->lam32, st->mu32, st->mat, Fb,
                                      Pt, Pm, i3, j3, st->E_avg, st->nu_avg, st->tol_NW, st->tol_CG, st->max_NW,
                                      st->tol_macro, st->max_iter_macro, st->P, st->sig, st->ws, st->ws_bytes,
                                      st->ws32, st->ws32_bytes, stream, info);
  return mgk_fs_solver_step(st->shape, st->precision, st->F, st->Fp, st->lam, st->mu, st->mat, Fb, Pt, Pm, i3, j3,
                            st->E_avg, st->nu_avg, st->tol_NW, st->tol_CG, st->max_NW, st->tol_macro,
                            st->max_iter_macro, st->P, st->sig, st->ws, st->ws_bytes, stream, info);
}

const int kPlaneAx[2] = {0, 2};  // in-plane axes of the (nx, 1, ny) grid

// plane problems: the in-plane Cauchy block (xx, xy, yx, yy) of the 3 x 3 field
int fs_sig2_sync(const mgk_kmc_fs_substep* c, long long NV, hipStream_t s) {
  if (!c->sig2) return MGK_OK;
  const double* sig = (const double*)c->step->sig;
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      MGK_HIP(hipMemcpyAsync(c->sig2 + (size_t)(2 * a + b) * NV, sig + (size_t)(3 * kPlaneAx[a] + kPlaneAx[b]) * NV,
                             sizeof(double) * NV, hipMemcpyDeviceToDevice, s));
  return MGK_OK;
}

// update_stresses of the finite-strain KMC state (mgkmc/aqs.py:276-330; 2D
// mgkmc/kmc_simulator.py:389-445): the step's macro BC, one step, F_macro and
// eps_macro from its F_bar, then the stability test.  0: done; 1: the
// reference's ValueError (no convergence, a NaN field or a mean stress above
// 20 GPa); < 0: an error to pass on.
int fs_update(const mgk_kmc_fs_substep* c, int d, long long NV, double* F_macro, double* eps_macro, int* host_flag,
              hipStream_t s, mgk_fs_info* info) {
  const int i = c->drv_i, j = c->drv_j;
  double Fb[9] = {0}, Pt[9] = {0};
  int Pm[9] = {0};
  for (int k = 0; k < d * d; ++k) Fb[k] = F_macro[k];
  Fb[i * d + j] = (i == j ? 1.0 : 0.0) + eps_macro[i * d + j];
  for (int a = 0; a < d; ++a)
    for (int b = 0; b < d; ++b)
      if (a != b && !(a == i && b == j)) Fb[a * d + b] = 0.0;
  for (int k = 0; k < d; ++k) {
    if (k == i && i == j) continue;
    int t = -1;
    for (int q = 0; q < c->n_targets; ++q)
      if (c->target_k[q] == k) t = q;
    if (t >= 0) {
      Pt[k * d + k] = c->target_val[t];
      Pm[k * d + k] = 1;
    } else {
      Fb[k * d + k] = 1.0;
    }
  }
  double Fb3[9] = {0}, Pt3[9] = {0};
  int Pm3[9] = {0}, i3 = i, j3 = j;
  if (d == 3) {
    for (int k = 0; k < 9; ++k) {
      Fb3[k] = Fb[k];
      Pt3[k] = Pt[k];
      Pm3[k] = Pm[k];
    }
  } else {  // the plane problem: in-plane block, out-of-plane F 1, P and mask 0
    for (int a = 0; a < 2; ++a)
      for (int b = 0; b < 2; ++b) {
        const int q = 3 * kPlaneAx[a] + kPlaneAx[b];
        Fb3[q] = Fb[2 * a + b];
        Pt3[q] = Pt[2 * a + b];
        Pm3[q] = Pm[2 * a + b];
      }
    Fb3[4] = 1.0;
    i3 = kPlaneAx[i];
    j3 = kPlaneAx[j];
  }
  const int rc = fs_step_once(c->step, Fb3, Pt3, Pm3, i3, j3, s, info);
  if (rc == MGK_ERR_SOLVER) return 1;
  if (rc != MGK_OK) return rc;
  double Fm[9], sm[9];
  for (int a = 0; a < d; ++a)
    for (int b = 0; b < d; ++b) {
      const int q = d == 3 ? 3 * a + b : 3 * kPlaneAx[a] + kPlaneAx[b];
      Fm[a * d + b] = info->F_bar[q];
      sm[a * d + b] = info->sig_mean[q];
    }
  int e = fs_sig2_sync(c, NV, s);
  if (e != MGK_OK) return e;
  for (int a = 0; a < d; ++a)
    for (int b = 0; b < d; ++b) {
      F_macro[a * d + b] = Fm[a * d + b];
      eps_macro[a * d + b] = a == b ? Fm[a * d + b] - 1.0 : Fm[a * d + b];
    }
  const long long NV3 = c->step->shape[0] * c->step->shape[1] * c->step->shape[2];
  e = mgk_any_nan2((const double*)c->step->sig, (const double*)c->step->F, 9 * (size_t)NV3, c->flag, host_flag, s);
  if (e != MGK_OK) return e;
  // np.max(np.abs(sig_mean)) propagates a NaN, and a NaN mean fails no comparison
  bool nan_mean = false;
  double mx = 0.0;
  for (int k = 0; k < d * d; ++k) {
    if (std::isnan(sm[k])) nan_mean = true;
    mx = std::fmax(mx, std::fabs(sm[k]));
  }
  if (*host_flag || (!nan_mean && mx / 1e9 > 20.0)) return 1;
  return 0;
}

}  // namespace

namespace {
int fs_flip_core(const mgk_kmc_loop* L, const mgk_kmc_fs_substep* c, int nflips, const long long* voxels,
                 const int* modes, double time, Mt& R, double* F_macro, double* eps_macro, int* substeps,
                 mgk_fs_info* info, hipStream_t s) {
  const mgk_fs_stepper* st = c->step;
  const int d = L->ndim, M = L->M;
  long long NV = 1;
  int shape_i[3] = {1, 1, 1};
  for (int i = 0; i < d; ++i) {
    NV *= L->shape[i];
    shape_i[i] = (int)L->shape[i];
  }
  const long long NV3 = st->shape[0] * st->shape[1] * st->shape[2];
  if (NV3 != NV) {
    mgk_set_error("mgk_kmc_fs_flip: the finite-strain grid does not hold the KMC grid's voxels", __FILE__,
                  __LINE__);
    return MGK_ERR_ARG;
  }
  // the backup's layout: every live buffer at a 256-byte aligned offset
  size_t off[12], total = 0;
  for (int k = 0; k < c->nlive; ++k) {
    off[k] = total;
    total += (c->live_bytes[k] + 255) / 256 * 256;
  }
  if (total > c->backup_bytes) {
    mgk_set_error("mgk_kmc_fs_flip: backup area too small", __FILE__, __LINE__);
    return MGK_ERR_WORKSPACE;
  }
  // pinned staging: redraws (3M), the flips' planar modes (2 per flip), the NaN flag
  double* host = redraw_stage(3 * (size_t)M + 2 * (size_t)nflips + 1);
  if (!host) {
    mgk_set_error("mgk_kmc_fs_flip: pinned host allocation failed", __FILE__, __LINE__);
    return MGK_ERR_HIP;
  }
  double* cm = host + 3 * (size_t)M;
  int* host_flag = (int*)(cm + 2 * (size_t)nflips);
  // the flipped modes C (mgkmc/aqs.py:900-905: the catalog entry before its redraw)
  for (int f = 0; f < nflips; ++f) {
    if (voxels[f] < 0 || voxels[f] >= NV || modes[f] < 0 || modes[f] >= M) {
      mgk_set_error("mgk_kmc_fs_flip: flip outside the grid or the catalog", __FILE__, __LINE__);
      return MGK_ERR_ARG;
    }
    for (int k = 0; k < 2; ++k)
      MGK_HIP(hipMemcpyAsync(cm + 2 * f + k, L->cat + ((size_t)k * M + modes[f]) * NV + voxels[f], sizeof(double),
                             hipMemcpyDeviceToHost, s));
  }
  MGK_HIP(hipStreamSynchronize(s));
  std::vector<double> C((size_t)nflips * d * d, 0.0);
  const double step = d == 3 ? 0.005 : 0.02;
  int n0 = 0;
  for (int f = 0; f < nflips; ++f) {
    double* Cf = C.data() + (size_t)f * d * d;
    Cf[0] = cm[2 * f];
    Cf[d + 1] = -cm[2 * f];
    Cf[1] = cm[2 * f + 1];
    Cf[d] = cm[2 * f + 1];
    double ss = 0.0;
    for (int k = 0; k < d * d; ++k) ss += Cf[k] * Cf[k];
    n0 = std::max(n0, std::max(2, (int)std::ceil(std::sqrt(0.5 * ss) / step)));
  }
  n0 = std::min(n0, 40);
  std::vector<int> schedule;
  if (d == 2) {
    schedule.push_back(1);
    n0 = std::max(n0, 2);
  }
  for (int k = 0; k < 20; ++k) schedule.push_back(n0 + k);
  for (int k = 0; k < c->nlive; ++k)
    if (c->live[k] && c->live_bytes[k])
      MGK_HIP(hipMemcpyAsync((char*)c->backup + off[k], c->live[k], c->live_bytes[k], hipMemcpyDeviceToDevice, s));
  double bfm[9], bem[9];
  std::memcpy(bfm, F_macro, sizeof(double) * d * d);
  std::memcpy(bem, eps_macro, sizeof(double) * d * d);
  double dF[9];
  for (int N : schedule) {
    for (int k = 0; k < c->nlive; ++k)
      if (c->live[k] && c->live_bytes[k])
        MGK_HIP(hipMemcpyAsync(c->live[k], (char*)c->backup + off[k], c->live_bytes[k], hipMemcpyDeviceToDevice, s));
    std::memcpy(F_macro, bfm, sizeof(double) * d * d);
    std::memcpy(eps_macro, bem, sizeof(double) * d * d);
    int e = fs_sig2_sync(c, NV, s);
    if (e != MGK_OK) return e;
    bool failed = false;
    for (int it = 1; it <= N && !failed; ++it) {
      for (int f = 0; f < nflips; ++f) {
        double A[9];
        for (int k = 0; k < d * d; ++k) A[k] = C[(size_t)f * d * d + k] / N;
        expm_small(A, d, dF);
        e = mgk_fp_left_mul((double*)st->Fp, NV, voxels[f], d, dF, s);
        if (e != MGK_OK) return e;
        if (it == N) {
          e = mgk_kmc_flip(d, shape_i, M, 0, voxels[f], modes[f], time, L->jp, L->jt, L->kp.cap, L->jn, L->cat,
                           L->epsp, L->eps_precision, L->epsp_shadow, L->soft, L->tlast, L->prev, nullptr, 0, s);
          if (e != MGK_OK) return e;
          redraw_draws(L, R, host);
          MGK_HIP(hipMemcpyAsync(L->redraw_dev, host, sizeof(double) * 3 * (size_t)M, hipMemcpyHostToDevice, s));
          e = mgk_kmc_set_voxel(NV, M, 2, voxels[f], L->redraw_directions ? L->redraw_dev : nullptr,
                                L->redraw_barriers ? L->redraw_dev + 2 * M : nullptr, L->cat, L->Q0, s);
          if (e != MGK_OK) return e;
          // the staging buffer is refilled by the next voxel's draws: wait for the upload
          MGK_HIP(hipStreamSynchronize(s));
        }
      }
      e = fs_update(c, d, NV, F_macro, eps_macro, host_flag, s, info);
      if (e < 0) return e;
      failed = e == 1;
    }
    if (!failed) {
      *substeps = N;
      return MGK_OK;
    }
  }
  mgk_set_error("Mechanical solver failed to converge under sub-stepping.", __FILE__, __LINE__);
  return MGK_ERR_SOLVER;
}
}  // namespace

extern "C" int mgk_kmc_fs_flip(const mgk_kmc_fs_substep* c, int nflips, const long long* voxels, const int* modes,
                               double time, mgk_mt_state* rng, double* F_macro, double* eps_macro, int* substeps,
                               mgk_fs_info* info, void* stream) {
  const mgk_kmc_loop* L = c ? c->kmc : nullptr;
  const mgk_fs_stepper* st = c ? c->step : nullptr;
  if (!c || !L || !st || nflips < 1 || !voxels || !modes || !rng || rng->pos < 0 || rng->pos > 624 || !F_macro ||
      !eps_macro || !substeps || !info || !c->flag || !c->backup || c->nlive < 0 || c->nlive > 12 ||
      (L->ndim != 2 && L->ndim != 3) || L->M < 1 || !L->cat || !L->Q0 || !L->soft || !L->tlast || !L->redraw_dev ||
      !st->F || !st->Fp || !st->sig || (L->ndim == 2) != (c->sig2 != nullptr) || c->drv_i < 0 ||
      c->drv_i >= L->ndim || c->drv_j < 0 || c->drv_j >= L->ndim || c->n_targets < 0 || c->n_targets > 3 ||
      st->precision != 0 || (!L->redraw_directions && !L->redraw_barriers)) {
    mgk_set_error("mgk_kmc_fs_flip: bad argument", __FILE__, __LINE__);
    return MGK_ERR_ARG;
  }
  Mt R{rng};
  return fs_flip_core(L, c, nflips, voxels, modes, time, R, F_macro, eps_macro, substeps, info,
                      (hipStream_t)stream);
}

// ---------------------------------------------------------------------------
// Native small-strain sub-stepping of the 2D nonlinear laws (mgkmc/
// kmc_simulator.py:263-385): one full flip first, then N sub-steps of mode / N.
// ---------------------------------------------------------------------------
namespace {

// elastic_run with the law (mgkmc/kmc_simulator.py:449-525): 0 done, 1 the
// reference's ValueError / FloatingPointError (retry), < 0 an error
int law_update(const mgk_kmc_loop* L, const mgk_law_stepper* st, const double* eps_macro, mgk_ls_info* info,
               hipStream_t s) {
  int rc;
  if (st->kind == 0) {
    const int have = (st->plane_stress && st->e33_have && *st->e33_have) ? 1 : 0;
    rc = mgk_ls_solve_landau2d(st->shape, st->lam, st->mu, st->lam0, st->mu0, st->landau, st->plane_stress,
                               eps_macro, (const double*)L->epsp, st->eps, st->sig, st->e33, 1, have, st->tau,
                               st->anderson_m, st->aa, 1, st->max_iter, st->tol, st->ws, st->ws_bytes, s, info);
    if (st->plane_stress && st->e33_have) *st->e33_have = 1;
    if (rc == MGK_ERR_NONFINITE) return 1;
  } else {
    rc = mgk_ls_solve_secant(2, st->shape, 0, st->lam, st->mu, st->lam0, st->mu0, st->secant, eps_macro,
                             L->epsp, st->eps, st->sig, st->tau, 0, st->max_iter, st->tol, st->ws, st->ws_bytes, s,
                             info);
  }
  if (rc < 0 && rc != MGK_ERR_NONFINITE) return rc;
  return unstable_mean(*info, 2) ? 1 : 0;
}

int law_flip_core(const mgk_kmc_loop* L, const mgk_kmc_law_substep* c, int nflips, const long long* voxels,
                  const int* modes, double time, Mt& R, const double* eps_macro, int* substeps, mgk_ls_info* info,
                  hipStream_t s) {
  const int d = 2, M = L->M;
  const long long NV = L->shape[0] * L->shape[1];
  int shape_i[3] = {(int)L->shape[0], (int)L->shape[1], 1};
  size_t off[12], total = 0;
  for (int k = 0; k < c->nlive; ++k) {
    off[k] = total;
    total += (c->live_bytes[k] + 255) / 256 * 256;
  }
  if (total > c->backup_bytes) {
    mgk_set_error("mgk_kmc_law_flip: backup area too small", __FILE__, __LINE__);
    return MGK_ERR_WORKSPACE;
  }
  for (int f = 0; f < nflips; ++f)
    if (voxels[f] < 0 || voxels[f] >= NV || modes[f] < 0 || modes[f] >= M) {
      mgk_set_error("mgk_kmc_law_flip: flip outside the grid or the catalog", __FILE__, __LINE__);
      return MGK_ERR_ARG;
    }
  double* host = redraw_stage(3 * (size_t)M + 2 * (size_t)nflips);
  if (!host) {
    mgk_set_error("mgk_kmc_law_flip: pinned host allocation failed", __FILE__, __LINE__);
    return MGK_ERR_HIP;
  }
  double* cm = host + 3 * (size_t)M;
  for (int k = 0; k < c->nlive; ++k)
    if (c->live[k] && c->live_bytes[k])
      MGK_HIP(hipMemcpyAsync((char*)c->backup + off[k], c->live[k], c->live_bytes[k], hipMemcpyDeviceToDevice, s));
  auto restore = [&]() -> int {
    for (int k = 0; k < c->nlive; ++k)
      if (c->live[k] && c->live_bytes[k])
        MGK_HIP(hipMemcpyAsync(c->live[k], (char*)c->backup + off[k], c->live_bytes[k], hipMemcpyDeviceToDevice, s));
    return MGK_OK;
  };
  std::vector<int> schedule{1};
  for (size_t si = 0; si < schedule.size(); ++si) {
    const int N = schedule[si];
    int e;
    if (N > 1 && (e = restore()) != MGK_OK) return e;
    bool failed = false;
    for (int it = 1; it <= N && !failed; ++it) {
      for (int f = 0; f < nflips; ++f) {
        e = mgk_kmc_flip(d, shape_i, M, 0, voxels[f], modes[f], time, L->jp, L->jt, L->kp.cap, L->jn, L->cat,
                         L->epsp, L->eps_precision, L->epsp_shadow, L->soft, L->tlast, L->prev, nullptr,
                         it < N ? -N : N, s);
        if (e != MGK_OK) return e;
        if (it == N) {
          redraw_draws(L, R, host);
          MGK_HIP(hipMemcpyAsync(L->redraw_dev, host, sizeof(double) * 3 * (size_t)M, hipMemcpyHostToDevice, s));
          e = mgk_kmc_set_voxel(NV, M, 2, voxels[f], L->redraw_directions ? L->redraw_dev : nullptr,
                                L->redraw_barriers ? L->redraw_dev + 2 * M : nullptr, L->cat, L->Q0, s);
          if (e != MGK_OK) return e;
          MGK_HIP(hipStreamSynchronize(s));
        }
      }
      e = law_update(L, c->step, eps_macro, info, s);
      if (e < 0) return e;
      failed = e == 1;
    }
    if (!failed) {
      *substeps = N;
      return MGK_OK;
    }
    if (N == 1) {
      // restore, then N_h from the modes as they were (mgkmc/kmc_simulator.py:290-300)
      if ((e = restore()) != MGK_OK) return e;
      for (int f = 0; f < nflips; ++f)
        for (int k = 0; k < 2; ++k)
          MGK_HIP(hipMemcpyAsync(cm + 2 * f + k, L->cat + ((size_t)k * M + modes[f]) * NV + voxels[f],
                                 sizeof(double), hipMemcpyDeviceToHost, s));
      MGK_HIP(hipStreamSynchronize(s));
      int nh = 2;
      for (int f = 0; f < nflips; ++f) {
        const double g0 = cm[2 * f], g1 = cm[2 * f + 1];
        const double ss = ((g0 * g0 + g1 * g1) + g1 * g1) + g0 * g0;  // C:C of [[g0, g1], [g1, -g0]]
        nh = std::max(nh, std::max(2, (int)std::ceil(std::sqrt(0.5 * ss) / 0.02)));
      }
      nh = std::min(nh, 40);
      for (int k = 0; k < 20; ++k) schedule.push_back(std::max(nh, 2) + k);
    }
  }
  mgk_set_error("Mechanical solver failed to converge under small-strain sub-stepping.", __FILE__, __LINE__);
  return MGK_ERR_SOLVER;
}

}  // namespace

extern "C" int mgk_kmc_law_flip(const mgk_kmc_loop* L, const mgk_kmc_law_substep* c, int nflips,
                                const long long* voxels, const int* modes, double time, mgk_mt_state* rng,
                                const double* eps_macro, int* substeps, mgk_ls_info* info, void* stream) {
  if (!law_ok(L, c) || c->step->kind == 2 || nflips < 1 || !voxels || !modes || !rng || rng->pos < 0 || rng->pos > 624 || !eps_macro ||
      !substeps || !info) {
    mgk_set_error("mgk_kmc_law_flip: bad argument", __FILE__, __LINE__);
    return MGK_ERR_ARG;
  }
  Mt R{rng};
  return law_flip_core(L, c, nflips, voxels, modes, time, R, eps_macro, substeps, info, (hipStream_t)stream);
}
