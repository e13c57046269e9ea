/*
 * mgkmc_hip.h -- C ABI of libmgkmc_hip.so, the MI355X (gfx950) equilibrium-solve
 * hot path of mgkmc.
 *
 * All field pointers are DEVICE pointers.  Tensor fields are structure-of-arrays,
 * component-major, voxel index (x*ny + y)*nz + z (2D: x*ny + y):
 *   3D symmetric fields: 6 components xx, yy, zz, yz, xz, xy
 *   2D symmetric fields: 3 components xx, yy, xy
 * precision: 0 = FP64, 1 = FP32 (for the type of every field pointer).
 * stream: a hipStream_t (NULL = default stream).
 * Return value: MGK_OK (0), MGK_NOT_CONVERGED (1, silent max_iter stop, as the
 * reference's LS loop), or a negative error; mgk_last_error() describes it.
 */
#ifndef MGKMC_HIP_H
#define MGKMC_HIP_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MGK_OK 0
#define MGK_NOT_CONVERGED 1
#define MGK_ERR_ARG -1
#define MGK_ERR_HIP -2
#define MGK_ERR_WORKSPACE -3
#define MGK_ERR_NONFINITE -4
#define MGK_ERR_SOLVER -5

int mgk_version(void);
const char* mgk_last_error(void);
int mgk_device_info(int* lds_per_block, int* cu_count);

/* Live per-kernel timing with HIP events on the launching stream (no
 * counterpart in the reference; used by bench.py).  kid: 0 ls_tau_xfwd,
 * 1 ls_fft_mid, 2 ls_pencil_gamma, 3 ls_xinv_update, 4 kmc_barrier_rates,
 * 5 kmc_select, 6 fs_constitutive, 7 fs_xfwd, 8 fs_fft_mid, 9 fs_pencil_proj,
 * 10 fs_xinv, 11 fs_vector, 12 ls_tau_xfwd_start, 13 ls_xinv_update_start (the
 * first LS iteration when it reads its starting field as a constant or in FP32;
 * 0 and 3 then time the later iterations).  total_bytes: algorithmic HBM bytes
 * of the timed launches. */
/* Streaming device copy of `bytes` (a multiple of 16, 16-byte aligned pointers) on
 * `stream`: the practical HBM read + write ceiling that bench.py reports beside the
 * kernels' roofline (no counterpart in the reference). */
int mgk_stream_copy(const void* src, void* dst, size_t bytes, void* stream);

void mgk_profile_enable(int on);
void mgk_profile_reset(void);
int mgk_profile_read(int kid, double* total_ms, double* total_bytes, long long* launches);

/* ---------------------------------------------------------------------------
 * Small-strain Lippmann-Schwinger fixed point.
 * Replaces spectral_solver_2d (mgkmc/linear_elastic_simulator.py:12-88) and
 * spectral_solver_3d (mgkmc/linear_elastic_simulator.py:274-376); also the body
 * of update_stress_fft_full (mgkmc/stz/update_fft.py:4-22).
 * ------------------------------------------------------------------------- */
typedef struct {
  int iters;           /* LS iterations executed                              */
  int status;          /* MGK_OK / MGK_NOT_CONVERGED / MGK_ERR_NONFINITE      */
  double diff;         /* last |eps_new-eps|_F / (|eps|_F + 1e-20)            */
  double eps_mean[9];  /* <eps>, full d x d row-major (2D: first 4 entries)   */
  double sig_mean[9];  /* <sig> of the symmetric stress                       */
} mgk_ls_info;

/* Workspace bytes needed by mgk_ls_solve for this grid and precision. */
int mgk_ls_workspace_bytes(int ndim, const int64_t* shape, int precision, size_t* bytes);

/* lam, mu: per-voxel Lame fields; lam0, mu0: reference medium (the reference
 * computes both from E.mean(), nu.mean()).  eps_bar: d*d row-major FP64 host
 * array.  eps_plastic: symmetric eigenstrain field or NULL.  eps: in (warm != 0:
 * initial guess, eps_init of the reference) / out.  sig: out or NULL.
 * Blocks until the solve has finished on `stream`. */
int mgk_ls_solve(int ndim, const int64_t* shape, int precision, const void* lam, const void* mu,
                 double lam0, double mu0, const double* eps_bar, const void* eps_plastic, void* eps,
                 void* sig, int warm, int max_iter, double tol, void* workspace, size_t workspace_bytes,
                 void* stream, mgk_ls_info* info);

/* No-allocation variant of mgk_ls_solve (no counterpart in the reference: for
 * callers that cannot allocate device memory -- code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments and
 * results as mgk_ls_solve without the workspace: the workspace and the FFT twiddle
 * tables live in a static device array of MGK_LS_NOALLOC_BYTES per GPU (held until
 * the process ends; its device address is looked up once per device).  Grids whose
 * workspace plus twiddles exceed it (for example 3D FP64 above ~100^3, FP32 above
 * 128^3) and axes longer than 4096 return MGK_ERR_WORKSPACE / MGK_ERR_ARG before any
 * launch.  Calls no hipMalloc, hipFree, hipHostMalloc or event creation.  All calls
 * on one GPU share the array: calls that overlap on one GPU (other streams or
 * threads) race.  Blocks until the solve has finished on `stream`. */
#define MGK_LS_NOALLOC_BYTES ((size_t)64 << 20)
int mgk_ls_solve_noalloc(int ndim, const int64_t* shape, int precision, const void* lam, const void* mu,
                         double lam0, double mu0, const double* eps_bar, const void* eps_plastic, void* eps,
                         void* sig, int warm, int max_iter, double tol, void* stream, mgk_ls_info* info);

/* FP32 form with FP64 refinement of the same solve (replaces the same reference
 * functions as mgk_ls_solve; the precision split has no reference counterpart).
 * Stage 1: FP32 fixed point on (lam32, mu32, epsp32) from eps32 (warm != 0: eps32 is
 * the initial guess) until diff < tol32, without the final stress.  Stage 2: eps32 is
 * widened into eps64 and the FP64 fixed point continues from it (warm start) until the
 * reference's stop test holds at tol in FP64; sig64 and info describe stage 2 (info->iters
 * = FP64 iterations), *iters32 the FP32 iterations.  epsp32/epsp64 are both NULL or both
 * given; ws32/ws64 are sized by mgk_ls_workspace_bytes for precision 1 and 0.  Blocks
 * until done. */
int mgk_ls_solve_refined(int ndim, const int64_t* shape, const void* lam32, const void* mu32, const void* lam64,
                         const void* mu64, double lam0, double mu0, const double* eps_bar, const void* epsp32,
                         const void* epsp64, void* eps32, void* eps64, void* sig64, int warm, int max_iter, double tol,
                         double tol32, void* ws32, size_t ws32_bytes, void* ws64, size_t ws64_bytes, void* stream,
                         mgk_ls_info* info, int* iters32);

/* Batched forms of mgk_ls_solve and mgk_ls_solve_refined (no counterpart in the
 * reference: the serving form, in which several resident simulations of one grid
 * shape are solved with one launch per sweep).  Each member carries its own fields,
 * reference medium (lam0, mu0), macro strain and workspace (sized as for the
 * unbatched call); the *32 fields are read by the refined form only.  A member's
 * fields, stress, iteration counts and status are bit-identical to the unbatched
 * call on the same inputs: it runs the same kernels on its own buffers and stops on
 * its own stop test, while the batch continues until every member has stopped.
 * info (and iters32) are arrays of nb.  The return value is the first negative
 * member status, else MGK_NOT_CONVERGED if any member did not converge, else
 * MGK_OK.  Blocks until done. */
#define MGK_LS_MAX_BATCH 8
typedef struct {
  const void* lam;           /* FP64 (or the precision of mgk_ls_solve_batch) */
  const void* mu;
  double lam0, mu0;
  const double* eps_bar;     /* d*d row-major, host                           */
  const void* eps_plastic;   /* or NULL                                       */
  void* eps;                 /* in (warm) / out                               */
  void* sig;                 /* out or NULL                                   */
  void* workspace;
  size_t workspace_bytes;
  const void* lam32;         /* refined form: FP32 copies and stage-1 field   */
  const void* mu32;
  const void* eps_plastic32;
  void* eps32;
  void* workspace32;
  size_t workspace32_bytes;
} mgk_ls_member;

int mgk_ls_solve_batch(int ndim, const int64_t* shape, int precision, int nb, const mgk_ls_member* members,
                       int warm, int max_iter, double tol, void* stream, mgk_ls_info* info);

int mgk_ls_solve_refined_batch(int ndim, const int64_t* shape, int nb, const mgk_ls_member* members, int warm,
                               int max_iter, double tol, double tol32, void* stream, mgk_ls_info* info,
                               int* iters32);

/* Landau small-strain model (3D).  Replaces spectral_solver_landau_3d, default
 * "dbfft" path (mgkmc/linear_elastic_simulator.py:1231-1350) with the stress of
 * stress_from_strain_landau_3d (mgkmc/elasticity.py:297-335, :811-861).  Constants in
 * Pa as given; cap_limit is the resolved E_cap (the reference's default when its
 * limit is None or <= 0: sqrt(0.2 <mu> / |g4|) if g4 < 0, else 999). */
typedef struct {
  double v1, v2, v3, g1, g2, g3, g4;
  int capping;     /* strain capping on/off                       */
  int cap_smooth;  /* 1 smooth (tanh), 0 piecewise                */
  double cap_limit, cap_power, cap_ratio;
} mgk_landau_params;

/* lam, mu per voxel; lam0, mu0 = their means (the reference medium of this solver);
 * tau_buf: 6*NV scratch values of the field precision; other arguments as mgk_ls_solve.
 * Returns MGK_ERR_NONFINITE when an update is not finite (the reference raises
 * FloatingPointError). */
int mgk_ls_solve_landau(const int64_t* shape, int precision, const void* lam, const void* mu, double lam0,
                        double mu0, const mgk_landau_params* params, const double* eps_bar, const void* eps_plastic,
                        void* eps, void* sig, void* tau_buf, int warm, int max_iter, double tol, void* workspace,
                        size_t workspace_bytes, void* stream, mgk_ls_info* info);

/* Landau small-strain model (2D, FP64).  Replaces spectral_solver_landau_2d, default
 * "dbfft" path with Anderson mixing (mgkmc/linear_elastic_simulator.py:952-1228) and the
 * stress of stress_from_strain_landau_2d (mgkmc/elasticity.py:487-638).  Fields 3
 * components (xx, yy, xy).  plane_stress: e33 from sigma33 = 0 (Newton, the reference's
 * global step count); e33 (NV values) holds the warm start when e33_persist and e33_have
 * (the reference's e33_state dict) and receives the last e33.  anderson_m in 0..5;
 * aa_buf: (2 (m + 1) + 1) * 3 * NV doubles (history and previous iterate), NULL if m = 0.
 * tau_buf: 3 NV doubles.  Other arguments as mgk_ls_solve_landau. */
int mgk_ls_solve_landau2d(const int64_t* shape, const double* lam, const double* mu, double lam0, double mu0,
                          const mgk_landau_params* params, int plane_stress, const double* eps_bar,
                          const double* eps_plastic, double* eps, double* sig, double* e33, int e33_persist,
                          int e33_have, double* tau_buf, int anderson_m, double* aa_buf, int warm, int max_iter,
                          double tol, void* workspace, size_t workspace_bytes, void* stream, mgk_ls_info* info);

/* The same model's Newton-CG solver: spectral_solver_landau_2d(solver="newton_cg")
 * (mgkmc/linear_elastic_simulator.py:1011-1117).  Newton steps on eps; each solves
 * A dx = b with BiCGSTAB (SciPy's recurrence, rtol = tol, maxiter 150, x0 = 0) where
 * A = G C_h G, G = -Gamma0 of (lam0, mu0) and C_h d = (sigma(eps_el + h d) -
 * sigma(eps_el - h d)) / 2h, h = 1e-7, every stress evaluation with a cold plane-stress
 * e33 Newton; b = -G C_h (eps_bar - <eps>) at the first step, -G sigma after it; eps +=
 * G dx (plus eps_bar - <eps> at the first step); stop when |dx| / (|eps| + 1e-20) < tol
 * after the first step (norms over the full 2x2 fields).  info->iters = Newton steps,
 * info->diff = the last |dx| / |eps|, *krylov_iters = BiCGSTAB iterations in total;
 * MGK_NOT_CONVERGED (silent) after max_iter steps, MGK_ERR_NONFINITE on a non-finite
 * stress or update.  sig: out (3 NV) or NULL.  Workspace: the bytes below. */
int mgk_ls_landau2d_ncg_workspace_bytes(const int64_t* shape, size_t* bytes);
int mgk_ls_solve_landau2d_ncg(const int64_t* shape, const double* lam, const double* mu, double lam0, double mu0,
                              const mgk_landau_params* params, int plane_stress, const double* eps_bar,
                              const double* eps_plastic, double* eps, double* sig, int warm, int max_iter, double tol,
                              void* workspace, size_t workspace_bytes, void* stream, mgk_ls_info* info,
                              int* krylov_iters);

/* Secant-modulus small-strain laws.  Replaces spectral_solver_secant_2d/3d
 * (mgkmc/linear_elastic_simulator.py:614-850; stress_from_strain_secant_2d/3d,
 * mgkmc/elasticity.py:641-810) for law 0, and spectral_solver_rose_2d
 * (mgkmc/linear_elastic_simulator.py:852-950; stress_from_strain_rose_2d,
 * mgkmc/elasticity.py:868-943) for law 1 (2D only).  p1, p2 = (d, k) for law 0,
 * (eta, mu_floor_fraction) for law 1; plane_stress selects the Rose e33 update. */
typedef struct {
  int law;           /* 0 exponential degradation, 1 Rose */
  int plane_stress;  /* Rose only                          */
  double p1, p2;
} mgk_secant_params;

/* lam, mu per voxel; (lam0, mu0) = the reference medium: (mean lam, mean mu) for
 * law 0, (mean lam, (1 + floor) mean mu / 2) for law 1.  Fields 6 (3D) or 3 (2D)
 * components; tau_buf: C*NV scratch values of the field precision.  The reference
 * has no non-finite check here: a non-finite update returns MGK_ERR_NONFINITE with
 * the fields as they are (non-finite), where the reference runs on to max_iter. */
int mgk_ls_solve_secant(int ndim, const int64_t* shape, int precision, const void* lam, const void* mu,
                        double lam0, double mu0, const mgk_secant_params* params, const double* eps_bar,
                        const void* eps_plastic, void* eps, void* sig, void* tau_buf, int warm, int max_iter,
                        double tol, void* workspace, size_t workspace_bytes, void* stream, mgk_ls_info* info);

/* Heat conduction of the local temperature field, the thermal extension of the KMC
 * simulators (heat_conducting_3d, mgkmc/aqs.py:462-478; heat_conducting_2d,
 * mgkmc/kmc_simulator.py:559-575): T <- F^-1[exp(-|k|^2 Ddt) F[T]] with k the
 * periodic wave numbers of box lengths lengths[ndim], then thermostat 1: T0 + (T - T0) *
 * bath_factor, 2: T - (<T> - T0), and the cap min(T, cap) when cap > 0.  T: FP64,
 * in place.  Workspace from mgk_heat_workspace_bytes. */
int mgk_heat_workspace_bytes(int ndim, const int64_t* shape, size_t* bytes);
int mgk_heat_step(int ndim, const int64_t* shape, double* T, const double* lengths, double Ddt, double T0,
                  int thermostat, double bath_factor, double cap, void* workspace, size_t workspace_bytes,
                  void* stream);

/* ---------------------------------------------------------------------------
 * KMC: barriers, rates, selection, compaction, flips.
 * Device layouts: Q, Q0 [m][v]; catalog [k][m][v] with k over 2 planar
 * parameters (gxx, gxy) (cat_kind 0) or d*d row-major entries (cat_kind 1);
 * sigma symmetric SoA FP64 (sig_kind 0), FP32 (1) or d*d row-major FP64 (2);
 * soft [2][v] = (g_p, g_t); tlast [v]; prev [d*d][v].
 * ------------------------------------------------------------------------- */
typedef struct {
  double volume;       /* barrier work volume (pixel^3)                        */
  double rate_volume;  /* rate prefactor volume (1 unless scale_rate_by_volume) */
  double nu0;          /* attempt frequency                                    */
  double temperature;  /* K, used when Tfield == NULL                          */
  double cap;          /* softening cap                                        */
  double tau;          /* transient softening decay time (inf allowed)         */
  double time;         /* current simulation time                              */
  int scheme;          /* 0 isotropic, 1 directional                           */
  int cascade;         /* 1: only q > 0 modes are events                       */
} mgk_kmc_params;

typedef struct {
  long long item;   /* selected voxel (or array index)                        */
  int mode;         /* selected mode within the voxel                         */
  int fallthrough;  /* nonzero if the running sum never reached r             */
  double r;         /* the draw U(0, total) actually used                     */
  double cum_before, cum_at;  /* running sum before / including the event    */
  double q;         /* barrier of the selected mode (is_instab test)          */
} mgk_kmc_sel;

int mgk_kmc_workspace_bytes(long long nvox, size_t* bytes);

/* Fused barrier + rate pass.  Replaces compute_barrier (mgkmc/stz/barriers.py:4-100;
 * 2D kmc_simulator_functions.py:100-160) followed by compute_rates
 * (mgkmc/stz/kmc.py:4-73; 2D kmc_simulator_functions.py:5-56): writes Q (if not
 * NULL), per-voxel FP64 rate sums and the total (returned if total_out). */
int mgk_kmc_barrier_rates(int ndim, long long nvox, int M, int sig_kind, int cat_kind, const void* sig,
                          const double* Q0, const double* cat, const double* soft, const double* tlast,
                          const double* prev, const double* Tfield, const mgk_kmc_params* params, double* Q,
                          void* workspace, void* stream, double* total_out);

/* Event selection after mgk_kmc_barrier_rates (same inputs): r = total * u, first
 * (voxel, mode) in (x,y,z,m) order whose running FP64 sum reaches r; the voxel's
 * barriers are recomputed bit-identically, so Q need not be stored.  Replaces
 * select_event + decode_index (mgkmc/stz/kmc.py:76-103; 2D
 * kmc_simulator_functions.py:58-76). */
int mgk_kmc_select(int ndim, long long nvox, int M, int sig_kind, int cat_kind, double u, const void* sig,
                   const double* Q0, const double* cat, const double* soft, const double* tlast,
                   const double* prev, const double* Tfield, const mgk_kmc_params* params, void* workspace,
                   void* stream, mgk_kmc_sel* out);

/* compute_rates on a Q field in the reference (v, m) layout: full per-mode
 * rates (-1 = not an event), compacted rates/indices in reference order, total
 * (mgkmc/stz/kmc.py:4-73). */
int mgk_kmc_rates_from_q(long long nvox, int M, const double* Qvm, const double* Tfield,
                         const mgk_kmc_params* params, double* rates_full, long long* idx_out, double* val_out,
                         void* workspace, void* stream, long long* count_out, double* total_out);

/* select_event over a rates array with a given draw r (mgkmc/stz/kmc.py:76-92). */
int mgk_kmc_select_array(long long n, const double* rates, double r, void* workspace, void* stream,
                         mgk_kmc_sel* out);

/* Indices of a[i] < thr in order (find_unstable, mgkmc/stz/cascade.py:4-41). */
int mgk_find_below(long long n, const double* a, double thr, long long* idx_out, void* stream,
                   long long* count_out);

/* apply_flip_soa (mgkmc/stz/cascade.py:43-113; 2D kmc_simulator_functions.py:183-218)
 * on device state; also sets prev (if not NULL) to the mode, as the callers do
 * (mgkmc/aqs.py:999, kmc_simulator.py:1026).  epsp_shadow (FP64 epsp only, may be
 * NULL): an FP32 copy of epsp whose flipped voxel receives the rounded new values.
 * eps_div: 1 for the flip itself; N > 1 adds mode / N to epsp (the last sub-step of a
 * sub-stepped flip, mgkmc/kmc_simulator.py:268-385, which also applies the softening,
 * time and prev updates); -N adds mode / N only (the earlier sub-steps); 0 leaves epsp
 * alone (may be NULL): the softening, time and prev updates of a finite-strain flip,
 * whose mode enters F_p instead (mgkmc/aqs.py:912-975). */
int mgk_kmc_flip(int ndim, const int* shape, int M, int cat_kind, long long v, int m, double time, double jp,
                 double jt, double gmax, double jn, const double* cat, void* epsp, int precision, float* epsp_shadow,
                 double* soft, double* tlast, double* prev, double* gamma_out, int eps_div, void* stream);

/* Redraw of one voxel's modes (mgkmc/aqs.py:1007-1019): newcat [m][k], newq0 [m]. */
int mgk_kmc_set_voxel(long long nvox, int M, int CK, long long v, const double* newcat, const double* newq0,
                      double* cat, double* Q0, void* stream);

/* Fast patching, the stress predictor that replaces a flip's re-solve in the
 * reference's KMC loops (mgkmc/aqs.py:1021-1047, mgkmc/kmc_simulator.py:764-774,
 * :1044-1066, :1094-1096).  In one pass over sig (symmetric SoA, sig_kind 0 FP64 or
 * 1 FP32; 2D on the (nx, 1, ny) LS grid), per voxel and in this order:
 * sig += sig_unit * a (if sig_unit, FP64, same layout: sigma_macro_unit over the
 * waiting time); for every offset of the periodic box of radius radius[] (one per
 * simulation axis, 0 on a flat axis) around voxel v that lands on the voxel, in
 * (dx, dy, dz) order, sig += sum_b coef[b] * K_b (nbasis = 5 in 3D, 2 in 2D, 0 for
 * no patch; kernels [b][c][2Rx+1][2Ry+1][2Rz+1] FP64); then sig += shift[c] (the
 * combined missing mean).  Products and sums are rounded separately, as numpy
 * does.  sig_mean (host, optional): the new field's mean, d*d row-major, after a
 * stream synchronisation.  Workspace from mgk_kmc_patch_workspace_bytes. */
int mgk_kmc_patch_workspace_bytes(size_t* bytes);
int mgk_kmc_patch(int ndim, const int64_t* shape, int sig_kind, void* sig, const double* sig_unit, double a,
                  long long v, const int* radius, int nbasis, const double* kernels, const double* coef,
                  const double* shift, void* workspace, size_t workspace_bytes, void* stream, double* sig_mean);

/* Initial modes drawn on the device (no counterpart in the reference, which draws
 * Q0 and the catalog on the host with the legacy stream: mgkmc/aqs.py:101-120,
 * kmc_simulator.py:117-127).  For callers that start many large simulations (the
 * serving loop).  Q0 [m][v] = clip(mean + std*N(0,1)); planar catalog [k][m][v],
 * k = gxx, gxy = 0.5*gamma0*N(0,1); Philox4x32-10 keyed by seed, counter (m, v, draw),
 * Box-Muller normals: deterministic in (seed, v, m). */
int mgk_kmc_init_modes(long long nvox, int M, uint64_t seed, double mean, double std_dev, int has_min,
                       double min_cutoff, int has_max, double max_cutoff, double gamma0, double* Q0, double* cat,
                       void* stream);

/* ---------------------------------------------------------------------------
 * Finite strain (de Geus Newton-Krylov).  Fields: 9 components F[3*i+j][v].
 * ------------------------------------------------------------------------- */
typedef struct {
  int model;       /* 0 svk, 1 landau, 2 murnaghan, 3 neo-Hookean (2D only:
                      the (nx, 1, ny) grid, plane 1 or 2)                  */
  int capping;     /* landau strain capping on/off                          */
  int cap_smooth;  /* 1 smooth (tanh), 0 piecewise                          */
  int plane;       /* 0 3D; a 2D problem on an (nx, 1, ny) grid: 1 plane strain,
                      2 plane stress (lam = the plane-stress lam*): svk, landau
                      (capped or not) and murnaghan at their out-of-plane E33
                      with the condensed in-plane tangent, neo-Hookean at its
                      out-of-plane stretch (finite_strain_simulator.py:486-1031);
                      the Newton stop test then uses the in-plane |F| and
                      Cauchy J = det F * F33 (finite_strain_simulator.py:
                      2356-2572).  A model outside 0..3 or a plane outside 0..2
                      returns MGK_ERR_ARG. */
  double v1, v2, v3, g1, g2, g3, g4;  /* Landau constants in Pa             */
  double Am, Bm, Cm;                  /* Murnaghan constants                */
  double cap_limit;                   /* resolved capping limit E_cap       */
  double cap_power, cap_ratio;        /* smooth power, tangent ratio        */
  const double* landau_field;         /* NULL, or per-voxel landau constants
                                         (7 x nvox device FP64 in Pa, v1 v2 v3 g1
                                         g2 g3 g4) overriding v1..g4, as the
                                         reference takes arrays
                                         (finite_strain_simulator.py:1678-1690) */
} mgk_fs_material;

typedef struct {
  int status;        /* MGK_OK or MGK_ERR_SOLVER                             */
  int fail_kind;     /* 1 line search, 2 Newton, 3 macro BC loop             */
  int newton_iters, krylov_iters, macro_iters;
  double max_err;    /* last macro stress error (Pa)                         */
  double F_mean[9], sig_mean[9], F_bar[9];
} mgk_fs_info;

int mgk_fs_workspace_bytes(const int64_t* shape, int precision, size_t* bytes);

/* One load step: replaces finite_strain_solver_step_3d, solver="newton_cg"
 * (mgkmc/finite_strain_simulator.py:2967-3168).  F in/out (device, 9 comps);
 * Fp optional; lam/mu per voxel (the reference reads C4[0,0,1,1], C4[0,1,0,1]);
 * F_bar, P_target row-major 9 (host); P_mask 9 ints (host).  BiCGSTAB follows
 * SciPy 1.15.3 (x0 = 0, maxiter 150, rtol = tol_CG). */
int mgk_fs_solver_step(const int64_t* shape, int precision, void* F, const void* Fp, const void* lam,
                       const void* mu, const mgk_fs_material* mat, const double* F_bar, const double* P_target,
                       const int* P_mask, int i_drv, int j_drv, double E_avg, double nu_avg, double tol_NW,
                       double tol_CG, int max_NW, double tol_macro, int max_iter_macro, void* P_out,
                       void* sig_out, void* workspace, size_t workspace_bytes, void* stream, mgk_fs_info* info);

/* One load step with the augmented-Lagrangian solver: replaces
 * finite_strain_solver_step_3d, solver="al" (_al_step_3d, mgkmc/finite_strain_simulator.py:
 * 2145-2242, inside the same macro loop, :3011-3165).  tol, max_iter = the reference's
 * tol_NW, max_NW; L0 = {<lam + 2 mu>, <lam>, <mu>}, the voxel means of C4_iiii, C4_iijj,
 * C4_ijij of the linear stiffness (its reference medium).  Other arguments as
 * mgk_fs_solver_step; newton_iters counts AL iterations. */
int mgk_fs_solver_step_al(const int64_t* shape, int precision, void* F, const void* Fp, const void* lam,
                          const void* mu, const mgk_fs_material* mat, const double* F_bar, const double* P_target,
                          const int* P_mask, int i_drv, int j_drv, double E_avg, double nu_avg, double tol,
                          int max_iter, double tol_macro, int max_iter_macro, const double* L0, void* P_out,
                          void* sig_out, void* workspace, size_t workspace_bytes, void* stream, mgk_fs_info* info);

/* One load step with the displacement-based FFT solver: replaces
 * finite_strain_solver_step_3d, solver="dbfft" (_dbfft_step_3d and
 * solve_dbfft_linear_system_3d, mgkmc/finite_strain_simulator.py:1212-1379,1508-1620, in the
 * macro loop :3011-3165).  tol, max_iter = the reference's tol_NW, max_NW; tol_CG as
 * mgk_fs_solver_step (BiCGSTAB maxiter 150, preconditioner (xi.<K4>.xi)^-1); pixel sets the
 * wave vectors 2 pi fftfreq(n, pixel).  Grid sizes and pixel values for which the reference's
 * Nyquist test misses its entry are rejected (MGK_ERR_ARG), and so is precision 1 (the
 * reference's stop test |b_hat| / |P| scales with sqrt(NV) times the rounding of P).  newton_iters counts DBFFT
 * Newton iterations; failures return MGK_ERR_SOLVER with the reference's messages. */
int mgk_fs_solver_step_dbfft(const int64_t* shape, int precision, void* F, const void* Fp, const void* lam,
                             const void* mu, const mgk_fs_material* mat, const double* F_bar,
                             const double* P_target, const int* P_mask, int i_drv, int j_drv, double E_avg,
                             double nu_avg, double tol, double tol_CG, int max_iter, double tol_macro,
                             int max_iter_macro, double pixel, void* P_out, void* sig_out, void* workspace,
                             size_t workspace_bytes, void* stream, mgk_fs_info* info);

/* FP32 form of the same step (replaces the same reference function; the precision
 * split has no reference counterpart): the Newton loop, P, the residual -G:P, the line
 * search and the macro BC update run in FP64 on F, lam, mu (FP64); each BiCGSTAB solve
 * runs in FP32 (vectors, tangent action, projection FFTs; Krylov scalars in FP64) on
 * lam32, mu32 with rtol = tol_CG, and its solution is widened into the FP64 update.
 * ws32: mgk_fs_workspace_bytes(shape, 1) bytes. */
int mgk_fs_solver_step_refined(const int64_t* shape, void* F, const void* Fp, const void* lam, const void* mu,
                               const void* lam32, const void* mu32, const mgk_fs_material* mat, const double* F_bar,
                               const double* P_target, const int* P_mask, int i_drv, int j_drv, double E_avg,
                               double nu_avg, double tol_NW, double tol_CG, int max_NW, double tol_macro,
                               int max_iter_macro, void* P_out, void* sig_out, void* workspace,
                               size_t workspace_bytes, void* ws32, size_t ws32_bytes, void* stream,
                               mgk_fs_info* info);

/* P(F) and optionally dP = K:dF (reference matvec convention) per voxel
 * (constitutive_hyperelastic_3d, mgkmc/finite_strain_simulator.py:1647-1885). */
int mgk_fs_constitutive(long long nvox, int precision, const void* F, const void* Fp, const void* lam,
                        const void* mu, const mgk_fs_material* mat, void* P, const void* dF, void* dP,
                        void* stream);

/* The 2D law (constitutive_hyperelastic_2d, mgkmc/finite_strain_simulator.py:591-1091) on
 * the (nx, 1, ny) embedding of the plane problem: mat->plane 1 (plane strain) or 2 (plane
 * stress, with lam = the plane-stress lam*), P and optionally dP as mgk_fs_constitutive,
 * F33 (optional, nvox values) the out-of-plane stretch the reference returns. */
int mgk_fs_constitutive_2d(long long nvox, int precision, const void* F, const void* Fp, const void* lam,
                           const void* mu, const mgk_fs_material* mat, void* P, const void* dF, void* dP, void* F33,
                           void* stream);

/* sigma = P F^T / det F (cauchy_from_P_3d, mgkmc/finite_strain_simulator.py:1892-1904). */
int mgk_fs_cauchy(long long nvox, int precision, const void* P, const void* F, void* sig, void* stream);

/* ---------------------------------------------------------------------------
 * Native KMC event loop.  Runs the per-event sequence of the reference's KMC loops
 * (mgkmc/aqs.py:855-1087, mgkmc/kmc_simulator.py:871-1101: small strain, linear
 * elasticity, KMC mode, planar catalogs) in C++: fused barrier + rate pass, waiting
 * time from the host stream, selection from the selection stream, flip, redraw of the
 * voxel's modes (2M normals, then M clipped Gaussian barriers), LS re-solve.  Random
 * streams are numpy legacy RandomState states, continued here and handed back.
 * ------------------------------------------------------------------------- */
typedef struct {
  uint32_t key[624];
  int pos;         /* index of the next word (624: regenerate first) */
  int has_gauss;   /* a cached second polar-method normal             */
  double gauss;
} mgk_mt_state;

/* n draws from the stream: kind 0 random_sample(), kind 1 loc + scale * normal. */
int mgk_mt_fill(mgk_mt_state* state, int kind, double loc, double scale, long long n, double* out);

struct mgk_kmc_patch_cfg;
struct mgk_kmc_fs_substep;
struct mgk_kmc_law_substep;

typedef struct {
  int ndim;
  int64_t shape[3];
  int M;
  /* KMC state (device; planar catalog [2][M][NV]) */
  double *Q0, *cat, *soft, *tlast, *prev; /* prev may be NULL (isotropic) */
  void* epsp;
  int eps_precision;
  float* epsp_shadow;                     /* FP32 copy of an FP64 epsp (refined LS form), or NULL */
  void* kmc_ws;                           /* mgk_kmc_workspace_bytes */
  double* redraw_dev;                     /* 3 M doubles of device scratch */
  mgk_kmc_params kp;                      /* time is taken from *time */
  double jp, jt, jn, stability_threshold;
  int redraw_directions, redraw_barriers;
  double gamma0, q_mean, q_std, q_min, q_max;
  int q_has_min, q_has_max;
  /* LS: precision 0 FP64, 1 FP32, 2 FP32 + FP64 refinement (mgk_ls_solve_refined) */
  int ls_precision;
  int sig_kind;                           /* of sig for the barrier pass */
  const void *lam, *mu;
  const float *lam32, *mu32;
  double lam0, mu0;
  void *eps, *sig;
  float* eps32;
  void* ws;
  size_t ws_bytes;
  void* ws32;
  size_t ws32_bytes;
  int warm, max_iter;
  double tol, tol32;
  int check_unstable;                     /* 2D elastic_run: NaN or |<sigma>| > 20 GPa stops the run */
  double strain_rate;
  double unit[9];                         /* d*d driving direction */
  /* fast patching (mgkmc/aqs.py:1021-1054, kmc_simulator.py:1044-1098), or NULL: the
   * event's re-solve is replaced by the predictor (mgk_kmc_patch with the flipped mode's
   * weights read on the device), a re-solve every sync_interval flips; 2D loops add the
   * macro-unit stress over the remaining increment time once a flip has happened */
  const struct mgk_kmc_patch_cfg* patch;
  /* finite strain (mgkmc/aqs.py:887-996, kmc_simulator.py:895-1000), or NULL: each event's
   * flip is the sub-stepped finite-strain flip (mgk_kmc_fs_flip with this description,
   * F_macro in/out through it) instead of the flip, redraw and LS re-solve; sig is the
   * Cauchy field the barrier pass reads (sig_kind 2; the in-plane block in 2D) and the
   * LS fields are not read */
  const struct mgk_kmc_fs_substep* fs;
  /* nonlinear small-strain law (mgkmc/kmc_simulator.py:263-385; 3D Landau mgkmc/aqs.py:
   * 330-351), or NULL: each event's flip and re-solve are the sub-stepped flip of
   * mgk_kmc_law_flip with this description (3D Landau: flip, redraw, cold Landau solve;
   * the LS fields above are the law solver's; ls_precision 0) */
  const struct mgk_kmc_law_substep* law;
} mgk_kmc_loop;

/* Fast-patching state of a native KMC loop (planar catalogs, sig_kind 0 or 1). */
typedef struct mgk_kmc_patch_cfg {
  int radius[3];                /* per simulation axis (0 on a flat axis)                */
  int nbasis;                   /* 5 (3D) or 2 (2D)                                      */
  const double* kernels;        /* device, as mgk_kmc_patch                              */
  double missing_mean[5][6];    /* per basis, symmetric components                      */
  const double* sig_unit;       /* device FP64 sigma_macro_unit                          */
  void* workspace;              /* mgk_kmc_patch_workspace_bytes                         */
  size_t workspace_bytes;
  int sync_interval;
  int* flips_since_sync;        /* host, in/out                                          */
  int* unit_live;               /* host, in/out: 1 once a patched flip has happened      */
  double* sig_mean;             /* host, out: d*d mean of the field after the last update */
  int* mean_from_patch;         /* host, out: 1 if that update was a predictor step      */
} mgk_kmc_patch_cfg;

typedef struct {
  long long v;
  int m;
  int is_instab;
  double time;
  int ls_iters, ls_iters32;
  double eps_macro[9];  /* macro strain after the event (d*d row-major)          */
  double sig_mean[9];   /* <sig> after the event's re-solve (d*d row-major)      */
  int substeps;         /* finite strain: the N of the event's sub-stepped flip  */
} mgk_kmc_event;

#define MGK_KMC_RUN_TIME 0       /* no event within *remaining: time advanced by it, *remaining = 0 */
#define MGK_KMC_RUN_COUNT 1      /* max_events events done                                         */
#define MGK_KMC_RUN_SEQ 2        /* *seq (consecutive is_instab events) exceeded max_seq           */
#define MGK_KMC_RUN_UNSTABLE -6  /* unstable stress after an event's re-solve (2D ValueError)      */

/* Runs events until one of the conditions above; *time, eps_macro (d*d), *remaining,
 * *seq and both streams are updated in place; events[0..*n_events) describe the
 * events; info / iters32 hold the last solve's result. */
int mgk_kmc_run(const mgk_kmc_loop* loop, mgk_mt_state* rng, mgk_mt_state* selection, double* time,
                double* eps_macro, double* remaining, int max_events, int max_seq, int* seq,
                mgk_kmc_event* events, int* n_events, mgk_ls_info* info, int* iters32, void* stream);

/* Native cascade loop: cascade batches until no mode is unstable (mgkmc/aqs.py:480-680,
 * mgkmc/kmc_simulator.py:577-794; small strain, linear elasticity, planar catalogs, at
 * least one of redraw_directions / redraw_barriers).  Per batch: the barriers of the
 * current state into Q (M*NV device doubles), the modes with Q < stability_threshold in
 * voxel-major order, sorted by Q (ties by that order) keeping the first mode per voxel,
 * their flips at `time` in that order with the redraws of their voxels from the host
 * stream `rng`, one LS re-solve at eps_macro (the loop's LS form).  Stops: no unstable
 * mode (MGK_CASCADE_DONE); after the batch that leaves eps_macro[eps_comp] >= eps_target
 * (eps_comp < 0: never; MGK_CASCADE_EPS_TARGET); once the batch count exceeds max_batches
 * (MGK_CASCADE_TRUNCATED if truncate, else MGK_CASCADE_DONE); the unstable-stress check of
 * 2D loops (MGK_CASCADE_UNSTABLE); before a batch whose flips would exceed cap records or
 * batch_cap batches (MGK_CASCADE_CAPACITY: call again to continue, the state is
 * consistent).  flips[0..*n_flips) hold (batch, v, m), batch_sizes[0..*n_batches) the
 * flips per batch, batch_sig (optional, 9 doubles per batch) each batch's <sig> after
 * its re-solve (the state a per-batch log reads); info / iters32 the last re-solve. */
typedef struct {
  int batch;
  long long v;
  int m;
} mgk_cascade_flip;

#define MGK_CASCADE_DONE 0
#define MGK_CASCADE_TRUNCATED 1
#define MGK_CASCADE_EPS_TARGET 2
#define MGK_CASCADE_UNSTABLE 3
#define MGK_CASCADE_CAPACITY 4

int mgk_kmc_cascade(const mgk_kmc_loop* loop, mgk_mt_state* rng, double* Q, double time, const double* eps_macro,
                    int max_batches, int truncate, int eps_comp, double eps_target, mgk_cascade_flip* flips, int cap,
                    int* n_flips, int* batch_sizes, double* batch_sig, int batch_cap, int* n_batches, int* reason,
                    mgk_ls_info* info, int* iters32, void* stream);

/* Macro relaxation after an elastic increment with stress targets, linear law, in the
 * loop's LS form (eps_macro d*d in/out; info holds the current field's mean on entry and
 * the last solve's on return).  3D (mgkmc/aqs.py:1105-1142): per iteration the error of
 * the n_targets components target_idx[2t], target_idx[2t+1] against the current mean;
 * below tol: *converged = 1; else eps_macro += (err - nu_avg tr(err) I) / E_avg with the
 * driving component (drv_i, drv_j) zeroed, and a re-solve.  2D (mgkmc/kmc_simulator.py:
 * 1106-1143): a re-solve first (unstable-stress check: MGK_KMC_RUN_UNSTABLE), the error
 * test, then per in-plane target in the given order the compliance step, replaced by
 * err / slope once two iterates exist and the secant slope is finite and above
 * 0.01 E_avg.  *iters: solves done; *max_err: the last error. */
int mgk_kmc_mixed_bc(const mgk_kmc_loop* loop, int n_targets, const int* target_idx, const double* target_val,
                     int drv_i, int drv_j, double E_avg, double nu_avg, double tol, int max_iter, double* eps_macro,
                     int* converged, double* max_err, int* iters, mgk_ls_info* info, int* iters32, void* stream);

/* Batched form of mgk_kmc_run (no counterpart in the reference: the serving form's
 * continuous batching).  nb (1..MGK_LS_MAX_BATCH) simulations of one grid shape and
 * one LS configuration (ls_precision, warm, max_iter, tol, tol32) step in lockstep on
 * `stream`: one synchronisation reads every member's total rate, one every member's
 * selection, and one batched LS solve (mgk_ls_solve_batch / _refined_batch)
 * re-equilibrates every member that had an event.  Each member runs mgk_kmc_run's
 * kernels and draws on its own buffers and streams, so its events, fields and stream
 * states are bit-identical to mgk_kmc_run's on the same inputs.  Per-member arrays:
 * loop, rng, selection, time, remaining, max_events (each 1..cap), seq, n_events,
 * status (the member's MGK_KMC_RUN_* stop reason or a negative error), info, iters32
 * (nb each), eps_macro (9 per member, d*d used), events (cap per member: member b's
 * at events + b*cap).  The call returns once every member has stopped: MGK_OK, or a
 * negative error if the batch itself failed (then every member still running
 * carries that error as its status). */
int mgk_kmc_run_batch(int nb, const mgk_kmc_loop* loop, mgk_mt_state* rng, mgk_mt_state* selection, double* time,
                      double* eps_macro, double* remaining, const int* max_events, int max_seq, int* seq,
                      mgk_kmc_event* events, int cap, int* n_events, int* status, mgk_ls_info* info, int* iters32,
                      void* stream);

/* One finite-strain load step as data: which of mgk_fs_solver_step (precision 0/1),
 * mgk_fs_solver_step_refined (mixed, solver 0), mgk_fs_solver_step_al and
 * mgk_fs_solver_step_dbfft runs, with their arguments (grid, fields, material,
 * tolerances, workspaces).  Plane problems use their (nx, 1, ny) grid. */
typedef struct mgk_fs_stepper {
  int solver;     /* 0 newton_cg, 1 al, 2 dbfft */
  int precision;  /* 0 FP64, 1 FP32 */
  int mixed;      /* newton_cg: FP64 Newton loop with FP32 Krylov solves */
  int64_t shape[3];
  void* F;
  void* Fp;
  const void* lam;
  const void* mu;
  const void* lam32;
  const void* mu32;
  const mgk_fs_material* mat;
  void* P;
  void* sig;
  void* ws;
  size_t ws_bytes;
  void* ws32;
  size_t ws32_bytes;
  const double* L0; /* al: the reference medium (3 doubles) */
  double E_avg, nu_avg, tol_NW, tol_CG, tol_macro, pixel;
  int max_NW, max_iter_macro;
} mgk_fs_stepper;

/* What mgk_kmc_fs_flip needs besides the flips.  kmc: the catalog, softening, event
 * times, redraw settings and redraw buffer (its LS fields are not read; planar
 * catalogs).  Macro BC of every step (build_finite_strain_bc_3d, mgkmc/
 * finite_strain_simulator.py:2852-2899): F_bar = F_macro with the driving component
 * (drv_i, drv_j, d-dimensional) set from eps_macro, the other off-diagonals 0, and
 * each diagonal either a stress target (target_k / target_val, P_kk) or 1.  live /
 * live_bytes: the device state restored before every sub-stepping attempt (F, P, sig,
 * F_p, softening, event times, prev, catalog, barriers), saved into `backup` (at least
 * their total size, each at a 256-byte aligned offset).  sig2: plane problems' in-plane
 * Cauchy block (4 x NV: xx, xy, yx, yy) refreshed after every step (NULL in 3D).  flag:
 * one device int (the NaN check). */
typedef struct mgk_kmc_fs_substep {
  const mgk_kmc_loop* kmc;
  const mgk_fs_stepper* step;
  int drv_i, drv_j;
  int n_targets;
  int target_k[3];
  double target_val[3];
  int nlive;
  void* live[12];
  size_t live_bytes[12];
  void* backup;
  size_t backup_bytes;
  double* sig2;
  int* flag;
  /* native event loops (mgk_kmc_loop.fs): F_macro (d*d, in/out) and the last step's record */
  double* F_macro;
  mgk_fs_info* info;
} mgk_kmc_fs_substep;

/* The finite-strain flips of one KMC event or one cascade batch with the reference's
 * sub-stepping (mgkmc/aqs.py:512-615,887-996; 2D mgkmc/kmc_simulator.py:605-700,
 * 895-1000): F_p <- expm(C/N) F_p at every flipped voxel (the in-plane block in 2D) in N
 * sub-steps, softening, prev and the redraw of each voxel at the last sub-step, one
 * finite-strain step after every sub-step; an attempt whose step fails to converge or
 * leaves a NaN in sig or F or a mean stress above 20 GPa is undone (live state, F_macro,
 * eps_macro) and retried with the next N: 3D N = n0 .. n0 + 19, n0 = min(max over the
 * flips of max(2, ceil(sqrt(C:C/2) / 0.005)), 40); 2D N = 1, then max(n0, 2) .. + 19
 * with the step 0.02.  Redraws are drawn from `rng` anew at every attempt.  F_macro and
 * eps_macro (d*d) follow the steps; *substeps = the N that succeeded; info = the last
 * step's.  Returns MGK_OK, MGK_ERR_SOLVER when every N failed ("Mechanical solver failed
 * to converge under sub-stepping."), MGK_ERR_NONFINITE at once when a plane-stress E33
 * iteration overflows (as the reference's FloatingPointError, not retried). */
int mgk_kmc_fs_flip(const mgk_kmc_fs_substep* cfg, int nflips, const long long* voxels, const int* modes,
                    double time, mgk_mt_state* rng, double* F_macro, double* eps_macro, int* substeps,
                    mgk_fs_info* info, void* stream);

/* One nonlinear small-strain solve as data (FP64): kind 0 mgk_ls_solve_landau2d (warm,
 * the E33 state persisting: *e33_have in/out, set after a plane-stress solve), kind 1
 * mgk_ls_solve_secant (2D secant or Rose law, cold), kind 2 mgk_ls_solve_landau (3D, cold;
 * the loop's own grid; its events are not sub-stepped: flip, redraw, solve, and a
 * non-finite update stops the loop with MGK_ERR_NONFINITE, mgkmc/aqs.py:330-351), with the
 * solver's fields and buffers and the reference's max_iter / tol. */
typedef struct mgk_law_stepper {
  int kind;
  int64_t shape[2];
  const double* lam;
  const double* mu;
  double lam0, mu0;
  const mgk_landau_params* landau;
  const mgk_secant_params* secant;
  int plane_stress;
  double* eps;
  double* sig;
  double* e33;
  int* e33_have;
  double* tau;
  int anderson_m;
  double* aa;
  void* ws;
  size_t ws_bytes;
  int max_iter;
  double tol;
} mgk_law_stepper;

/* The law's sub-stepped flip (mgk_kmc_law_flip): live / live_bytes / backup as in
 * mgk_kmc_fs_substep (eigenstrain, softening, event times, prev, catalog, barriers, eps,
 * sig). */
typedef struct mgk_kmc_law_substep {
  const mgk_law_stepper* step;
  int nlive;
  void* live[12];
  size_t live_bytes[12];
  void* backup;
  size_t backup_bytes;
} mgk_kmc_law_substep;

/* The flips of one KMC event or cascade batch of a 2D nonlinear-law simulation with the
 * reference's small-strain sub-stepping (mgkmc/kmc_simulator.py:263-385): one full flip
 * and solve first; if the solve fails (a non-finite Landau update, a NaN mean stress or
 * one above 20 GPa) the live state is restored and the flips retried as N sub-steps of
 * mode / N each followed by a solve, softening, prev and redraws at the last, N =
 * max(N_h, 2) .. + 19, N_h = min(max over the flips of max(2, ceil(sqrt(C:C/2) /
 * 0.02)), 40).  Redraws come from `rng`, anew at every attempt.  *substeps = the N that
 * succeeded; info = the last solve's.  MGK_ERR_SOLVER when every N failed ("Mechanical
 * solver failed to converge under small-strain sub-stepping."). */
int mgk_kmc_law_flip(const mgk_kmc_loop* loop, const mgk_kmc_law_substep* cfg, int nflips, const long long* voxels,
                     const int* modes, double time, mgk_mt_state* rng, const double* eps_macro, int* substeps,
                     mgk_ls_info* info, void* stream);

#ifdef __cplusplus
}
#endif
#endif
