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
 * 10 fs_xinv, 11 fs_vector.  total_bytes: algorithmic HBM bytes of the timed
 * launches. */
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

#ifdef __cplusplus
}
#endif
#endif
