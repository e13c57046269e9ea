// Device pieces of the native finite-strain flip (mgk_kmc_fs_flip in
// mgk_loop.cpp): the plastic deformation update of one voxel and the NaN test
// of the step's fields (mgkmc/aqs.py:931-935 and :321; 2D
// mgkmc/kmc_simulator.py:950-954 and :436).
#include "mgk_common.h"
#include <algorithm>
#include "mgk_host.h"

namespace {

struct Mat3 {
  double a[9];
};

// F_p <- dF F_p at voxel v.  d = 3: the nine components; d = 2: the in-plane
// block (axes 0 and 2 of the (nx, 1, ny) grid), dF 2 x 2 in dF.a[0, 1, 3, 4].
// One thread; products and sums rounded separately, summed in k order.
__global__ void k_fp_left_mul(double* __restrict__ Fp, long long NV, long long v, int d, Mat3 dF) {
#pragma clang fp contract(off)
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const int ax3[3] = {0, 1, 2}, ax2[2] = {0, 2};
  const int* ax = d == 3 ? ax3 : ax2;
  double f[3][3], o[3][3];
  for (int a = 0; a < d; ++a)
    for (int b = 0; b < d; ++b) f[a][b] = Fp[(size_t)(3 * ax[a] + ax[b]) * NV + v];
  for (int i = 0; i < d; ++i)
    for (int j = 0; j < d; ++j) {
      double s = dF.a[3 * i] * f[0][j];
      for (int k = 1; k < d; ++k) s = s + dF.a[3 * i + k] * f[k][j];
      o[i][j] = s;
    }
  for (int a = 0; a < d; ++a)
    for (int b = 0; b < d; ++b) Fp[(size_t)(3 * ax[a] + ax[b]) * NV + v] = o[a][b];
}

__global__ void k_flag_zero(int* flag) {
  if (threadIdx.x == 0 && blockIdx.x == 0) flag[0] = 0;
}

// flag = 1 when a or b holds a NaN (plain vector stores of the same value)
__global__ __launch_bounds__(MGK_NT) void k_any_nan2(const double* __restrict__ a, const double* __restrict__ b,
                                                     size_t n, int* __restrict__ flag) {
  bool hit = false;
  for (size_t i = blockIdx.x * (size_t)MGK_NT + threadIdx.x; i < n; i += (size_t)gridDim.x * MGK_NT)
    hit = hit || isnan(a[i]) || isnan(b[i]);
  if (hit) flag[0] = 1;
}

}  // namespace

int mgk_fp_left_mul(double* Fp, long long NV, long long v, int d, const double* dF, hipStream_t s) {
  Mat3 m{};
  for (int i = 0; i < d; ++i)
    for (int j = 0; j < d; ++j) m.a[3 * i + j] = dF[d * i + j];
  k_fp_left_mul<<<1, 64, 0, s>>>(Fp, NV, v, d, m);
  MGK_HIP(hipGetLastError());
  return MGK_OK;
}

int mgk_any_nan2(const double* a, const double* b, size_t n, int* flag_dev, int* host, hipStream_t s) {
  k_flag_zero<<<1, 64, 0, s>>>(flag_dev);
  const int nb = (int)std::min<size_t>((n + MGK_NT - 1) / MGK_NT, 4096);
  k_any_nan2<<<nb, MGK_NT, 0, s>>>(a, b, n, flag_dev);
  MGK_HIP(hipGetLastError());
  MGK_HIP(hipMemcpyAsync(host, flag_dev, sizeof(int), hipMemcpyDeviceToHost, s));
  MGK_HIP(hipStreamSynchronize(s));
  return MGK_OK;
}
