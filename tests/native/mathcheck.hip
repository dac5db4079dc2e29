// mathcheck.hip -- evaluates the chain's own log / exp / pow (chain.h) on the
// host and on the device for accuracy tests.  Not part of the model library:
// built as libmhmcheck.so next to this file (mhm_amd/build.py).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "chain.h"

namespace {

// which: 0 = log_pos(x), 1 = exp_fin(x), 2 = exp_ylogx(y, x),
// 3 = exp_ylogx_v<4> over the four pairs i & ~3 .. (i & ~3) + 3 (element i;
// pairs past n read as (1, 1)), 4 = div_rcp(x, y, 1 / y), 5 = x / y,
// 6 = pow23(x)
__host__ __device__ inline double eval(int which, const double *x, const double *y, int64_t i,
                                       int64_t n) {
  if (which == 0) return mhmhip::log_pos(x[i]);
  if (which == 1) return mhmhip::exp_fin(x[i]);
  if (which == 2) return mhmhip::exp_ylogx(y[i], x[i]);
  if (which == 4) return mhmhip::div_rcp(x[i], y[i], 1.0 / y[i]);
  if (which == 5) return x[i] / y[i];
  if (which == 6) return mhmhip::pow23(x[i]);
  mhmhip::DV<4> yy, xx;
  const int64_t b = i & ~(int64_t)3;
  for (int k = 0; k < 4; ++k) {
    yy.v[k] = b + k < n ? y[b + k] : 1.0;
    xx.v[k] = b + k < n ? x[b + k] : 1.0;
  }
  return mhmhip::exp_ylogx_v<4>(yy, xx).v[i - b];
}

__global__ void eval_kernel(int which, const double *x, const double *y, double *out, int64_t n) {
  mhmhip::mathtab_fill();
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = eval(which, x, y, i, n);
}

}  // namespace

extern "C" int mhmhip_check_eval_host(int which, const double *x, const double *y, double *out,
                                      int64_t n) {
  if (which < 0 || which > 6 || n < 0 || !x || !y || !out) return -1;
  for (int64_t i = 0; i < n; ++i) out[i] = eval(which, x, y, i, n);
  return 0;
}

// Same evaluation on device 0: copies in, one launch over n lanes, copies out.
extern "C" int mhmhip_check_eval_device(int which, const double *x, const double *y, double *out,
                                        int64_t n) {
  if (which < 0 || which > 6 || n <= 0 || !x || !y || !out) return -1;
  double *d = nullptr;
  const size_t bytes = sizeof(double) * (size_t)n;
  if (hipMalloc(&d, 3 * bytes) != hipSuccess) return -2;
  int rc = 0;
  if (hipMemcpy(d, x, bytes, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(d + n, y, bytes, hipMemcpyHostToDevice) != hipSuccess) {
    rc = -3;
  } else {
    const int block = 256;
    const int64_t grid = (n + block - 1) / block;
    hipLaunchKernelGGL(eval_kernel, dim3((unsigned)grid), dim3(block), 0, 0, which, d, d + n,
                       d + 2 * n, n);
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess ||
        hipMemcpy(out, d + 2 * n, bytes, hipMemcpyDeviceToHost) != hipSuccess)
      rc = -4;
  }
  (void)hipFree(d);
  return rc;
}
