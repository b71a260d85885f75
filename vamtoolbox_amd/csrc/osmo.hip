// OSMO's element-wise steps for gfx950 (vamtoolbox/optimizer/OSMO.py:36-110):
// the model update of the void and gel steps and the max normalisation with
// its lower bound, each one pass over the array (the projections are the
// parallel-beam kernels).  Float32 with the reference's operations in its
// order; NaN passes the bounds as the reference's masked assignments let it.
#include <algorithm>

#include "vamx_internal.h"

namespace vamx {

namespace {

// void (gel = 0): x_model -= max(x - d, 0) on the masked voxels (OSMO.py:54-56)
// gel  (gel = 1): x_model += max(d - x, 0) on the masked voxels (OSMO.py:92-94)
__global__ __launch_bounds__(256) void osmo_model_update(const float* __restrict__ x, float* __restrict__ xm,
                                                         const unsigned char* __restrict__ mask, long long n,
                                                         float d, int gel) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    if (!mask[i]) continue;
    float diff = gel ? d - x[i] : x[i] - d;
    diff = diff < 0.0f ? 0.0f : diff;   // diff[diff < 0] = 0 (NaN stays)
    xm[i] = gel ? xm[i] + diff : xm[i] - diff;
  }
}

// v = v / m, then v[v < lo] = lo (OSMO.py:60-65, 67-69)
__global__ __launch_bounds__(256) void scale_clamp(float* __restrict__ v, long long n, const float* __restrict__ m,
                                                   float lo) {
  const float mv = *m;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const float s = v[i] / mv;
    v[i] = s < lo ? lo : s;
  }
}

long long ew_blocks(long long n) { return std::min<long long>(4096, std::max<long long>(1, (n + 255) / 256)); }

}  // namespace

}  // namespace vamx

using namespace vamx;

extern "C" int vamx_osmo_model_update(const float* x, float* x_model, const uint8_t* mask, int64_t n, float d,
                                      int32_t gel, void* stream) {
  if (!x || !x_model || !mask) return set_error(VAMX_ERR_ARG, "vamx_osmo_model_update: null argument");
  if (n < 0) return set_error(VAMX_ERR_ARG, "vamx_osmo_model_update: negative size");
  if (n == 0) return VAMX_OK;
  hipLaunchKernelGGL(osmo_model_update, dim3((unsigned)ew_blocks(n)), dim3(256), 0, (hipStream_t)stream, x, x_model,
                     (const unsigned char*)mask, (long long)n, d, gel ? 1 : 0);
  VAMX_HIP(hipGetLastError());
  return VAMX_OK;
}

extern "C" int vamx_scale_clamp(float* v, int64_t n, const float* m, float lo, void* stream) {
  if (!v || !m) return set_error(VAMX_ERR_ARG, "vamx_scale_clamp: null argument");
  if (n < 0) return set_error(VAMX_ERR_ARG, "vamx_scale_clamp: negative size");
  if (n == 0) return VAMX_OK;
  hipLaunchKernelGGL(scale_clamp, dim3((unsigned)ew_blocks(n)), dim3(256), 0, (hipStream_t)stream, v, (long long)n, m,
                     lo);
  VAMX_HIP(hipGetLastError());
  return VAMX_OK;
}
