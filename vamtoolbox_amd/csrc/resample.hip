// Bilinear resampling of a sinogram at arbitrary (detector, angle) points,
// every z slice at once (gfx950): the resampling step of the fan-beam rebinning
// (vamtoolbox/geometry.py:736-944: scipy.interpolate.interpn(method="linear",
// bounds_error=False, fill_value=0) at line 805, then the differential-area
// factor and the inverse Fresnel transmission at 814-821, in that order).
// The host resolves each point to its grid cell and weights; one thread per
// (point, z), z fastest, so the four taps and the store are coalesced along z.
#include "vamx_internal.h"

namespace vamx {

__global__ __launch_bounds__(256) void sino_resample_kernel(const float* __restrict__ sino, int n_r, int n_a, int nz,
                                                            const int32_t* __restrict__ cell,
                                                            const float* __restrict__ wts,
                                                            const float* __restrict__ fac, long long m,
                                                            float* __restrict__ out) {
  const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= m * nz) return;
  const int z = (int)(tid % nz);
  const long long p = tid / nz;
  const int i0 = cell[2 * p], j0 = cell[2 * p + 1];
  float v;
  if (i0 < 0 || i0 > n_r - 2 || j0 < 0 || j0 > n_a - 2) {
    v = i0 == -1 ? 0.0f : __builtin_nanf("");   // -1: outside the grid (fill value); else an undefined sample
  } else {
    const float wx = wts[2 * p], wa = wts[2 * p + 1];
    const size_t zs = (size_t)nz;
    const float* b = sino + ((size_t)i0 * n_a + j0) * zs + z;
    const size_t di = (size_t)n_a * zs, dj = zs;
    const float v00 = b[0], v01 = b[dj], v10 = b[di], v11 = b[di + dj];
    const float top = fmaf(wa, v01 - v00, v00), bot = fmaf(wa, v11 - v10, v10);
    v = fmaf(wx, bot - top, top);
  }
  out[tid] = fac[2 * p + 1] * (v * fac[2 * p]);
}

}  // namespace vamx

using namespace vamx;

extern "C" int vamx_sino_resample(const float* sino, int32_t n_r, int32_t n_a, int32_t nz, const int32_t* cell,
                                  const float* weights, const float* factors, int64_t m, float* out, void* stream) {
  if (!sino || !cell || !weights || !factors || !out)
    return set_error(VAMX_ERR_ARG, "vamx_sino_resample: null buffer");
  if (n_r < 2 || n_a < 2 || nz < 1 || m < 0) return set_error(VAMX_ERR_ARG, "vamx_sino_resample: invalid shape");
  if (m == 0) return VAMX_OK;
  hipLaunchKernelGGL(sino_resample_kernel, dim3(blocks_for(m * nz, 256)), dim3(256), 0, (hipStream_t)stream, sino,
                     (int)n_r, (int)n_a, (int)nz, cell, weights, factors, (long long)m, out);
  VAMX_HIP(hipGetLastError());
  return VAMX_OK;
}
