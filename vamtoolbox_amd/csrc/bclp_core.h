// BCLP per-voxel epilogue shared by the parallel-beam (bclp.hip) and the
// ray-tracer (raytrace.hip) fused iterations: the material response and its
// derivative (vamtoolbox/response.py:191-291), the band error and the Lp
// weighting of one voxel (vamtoolbox/optimizer/BCLP.py:277-367), in float64
// as the reference computes them in NumPy.
#pragma once

#include "vamx_internal.h"

namespace vamx {

struct RespDev {
  int kind;
  double A, K, B, M, nu, lin_M, lin_C;
  int n_pts;
  const double* tf;
  const double* tm;
  const double* td;
};

struct BclpDev {
  double p, q, eps, weight, dvol;
  const float* eps_arr;
  const float* w_arr;
};

// np.interp(x, xp, fp, left, right) for increasing xp (numpy's branch order)
__device__ inline double np_interp(double x, const double* __restrict__ xp, const double* __restrict__ fp, int n,
                            double left, double right) {
  if (x > xp[n - 1]) return right;
  if (x < xp[0]) return left;
  if (x == xp[n - 1]) return fp[n - 1];
  int lo = 0, hi = n - 1;  // xp[lo] <= x < xp[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (xp[mid] <= x) lo = mid; else hi = mid;
  }
  if (xp[lo] == x) return fp[lo];
  const double slope = (fp[lo + 1] - fp[lo]) / (xp[lo + 1] - xp[lo]);
  return slope * (x - xp[lo]) + fp[lo];
}

// response map and derivative (response.py:191-218, 233-253, 256-279)
__device__ inline void response_eval(const RespDev& r, double f, double& m, double& dm) {
  if (r.kind == VAMX_RESP_GLF) {
    const double e = exp(-r.B * (f - r.M));
    if (r.nu == 1.0) {
      m = r.A + (r.K - r.A) / (1.0 + e);
      const double c1 = 1.0 / (r.K - r.A);
      const double d = m - r.A;
      dm = c1 * r.B * (d * d) * e;
    } else {
      m = r.A + (r.K - r.A) / pow(1.0 + e, 1.0 / r.nu);
      const double c1 = pow(1.0 / (r.K - r.A), r.nu);
      const double c2 = r.B / r.nu;
      const double c3 = pow(m - r.A, r.nu + 1.0);
      dm = c1 * c2 * c3 * e;
    }
  } else if (r.kind == VAMX_RESP_LINEAR) {
    m = r.lin_M * f + r.lin_C;
    dm = r.lin_M;
  } else if (r.kind == VAMX_RESP_IDENTITY) {
    m = f;
    dm = 1.0;
  } else {
    const int n = r.n_pts;
    m = np_interp(f, r.tf, r.tm, n, r.tm[0], r.tm[n - 1]);
    dm = np_interp(f, r.tf, r.td, n, r.td[0], r.td[n - 1]);
  }
}

// BCLP epilogue for one voxel: returns the loss integrand, writes the operand
__device__ __forceinline__ double bclp_voxel(const RespDev& r, const BclpDev& b, float dose, float target,
                                             size_t idx, float& operand) {
  double m, dm;
  response_eval(r, (double)dose, m, dm);
  const double eT = m - (double)target;                       // BCLP.py:298-300
  const double eps = b.eps_arr ? (double)b.eps_arr[idx] : b.eps;
  const double eb = fabs(eT) - eps;                            // 305-307
  const bool v = eb > 0.0;                                     // 311
  if (!v) {
    operand = 0.0f;
    return 0.0;
  }
  const double w = b.w_arr ? (double)b.w_arr[idx] : b.weight;
  const double sgn = eT > 0.0 ? 1.0 : (eT < 0.0 ? -1.0 : 0.0);
  const double ab = fabs(eb);
  if (b.p == 2.0) {
    operand = (float)(w * ab * sgn * dm);
    return w * (ab * ab);
  }
  operand = (float)(w * pow(ab, b.p - 1.0) * sgn * dm);  // 361-367
  return w * pow(ab, b.p);                               // 340-342
}

__device__ __forceinline__ double block_sum_256(double v, double* sh) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) sh[t] += sh[t + s];
    __syncthreads();
  }
  return sh[0];
}


// loss partials -> loss log / scalar and the gradient scale q L^((q-p)/q)
// (one block, fixed order; defined in bclp.hip)
int bclp_reduce_launch(const double* partials, int nparts, const BclpDev& b, double* loss_dst, int32_t* counter,
                       double* loss_log, double* scale_out, hipStream_t s);
// the response model of a C-ABI call in device form
int resp_fill_dev(const vamx_response* resp, RespDev& r);
// the response and loss parameters of a C-ABI call in device form
int bclp_fill_devs(const vamx_response* resp, const vamx_bclp_params* prm, const float* eps_arr, const float* w_arr,
                   RespDev& r, BclpDev& b);

}  // namespace vamx
