// Curved-ray tracing propagator for graded-index media (gfx950).
//
// Reference: vamtoolbox/projector/pyTorchRayTrace.py
//   ray setup           RayState.setupRaysParallel / resetRaysIterateToInitial (1515-1700)
//   stepping            RayTraceSolver._forwardSymplecticEuler / _forwardEuler /
//                       _leapfrog_init (738-832), parametrisations (692-735)
//   interaction         deposit / integrate (835-898), adjacent-voxel weights (939-1143),
//                       expressPositionInArrayIndices (1253-1260)
//   exit test           exitCheck (1343-1424), every num_step_per_exit_check steps (348-363)
//   loops               depositEnergyUntilExit / integrateEnergyUntilExit (378-571)
// Media: vamtoolbox/medium.py -- grid_sample trilinear, align_corners=True,
// border padding for n / grad n, zeros padding for an interpolated
// attenuation grid (161-211); analytic cylinder / ball attenuation (895-915).
//
// One lane traces one ray for its whole life (no per-step launches, no global
// active-set reductions); adjacent lanes trace adjacent rays (same azimuth and
// lateral position, neighbouring z) so their step counts and memory accesses
// coincide.  The ray state and all geometry are float64 (the field is stored
// as float32 (n, dn/dx, dn/dy, dn/dz) per voxel and fetched as one 16-byte
// load per tap).  RK4 (classical, on the (x, v) system of the chosen
// parametrisation) is provided in addition to the reference's steppers.
//
// forward  g = W (alpha_abs f): gather along the ray, one output per lane.
// backward dose = alpha_abs (W^T g): scatter along the ray into a 64-bit
//          fixed-point accumulator with integer atomics; integer addition is
//          associative, so the result is bitwise deterministic.  The fixed-point
//          scale is derived on the device from max|g|.
// Interaction taps use floor and floor+1 with weights 1-d and d (standard
// trilinear); the reference selects floor and ceil, which counts a tap twice
// when a coordinate is exactly integral.
#include "vamx_internal.h"

namespace vamx {

struct d3 {
  double x, y, z;
};
__device__ __forceinline__ d3 mk(double x, double y, double z) { return {x, y, z}; }
__device__ __forceinline__ d3 add(d3 a, d3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ d3 sub(d3 a, d3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ d3 scl(double s, d3 a) { return {s * a.x, s * a.y, s * a.z}; }

// trilinear sample of the float4 field, grid_sample(align_corners=True,
// padding_mode="border"): index coordinates clamped to [0, size-1]
__device__ __forceinline__ void field_sample(const RtDev& g, d3 x, double& n, d3& gn) {
  double ix = (x.x - g.xmin) * g.inv_dx, iy = (x.y - g.ymin) * g.inv_dy;
  double iz = g.nz > 1 ? (x.z - g.zmin) * g.inv_dz : 0.0;
  ix = fmin(fmax(ix, 0.0), (double)(g.nx - 1));
  iy = fmin(fmax(iy, 0.0), (double)(g.ny - 1));
  iz = fmin(fmax(iz, 0.0), (double)(g.nz - 1));
  const double fx = floor(ix), fy = floor(iy), fz = floor(iz);
  const double wx = ix - fx, wy = iy - fy, wz = iz - fz;
  const int x0 = (int)fx, y0 = (int)fy, z0 = (int)fz;
  double an = 0.0, ax = 0.0, ay = 0.0, az = 0.0;
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const int xi = x0 + (c & 1), yi = y0 + ((c >> 1) & 1), zi = z0 + ((c >> 2) & 1);
    const double w = ((c & 1) ? wx : 1.0 - wx) * (((c >> 1) & 1) ? wy : 1.0 - wy) * (((c >> 2) & 1) ? wz : 1.0 - wz);
    if (w == 0.0 || xi >= g.nx || yi >= g.ny || zi >= g.nz) continue;
    const float4 v = g.field[((size_t)xi * g.ny + yi) * g.nz + zi];
    an += w * (double)v.x;
    ax += w * (double)v.y;
    ay += w * (double)v.z;
    az += w * (double)v.w;
  }
  n = an;
  gn = mk(ax, ay, az);
}

// attenuation coefficient at x (medium.py AttenuationModel.alpha)
__device__ __forceinline__ double att_alpha(const RtDev& g, d3 x) {
  if (g.att_kind == VAMX_RT_ATT_CYLINDER) return sqrt(x.x * x.x + x.y * x.y) < g.att_radius ? g.att_alpha : 0.0;
  if (g.att_kind == VAMX_RT_ATT_BALL) return sqrt(x.x * x.x + x.y * x.y + x.z * x.z) < g.att_radius ? g.att_alpha : 0.0;
  if (g.att_kind == VAMX_RT_ATT_GRID) {
    // grid_sample(align_corners=True, padding_mode="zeros")
    const double ix = (x.x - g.xmin) * g.inv_dx, iy = (x.y - g.ymin) * g.inv_dy;
    const double iz = g.nz > 1 ? (x.z - g.zmin) * g.inv_dz : 0.0;
    const double fx = floor(ix), fy = floor(iy), fz = floor(iz);
    const double wx = ix - fx, wy = iy - fy, wz = iz - fz;
    const int x0 = (int)fx, y0 = (int)fy, z0 = (int)fz;
    double a = 0.0;
    for (int c = 0; c < 8; ++c) {
      const int xi = x0 + (c & 1), yi = y0 + ((c >> 1) & 1), zi = z0 + ((c >> 2) & 1);
      const double w = ((c & 1) ? wx : 1.0 - wx) * (((c >> 1) & 1) ? wy : 1.0 - wy) * (((c >> 2) & 1) ? wz : 1.0 - wz);
      if (w == 0.0 || xi < 0 || yi < 0 || zi < 0 || xi >= g.nx || yi >= g.ny || zi >= g.nz) continue;
      a += w * (double)g.att_grid[((size_t)xi * g.ny + yi) * g.nz + zi];
    }
    return a;
  }
  return 0.0;
}

// right-hand side of the ray equation in parametrisation PARAM
// (pyTorchRayTrace.py:695-735)
template <int PARAM>
__device__ __forceinline__ void rhs(const RtDev& g, d3 x, d3 v, d3& dx, d3& dv) {
  double n;
  d3 gn;
  field_sample(g, x, n, gn);
  if (PARAM == VAMX_RT_CANONICAL) {
    dx = v;
    dv = scl(n, gn);
  } else if (PARAM == VAMX_RT_PHYSICAL) {
    dx = scl(1.0 / n, v);
    dv = gn;
  } else {
    dx = scl(1.0 / (n * n), v);
    dv = scl(1.0 / n, gn);
  }
}

template <int PARAM>
__device__ __forceinline__ d3 accel(double n, d3 gn) {
  if (PARAM == VAMX_RT_CANONICAL) return scl(n, gn);
  if (PARAM == VAMX_RT_PHYSICAL) return gn;
  return scl(1.0 / n, gn);
}

// one step of solver SOLVER: (x, v) -> (x', v')
template <int SOLVER, int PARAM>
__device__ __forceinline__ void advance(const RtDev& g, d3& x, d3& v) {
  const double h = g.step;
  if (SOLVER == VAMX_RT_RK4) {
    // x' = x + h/6 ((k1 + 2 k2) + (2 k3 + k4)), likewise v; the two bracketed
    // sums are formed as soon as their terms exist, so only they stay live
    d3 kx, kv;
    rhs<PARAM>(g, x, v, kx, kv);
    d3 sx = kx, sv = kv;                                   // k1
    rhs<PARAM>(g, add(x, scl(0.5 * h, kx)), add(v, scl(0.5 * h, kv)), kx, kv);
    sx = add(sx, scl(2.0, kx));                            // k1 + 2 k2
    sv = add(sv, scl(2.0, kv));
    rhs<PARAM>(g, add(x, scl(0.5 * h, kx)), add(v, scl(0.5 * h, kv)), kx, kv);
    d3 tx = scl(2.0, kx), tv = scl(2.0, kv);               // 2 k3
    rhs<PARAM>(g, add(x, scl(h, kx)), add(v, scl(h, kv)), kx, kv);
    tx = add(tx, kx);                                      // 2 k3 + k4
    tv = add(tv, kv);
    x = add(x, scl(h / 6.0, add(sx, tx)));
    v = add(v, scl(h / 6.0, add(sv, tv)));
    return;
  }
  double n;
  d3 gn;
  field_sample(g, x, n, gn);
  const d3 vnew = add(v, scl(h, accel<PARAM>(n, gn)));
  const d3 vx = (SOLVER == VAMX_RT_FORWARD_EULER) ? v : vnew;  // symplectic uses v_{i+1}
  d3 dx;
  if (PARAM == VAMX_RT_CANONICAL) dx = vx;
  else if (PARAM == VAMX_RT_PHYSICAL) dx = scl(1.0 / n, vx);
  else dx = scl(1.0 / (n * n), vx);
  x = add(x, scl(h, dx));
  v = vnew;
}

// initial state of ray (t, a, zr) (setupRaysParallel, pyTorchRayTrace.py:1638-1666)
template <int SOLVER, int PARAM>
__device__ __forceinline__ void ray_init(const RtDev& g, int t, int a, int zr, d3& x, d3& v) {
  const double ct = g.ray_cs[2 * a], st = g.ray_cs[2 * a + 1];
  const double ci = cos(g.incl), si = sin(g.incl);
  const double G0 = g.ray_t[t], G2 = g.ray_z[zr], D = g.radial_offset;
  x.x = D * ct * ci + (-G0 * st - G2 * ct * si);
  x.y = D * st * ci + (G0 * ct - G2 * st * si);
  x.z = D * si + G2 * ci;
  v = mk(-ct * ci, -st * ci, -si);
  if (SOLVER == VAMX_RT_LEAPFROG) {
    // _leapfrog_init: v -= dv(x, v) h / 2 (pyTorchRayTrace.py:817-828)
    double n;
    d3 gn;
    field_sample(g, x, n, gn);
    v = sub(v, scl(g.step / 2.0, accel<PARAM>(n, gn)));
  }
}

// exit test (pyTorchRayTrace.py:1391-1420)
__device__ __forceinline__ bool exited(const RtDev& g, d3 x, d3 v) {
  const double vz = g.nz > 1 ? g.dz : INFINITY;
  const bool pos = (x.x > g.xmax + g.dx && v.x > 0.0) || (x.y > g.ymax + g.dy && v.y > 0.0) ||
                   (x.z > g.zmax + vz && v.z > 0.0);
  const bool neg = (x.x < g.xmin - g.dx && v.x < 0.0) || (x.y < g.ymin - g.dy && v.y < 0.0) ||
                   (x.z < g.zmin - vz && v.z < 0.0);
  return pos || neg;
}

// interaction taps at position x: index coordinates and base weight
struct Taps {
  int x0, y0, z0;
  double wx, wy, wz;
  int nzt;  // 2 in 3-D, 1 in 2-D
};
__device__ __forceinline__ Taps taps_at(const RtDev& g, d3 x) {
  Taps t;
  const double ix = (x.x - g.xmin) * g.inv_dx, iy = (x.y - g.ymin) * g.inv_dy;
  const double fx = floor(ix), fy = floor(iy);
  t.x0 = (int)fmax(fmin(fx, 2.0e9), -2.0e9);
  t.y0 = (int)fmax(fmin(fy, 2.0e9), -2.0e9);
  t.wx = ix - fx;
  t.wy = iy - fy;
  if (g.nz > 1) {
    const double iz = (x.z - g.zmin) * g.inv_dz;
    const double fz = floor(iz);
    t.z0 = (int)fmax(fmin(fz, 2.0e9), -2.0e9);
    t.wz = iz - fz;
    t.nzt = 2;
  } else {
    t.z0 = 0;
    t.wz = 0.0;
    t.nzt = 1;
  }
  return t;
}

// what a trace does at each interaction tap
enum RtMode { RT_INTEGRATE = 0, RT_DEPOSIT = 1, RT_COUNT = 2, RT_RECORD = 3 };

struct RtIo {
  const float* feff;              // integrate: alpha_abs * f
  float* gout;                    // integrate: one value per ray
  const float* gin;               // deposit: ray energies
  unsigned long long* acc;        // deposit: fixed-point accumulator
  const double* fx_scale;         // deposit: fixed-point scale
  unsigned long long* stats;      // [2] step statistics
  long long* counts;              // count: taps per ray (out); record: first slot per ray (in)
  int32_t* rec_vox;               // record: voxel of each tap
  float* rec_w;                   // record: weight of each tap
};

template <int MODE, int SOLVER, int PARAM>
__global__ __launch_bounds__(256) void rt_trace_kernel(RtDev g, RtIo io) {
  const long long nrays = (long long)g.nt * g.na * g.nzr;
  const long long rid = (long long)xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  // every lane runs to the end (no early return): the step statistics are
  // summed across the wavefront before one atomic per wavefront
  bool live = rid < nrays;
  const long long r = live ? rid : 0;
  const int zr = (int)(r % g.nzr);
  const long long ta = r / g.nzr;
  const int a = (int)(ta % g.na), t = (int)(ta / g.na);
  double E0 = (double)g.energy0;
  if (MODE == RT_DEPOSIT && live) {
    E0 = (double)io.gin[r];
    live = E0 != 0.0;  // a ray of zero energy contributes nothing and is not traced
  }
  const double scale = MODE == RT_DEPOSIT ? *io.fx_scale : 0.0;
  long long slot = (MODE == RT_RECORD && live) ? io.counts[r] : 0;   // record: this ray's first slot
  long long ntaps = 0;
  d3 x, v;
  ray_init<SOLVER, PARAM>(g, t, a, zr, x, v);
  const double hw = g.step / g.dx;  // step_size_idx_unit (pyTorchRayTrace.py:840-842)
  double A = 0.0, integral = 0.0;
  int step = 0;
  bool active = live;
  if (live) {
    for (; step < g.max_steps; ++step) {
      const d3 xo = x;
      advance<SOLVER, PARAM>(g, x, v);
      const d3 d = sub(x, xo);
      const double ds = sqrt(d.x * d.x + d.y * d.y + d.z * d.z);
      A += att_alpha(g, xo) * ds;                 // attenuance (767-769)
      const double E = E0 * exp(-A);              // ray energy (770-774)
      const Taps tp = taps_at(g, x);
      if (tp.x0 >= -1 && tp.x0 < g.nx && tp.y0 >= -1 && tp.y0 < g.ny && tp.z0 >= -1 && tp.z0 < g.nz) {
        const double base = hw * E;
#pragma unroll
        for (int c = 0; c < 8; ++c) {
          const int dzc = (c >> 2) & 1;
          if (dzc >= tp.nzt) continue;
          const int xi = tp.x0 + (c & 1), yi = tp.y0 + ((c >> 1) & 1), zi = tp.z0 + dzc;
          if (xi < 0 || yi < 0 || zi < 0 || xi >= g.nx || yi >= g.ny || zi >= g.nz) continue;
          const double w = ((c & 1) ? tp.wx : 1.0 - tp.wx) * (((c >> 1) & 1) ? tp.wy : 1.0 - tp.wy) *
                           (tp.nzt == 1 ? 1.0 : (dzc ? tp.wz : 1.0 - tp.wz));
          if (w == 0.0) continue;
          const size_t vi = ((size_t)xi * g.ny + yi) * g.nz + zi;
          if (MODE == RT_DEPOSIT) {
            const long long q = llrint(base * w * scale);
            atomicAdd(io.acc + vi, (unsigned long long)q);
          } else if (MODE == RT_INTEGRATE) {
            integral += (double)io.feff[vi] * (base * w);
          } else if (MODE == RT_COUNT) {
            ++ntaps;
          } else {
            io.rec_vox[slot] = (int32_t)vi;
            io.rec_w[slot] = (float)(base * w);
            ++slot;
          }
        }
      }
      if (step % g.exit_every == 0 && exited(g, x, v)) {
        active = false;
        ++step;
        break;
      }
    }
  }
  if (MODE == RT_INTEGRATE && rid < nrays) io.gout[rid] = (float)integral;
  if (MODE == RT_COUNT && rid < nrays) io.counts[rid] = ntaps;
  // wavefront sums of the step count and of the rays still active at max_steps
  unsigned long long st = live ? (unsigned long long)step : 0ull;
  unsigned long long un = active ? 1ull : 0ull;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    st += __shfl_xor(st, off);
    un += __shfl_xor(un, off);
  }
  if ((threadIdx.x & 63) == 0) {
    atomicAdd(io.stats, st);
    if (un) atomicAdd(io.stats + 1, un);
  }
}

// f_eff = alpha_abs * f  (pyTorchRayTrace.py:515-519)
__global__ __launch_bounds__(256) void rt_feff_kernel(const float* __restrict__ f, const float* __restrict__ ab,
                                                      float* __restrict__ out, long long nv) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nv) out[i] = ab ? f[i] * ab[i] : f[i];
}

// max |g| for the fixed-point scale, pass 1: block b reduces a contiguous
// range of g into part[b] (max is order-independent, so the result is exact)
constexpr int kGmaxParts = 512;

__global__ __launch_bounds__(256) void rt_gmax_part(const float* __restrict__ gin, long long n,
                                                    float* __restrict__ part) {
  __shared__ float sh[256];
  const long long chunk = (n + kGmaxParts - 1) / kGmaxParts;
  const long long i0 = (long long)blockIdx.x * chunk, i1 = i0 + chunk < n ? i0 + chunk : n;
  float m = 0.0f;
  for (long long i = i0 + threadIdx.x; i < i1; i += 256) m = fmaxf(m, fabsf(gin[i]));
  sh[threadIdx.x] = m;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] = fmaxf(sh[threadIdx.x], sh[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// pass 2 (one block): the scale 2^e and the cleared step statistics
__global__ __launch_bounds__(256) void rt_gmax_final(const float* __restrict__ part, double na,
                                                     double* __restrict__ scale_out,
                                                     unsigned long long* __restrict__ acc_zero_stats) {
  __shared__ float sh[256];
  float m = 0.0f;
  for (int i = threadIdx.x; i < kGmaxParts; i += 256) m = fmaxf(m, part[i]);
  sh[threadIdx.x] = m;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] = fmaxf(sh[threadIdx.x], sh[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    // per-voxel total weight is bounded by ~ (number of angles) * 4; keep the
    // accumulated magnitude below 2^60
    const double bound = (double)sh[0] * fmax(na, 1.0) * 16.0;
    int e = 40;
    if (bound > 0.0) e = 60 - (int)ceil(log2(bound));
    e = e < 0 ? 0 : (e > 62 ? 62 : e);
    *scale_out = ldexp(1.0, e);
    acc_zero_stats[0] = 0ull;
    acc_zero_stats[1] = 0ull;
  }
}

__global__ __launch_bounds__(256) void rt_zero_u64(unsigned long long* __restrict__ p, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

__global__ void rt_zero_stats(unsigned long long* __restrict__ s) {
  s[0] = 0ull;
  s[1] = 0ull;
}

// dose = alpha_abs * acc / scale  (pyTorchRayTrace.py:456-462)
__global__ __launch_bounds__(256) void rt_finalize_kernel(const unsigned long long* __restrict__ acc,
                                                          const float* __restrict__ ab,
                                                          const double* __restrict__ fx_scale,
                                                          float* __restrict__ dose, long long nv) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nv) return;
  const double v = (double)(long long)acc[i] / *fx_scale;
  dose[i] = (float)(ab ? v * (double)ab[i] : v);
}

// compile-time solver and parametrisation: each instantiation keeps only its
// own stepper's registers live
template <int MODE, int SOLVER>
int launch_trace_s(const RtDev& g, dim3 grid, hipStream_t s, const RtIo& io) {
  if (g.param == VAMX_RT_CANONICAL)
    hipLaunchKernelGGL((rt_trace_kernel<MODE, SOLVER, VAMX_RT_CANONICAL>), grid, dim3(256), 0, s, g, io);
  else if (g.param == VAMX_RT_PHYSICAL)
    hipLaunchKernelGGL((rt_trace_kernel<MODE, SOLVER, VAMX_RT_PHYSICAL>), grid, dim3(256), 0, s, g, io);
  else
    hipLaunchKernelGGL((rt_trace_kernel<MODE, SOLVER, VAMX_RT_OPTICAL>), grid, dim3(256), 0, s, g, io);
  return VAMX_OK;
}

template <int MODE>
int launch_trace(const RtDev& g, dim3 grid, hipStream_t s, const RtIo& io) {
  switch (g.solver) {
    case VAMX_RT_SYMPLECTIC_EULER: return launch_trace_s<MODE, VAMX_RT_SYMPLECTIC_EULER>(g, grid, s, io);
    case VAMX_RT_FORWARD_EULER: return launch_trace_s<MODE, VAMX_RT_FORWARD_EULER>(g, grid, s, io);
    case VAMX_RT_LEAPFROG: return launch_trace_s<MODE, VAMX_RT_LEAPFROG>(g, grid, s, io);
    case VAMX_RT_RK4: return launch_trace_s<MODE, VAMX_RT_RK4>(g, grid, s, io);
    default: return set_error(VAMX_ERR_ARG, "ray tracer: unknown solver");
  }
}

}  // namespace vamx

using namespace vamx;

extern "C" int vamx_rt_create(const vamx_rt_geom* geom, vamx_plan** out_plan) {
  if (!geom || !out_plan) return set_error(VAMX_ERR_ARG, "vamx_rt_create: null argument");
  *out_plan = nullptr;
  if (geom->nx < 2 || geom->ny < 2 || geom->nz < 1 || geom->nt < 1 || geom->n_angles < 1 || geom->nzr < 1)
    return set_error(VAMX_ERR_ARG, "vamx_rt_create: invalid grid or ray counts");
  if (!geom->angles_deg || !geom->n_grad) return set_error(VAMX_ERR_ARG, "vamx_rt_create: null angles or medium");
  if (!(geom->dx > 0.0) || !(geom->dy > 0.0) || (geom->nz > 1 && !(geom->dz > 0.0)) || !(geom->step > 0.0))
    return set_error(VAMX_ERR_ARG, "vamx_rt_create: voxel sizes and step must be > 0");
  if (geom->max_steps < 1 || geom->exit_every < 1) return set_error(VAMX_ERR_ARG, "vamx_rt_create: step limits");
  if (geom->solver < 0 || geom->solver > 3 || geom->param < 0 || geom->param > 2)
    return set_error(VAMX_ERR_ARG, "vamx_rt_create: unknown solver or parametrisation");
  if (geom->att_kind == VAMX_RT_ATT_GRID && !geom->att_grid)
    return set_error(VAMX_ERR_ARG, "vamx_rt_create: attenuation grid missing");
  vamx_plan* p = new vamx_plan();
  p->kind = PLAN_RAYTRACE;
  if (hipGetDevice(&p->device) != hipSuccess) {
    delete p;
    return set_error(VAMX_ERR_HIP, "vamx_rt_create: hipGetDevice failed");
  }
  RtDev& g = p->rt;
  g.nx = geom->nx; g.ny = geom->ny; g.nz = geom->nz;
  g.dx = geom->dx; g.dy = geom->dy; g.dz = geom->nz > 1 ? geom->dz : 1.0;
  // grid indices are taken as (x - min) * (1 / d): a multiply instead of a
  // float64 division per axis and sample (differs from the division by at
  // most an ulp; trilinear sampling is continuous across voxel faces)
  g.inv_dx = 1.0 / g.dx; g.inv_dy = 1.0 / g.dy; g.inv_dz = 1.0 / g.dz;
  g.xmin = geom->xmin; g.ymin = geom->ymin; g.zmin = geom->zmin;
  g.xmax = g.xmin + (g.nx - 1) * g.dx;
  g.ymax = g.ymin + (g.ny - 1) * g.dy;
  g.zmax = g.nz > 1 ? g.zmin + (g.nz - 1) * g.dz : g.zmin;
  g.span_x = g.xmax - g.xmin; g.span_y = g.ymax - g.ymin; g.span_z = g.zmax - g.zmin;
  g.nt = geom->nt; g.na = geom->n_angles; g.nzr = geom->nzr;
  g.incl = geom->incl_deg * (M_PI / 180.0);
  g.step = geom->step;
  g.max_steps = geom->max_steps;
  g.exit_every = geom->exit_every;
  g.solver = geom->solver;
  g.param = geom->param;
  g.energy0 = (float)geom->ray_energy0;
  g.field = (const float4*)geom->n_grad;
  g.att_kind = geom->att_kind;
  g.att_alpha = geom->att_alpha;
  g.att_radius = geom->att_radius;
  g.att_grid = geom->att_grid;
  g.abs_grid = geom->abs_grid;
  // ray coordinates (setupRaysParallel, pyTorchRayTrace.py:1545-1574):
  // sino_n0 = linspace(min(xmin, ymin), max(xmax, ymax), nt),
  // sino_n2 = linspace(zmin, zmax, nzr), D = sqrt(max_n0^2 + max_n2^2)
  const double n0min = std::min(g.xmin, g.ymin), n0max = std::max(g.xmax, g.ymax);
  std::vector<double> tab((size_t)g.nt + g.nzr + 2 * (size_t)g.na);
  for (int i = 0; i < g.nt; ++i)
    tab[(size_t)i] = g.nt == 1 ? n0min : n0min + (n0max - n0min) * ((double)i / (double)(g.nt - 1));
  if (g.nt > 1) tab[(size_t)g.nt - 1] = n0max;
  for (int i = 0; i < g.nzr; ++i)
    tab[(size_t)g.nt + i] = g.nzr == 1 ? g.zmin : g.zmin + (g.zmax - g.zmin) * ((double)i / (double)(g.nzr - 1));
  if (g.nzr > 1) tab[(size_t)g.nt + g.nzr - 1] = g.zmax;
  for (int i = 0; i < g.na; ++i) {
    const double ang = geom->angles_deg[i] * M_PI / 180.0;
    tab[(size_t)g.nt + g.nzr + 2 * i] = std::cos(ang);
    tab[(size_t)g.nt + g.nzr + 2 * i + 1] = std::sin(ang);
  }
  const double n2max = g.zmax;  // np.amax of the target's z coordinate vector
  g.radial_offset = std::sqrt(n0max * n0max + n2max * n2max);
  const size_t nv = (size_t)g.nx * g.ny * g.nz;
  hipError_t e = hipMalloc((void**)&p->d_rt_tables, tab.size() * sizeof(double));
  if (e == hipSuccess) e = hipMemcpy(p->d_rt_tables, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMalloc((void**)&p->d_accum, nv * sizeof(unsigned long long));
  // [0]: scale (double); from byte 64: kGmaxParts per-block maxima (float)
  if (e == hipSuccess) e = hipMalloc((void**)&p->d_gmax, 64 + kGmaxParts * sizeof(float));
  if (e == hipSuccess) e = hipMalloc((void**)&p->d_stats, 2 * sizeof(unsigned long long));
  if (e != hipSuccess) {
    vamx_destroy(p);
    return hip_check(e, "vamx_rt_create: allocation / upload");
  }
  g.ray_t = p->d_rt_tables;
  g.ray_z = p->d_rt_tables + g.nt;
  g.ray_cs = p->d_rt_tables + g.nt + g.nzr;
  *out_plan = p;
  return VAMX_OK;
}

extern "C" int vamx_rt_integrate(vamx_plan* plan, const float* f, float* gout, void* stream) {
  if (!check_plan(plan, PLAN_RAYTRACE)) return set_error(VAMX_ERR_STATE, "vamx_rt_integrate: not a ray-trace plan");
  if (!f || !gout) return set_error(VAMX_ERR_ARG, "vamx_rt_integrate: null buffer");
  hipStream_t s = (hipStream_t)stream;
  const RtDev& g = plan->rt;
  const long long nv = (long long)g.nx * g.ny * g.nz;
  float* feff = (float*)plan->d_accum;  // the accumulator doubles as f_eff storage
  hipLaunchKernelGGL(rt_feff_kernel, dim3(blocks_for(nv, 256)), dim3(256), 0, s, f, g.abs_grid, feff, nv);
  hipLaunchKernelGGL(rt_zero_stats, dim3(1), dim3(1), 0, s, plan->d_stats);
  const long long nrays = (long long)g.nt * g.na * g.nzr;
  RtIo io{};
  io.feff = feff;
  io.gout = gout;
  io.stats = plan->d_stats;
  const int rc = launch_trace<RT_INTEGRATE>(g, dim3(blocks_for(nrays, 256)), s, io);
  if (rc) return rc;
  VAMX_HIP(hipGetLastError());
  plan->last_stream = s;
  return VAMX_OK;
}

extern "C" int vamx_rt_deposit(vamx_plan* plan, const float* gin, float* dose, void* stream) {
  if (!check_plan(plan, PLAN_RAYTRACE)) return set_error(VAMX_ERR_STATE, "vamx_rt_deposit: not a ray-trace plan");
  if (!gin || !dose) return set_error(VAMX_ERR_ARG, "vamx_rt_deposit: null buffer");
  hipStream_t s = (hipStream_t)stream;
  const RtDev& g = plan->rt;
  const long long nv = (long long)g.nx * g.ny * g.nz;
  const long long nrays = (long long)g.nt * g.na * g.nzr;
  double* scale = (double*)plan->d_gmax;
  float* gmax_part = (float*)((char*)plan->d_gmax + 64);
  hipLaunchKernelGGL(rt_gmax_part, dim3(kGmaxParts), dim3(256), 0, s, gin, nrays, gmax_part);
  hipLaunchKernelGGL(rt_gmax_final, dim3(1), dim3(256), 0, s, gmax_part, (double)g.na, scale, plan->d_stats);
  hipLaunchKernelGGL(rt_zero_u64, dim3(blocks_for(nv, 256)), dim3(256), 0, s, plan->d_accum, nv);
  RtIo io{};
  io.gin = gin;
  io.acc = plan->d_accum;
  io.fx_scale = scale;
  io.stats = plan->d_stats;
  const int rc = launch_trace<RT_DEPOSIT>(g, dim3(blocks_for(nrays, 256)), s, io);
  if (rc) return rc;
  hipLaunchKernelGGL(rt_finalize_kernel, dim3(blocks_for(nv, 256)), dim3(256), 0, s, plan->d_accum, g.abs_grid,
                     (const double*)scale, dose, nv);
  VAMX_HIP(hipGetLastError());
  plan->last_stream = s;
  return VAMX_OK;
}

extern "C" int vamx_rt_record_counts(vamx_plan* plan, int64_t* counts, void* stream) {
  if (!check_plan(plan, PLAN_RAYTRACE))
    return set_error(VAMX_ERR_STATE, "vamx_rt_record_counts: not a ray-trace plan");
  if (!counts) return set_error(VAMX_ERR_ARG, "vamx_rt_record_counts: null buffer");
  hipStream_t s = (hipStream_t)stream;
  const RtDev& g = plan->rt;
  const long long nrays = (long long)g.nt * g.na * g.nzr;
  hipLaunchKernelGGL(rt_zero_stats, dim3(1), dim3(1), 0, s, plan->d_stats);
  RtIo io{};
  io.counts = (long long*)counts;
  io.stats = plan->d_stats;
  const int rc = launch_trace<RT_COUNT>(g, dim3(blocks_for(nrays, 256)), s, io);
  if (rc) return rc;
  VAMX_HIP(hipGetLastError());
  plan->last_stream = s;
  return VAMX_OK;
}

extern "C" int vamx_rt_record(vamx_plan* plan, const int64_t* offsets, int32_t* voxel, float* weight,
                              void* stream) {
  if (!check_plan(plan, PLAN_RAYTRACE)) return set_error(VAMX_ERR_STATE, "vamx_rt_record: not a ray-trace plan");
  if (!offsets || !voxel || !weight) return set_error(VAMX_ERR_ARG, "vamx_rt_record: null buffer");
  hipStream_t s = (hipStream_t)stream;
  const RtDev& g = plan->rt;
  const long long nrays = (long long)g.nt * g.na * g.nzr;
  hipLaunchKernelGGL(rt_zero_stats, dim3(1), dim3(1), 0, s, plan->d_stats);
  RtIo io{};
  io.counts = (long long*)offsets;
  io.rec_vox = voxel;
  io.rec_w = weight;
  io.stats = plan->d_stats;
  const int rc = launch_trace<RT_RECORD>(g, dim3(blocks_for(nrays, 256)), s, io);
  if (rc) return rc;
  VAMX_HIP(hipGetLastError());
  plan->last_stream = s;
  return VAMX_OK;
}

extern "C" int vamx_rt_stats(vamx_plan* plan, int64_t* steps) {
  if (!check_plan(plan, PLAN_RAYTRACE)) return set_error(VAMX_ERR_STATE, "vamx_rt_stats: not a ray-trace plan");
  if (!steps) return set_error(VAMX_ERR_ARG, "vamx_rt_stats: null output");
  VAMX_HIP(hipStreamSynchronize(plan->last_stream));
  unsigned long long h[2];
  VAMX_HIP(hipMemcpy(h, plan->d_stats, sizeof(h), hipMemcpyDeviceToHost));
  steps[0] = (int64_t)h[0];
  steps[1] = (int64_t)h[1];
  return VAMX_OK;
}
