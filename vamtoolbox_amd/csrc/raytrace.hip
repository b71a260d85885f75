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
#include "bclp_core.h"
#include "vamx_internal.h"

namespace vamx {

struct d3 {
  double x, y, z;
};
__device__ __forceinline__ d3 mk(double x, double y, double z) { return {x, y, z}; }
__device__ __forceinline__ d3 add(d3 a, d3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ d3 sub(d3 a, d3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ d3 scl(double s, d3 a) { return {s * a.x, s * a.y, s * a.z}; }

// Trilinear interpolation in polynomial form: from the corner values
// t[c] (c = x + 2 y + 4 z bits) of one cell,
//   v(x, y, z) = (a + y b) + z (c + y d),  a = c0 + x cx, b = cy + x cxy,
//                c = cz + x cxz, d = cyz + x cxyz,
// which is 7 fused multiply-adds per evaluation and is the same function as
// the weighted sum of the corners with weights (1-x or x)(1-y or y)(1-z or z).
struct Tri {
  float c0, cx, cy, cxy, cz, cxz, cyz, cxyz;
};
__device__ __forceinline__ Tri tri_coef(const float* t) {
  Tri r;
  r.c0 = t[0];
  r.cx = t[1] - t[0];
  r.cy = t[2] - t[0];
  r.cxy = (t[3] - t[2]) - r.cx;
  r.cz = t[4] - t[0];
  r.cxz = (t[5] - t[4]) - r.cx;
  r.cyz = (t[6] - t[4]) - r.cy;
  r.cxyz = ((t[7] - t[6]) - (t[5] - t[4])) - r.cxy;
  return r;
}
__device__ __forceinline__ float tri_eval(const Tri& r, float x, float y, float z) {
  const float a = fmaf(x, r.cx, r.c0), b = fmaf(x, r.cxy, r.cy);
  const float c = fmaf(x, r.cxz, r.cz), d = fmaf(x, r.cxyz, r.cyz);
  return fmaf(z, fmaf(y, d, c), fmaf(y, b, a));
}

// One grid cell of the float4 medium, held in registers as the trilinear
// coefficients of its four channels.  The samples of a ray lie a quarter or
// half voxel apart (RK4 samples x, twice x + h/2 and x + h, and x + h starts
// the next step), so most samples fall in the cell of the previous one: the
// 8 taps are loaded only when the cell changes.  Adjacent lanes trace
// adjacent rays, so a wavefront's lanes mostly change cells at the same
// sample.
struct FieldCell {
  int key;       // linear index of the cell's corner voxel; -1 = nothing held
  Tri ch[4];     // n, dn/dx, dn/dy, dn/dz
};

// trilinear sample of the float4 field, grid_sample(align_corners=True,
// padding_mode="border"): index coordinates clamped to [0, size-1].  The
// position and the cell choice are float64; the interpolation is float32
// (the field is stored in float32).  A tap past the last voxel along an axis
// (coordinate exactly size - 1, weight 0) reads the last voxel.
// analytical Luneburg lens at physical position (x, y, z) (medium.py:486-549):
// n = n_sur (2 - (r/R)^p)^(1/p) and grad n = -(n_sur/R)^p n^(1-p) r^(p-2) x
// for r < R; n = n_sur and grad n = 0 outside
__device__ __forceinline__ void lune_sample(const RtDev& g, double x, double y, double z, float& n, float& gx, float& gy,
                                         float& gz) {
  const double r = sqrt(x * x + y * y + z * z);
  if (r < g.lune_R) {
    double nn, f;
    if (g.lune_p == 2.0) {
      const double q = r / g.lune_R;
      nn = g.lune_nsur * sqrt(2.0 - q * q);
      f = -g.lune_c / nn;
    } else {
      nn = g.lune_nsur * pow(2.0 - pow(r / g.lune_R, g.lune_p), 1.0 / g.lune_p);
      f = -g.lune_c * pow(nn, 1.0 - g.lune_p) * pow(r, g.lune_p - 2.0);
    }
    n = (float)nn;
    gx = (float)(f * x);
    gy = (float)(f * y);
    gz = (float)(f * z);
  } else {
    n = (float)g.lune_nsur;
    gx = gy = gz = 0.0f;
  }
}

template <bool LUNE>
__device__ __forceinline__ void field_sample_idx(const RtDev& g, double ix, double iy, double iz, FieldCell& fc,
                                                 float& n, float& gx, float& gy, float& gz) {
  if constexpr (LUNE) {
    lune_sample(g, g.xmin + ix * g.dx, g.ymin + iy * g.dy, g.nz > 1 ? g.zmin + iz * g.dz : g.zmin, n, gx, gy, gz);
    return;
  }
  ix = fmin(fmax(ix, 0.0), g.ixmax);
  iy = fmin(fmax(iy, 0.0), g.iymax);
  iz = fmin(fmax(iz, 0.0), g.izmax);
  const double fx = floor(ix), fy = floor(iy), fz = floor(iz);
  const int x0 = (int)fx, y0 = (int)fy, z0 = (int)fz;
  const int key = (x0 * g.ny + y0) * g.nz + z0;
  if (key != fc.key) {
    fc.key = key;
    const int sx = x0 + 1 < g.nx ? g.ny * g.nz : 0;
    const int sy = y0 + 1 < g.ny ? g.nz : 0;
    const int sz = z0 + 1 < g.nz ? 1 : 0;
    const float4* b = g.field + key;
    float4 t[8];
    t[0] = b[0];
    t[1] = b[sx];
    t[2] = b[sy];
    t[3] = b[sx + sy];
    t[4] = b[sz];
    t[5] = b[sx + sz];
    t[6] = b[sy + sz];
    t[7] = b[sx + sy + sz];
    float q[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) q[c] = t[c].x;
    fc.ch[0] = tri_coef(q);
#pragma unroll
    for (int c = 0; c < 8; ++c) q[c] = t[c].y;
    fc.ch[1] = tri_coef(q);
#pragma unroll
    for (int c = 0; c < 8; ++c) q[c] = t[c].z;
    fc.ch[2] = tri_coef(q);
#pragma unroll
    for (int c = 0; c < 8; ++c) q[c] = t[c].w;
    fc.ch[3] = tri_coef(q);
  }
  const float wx = (float)(ix - fx), wy = (float)(iy - fy), wz = (float)(iz - fz);
  n = tri_eval(fc.ch[0], wx, wy, wz);
  gx = tri_eval(fc.ch[1], wx, wy, wz);
  gy = tri_eval(fc.ch[2], wx, wy, wz);
  gz = tri_eval(fc.ch[3], wx, wy, wz);
}

template <bool LUNE>
__device__ __forceinline__ void field_sample(const RtDev& g, d3 x, FieldCell& fc, double& n, d3& gn) {
  float nf, gx, gy, gz;
  if constexpr (LUNE)
    lune_sample(g, x.x, x.y, x.z, nf, gx, gy, gz);
  else
    field_sample_idx<false>(g, (x.x - g.xmin) * g.inv_dx, (x.y - g.ymin) * g.inv_dy,
                     g.nz > 1 ? (x.z - g.zmin) * g.inv_dz : 0.0, fc, nf, gx, gy, gz);
  n = (double)nf;
  gn = mk((double)gx, (double)gy, (double)gz);
}

// attenuation coefficient at x (medium.py AttenuationModel.alpha); the sharp
// boundary |x| < R is tested as |x|^2 < R^2
__device__ __forceinline__ double att_alpha(const RtDev& g, d3 x) {
  if (g.att_kind == VAMX_RT_ATT_CYLINDER)
    return x.x * x.x + x.y * x.y < g.att_r2 ? g.att_alpha : 0.0;
  if (g.att_kind == VAMX_RT_ATT_BALL)
    return x.x * x.x + x.y * x.y + x.z * x.z < g.att_r2 ? g.att_alpha : 0.0;
  if (g.att_kind == VAMX_RT_ATT_GRID) {
    // grid_sample(align_corners=True, padding_mode="zeros")
    const double ix = (x.x - g.xmin) * g.inv_dx, iy = (x.y - g.ymin) * g.inv_dy;
    const double iz = g.nz > 1 ? (x.z - g.zmin) * g.inv_dz : 0.0;
    const double fx = floor(ix), fy = floor(iy), fz = floor(iz);
    const double wx = ix - fx, wy = iy - fy, wz = iz - fz;
    const int x0 = (int)fmax(fmin(fx, 2.0e9), -2.0e9), y0 = (int)fmax(fmin(fy, 2.0e9), -2.0e9);
    const int z0 = (int)fmax(fmin(fz, 2.0e9), -2.0e9);
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
template <int PARAM, bool LUNE>
__device__ __forceinline__ void rhs(const RtDev& g, d3 x, d3 v, FieldCell& fc, d3& dx, d3& dv) {
  double n;
  d3 gn;
  field_sample<LUNE>(g, x, fc, n, gn);
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

__device__ __forceinline__ d3 fma3(double s, d3 a, d3 b) { return {fma(s, a.x, b.x), fma(s, a.y, b.y), fma(s, a.z, b.z)}; }

// one step of solver SOLVER: (x, v) -> (x', v')
template <int SOLVER, int PARAM, bool LUNE>
__device__ __forceinline__ void advance(const RtDev& g, d3& x, d3& v, FieldCell& fc) {
  const double h = g.step;
  if (SOLVER == VAMX_RT_RK4) {
    // x' = x + h/6 (k1 + 2 k2 + 2 k3 + k4), likewise v.  x and v stay
    // float64; the stage derivatives k, their running sums and the stage
    // offsets c k (at most h apart from x) are float32, and the stage
    // sample positions are formed in grid-index units as the float64 index
    // of x plus the float32 offset.  The four stages are one loop (not
    // unrolled), so the field sample -- and the cell it holds in registers --
    // exists once in the code and keeps one set of registers.
    const float hf = (float)h;
    const double bx = (x.x - g.xmin) * g.inv_dx, by = (x.y - g.ymin) * g.inv_dy;
    const double bz = g.nz > 1 ? (x.z - g.zmin) * g.inv_dz : 0.0;
    const float idx = (float)g.inv_dx, idy = (float)g.inv_dy, idz = g.nz > 1 ? (float)g.inv_dz : 0.0f;
    float kx0 = 0.0f, kx1 = 0.0f, kx2 = 0.0f, kv0 = 0.0f, kv1 = 0.0f, kv2 = 0.0f;
    float sx0 = 0.0f, sx1 = 0.0f, sx2 = 0.0f, sv0 = 0.0f, sv1 = 0.0f, sv2 = 0.0f;
#pragma unroll 1
    for (int st = 0; st < 4; ++st) {
      const float c = st == 0 ? 0.0f : (st == 3 ? hf : 0.5f * hf);  // stage offset along the previous k
      const float wk = (st == 0 || st == 3) ? 1.0f : 2.0f;          // weight of this k in the sum
      float n, gx, gy, gz;
      field_sample_idx<LUNE>(g, bx + (double)((c * kx0) * idx), by + (double)((c * kx1) * idy),
                       bz + (double)((c * kx2) * idz), fc, n, gx, gy, gz);
      const float vs0 = (float)(v.x + (double)(c * kv0));
      const float vs1 = (float)(v.y + (double)(c * kv1));
      const float vs2 = (float)(v.z + (double)(c * kv2));
      if (PARAM == VAMX_RT_CANONICAL) {
        kx0 = vs0; kx1 = vs1; kx2 = vs2;
        kv0 = n * gx; kv1 = n * gy; kv2 = n * gz;
      } else if (PARAM == VAMX_RT_PHYSICAL) {
        const float r = __builtin_amdgcn_rcpf(n);
        kx0 = vs0 * r; kx1 = vs1 * r; kx2 = vs2 * r;
        kv0 = gx; kv1 = gy; kv2 = gz;
      } else {
        const float r = __builtin_amdgcn_rcpf(n), r2 = r * r;
        kx0 = vs0 * r2; kx1 = vs1 * r2; kx2 = vs2 * r2;
        kv0 = gx * r; kv1 = gy * r; kv2 = gz * r;
      }
      sx0 = fmaf(wk, kx0, sx0); sx1 = fmaf(wk, kx1, sx1); sx2 = fmaf(wk, kx2, sx2);
      sv0 = fmaf(wk, kv0, sv0); sv1 = fmaf(wk, kv1, sv1); sv2 = fmaf(wk, kv2, sv2);
    }
    const float h6 = hf / 6.0f;
    x = add(x, mk((double)(h6 * sx0), (double)(h6 * sx1), (double)(h6 * sx2)));
    v = add(v, mk((double)(h6 * sv0), (double)(h6 * sv1), (double)(h6 * sv2)));
    return;
  }
  double n;
  d3 gn;
  field_sample<LUNE>(g, x, fc, n, gn);
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
template <int SOLVER, int PARAM, bool LUNE>
__device__ __forceinline__ void ray_init(const RtDev& g, int t, int a, int zr, d3& x, d3& v, FieldCell& fc) {
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
    field_sample<LUNE>(g, x, fc, n, gn);
    v = sub(v, scl(g.step / 2.0, accel<PARAM>(n, gn)));
  }
}

// exit test (pyTorchRayTrace.py:1391-1420)
__device__ __forceinline__ bool exited(const RtDev& g, d3 x, d3 v) {
  const bool pos = (x.x > g.exit_hi_x && v.x > 0.0) || (x.y > g.exit_hi_y && v.y > 0.0) ||
                   (x.z > g.exit_hi_z && v.z > 0.0);
  const bool neg = (x.x < g.exit_lo_x && v.x < 0.0) || (x.y < g.exit_lo_y && v.y < 0.0) ||
                   (x.z < g.exit_lo_z && v.z < 0.0);
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
  // integrate with the BCLP update fused into the ray's output: instead of
  // writing the integral, g[ray] <- clip(g[ray] - lr * scale * integral)
  float* upd_g;
  const double* upd_scale;
  float upd_lr, upd_glb, upd_gub;
};

// Interaction taps of one cell (corner x0, y0, z0, possibly -1 on an axis):
// integrate holds f_eff at its taps; deposit holds, per tap, the float32 sum
// of the weights of the ray's consecutive steps in the cell, written out as
// one fixed-point integer atomic per tap when the ray moves to another cell
// and at the end of the ray (integer addition is associative, so the
// accumulator does not depend on the order of the atomics).
struct TapCell {
  int key;           // cell key ((x0+1) (ny+1) + y0+1) (nz+1) + z0+1; -1 = nothing held
  int x0, y0, z0;
  Tri f;             // integrate: f_eff at the taps (0 outside the grid), trilinear coefficients
  float q[8];        // deposit: summed tap weights
};

// writes out the taps of `mask` (bit c = tap c) and clears them
__device__ __forceinline__ void tap_flush(const RtDev& g, const RtIo& io, TapCell& tc, double scale,
                                          unsigned mask = 0xffu) {
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    if (!((mask >> c) & 1u)) continue;
    const long long q = llrint((double)tc.q[c] * scale);
    tc.q[c] = 0.0f;
    if (q == 0) continue;   // taps outside the grid carry weight 0
    const int xi = tc.x0 + (c & 1), yi = tc.y0 + ((c >> 1) & 1), zi = tc.z0 + ((c >> 2) & 1);
    atomicAdd(io.acc + ((size_t)xi * g.ny + yi) * g.nz + zi, (unsigned long long)q);
  }
}

// Moves the held cell by d (-1, 0 or +1) along axis AX (bit AX of a tap
// index): the face the new cell shares with the old one keeps its sums, the
// other face is written out.  |d| > 1: everything is written out.
template <int AX>
__device__ __forceinline__ void tap_shift(const RtDev& g, const RtIo& io, TapCell& tc, double scale, int d) {
  constexpr unsigned lo = AX == 0 ? 0x55u : (AX == 1 ? 0x33u : 0x0fu);   // taps with bit AX = 0
  constexpr int bit = 1 << AX;
  if (d == 0) return;
  if (d == 1) {
    tap_flush(g, io, tc, scale, lo);
#pragma unroll
    for (int c = 0; c < 8; ++c)
      if (!(c & bit)) {
        tc.q[c] = tc.q[c | bit];
        tc.q[c | bit] = 0.0f;
      }
  } else if (d == -1) {
    tap_flush(g, io, tc, scale, lo ^ 0xffu);
#pragma unroll
    for (int c = 0; c < 8; ++c)
      if (!(c & bit)) {
        tc.q[c | bit] = tc.q[c];
        tc.q[c] = 0.0f;
      }
  } else {
    tap_flush(g, io, tc, scale);
  }
  if (AX == 0) tc.x0 += d;
  else if (AX == 1) tc.y0 += d;
  else tc.z0 += d;
}

// at most 128 VGPRs: 4 wavefronts per SIMD
template <int MODE, int SOLVER, int PARAM, bool LUNE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void rt_trace_kernel(RtDev g, RtIo io) {
  const long long nrays = (long long)g.nt * g.na * g.nzr;
  const long long rid = (long long)xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  // every lane runs to the end (no early return): the step statistics are
  // summed across the wavefront before one atomic per wavefront
  const bool valid = rid < nrays;
  bool live = valid;
  // launch order: z fastest, then t, then the angle, so that the rays in
  // flight on one XCD (a contiguous range of blocks) are sheets of a few
  // angles sweeping the volume together (their field reads and deposit
  // atomics stay in a slab of cells); r is the ray's index in the sinogram
  // (t, angle, z)
  const long long q = valid ? rid : 0;
  const int zr = (int)(q % g.nzr);
  const long long qa = q / g.nzr;
  const int t = (int)(qa % g.nt), a = (int)(qa / g.nt);
  const long long r = ((long long)t * g.na + a) * g.nzr + zr;
  double E0 = (double)g.energy0;
  if (MODE == RT_DEPOSIT && live) {
    E0 = (double)io.gin[r];
    live = E0 != 0.0;  // a ray of zero energy contributes nothing and is not traced
  }
  const double scale = MODE == RT_DEPOSIT ? *io.fx_scale : 0.0;
  long long slot = (MODE == RT_RECORD && live) ? io.counts[r] : 0;   // record: this ray's first slot
  long long ntaps = 0;
  FieldCell fc;
  fc.key = -1;
  TapCell tc;
  tc.key = -1;
  tc.x0 = tc.y0 = tc.z0 = 0;
  tc.f = Tri{0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int c = 0; c < 8; ++c) tc.q[c] = 0.0f;
  d3 x, v;
  ray_init<SOLVER, PARAM, LUNE>(g, t, a, zr, x, v, fc);
  const double hw = g.hw;
  const float e0f = (float)E0;
  double A = 0.0, integral = 0.0;
  int step = 0, until_exit = 1;
  bool active = live;
  if (live) {
    for (; step < g.max_steps; ++step) {
      const d3 xo = x;
      advance<SOLVER, PARAM, LUNE>(g, x, v, fc);
      const d3 d = sub(x, xo);
      const double al = att_alpha(g, xo);
      // attenuance (767-769) and ray energy (770-774); step length and
      // exponential in float32
      if (al != 0.0) A += al * (double)__builtin_amdgcn_sqrtf((float)(d.x * d.x + d.y * d.y + d.z * d.z));
      const double E = (double)(e0f * __expf(-(float)A));
      const Taps tp = taps_at(g, x);
      if (tp.x0 >= -1 && tp.x0 < g.nx && tp.y0 >= -1 && tp.y0 < g.ny && tp.z0 >= -1 && tp.z0 < g.nz) {
        const double base = hw * E;
        // tap weights, 0 for taps outside the grid (and, in 2-D, above the plane)
        const float wx = (float)tp.wx, wy = (float)tp.wy, wz = tp.nzt == 1 ? 0.0f : (float)tp.wz;
        const float ux = 1.0f - wx, uy = 1.0f - wy, uz = 1.0f - wz;
        const bool vx0 = tp.x0 >= 0, vx1 = tp.x0 + 1 < g.nx, vy0 = tp.y0 >= 0, vy1 = tp.y0 + 1 < g.ny;
        const bool vz0 = tp.z0 >= 0, vz1 = tp.nzt == 2 && tp.z0 + 1 < g.nz;
        float w[8];
#pragma unroll
        for (int c = 0; c < 8; ++c) {
          const bool ok = ((c & 1) ? vx1 : vx0) && (((c >> 1) & 1) ? vy1 : vy0) && (((c >> 2) & 1) ? vz1 : vz0);
          const float ww = (((c & 1) ? wx : ux) * (((c >> 1) & 1) ? wy : uy)) * (((c >> 2) & 1) ? wz : uz);
          w[c] = ok ? ww : 0.0f;
        }
        const int key = ((tp.x0 + 1) * (g.ny + 1) + (tp.y0 + 1)) * (g.nz + 1) + (tp.z0 + 1);
        if (MODE == RT_INTEGRATE) {
          if (key != tc.key) {
            tc.key = key;
            float fv[8];
#pragma unroll
            for (int c = 0; c < 8; ++c) {
              // taps outside the grid read an in-grid voxel and hold 0
              const int xi = min(max(tp.x0 + (c & 1), 0), g.nx - 1);
              const int yi = min(max(tp.y0 + ((c >> 1) & 1), 0), g.ny - 1);
              const int zi = min(max(tp.z0 + ((c >> 2) & 1), 0), g.nz - 1);
              const float f = io.feff[((size_t)xi * g.ny + yi) * g.nz + zi];
              const bool ok = ((c & 1) ? vx1 : vx0) && (((c >> 1) & 1) ? vy1 : vy0) && (((c >> 2) & 1) ? vz1 : vz0);
              fv[c] = ok ? f : 0.0f;
            }
            tc.f = tri_coef(fv);
          }
          integral += base * (double)tri_eval(tc.f, wx, wy, wz);
        } else if (MODE == RT_DEPOSIT) {
          if (key != tc.key) {
            // the taps shared with the previous cell keep accumulating: about
            // half of the atomics of a flush per cell change
            if (tc.key >= 0) {
              tap_shift<0>(g, io, tc, scale, tp.x0 - tc.x0);
              tap_shift<1>(g, io, tc, scale, tp.y0 - tc.y0);
              tap_shift<2>(g, io, tc, scale, tp.z0 - tc.z0);
            }
            tc.key = key;
            tc.x0 = tp.x0;
            tc.y0 = tp.y0;
            tc.z0 = tp.z0;
          }
          const float bf = (float)base;
#pragma unroll
          for (int c = 0; c < 8; ++c) tc.q[c] = fmaf(bf, w[c], tc.q[c]);
        } else {
#pragma unroll
          for (int c = 0; c < 8; ++c) {
            if (w[c] == 0.0f) continue;
            if (MODE == RT_COUNT) {
              ++ntaps;
            } else {
              const int xi = tp.x0 + (c & 1), yi = tp.y0 + ((c >> 1) & 1), zi = tp.z0 + ((c >> 2) & 1);
              io.rec_vox[slot] = (int32_t)(((size_t)xi * g.ny + yi) * g.nz + zi);
              io.rec_w[slot] = (float)(base * (double)w[c]);
              ++slot;
            }
          }
        }
      }
      // the exit test runs at steps 0, exit_every, 2 exit_every, ...
      if (--until_exit == 0) {
        until_exit = g.exit_every;
        if (exited(g, x, v)) {
          active = false;
          ++step;
          break;
        }
      }
    }
    if (MODE == RT_DEPOSIT && tc.key >= 0) tap_flush(g, io, tc, scale);
  }
  if (MODE == RT_INTEGRATE && valid) {
    if (io.upd_g) {
      // BCLP.py:486-487, 532: g <- clip(g - lr * scale * gradient)
      const float grad = (float)(*io.upd_scale) * (float)integral;
      io.upd_g[r] = fminf(fmaxf(io.upd_g[r] - io.upd_lr * grad, io.upd_glb), io.upd_gub);
    } else {
      io.gout[r] = (float)integral;
    }
  }
  if (MODE == RT_COUNT && valid) io.counts[r] = ntaps;
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

// Deposit finalisation fused with the BCLP epilogue (the ray tracer's
// backward followed by BCLPNorm's loss and operand, BCLP.py:277-367):
// dose = alpha_abs acc / scale, the response, band and Lp weighting per
// voxel, the operand (times ab_op when given: alpha_abs makes it the f_eff
// that the gradient's integrate pass reads) and per-block loss partials.
__global__ __launch_bounds__(256) void rt_bclp_finalize_kernel(const unsigned long long* __restrict__ acc,
                                                               const float* __restrict__ ab,
                                                               const float* __restrict__ ab_op,
                                                               const double* __restrict__ fx_scale, RespDev r,
                                                               BclpDev b, const float* __restrict__ target,
                                                               float* __restrict__ dose,
                                                               float* __restrict__ operand_eff,
                                                               double* __restrict__ partials, long long nv) {
  __shared__ double sh[256];
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  double part = 0.0;
  if (i < nv) {
    const double v = (double)(long long)acc[i] / *fx_scale;
    const float d = (float)(ab ? v * (double)ab[i] : v);
    dose[i] = d;
    float o;
    part = bclp_voxel(r, b, d, target[i], (size_t)i, o);
    operand_eff[i] = ab_op ? o * ab_op[i] : o;
  }
  const double sum = block_sum_256(part, sh);
  if (threadIdx.x == 0) partials[blockIdx.x] = sum;
}

// compile-time solver, parametrisation and index model: each instantiation
// keeps only its own stepper's and field's registers live
template <int MODE, int SOLVER, bool LUNE>
int launch_trace_s(const RtDev& g, dim3 grid, hipStream_t s, const RtIo& io) {
  if (g.param == VAMX_RT_CANONICAL)
    hipLaunchKernelGGL((rt_trace_kernel<MODE, SOLVER, VAMX_RT_CANONICAL, LUNE>), grid, dim3(256), 0, s, g, io);
  else if (g.param == VAMX_RT_PHYSICAL)
    hipLaunchKernelGGL((rt_trace_kernel<MODE, SOLVER, VAMX_RT_PHYSICAL, LUNE>), grid, dim3(256), 0, s, g, io);
  else
    hipLaunchKernelGGL((rt_trace_kernel<MODE, SOLVER, VAMX_RT_OPTICAL, LUNE>), grid, dim3(256), 0, s, g, io);
  return VAMX_OK;
}

template <int MODE, bool LUNE>
int launch_trace_l(const RtDev& g, dim3 grid, hipStream_t s, const RtIo& io) {
  switch (g.solver) {
    case VAMX_RT_SYMPLECTIC_EULER: return launch_trace_s<MODE, VAMX_RT_SYMPLECTIC_EULER, LUNE>(g, grid, s, io);
    case VAMX_RT_FORWARD_EULER: return launch_trace_s<MODE, VAMX_RT_FORWARD_EULER, LUNE>(g, grid, s, io);
    case VAMX_RT_LEAPFROG: return launch_trace_s<MODE, VAMX_RT_LEAPFROG, LUNE>(g, grid, s, io);
    case VAMX_RT_RK4: return launch_trace_s<MODE, VAMX_RT_RK4, LUNE>(g, grid, s, io);
    default: return set_error(VAMX_ERR_ARG, "ray tracer: unknown solver");
  }
}

template <int MODE>
int launch_trace(const RtDev& g, dim3 grid, hipStream_t s, const RtIo& io) {
  return g.index_kind == VAMX_RT_INDEX_LUNEBURG ? launch_trace_l<MODE, true>(g, grid, s, io)
                                                : launch_trace_l<MODE, false>(g, grid, s, io);
}

}  // namespace vamx

using namespace vamx;

extern "C" int vamx_rt_create(const vamx_rt_geom* geom, vamx_plan** out_plan) {
  if (!geom || !out_plan) return set_error(VAMX_ERR_ARG, "vamx_rt_create: null argument");
  *out_plan = nullptr;
  if (geom->nx < 2 || geom->ny < 2 || geom->nz < 1 || geom->nt < 1 || geom->n_angles < 1 || geom->nzr < 1)
    return set_error(VAMX_ERR_ARG, "vamx_rt_create: invalid grid or ray counts");
  if (geom->index_kind != VAMX_RT_INDEX_GRID && geom->index_kind != VAMX_RT_INDEX_LUNEBURG)
    return set_error(VAMX_ERR_ARG, "vamx_rt_create: unknown index kind");
  if (!geom->angles_deg || (geom->index_kind == VAMX_RT_INDEX_GRID && !geom->n_grad))
    return set_error(VAMX_ERR_ARG, "vamx_rt_create: null angles or medium");
  if (geom->index_kind == VAMX_RT_INDEX_LUNEBURG &&
      (!(geom->index_R > 0.0) || !(geom->index_p > 0.0) || !(geom->index_n_sur > 0.0)))
    return set_error(VAMX_ERR_ARG, "vamx_rt_create: Luneburg lens needs R, p, n_sur > 0");
  if (!(geom->dx > 0.0) || !(geom->dy > 0.0) || (geom->nz > 1 && !(geom->dz > 0.0)) || !(geom->step > 0.0))
    return set_error(VAMX_ERR_ARG, "vamx_rt_create: voxel sizes and step must be > 0");
  if (geom->max_steps < 1 || geom->exit_every < 1) return set_error(VAMX_ERR_ARG, "vamx_rt_create: step limits");
  if (geom->solver < 0 || geom->solver > 3 || geom->param < 0 || geom->param > 2)
    return set_error(VAMX_ERR_ARG, "vamx_rt_create: unknown solver or parametrisation");
  // cell keys and tap indices are 32-bit
  if ((double)(geom->nx + 1) * (geom->ny + 1) * (geom->nz + 1) >= 2147483647.0)
    return set_error(VAMX_ERR_ARG, "vamx_rt_create: grid too large (at most 2^31 cells)");
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
  g.ixmax = g.nx - 1; g.iymax = g.ny - 1; g.izmax = g.nz - 1;
  g.nt = geom->nt; g.na = geom->n_angles; g.nzr = geom->nzr;
  g.incl = geom->incl_deg * (M_PI / 180.0);
  g.step = geom->step;
  g.max_steps = geom->max_steps;
  g.exit_every = geom->exit_every;
  g.solver = geom->solver;
  g.param = geom->param;
  g.energy0 = (float)geom->ray_energy0;
  g.field = (const float4*)geom->n_grad;
  g.index_kind = geom->index_kind;
  g.lune_R = geom->index_R;
  g.lune_p = geom->index_p;
  g.lune_nsur = geom->index_n_sur;
  g.lune_c = geom->index_kind == VAMX_RT_INDEX_LUNEBURG ? std::pow(geom->index_n_sur / geom->index_R, geom->index_p)
                                                         : 0.0;
  g.att_kind = geom->att_kind;
  g.att_alpha = geom->att_alpha;
  g.att_radius = geom->att_radius;
  g.att_r2 = g.att_radius * g.att_radius;
  g.hw = g.step / g.dx;   // step_size_idx_unit (pyTorchRayTrace.py:840-842)
  {
    const double vz = g.nz > 1 ? g.dz : INFINITY;   // exitCheck (pyTorchRayTrace.py:1391-1420)
    g.exit_lo_x = g.xmin - g.dx; g.exit_hi_x = g.xmax + g.dx;
    g.exit_lo_y = g.ymin - g.dy; g.exit_hi_y = g.ymax + g.dy;
    g.exit_lo_z = g.zmin - vz; g.exit_hi_z = g.zmax + vz;
  }
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

namespace {

struct RtBclpWs {
  size_t partials_off, scale_off, operand_off, total;
  int nparts;
};

RtBclpWs rt_bclp_ws(const vamx_plan* p) {
  const RtDev& g = p->rt;
  const long long nv = (long long)g.nx * g.ny * g.nz;
  RtBclpWs w;
  w.nparts = (int)blocks_for(nv, 256);
  w.partials_off = 0;
  w.scale_off = ((size_t)w.nparts * sizeof(double) + 255) / 256 * 256;
  w.operand_off = w.scale_off + 256;
  w.total = w.operand_off + (size_t)nv * sizeof(float);
  return w;
}

}  // namespace

extern "C" int vamx_rt_bclp_workspace_bytes(vamx_plan* plan, int64_t* bytes) {
  if (!check_plan(plan, PLAN_RAYTRACE))
    return set_error(VAMX_ERR_STATE, "vamx_rt_bclp_workspace_bytes: not a ray-trace plan");
  if (!bytes) return set_error(VAMX_ERR_ARG, "vamx_rt_bclp_workspace_bytes: null output");
  *bytes = (int64_t)rt_bclp_ws(plan).total;
  return VAMX_OK;
}

extern "C" int vamx_rt_bclp_step(vamx_plan* plan, const vamx_response* resp, const vamx_bclp_params* prm, float* g,
                                 const float* target, const float* eps_arr, const float* weight_arr, float* dose,
                                 double* loss_log, int32_t* iter_counter, void* workspace, void* stream) {
  if (!check_plan(plan, PLAN_RAYTRACE)) return set_error(VAMX_ERR_STATE, "vamx_rt_bclp_step: not a ray-trace plan");
  if (!g || !target || !dose || !loss_log || !iter_counter || !workspace)
    return set_error(VAMX_ERR_ARG, "vamx_rt_bclp_step: null buffer");
  RespDev r;
  BclpDev b;
  int rc = bclp_fill_devs(resp, prm, eps_arr, weight_arr, r, b);
  if (rc) return rc;
  if (prm->grad_op != 0)
    return set_error(VAMX_ERR_ARG, "vamx_rt_bclp_step: the ray tracer's forward is already the exact transpose "
                                   "(grad_op must be 0)");
  hipStream_t s = (hipStream_t)stream;
  const RtDev& gd = plan->rt;
  const long long nv = (long long)gd.nx * gd.ny * gd.nz;
  const long long nrays = (long long)gd.nt * gd.na * gd.nzr;
  const RtBclpWs w = rt_bclp_ws(plan);
  char* ws = (char*)workspace;
  double* partials = (double*)(ws + w.partials_off);
  double* scale = (double*)(ws + w.scale_off);
  float* operand_eff = (float*)(ws + w.operand_off);
  // dose = alpha_abs W^T g (the deposit, pyTorchRayTrace.py:109-133)
  double* fx_scale = (double*)plan->d_gmax;
  float* gmax_part = (float*)((char*)plan->d_gmax + 64);
  hipLaunchKernelGGL(rt_gmax_part, dim3(kGmaxParts), dim3(256), 0, s, g, nrays, gmax_part);
  hipLaunchKernelGGL(rt_gmax_final, dim3(1), dim3(256), 0, s, gmax_part, (double)gd.na, fx_scale, plan->d_stats);
  hipLaunchKernelGGL(rt_zero_u64, dim3(blocks_for(nv, 256)), dim3(256), 0, s, plan->d_accum, nv);
  RtIo io{};
  io.gin = g;
  io.acc = plan->d_accum;
  io.fx_scale = fx_scale;
  io.stats = plan->d_stats;
  rc = launch_trace<RT_DEPOSIT>(gd, dim3(blocks_for(nrays, 256)), s, io);
  if (rc) return rc;
  // response, band, Lp weighting, operand and loss partials; loss and scale
  hipLaunchKernelGGL(rt_bclp_finalize_kernel, dim3(w.nparts), dim3(256), 0, s, plan->d_accum, gd.abs_grid,
                     gd.abs_grid, (const double*)fx_scale, r, b, target, dose, operand_eff, partials, nv);
  VAMX_HIP(hipGetLastError());
  rc = bclp_reduce_launch(partials, w.nparts, b, nullptr, iter_counter, loss_log, scale, s);
  if (rc) return rc;
  // gradient W (alpha_abs operand) fused with the update of g (BCLP.py:353-375, 486-487)
  hipLaunchKernelGGL(rt_zero_stats, dim3(1), dim3(1), 0, s, plan->d_stats);
  RtIo iu{};
  iu.feff = operand_eff;
  iu.stats = plan->d_stats;
  iu.upd_g = g;
  iu.upd_scale = scale;
  iu.upd_lr = (float)prm->learning_rate;
  iu.upd_glb = (float)prm->glb;
  iu.upd_gub = (float)prm->gub;
  rc = launch_trace<RT_INTEGRATE>(gd, dim3(blocks_for(nrays, 256)), s, iu);
  if (rc) return rc;
  VAMX_HIP(hipGetLastError());
  plan->last_stream = s;
  return VAMX_OK;
}

// The loss half of vamx_rt_bclp_step for an autograd node: dose, operand
// (without alpha_abs: vamx_rt_integrate applies it) and L(g); g is not changed.
extern "C" int vamx_rt_bclp_eval(vamx_plan* plan, const vamx_response* resp, const vamx_bclp_params* prm,
                                 const float* g, const float* target, const float* eps_arr, const float* weight_arr,
                                 float* dose, float* operand, double* loss_out, void* workspace, void* stream) {
  if (!check_plan(plan, PLAN_RAYTRACE)) return set_error(VAMX_ERR_STATE, "vamx_rt_bclp_eval: not a ray-trace plan");
  if (!g || !target || !dose || !loss_out || !workspace)
    return set_error(VAMX_ERR_ARG, "vamx_rt_bclp_eval: null buffer");
  RespDev r;
  BclpDev b;
  int rc = bclp_fill_devs(resp, prm, eps_arr, weight_arr, r, b);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  const RtDev& gd = plan->rt;
  const long long nv = (long long)gd.nx * gd.ny * gd.nz;
  const long long nrays = (long long)gd.nt * gd.na * gd.nzr;
  const RtBclpWs w = rt_bclp_ws(plan);
  char* ws = (char*)workspace;
  double* partials = (double*)(ws + w.partials_off);
  double* scale = (double*)(ws + w.scale_off);
  float* op = operand ? operand : (float*)(ws + w.operand_off);
  double* fx_scale = (double*)plan->d_gmax;
  float* gmax_part = (float*)((char*)plan->d_gmax + 64);
  hipLaunchKernelGGL(rt_gmax_part, dim3(kGmaxParts), dim3(256), 0, s, g, nrays, gmax_part);
  hipLaunchKernelGGL(rt_gmax_final, dim3(1), dim3(256), 0, s, gmax_part, (double)gd.na, fx_scale, plan->d_stats);
  hipLaunchKernelGGL(rt_zero_u64, dim3(blocks_for(nv, 256)), dim3(256), 0, s, plan->d_accum, nv);
  RtIo io{};
  io.gin = g;
  io.acc = plan->d_accum;
  io.fx_scale = fx_scale;
  io.stats = plan->d_stats;
  rc = launch_trace<RT_DEPOSIT>(gd, dim3(blocks_for(nrays, 256)), s, io);
  if (rc) return rc;
  hipLaunchKernelGGL(rt_bclp_finalize_kernel, dim3(w.nparts), dim3(256), 0, s, plan->d_accum, gd.abs_grid,
                     (const float*)nullptr, (const double*)fx_scale, r, b, target, dose, op, partials, nv);
  VAMX_HIP(hipGetLastError());
  rc = bclp_reduce_launch(partials, w.nparts, b, loss_out, nullptr, nullptr, scale, s);
  if (rc) return rc;
  plan->last_stream = s;
  return VAMX_OK;
}
