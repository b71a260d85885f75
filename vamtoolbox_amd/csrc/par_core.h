// Device-level cores of the parallel-beam operators, shared by the plain
// kernels (parallel.hip), the fused BCLP kernels (bclp.hip) and the batched
// launch (batched.hip).
//
// Discretisation: the reference's CPU projector "D1"
// (vamtoolbox/projector/Projector3DParallel.py:113-207) plus the detector
// offset / absorption / inclination extensions.  Layouts: volume (n, n, nz)
// and sinogram (n, n_angles, nz), both z-contiguous; a lane owns one ray or
// voxel column and ZV consecutive z values, so sample weights are identical
// across a wavefront and every tap is a coalesced load.
//
// Geometry (sample coordinates, floor/ceil decisions, weights) is evaluated
// in float64 with the reference's operation order (translation units are
// compiled with -ffp-contract=off) so that every discontinuous decision
// matches the reference; data arithmetic is float32 with explicit fused
// multiply-adds.  No core uses atomics.
#pragma once

#include "vamx_internal.h"

namespace vamx {

template <int ZV> struct VecT;
template <> struct VecT<1> { using T = float; };
template <> struct VecT<2> { using T = float2; };
template <> struct VecT<4> { using T = float4; };

template <int ZV> __device__ __forceinline__ typename VecT<ZV>::T vzero();
template <> __device__ __forceinline__ float vzero<1>() { return 0.0f; }
template <> __device__ __forceinline__ float2 vzero<2>() { return make_float2(0.f, 0.f); }
template <> __device__ __forceinline__ float4 vzero<4>() { return make_float4(0.f, 0.f, 0.f, 0.f); }

template <int ZV> __device__ __forceinline__ typename VecT<ZV>::T vload(const float* p);
template <> __device__ __forceinline__ float vload<1>(const float* p) { return *p; }
template <> __device__ __forceinline__ float2 vload<2>(const float* p) { return *reinterpret_cast<const float2*>(p); }
template <> __device__ __forceinline__ float4 vload<4>(const float* p) { return *reinterpret_cast<const float4*>(p); }

// Buffer loads: a wave-uniform base in a buffer descriptor (scalar
// registers) plus a per-lane byte offset and a scalar byte offset, so the
// per-load address arithmetic is scalar work.  The descriptor covers
// `bytes` from the base (raw buffer, no stride); word 3 is the gfx9 raw
// 32-bit data format.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const float* base, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc((void*)base, (short)0, (int)bytes, 0x00020000);
}
template <int ZV>
__device__ __forceinline__ typename VecT<ZV>::T bload(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff);
template <>
__device__ __forceinline__ float bload<1>(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
  return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0));
}
template <>
__device__ __forceinline__ float2 bload<2>(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
  const auto v = __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, (int)soff, 0);
  return make_float2(__uint_as_float(v[0]), __uint_as_float(v[1]));
}
template <>
__device__ __forceinline__ float4 bload<4>(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
  const auto v = __builtin_amdgcn_raw_buffer_load_b128(r, (int)voff, (int)soff, 0);
  return make_float4(__uint_as_float(v[0]), __uint_as_float(v[1]), __uint_as_float(v[2]), __uint_as_float(v[3]));
}

// Lane z layouts of the wavefront kernels: contiguous (a lane owns ZV
// consecutive z) or split (SPLIT, ZV = 2: a lane owns z and z + 64, so one
// wavefront covers 128 z with dword loads when nz is odd and no wider load
// is aligned).  zb / zb2 are the byte offsets of the lane's (clamped) z.
template <int ZV, bool SPLIT>
__device__ __forceinline__ typename VecT<ZV>::T tload(__amdgpu_buffer_rsrc_t r, uint32_t zb, uint32_t zb2,
                                                      uint32_t soff) {
  if constexpr (SPLIT) {
    static_assert(ZV == 2, "the split layout holds two z values per lane");
    return make_float2(__uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, (int)zb, (int)soff, 0)),
                       __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, (int)zb2, (int)soff, 0)));
  } else {
    return bload<ZV>(r, zb, soff);
  }
}

template <int ZV> __device__ __forceinline__ void vstore(float* p, typename VecT<ZV>::T v);
template <> __device__ __forceinline__ void vstore<1>(float* p, float v) { *p = v; }
template <> __device__ __forceinline__ void vstore<2>(float* p, float2 v) { *reinterpret_cast<float2*>(p) = v; }
template <> __device__ __forceinline__ void vstore<4>(float* p, float4 v) { *reinterpret_cast<float4*>(p) = v; }

__device__ __forceinline__ float vfma(float a, float x, float y) { return __builtin_fmaf(a, x, y); }
__device__ __forceinline__ float4 vfma(float a, float4 x, float4 y) {
  return make_float4(__builtin_fmaf(a, x.x, y.x), __builtin_fmaf(a, x.y, y.y), __builtin_fmaf(a, x.z, y.z),
                     __builtin_fmaf(a, x.w, y.w));
}
__device__ __forceinline__ float2 vfma(float a, float2 x, float2 y) {
  return make_float2(__builtin_fmaf(a, x.x, y.x), __builtin_fmaf(a, x.y, y.y));
}
__device__ __forceinline__ float vmul(float a, float x) { return a * x; }
__device__ __forceinline__ float2 vmul(float a, float2 x) { return make_float2(a * x.x, a * x.y); }
__device__ __forceinline__ float4 vmul(float a, float4 x) { return make_float4(a * x.x, a * x.y, a * x.z, a * x.w); }
__device__ __forceinline__ float vsub(float a, float b) { return a - b; }
__device__ __forceinline__ float4 vsub(float4 a, float4 b) {
  return make_float4(a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w);
}
__device__ __forceinline__ float2 vsub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float vadd(float a, float b) { return a + b; }
__device__ __forceinline__ float2 vadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float4 vadd(float4 a, float4 b) {
  return make_float4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w);
}
__device__ __forceinline__ float clampf(float v, float lo, float hi) { return fminf(fmaxf(v, lo), hi); }
__device__ __forceinline__ float vclamp(float v, float lo, float hi) { return clampf(v, lo, hi); }
__device__ __forceinline__ float2 vclamp(float2 v, float2 lo, float2 hi) {
  return make_float2(clampf(v.x, lo.x, hi.x), clampf(v.y, lo.y, hi.y));
}
__device__ __forceinline__ float4 vclamp(float4 v, float4 lo, float4 hi) {
  return make_float4(clampf(v.x, lo.x, hi.x), clampf(v.y, lo.y, hi.y), clampf(v.z, lo.z, hi.z),
                     clampf(v.w, lo.w, hi.w));
}
__device__ __forceinline__ float vget(float v, int) { return v; }
__device__ __forceinline__ float vget(float2 v, int q) { return q == 0 ? v.x : v.y; }
__device__ __forceinline__ float vget(float4 v, int q) { return q == 0 ? v.x : (q == 1 ? v.y : (q == 2 ? v.z : v.w)); }
__device__ __forceinline__ void vset(float& v, int, float x) { v = x; }
__device__ __forceinline__ void vset(float2& v, int q, float x) {
  if (q == 0) v.x = x; else v.y = x;
}
__device__ __forceinline__ void vset(float4& v, int q, float x) {
  if (q == 0) v.x = x; else if (q == 1) v.y = x; else if (q == 2) v.z = x; else v.w = x;
}

// absorption weight of a sample at lateral coordinate t and along-ray s:
// exp(-mu * max(0, sqrt(R^2 - t^2) - s) / cos(phi)) inside |t| < R, else 0
// (per-ray form of astra3Dabsorption.getExpabs, astra3Dabsorption.py:129-145)
__device__ __forceinline__ float abs_weight(const ParDev& g, double t, double s) {
  if (g.mu == 0.0) return 1.0f;
  if (!(fabs(t) < g.radius)) return 0.0f;
  const double half = sqrt(g.radius * g.radius - t * t);
  const double path = fmax(0.0, half - s) * g.inv_cos;
  return expf((float)(-g.mu * path));
}

// np.interp(t, xp, fp, left=0, right=0) with xp[k] = (k - c) - det_offset:
// value = fp[j] + w (fp[j1] - fp[j]); numpy compares t with the end nodes and
// returns fp[-1] exactly at the last node.
__device__ __forceinline__ bool interp_node(const ParDev& g, double t, int& j, int& j1, float& w) {
  if (!(t >= g.xp0 && t <= g.xp_last)) return false;
  const double u = t - g.xp0;
  int jj = (int)floor(u);
  jj = jj < 0 ? 0 : (jj > g.n - 1 ? g.n - 1 : jj);
  j = jj;
  j1 = jj + 1 < g.n ? jj + 1 : g.n - 1;
  w = (jj == g.n - 1) ? 0.0f : (float)(t - ((double)(jj - g.c) - g.det_offset));
  return true;
}

// weight that a sample at `coord` gives to integer position idx
// (skimage bilinear: floor tap 1-d, ceil tap d; floor == ceil -> 1 + 0)
__device__ __forceinline__ float tap_weight_axis(double coord, int idx) {
  const double f = floor(coord), cl = ceil(coord);
  const double d = coord - f;
  double w = 0.0;
  if ((int)f == idx) w += 1.0 - d;
  if ((int)cl == idx) w += d;
  return (float)w;
}

// ---------------------------------------------------------------------------
// forward projection of ray (k, i) at z .. z+ZV-1, zero inclination
// (Projector3DParallel.py:129-156: warp with matrix R, clip, column sum)
// ---------------------------------------------------------------------------
template <int ZV, bool CLIP>
__device__ __forceinline__ typename VecT<ZV>::T fp_ray(const ParDev& g, const float* __restrict__ vol,
                                                        const float* __restrict__ minmax, int k, int i, int z) {
  using V = typename VecT<ZV>::T;
  const int n = g.n;
  const double* M = g.mats + 6 * i;
  const double M01 = M[1], M02 = M[2], M11 = M[4], M12 = M[5];
  const double kp = (double)k - g.det_offset;
  const double a0 = M[0] * kp, a1 = M[3] * kp;
  const double t = (double)(k - g.c) - g.det_offset;
  V lo = vzero<ZV>(), hi = vzero<ZV>();
  if constexpr (CLIP) {
    lo = vload<ZV>(minmax + z);
    hi = vload<ZV>(minmax + g.nz + z);
  }
  V acc = vzero<ZV>();
  const size_t zstride = (size_t)g.nz;
  // occlusion: slice z+q drops the samples with y = r - c > lim[q]
  // (Projector3DParallel.py:147-153); lim = +inf where no shadow applies
  float lim[ZV];
#pragma unroll
  for (int q = 0; q < ZV; ++q) {
    lim[q] = INFINITY;
    if (g.occ && g.occ_active[(size_t)i * g.nz + z + q]) {
      const float o = g.occ[((size_t)k * g.na + i) * zstride + z + q];
      if (!isnan(o)) lim[q] = o;
    }
  }
  for (int r = 0; r < n; ++r) {
    const double rr = (double)r;
    const double col = (a0 + M01 * rr) + M02;
    const double row = (a1 + M11 * rr) + M12;
    if (!(row > -1.0 && row < (double)n && col > -1.0 && col < (double)n)) continue;
    const double fr = floor(row), fc = floor(col);
    const double dr = row - fr, dc = col - fc;
    const int r0 = (int)fr, c0 = (int)fc;
    const int r1 = (int)ceil(row), c1 = (int)ceil(col);
    const bool vr0 = r0 >= 0, vr1 = r1 < n, vc0 = c0 >= 0, vc1 = c1 < n;
    const float f1dc = (float)(1.0 - dc), fdc = (float)dc;
    const float f1dr = (float)(1.0 - dr), fdr = (float)dr;
    V tl = vzero<ZV>(), tr = vzero<ZV>(), bl = vzero<ZV>(), br = vzero<ZV>();
    if (vr0 && vc0) tl = vload<ZV>(vol + ((size_t)r0 * n + c0) * zstride + z);
    if (vr0 && vc1) tr = vload<ZV>(vol + ((size_t)r0 * n + c1) * zstride + z);
    if (vr1 && vc0) bl = vload<ZV>(vol + ((size_t)r1 * n + c0) * zstride + z);
    if (vr1 && vc1) br = vload<ZV>(vol + ((size_t)r1 * n + c1) * zstride + z);
    const V top = vfma(fdc, tr, vmul(f1dc, tl));
    const V bottom = vfma(fdc, br, vmul(f1dc, bl));
    V v = vfma(fdr, bottom, vmul(f1dr, top));
    if constexpr (CLIP) {
      // skimage keeps exact-cval (0) outputs when 0 lies outside the slice
      // range; an output is exactly 0 iff no in-range tap has positive weight.
      const bool colw = vc0 || (dc > 0.0 && vc1);
      const bool anyin = (vr0 && colw) || (dr > 0.0 && vr1 && colw);
      if (anyin) v = vclamp(v, lo, hi);
    }
    if (g.occ) {
      const float yr = (float)(r - g.c);
#pragma unroll
      for (int q = 0; q < ZV; ++q)
        if (yr > lim[q]) vset(v, q, 0.0f);
    }
    acc = vfma(abs_weight(g, t, (double)(r - g.c)), v, acc);
  }
  return acc;
}

// forward projection with inclination, one detector row j
__device__ __forceinline__ float fp_ray_incl(const ParDev& g, const float* __restrict__ vol, int k, int i, int j) {
  const int n = g.n, nz = g.nz;
  const double* M = g.mats + 6 * i;
  const double kp = (double)k - g.det_offset;
  const double a0 = M[0] * kp, a1 = M[3] * kp;
  const double t = (double)(k - g.c) - g.det_offset;
  float acc = 0.0f;
  for (int r = 0; r < n; ++r) {
    const double rr = (double)r;
    const double col = (a0 + M[1] * rr) + M[2];
    const double row = (a1 + M[4] * rr) + M[5];
    if (!(row > -1.0 && row < (double)n && col > -1.0 && col < (double)n)) continue;
    const double s = (double)(r - g.c);
    const double zpos = (double)j + s * g.tan_phi;
    const double fz = floor(zpos);
    const double wz1 = zpos - fz;
    const int z0 = (int)fz;
    const double fr = floor(row), fc = floor(col);
    const double dr = row - fr, dc = col - fc;
    const int r0 = (int)fr, c0 = (int)fc, r1 = (int)ceil(row), c1 = (int)ceil(col);
    const bool vr0 = r0 >= 0, vr1 = r1 < n, vc0 = c0 >= 0, vc1 = c1 < n;
    const float f1dc = (float)(1.0 - dc), fdc = (float)dc;
    const float f1dr = (float)(1.0 - dr), fdr = (float)dr;
    float v = 0.0f;
    for (int q = 0; q < 2; ++q) {
      const int zz = z0 + q;
      if (zz < 0 || zz >= nz) continue;
      const float wz = (float)(q == 0 ? 1.0 - wz1 : wz1);
      float tl = 0.f, tr = 0.f, bl = 0.f, br = 0.f;
      if (vr0 && vc0) tl = vol[((size_t)r0 * n + c0) * nz + zz];
      if (vr0 && vc1) tr = vol[((size_t)r0 * n + c1) * nz + zz];
      if (vr1 && vc0) bl = vol[((size_t)r1 * n + c0) * nz + zz];
      if (vr1 && vc1) br = vol[((size_t)r1 * n + c1) * nz + zz];
      const float top = __builtin_fmaf(fdc, tr, f1dc * tl);
      const float bottom = __builtin_fmaf(fdc, br, f1dc * bl);
      v = __builtin_fmaf(wz, __builtin_fmaf(fdr, bottom, f1dr * top), v);
    }
    acc = __builtin_fmaf(abs_weight(g, t, s), v, acc);
  }
  return acc;
}

// ---------------------------------------------------------------------------
// back projection at voxel column pix, z .. z+ZV-1 (no circle clip here)
// (Projector3DParallel.py:175-205)
// ---------------------------------------------------------------------------
// Occlusion of a voxel at detector coordinate t, along-ray coordinate s,
// for angle i and slices z..z+ZV-1: the voxel gets nothing from this angle
// when s > floor(np.interp(t, xp, occ column, left=nan, right=nan))
// (Projector3DParallel.py:192-199, 209-216).  np.interp returns fp[j]
// exactly when t equals a node xp[j] (also the last one) and otherwise
// evaluates slope * (t - xp[j]) + fp[j] on the interval containing t (NaN if
// a node is NaN); NaN compares false.  t is inside [xp[0], xp[n-1]] here
// (interp_node accepted it).
template <int ZV>
__device__ __forceinline__ void occ_shadow_mask(const ParDev& g, int i, int j, double t, double s, int z,
                                                typename VecT<ZV>::T& val) {
  const size_t zs = (size_t)g.nz;
  const float* oc = g.occ + (size_t)i * zs + z;
  const size_t step_j = (size_t)g.na * zs;
  // numpy's interval: xp[jn] <= t < xp[jn + 1], found by comparing t with
  // the node values (floor(t - xp[0]) can be off by one when t is within an
  // ulp of a node)
  int jn = j;
  if (jn > 0 && t < (double)(jn - g.c) - g.det_offset) {
    --jn;
  } else if (jn < g.n - 1 && t >= (double)(jn + 1 - g.c) - g.det_offset) {
    ++jn;
  }
#pragma unroll
  for (int q = 0; q < ZV; ++q) {
    const double x0 = (double)(jn - g.c) - g.det_offset;
    const double o0 = (double)oc[(size_t)jn * step_j + q];
    double o = o0;
    if (jn < g.n - 1 && t != x0) {
      const double o1 = (double)oc[(size_t)(jn + 1) * step_j + q];
      const double x1 = (double)(jn + 1 - g.c) - g.det_offset;
      const double slope = (o1 - o0) / (x1 - x0);
      o = slope * (t - x0) + o0;
    }
    if (s > floor(o)) vset(val, q, 0.0f);
  }
}

template <int ZV>
__device__ __forceinline__ typename VecT<ZV>::T bp_voxel(const ParDev& g, const float* __restrict__ sino, int pix,
                                                          int z) {
  using V = typename VecT<ZV>::T;
  const int row = pix / g.n, col = pix % g.n;
  const double x = (double)(col - g.c), y = (double)(row - g.c);
  const size_t zstride = (size_t)g.nz;
  V acc = vzero<ZV>();
  for (int i = 0; i < g.na; ++i) {
    const double co = g.cs[2 * i], sn = g.cs[2 * i + 1];
    const double t = x * co - y * sn;
    int j, j1;
    float w;
    if (!interp_node(g, t, j, j1, w)) continue;
    const V b0 = vload<ZV>(sino + ((size_t)j * g.na + i) * zstride + z);
    const V b1 = vload<ZV>(sino + ((size_t)j1 * g.na + i) * zstride + z);
    V val = vfma(w, vsub(b1, b0), b0);
    if (g.mu != 0.0) val = vmul(abs_weight(g, t, x * sn + y * co), val);
    if (g.occ) occ_shadow_mask<ZV>(g, i, j, t, x * sn + y * co, z, val);
    acc = vadd(acc, val);
  }
  return acc;
}

__device__ __forceinline__ float bp_voxel_incl(const ParDev& g, const float* __restrict__ sino, int pix, int z) {
  const int row = pix / g.n, col = pix % g.n;
  const double x = (double)(col - g.c), y = (double)(row - g.c);
  float acc = 0.0f;
  for (int i = 0; i < g.na; ++i) {
    const double co = g.cs[2 * i], sn = g.cs[2 * i + 1];
    const double t = x * co - y * sn;
    int j, j1;
    float w;
    if (!interp_node(g, t, j, j1, w)) continue;
    const double s = x * sn + y * co;
    const double vpos = (double)z - s * g.tan_phi;
    const double fz = floor(vpos);
    const double wz1 = vpos - fz;
    const int q0 = (int)fz;
    float val = 0.0f;
    for (int q = 0; q < 2; ++q) {
      const int jz = q0 + q;
      if (jz < 0 || jz >= g.nz) continue;
      const float wz = (float)(q == 0 ? 1.0 - wz1 : wz1);
      const float b0 = sino[((size_t)j * g.na + i) * g.nz + jz];
      const float b1 = sino[((size_t)j1 * g.na + i) * g.nz + jz];
      val = __builtin_fmaf(wz, __builtin_fmaf(w, b1 - b0, b0), val);
    }
    acc = __builtin_fmaf(abs_weight(g, t, s), val, acc);
  }
  return acc;
}

// ---------------------------------------------------------------------------
// exact transpose of the forward projection (linear part): voxel-driven
// gather over the <= 4x4 candidate samples (k, r) per angle whose bilinear
// footprint covers the voxel
// ---------------------------------------------------------------------------
template <int ZV>
__device__ __forceinline__ typename VecT<ZV>::T fpT_voxel(const ParDev& g, const float* __restrict__ sino, int pix,
                                                           int z) {
  using V = typename VecT<ZV>::T;
  const int n = g.n;
  const int row = pix / n, col = pix % n;
  const double X = (double)(col - g.c), Y = (double)(row - g.c);
  const size_t zstride = (size_t)g.nz;
  V acc = vzero<ZV>();
  for (int i = 0; i < g.na; ++i) {
    const double* M = g.mats + 6 * i;
    const double co = M[0], sn = M[1];
    const double kstar = co * X - sn * Y + (double)g.c + g.det_offset;
    const double rstar = sn * X + co * Y + (double)g.c;
    const int kb = (int)floor(kstar) - 1, rb = (int)floor(rstar) - 1;
    for (int kk = kb; kk < kb + 4; ++kk) {
      if (kk < 0 || kk >= n) continue;
      const double kp = (double)kk - g.det_offset;
      const double a0 = M[0] * kp, a1 = M[3] * kp;
      const double t = (double)(kk - g.c) - g.det_offset;
      float wk = 0.0f;
      for (int r = rb; r < rb + 4; ++r) {
        if (r < 0 || r >= n) continue;
        const double rr = (double)r;
        const double colS = (a0 + M[1] * rr) + M[2];
        const double rowS = (a1 + M[4] * rr) + M[5];
        if (!(fabs(rowS - (double)row) < 1.0 && fabs(colS - (double)col) < 1.0)) continue;
        const float wr = tap_weight_axis(rowS, row);
        const float wc = tap_weight_axis(colS, col);
        wk = __builtin_fmaf(wr * wc, abs_weight(g, t, (double)(r - g.c)), wk);
      }
      if (wk != 0.0f) acc = vfma(wk, vload<ZV>(sino + ((size_t)kk * g.na + i) * zstride + z), acc);
    }
  }
  return acc;
}

__device__ __forceinline__ float fpT_voxel_incl(const ParDev& g, const float* __restrict__ sino, int pix, int z) {
  const int n = g.n, nz = g.nz;
  const int row = pix / n, col = pix % n;
  const double X = (double)(col - g.c), Y = (double)(row - g.c);
  float acc = 0.0f;
  for (int i = 0; i < g.na; ++i) {
    const double* M = g.mats + 6 * i;
    const double co = M[0], sn = M[1];
    const double kstar = co * X - sn * Y + (double)g.c + g.det_offset;
    const double rstar = sn * X + co * Y + (double)g.c;
    const int kb = (int)floor(kstar) - 1, rb = (int)floor(rstar) - 1;
    for (int kk = kb; kk < kb + 4; ++kk) {
      if (kk < 0 || kk >= n) continue;
      const double kp = (double)kk - g.det_offset;
      const double a0 = M[0] * kp, a1 = M[3] * kp;
      const double t = (double)(kk - g.c) - g.det_offset;
      for (int r = rb; r < rb + 4; ++r) {
        if (r < 0 || r >= n) continue;
        const double rr = (double)r;
        const double colS = (a0 + M[1] * rr) + M[2];
        const double rowS = (a1 + M[4] * rr) + M[5];
        if (!(fabs(rowS - (double)row) < 1.0 && fabs(colS - (double)col) < 1.0)) continue;
        const float wxy = tap_weight_axis(rowS, row) * tap_weight_axis(colS, col);
        if (wxy == 0.0f) continue;
        const double s = (double)(r - g.c);
        const float wa = abs_weight(g, t, s) * wxy;
        // detector rows j whose sample z position j + s tan(phi) has a tap on z
        const int jc = (int)floor((double)z - s * g.tan_phi);
        for (int j = jc - 1; j <= jc + 2; ++j) {
          if (j < 0 || j >= nz) continue;
          const float wz = tap_weight_axis((double)j + s * g.tan_phi, z);
          if (wz == 0.0f) continue;
          acc = __builtin_fmaf(wa * wz, sino[((size_t)kk * g.na + i) * nz + j], acc);
        }
      }
    }
  }
  return acc;
}

// ---------------------------------------------------------------------------
// exact transpose of the back projection: sinogram-driven gather over the
// voxels whose detector coordinate lies within one bin of bin k
// ---------------------------------------------------------------------------
__device__ __forceinline__ void bpT_range(const ParDev& g, double co, double sn, double xk, int a, bool rows_major,
                                          int& b0, int& b1) {
  double lo, hi;
  if (rows_major) {
    const double Y = (double)(a - g.c);
    const double e0 = (xk - 1.0 + Y * sn) / co, e1 = (xk + 1.0 + Y * sn) / co;
    lo = fmin(e0, e1);
    hi = fmax(e0, e1);
  } else {
    const double X = (double)(a - g.c);
    const double e0 = (X * co - (xk - 1.0)) / sn, e1 = (X * co - (xk + 1.0)) / sn;
    lo = fmin(e0, e1);
    hi = fmax(e0, e1);
  }
  b0 = (int)floor(lo) - 1 + g.c;
  b1 = (int)ceil(hi) + 1 + g.c;
  if (b0 < 0) b0 = 0;
  if (b1 > g.n - 1) b1 = g.n - 1;
}

template <int ZV>
__device__ __forceinline__ typename VecT<ZV>::T bpT_bin(const ParDev& g, const float* __restrict__ vol, int k, int i,
                                                         int z, int clipc) {
  using V = typename VecT<ZV>::T;
  const int n = g.n;
  const size_t zstride = (size_t)g.nz;
  const double co = g.cs[2 * i], sn = g.cs[2 * i + 1];
  const double xk = (double)(k - g.c) - g.det_offset;
  const bool rows_major = fabs(co) >= fabs(sn);
  V acc = vzero<ZV>();
  for (int a = 0; a < n; ++a) {
    int b0, b1;
    bpT_range(g, co, sn, xk, a, rows_major, b0, b1);
    for (int bb = b0; bb <= b1; ++bb) {
      const int row = rows_major ? a : bb;
      const int col = rows_major ? bb : a;
      const int pix = row * n + col;
      if (clipc && g.outside[pix]) continue;
      const double x = (double)(col - g.c), y = (double)(row - g.c);
      const double t = x * co - y * sn;
      int j, j1;
      float w;
      if (!interp_node(g, t, j, j1, w)) continue;
      float wk = 0.0f;
      if (j == k) wk += 1.0f - w;
      if (j1 == k) wk += w;
      if (wk == 0.0f) continue;
      if (g.mu != 0.0) wk *= abs_weight(g, t, x * sn + y * co);
      acc = vfma(wk, vload<ZV>(vol + (size_t)pix * zstride + z), acc);
    }
  }
  return acc;
}

__device__ __forceinline__ float bpT_bin_incl(const ParDev& g, const float* __restrict__ vol, int k, int i, int jz,
                                              int clipc) {
  const int n = g.n, nz = g.nz;
  const double co = g.cs[2 * i], sn = g.cs[2 * i + 1];
  const double xk = (double)(k - g.c) - g.det_offset;
  const bool rows_major = fabs(co) >= fabs(sn);
  float acc = 0.0f;
  for (int a = 0; a < n; ++a) {
    int b0, b1;
    bpT_range(g, co, sn, xk, a, rows_major, b0, b1);
    for (int bb = b0; bb <= b1; ++bb) {
      const int row = rows_major ? a : bb;
      const int col = rows_major ? bb : a;
      const int pix = row * n + col;
      if (clipc && g.outside[pix]) continue;
      const double x = (double)(col - g.c), y = (double)(row - g.c);
      const double t = x * co - y * sn;
      int j, j1;
      float w;
      if (!interp_node(g, t, j, j1, w)) continue;
      float wk = 0.0f;
      if (j == k) wk += 1.0f - w;
      if (j1 == k) wk += w;
      if (wk == 0.0f) continue;
      const double s = x * sn + y * co;
      wk *= abs_weight(g, t, s);
      // voxels z whose detector-row position z - s tan(phi) has a tap on jz
      const int zc = (int)floor((double)jz + s * g.tan_phi);
      for (int z = zc - 1; z <= zc + 2; ++z) {
        if (z < 0 || z >= nz) continue;
        const float wz = tap_weight_axis((double)z - s * g.tan_phi, jz);
        if (wz == 0.0f) continue;
        acc = __builtin_fmaf(wk * wz, vol[(size_t)pix * nz + z], acc);
      }
    }
  }
  return acc;
}


// ---------------------------------------------------------------------------
// wavefront-uniform cores
//
// When a wavefront's 64 lanes cover 64*ZV consecutive z values of ONE ray (or
// voxel column), the geometry of a sample (or angle) is the same for every
// lane.  Lane l then evaluates the geometry of sample r0 + l (or angle
// i0 + l) -- 64 samples in parallel, float64, the reference's operation order
// -- and the inner loop walks the samples with an in-range tap (a ballot
// mask), broadcasting each sample's tap base and weights with readlane into
// scalar registers: per sample the vector units only issue the tap loads and
// the fused multiply-adds.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int rl_i(int v, int lane) { return __builtin_amdgcn_readlane(v, lane); }

// ---------------------------------------------------------------------------
// Exact transposes, wavefront-uniform (zero inclination).  A wavefront owns
// one detector bin (B^T) or one voxel column (A^T) and 64*ZV consecutive z.
// Lane l evaluates, in float64 and with exactly the per-lane cores'
// arithmetic, the taps of row a0 + l of the bin's strip (B^T) or of angle
// i0 + l (A^T): at most four (voxel or detector bin, weight) pairs with a
// nonzero weight, in the per-lane cores' order.  They pass through a
// per-wavefront LDS buffer; the wave then walks the rows (angles) in order
// with one coalesced load per tap, so each z sums the same terms in the same
// order as bpT_bin / fpT_voxel: the results are bitwise equal.
// ---------------------------------------------------------------------------
struct AdjSlot {
  float4* w;   // [64] tap weights (0 = no tap)
  int4* o;     // [64] tap voxel column (B^T) or sinogram row (A^T), -1 = none
};
__device__ __forceinline__ AdjSlot adj_lds_slot() {
  __shared__ float4 w[4][64];
  __shared__ int4 o[4][64];
  const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)) & 3;
  return {w[wv], o[wv]};
}

__device__ __forceinline__ void adj_put(float4& w, int4& o, int m, float wk, int off) {
  if (m == 0) { w.x = wk; o.x = off; }
  else if (m == 1) { w.y = wk; o.y = off; }
  else if (m == 2) { w.z = wk; o.z = off; }
  else { w.w = wk; o.w = off; }
}

__device__ __forceinline__ void adj_putw(float4& w, int m, float wk) {
  if (m == 0) w.x = wk;
  else if (m == 1) w.y = wk;
  else if (m == 2) w.z = wk;
  else w.w = wk;
}

template <int ZV>
__device__ __forceinline__ typename VecT<ZV>::T adj_taps(const float* __restrict__ base, size_t stride, int zl,
                                                         const AdjSlot& sl, int rows, typename VecT<ZV>::T acc) {
  for (int r = 0; r < rows; ++r) {
    const float4 w = sl.w[r];
    const int4 o = sl.o[r];
    const int o0 = __builtin_amdgcn_readfirstlane(o.x), o1 = __builtin_amdgcn_readfirstlane(o.y);
    const int o2 = __builtin_amdgcn_readfirstlane(o.z), o3 = __builtin_amdgcn_readfirstlane(o.w);
    if (o0 >= 0) acc = vfma(w.x, vload<ZV>(base + (size_t)o0 * stride + zl), acc);
    if (o1 >= 0) acc = vfma(w.y, vload<ZV>(base + (size_t)o1 * stride + zl), acc);
    if (o2 >= 0) acc = vfma(w.z, vload<ZV>(base + (size_t)o2 * stride + zl), acc);
    if (o3 >= 0) acc = vfma(w.w, vload<ZV>(base + (size_t)o3 * stride + zl), acc);
  }
  return acc;
}

// B^T of bin (k, i): the voxels of each row (column, when |sin| > |cos|) of
// the strip |t - t_k| < 1, rows in order (bpT_bin)
template <int ZV>
__device__ __forceinline__ typename VecT<ZV>::T bpT_bin_wave(const ParDev& g, const float* __restrict__ vol, int k,
                                                              int i, int zl, int lane, int clipc) {
  using V = typename VecT<ZV>::T;
  const int n = g.n;
  const double co = g.cs[2 * i], sn = g.cs[2 * i + 1];
  const double xk = (double)(k - g.c) - g.det_offset;
  const bool rows_major = fabs(co) >= fabs(sn);
  const AdjSlot sl = adj_lds_slot();
  V acc = vzero<ZV>();
  for (int a0 = 0; a0 < n; a0 += 64) {
    const int a = a0 + lane;
    float4 w = make_float4(0.f, 0.f, 0.f, 0.f);
    int4 o = make_int4(-1, -1, -1, -1);
    if (a < n) {
      int b0, b1;
      bpT_range(g, co, sn, xk, a, rows_major, b0, b1);
      int m = 0;
      for (int bb = b0; bb <= b1 && m < 4; ++bb) {
        const int row = rows_major ? a : bb;
        const int col = rows_major ? bb : a;
        const int pix = row * n + col;
        if (clipc && g.outside[pix]) continue;
        const double x = (double)(col - g.c), y = (double)(row - g.c);
        const double t = x * co - y * sn;
        int j, j1;
        float wj;
        if (!interp_node(g, t, j, j1, wj)) continue;
        float wk = 0.0f;
        if (j == k) wk += 1.0f - wj;
        if (j1 == k) wk += wj;
        if (wk == 0.0f) continue;
        if (g.mu != 0.0) wk *= abs_weight(g, t, x * sn + y * co);
        adj_put(w, o, m, wk, pix);
        ++m;
      }
    }
    __builtin_amdgcn_wave_barrier();
    sl.w[lane] = w;
    sl.o[lane] = o;
    __builtin_amdgcn_wave_barrier();
    acc = adj_taps<ZV>(vol, (size_t)g.nz, zl, sl, n - a0 < 64 ? n - a0 : 64, acc);
  }
  return acc;
}

// weight of voxel (x, y) in B^T of bin k at the angle (co, sn): bpT_bin's
// arithmetic, 0 when the voxel is outside the bin's strip
__device__ __forceinline__ float bpT_weight(const ParDev& g, double co, double sn, int k, double x, double y) {
  const double t = x * co - y * sn;
  int j, j1;
  float wj;
  if (!interp_node(g, t, j, j1, wj)) return 0.0f;
  float wk = 0.0f;
  if (j == k) wk += 1.0f - wj;
  if (j1 == k) wk += wj;
  if (wk == 0.0f) return 0.0f;
  if (g.mu != 0.0) wk *= abs_weight(g, t, x * sn + y * co);
  return wk;
}

// B^T of bin (k, i) and of the opposite bin (k2, i2): angle i2 is angle
// i + 180 deg and k2 = 2c + 2 det_offset - k, so t2 = -t and the two strips
// hold the same voxels.  Each voxel column of the strip is loaded once and
// feeds both sums.  The partner's weights use its own angle's cos / sin; both
// sums walk the primary's rows (columns) in order, which is the partner's own
// order whenever the two angles agree on the row direction (all but exact
// 45-degree multiples), so each sum is bpT_bin_wave's to the bit there.  Per
// row the two strips have at most 3 voxels together (|t - t_k| < 1 on a row
// where t moves >= 1/sqrt(2) per voxel; the partner's strip differs by the
// pairing tolerance only), within the 4 slots.
struct AdjSlot2 {
  float4* w;   // [64] primary tap weights
  float4* w2;  // [64] partner tap weights
  int4* o;     // [64] tap voxel column, -1 = none
};
__device__ __forceinline__ AdjSlot2 adj_lds_slot2() {
  __shared__ float4 w[4][64];
  __shared__ float4 w2[4][64];
  __shared__ int4 o[4][64];
  const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)) & 3;
  return {w[wv], w2[wv], o[wv]};
}

template <int ZV>
__device__ __forceinline__ void bpT_pair_wave(const ParDev& g, const float* __restrict__ vol, int k, int i, int k2,
                                              int i2, int zl, int lane, int clipc, typename VecT<ZV>::T& acc,
                                              typename VecT<ZV>::T& acc2) {
  const int n = g.n;
  const size_t stride = (size_t)g.nz;
  const double co = g.cs[2 * i], sn = g.cs[2 * i + 1];
  const double co2 = g.cs[2 * i2], sn2 = g.cs[2 * i2 + 1];
  const double xk = (double)(k - g.c) - g.det_offset;
  const bool rows_major = fabs(co) >= fabs(sn);
  const AdjSlot2 sl = adj_lds_slot2();
  acc = vzero<ZV>();
  acc2 = vzero<ZV>();
  for (int a0 = 0; a0 < n; a0 += 64) {
    const int a = a0 + lane;
    float4 w = make_float4(0.f, 0.f, 0.f, 0.f), w2 = w;
    int4 o = make_int4(-1, -1, -1, -1);
    if (a < n) {
      int b0, b1;
      bpT_range(g, co, sn, xk, a, rows_major, b0, b1);
      int m = 0;
      for (int bb = b0; bb <= b1 && m < 4; ++bb) {
        const int row = rows_major ? a : bb;
        const int col = rows_major ? bb : a;
        const int pix = row * n + col;
        if (clipc && g.outside[pix]) continue;
        const double x = (double)(col - g.c), y = (double)(row - g.c);
        const float wk = bpT_weight(g, co, sn, k, x, y);
        const float wk2 = bpT_weight(g, co2, sn2, k2, x, y);
        if (wk == 0.0f && wk2 == 0.0f) continue;
        adj_put(w, o, m, wk, pix);
        adj_putw(w2, m, wk2);
        ++m;
      }
    }
    __builtin_amdgcn_wave_barrier();
    sl.w[lane] = w;
    sl.w2[lane] = w2;
    sl.o[lane] = o;
    __builtin_amdgcn_wave_barrier();
    const int rows = n - a0 < 64 ? n - a0 : 64;
    for (int r = 0; r < rows; ++r) {
      const float4 wr = sl.w[r], wr2 = sl.w2[r];
      const int4 orr = sl.o[r];
      const int o0 = __builtin_amdgcn_readfirstlane(orr.x), o1 = __builtin_amdgcn_readfirstlane(orr.y);
      const int o2 = __builtin_amdgcn_readfirstlane(orr.z), o3 = __builtin_amdgcn_readfirstlane(orr.w);
      if (o0 >= 0) {
        const typename VecT<ZV>::T v = vload<ZV>(vol + (size_t)o0 * stride + zl);
        acc = vfma(wr.x, v, acc);
        acc2 = vfma(wr2.x, v, acc2);
      }
      if (o1 >= 0) {
        const typename VecT<ZV>::T v = vload<ZV>(vol + (size_t)o1 * stride + zl);
        acc = vfma(wr.y, v, acc);
        acc2 = vfma(wr2.y, v, acc2);
      }
      if (o2 >= 0) {
        const typename VecT<ZV>::T v = vload<ZV>(vol + (size_t)o2 * stride + zl);
        acc = vfma(wr.z, v, acc);
        acc2 = vfma(wr2.z, v, acc2);
      }
      if (o3 >= 0) {
        const typename VecT<ZV>::T v = vload<ZV>(vol + (size_t)o3 * stride + zl);
        acc = vfma(wr.w, v, acc);
        acc2 = vfma(wr2.w, v, acc2);
      }
    }
  }
}

// A^T of voxel column pix: the detector bins of each angle whose samples
// tap the voxel, angles in order (fpT_voxel)
template <int ZV>
__device__ __forceinline__ typename VecT<ZV>::T fpT_voxel_wave(const ParDev& g, const float* __restrict__ sino,
                                                                int pix, int zl, int lane) {
  using V = typename VecT<ZV>::T;
  const int n = g.n;
  const int row = pix / n, col = pix % n;
  const double X = (double)(col - g.c), Y = (double)(row - g.c);
  const AdjSlot sl = adj_lds_slot();
  V acc = vzero<ZV>();
  for (int i0 = 0; i0 < g.na; i0 += 64) {
    const int i = i0 + lane;
    float4 w = make_float4(0.f, 0.f, 0.f, 0.f);
    int4 o = make_int4(-1, -1, -1, -1);
    if (i < g.na) {
      const double* M = g.mats + 6 * i;
      const double co = M[0], sn = M[1];
      const double kstar = co * X - sn * Y + (double)g.c + g.det_offset;
      const double rstar = sn * X + co * Y + (double)g.c;
      const int kb = (int)floor(kstar) - 1, rb = (int)floor(rstar) - 1;
      int m = 0;
      for (int kk = kb; kk < kb + 4; ++kk) {
        if (kk < 0 || kk >= n) continue;
        const double kp = (double)kk - g.det_offset;
        const double a0 = M[0] * kp, a1 = M[3] * kp;
        const double t = (double)(kk - g.c) - g.det_offset;
        float wk = 0.0f;
        for (int r = rb; r < rb + 4; ++r) {
          if (r < 0 || r >= n) continue;
          const double rr = (double)r;
          const double colS = (a0 + M[1] * rr) + M[2];
          const double rowS = (a1 + M[4] * rr) + M[5];
          if (!(fabs(rowS - (double)row) < 1.0 && fabs(colS - (double)col) < 1.0)) continue;
          const float wr = tap_weight_axis(rowS, row);
          const float wc = tap_weight_axis(colS, col);
          wk = __builtin_fmaf(wr * wc, abs_weight(g, t, (double)(r - g.c)), wk);
        }
        if (wk != 0.0f) {
          adj_put(w, o, m, wk, kk * g.na + i);
          ++m;
        }
      }
    }
    __builtin_amdgcn_wave_barrier();
    sl.w[lane] = w;
    sl.o[lane] = o;
    __builtin_amdgcn_wave_barrier();
    acc = adj_taps<ZV>(sino, (size_t)g.nz, zl, sl, g.na - i0 < 64 ? g.na - i0 : 64, acc);
  }
  return acc;
}
__device__ __forceinline__ float rl_f(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// slot weights of one axis: the taps floor / floor+1 (skimage's floor / ceil
// with weights 1-d / d) mapped onto the pair (b, b+1) with b in [0, n-2]
// The warp clip of a slice is a no-op when 0 lies in [min, max] of the slice:
// bilinear samples are convex combinations of the slice's values and the zero
// padding.  Wave-uniform test over the wave's 64*ZV slices.
// z of the loads of a lane whose first z is z: lanes past nz load a copy of
// the last valid z and store nothing
template <int ZV, bool SPLIT>
__device__ __forceinline__ void wave_zload(int z, int nz, int& zl, int& zl2) {
  if constexpr (SPLIT) {
    zl = z < nz - 1 ? z : nz - 1;
    zl2 = z + 64 < nz - 1 ? z + 64 : nz - 1;
  } else {
    zl = z < nz - ZV ? z : nz - ZV;
    zl2 = 0;
  }
}
// store of a lane's z values into the z column at col
template <int ZV, bool SPLIT>
__device__ __forceinline__ void wave_zstore(float* col, int z, int nz, typename VecT<ZV>::T v) {
  if constexpr (SPLIT) {
    if (z < nz) col[z] = v.x;
    if (z + 64 < nz) col[z + 64] = v.y;
  } else {
    if (z < nz) vstore<ZV>(col + z, v);
  }
}

// the slice range of the lane's z values (contiguous, or z and z2 split)
template <int ZV, bool SPLIT = false>
__device__ __forceinline__ typename VecT<ZV>::T zvals(const float* __restrict__ a, int z, int z2) {
  if constexpr (SPLIT) return make_float2(a[z], a[z2]);
  else return vload<ZV>(a + z);
}
template <int ZV, bool SPLIT = false>
__device__ __forceinline__ bool wave_clip_active(const float* __restrict__ minmax, int nz, int z, int z2 = 0) {
  const typename VecT<ZV>::T lo = zvals<ZV, SPLIT>(minmax, z, z2), hi = zvals<ZV, SPLIT>(minmax + nz, z, z2);
  bool need = false;
#pragma unroll
  for (int q = 0; q < ZV; ++q) need |= !(vget(lo, q) <= 0.0f && 0.0f <= vget(hi, q));
  return __ballot(need) != 0ull;
}

__device__ __forceinline__ void axis_slots(double coord, int n, int& b, double& s0, double& s1) {
  const double f = floor(coord);
  const double d = coord - f;
  const int fi = (int)f;
  if (fi < 0) {          // floor = -1: only floor+1 = 0 is a tap
    b = 0; s0 = d; s1 = 0.0;
  } else if (fi >= n - 1) {  // floor = n-1: floor+1 = n is outside
    b = n - 2; s0 = 0.0; s1 = 1.0 - d;
  } else {
    b = fi; s0 = 1.0 - d; s1 = d;
  }
}

// forward projection of ray (k, i), lanes = z; requires n >= 2.  REUSE keeps
// the previous sample's four taps in registers: between consecutive samples
// the floor corner b moves by at most one pixel per axis, and when it moves
// along one axis only, two of the four taps are the previous sample's.  b is
// wave-uniform, so the case test is a scalar branch.  Lanes evaluate the
// geometry of 64 samples at once; the per-sample loop takes the cell from
// the sample's lane and its weights from the LDS buffer (fp_lds_slot).
//
// PAIR also projects the opposite ray, k2 = 2c + D - k at angle i + 180 deg,
// into *acc2: its sample r2 lies on sample r = 2c - r2 of ray k (rotating by
// 180 deg negates the matrix's linear part, and D = 2 det_offset maps the
// detector coordinate t to -t), so the taps are loaded once and weighted
// twice.  Ray k's samples are r = 0..n-1, the partner's r = 2c-n+1..2c; the
// loop runs over their union, each ray weighting only its own samples (the
// other's extra end sample adds an exact zero).  Each sample is interpolated
// once and multiplied by each ray's absorption weight (unpaired, the weight
// folds into the taps' weights: the sums agree to float32 rounding).  The
// partner's samples are summed in reverse order, at positions computed from
// angle i's matrix -- equal to the partner's own to ~1e-13 pixel, which moves
// the result by as little since the unclipped sample is continuous in its
// position.  The warp clip is not: a sample a
// rounding error inside the -1 or n border clips up to the slice minimum
// while one on the border stays 0, so PAIR excludes CLIP and clipped waves
// project the two rays separately.
// Per-wavefront LDS buffers through which the wavefront cores hand each
// lane's sample weights to the whole wavefront: a broadcast LDS read per
// sample replaces one cross-lane move per weight.  One buffer per kernel
// (these functions are not templates), for blocks of at most 4 wavefronts.
struct FpSlot {
  float4* w;   // [64] the four bilinear tap weights of each lane's sample
  float2* a;   // [64] its absorption weights (ray, partner ray)
};
__device__ __forceinline__ FpSlot fp_lds_slot() {
  __shared__ float4 w[4][64];
  __shared__ float2 a[4][64];
  const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)) & 3;
  return {w[wv], a[wv]};
}
__device__ __forceinline__ float4* bp_lds_slot() {
  __shared__ float4 w[4][2][16];   // weights of the lower / upper row, 16 angles x 4 pixels
  return w[__builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)) & 3][0];
}

template <int ZV, bool CLIP, bool REUSE, bool PAIR = false, bool SPLIT = false>
__device__ __forceinline__ typename VecT<ZV>::T fp_ray_wave(const ParDev& g, const float* __restrict__ vol,
                                                             const float* __restrict__ minmax, int k, int i, int z,
                                                             int lane, typename VecT<ZV>::T* acc2 = nullptr,
                                                             int z2 = 0) {
  using V = typename VecT<ZV>::T;
  static_assert(!(PAIR && CLIP), "the clipped sample is discontinuous at the border: no pairing");
  const int n = g.n;
  const double* M = g.mats + 6 * i;
  const double M01 = M[1], M02 = M[2], M11 = M[4], M12 = M[5];
  const double kp = (double)k - g.det_offset;
  const double a0 = M[0] * kp, a1 = M[3] * kp;
  const double t = (double)(k - g.c) - g.det_offset;
  V lo = vzero<ZV>(), hi = vzero<ZV>();
  if constexpr (CLIP) {
    lo = zvals<ZV, SPLIT>(minmax, z, z2);
    hi = zvals<ZV, SPLIT>(minmax + g.nz, z, z2);
  }
  const size_t nzs = (size_t)g.nz;
  const size_t step_c = nzs, step_r = (size_t)n * nzs;
  // taps are buffer loads from a per-sample descriptor at the cell's corner
  // voxel column: lane byte offset 4z, scalar byte offsets of the other
  // three corners; the descriptor spans the two rows of the cell
  const uint32_t zb = 4u * (uint32_t)z, zb2 = 4u * (uint32_t)z2;
  const uint32_t oc = (uint32_t)(4 * step_c), orr = (uint32_t)(4 * step_r), orc = oc + orr;
  const uint32_t span = orc + oc;
  V acc = vzero<ZV>(), accp = vzero<ZV>();
  // Tap registers by absolute parity: T<a><b> holds the tap of the current
  // cell whose volume row is even (a = 0) or odd (a = 1) and whose column is
  // even (b = 0) or odd (b = 1).  When the cell moves by one row or column
  // the taps it keeps stay where they are and the new ones replace exactly
  // the ones that left (same parity), so the window never shifts between
  // registers; each lane delivers its weights in the same parity order.
  V T00 = vzero<ZV>(), T01 = vzero<ZV>(), T10 = vzero<ZV>(), T11 = vzero<ZV>();
  const FpSlot sl = fp_lds_slot();
  int prev = -2 * n - 2;   // no taps held
  const int r_end = PAIR ? (2 * g.c + 1 > n ? 2 * g.c + 1 : n) : n;
  const int r2_lo = 2 * g.c - n + 1;   // partner's first sample (PAIR)
  for (int r0 = 0; r0 < r_end; r0 += 64) {
    const int r = r0 + lane;
    int base = 0;
    float w00 = 0.f, w01 = 0.f, w10 = 0.f, w11 = 0.f, aw = 0.f;
    float aw2 = 0.f;
    bool valid = false, anyin = false;
    if (r < r_end) {
      const double rr = (double)r;
      const double col = (a0 + M01 * rr) + M02;
      const double row = (a1 + M11 * rr) + M12;
      if (row > -1.0 && row < (double)n && col > -1.0 && col < (double)n) {
        int br, bc;
        double r0w, r1w, c0w, c1w;
        axis_slots(row, n, br, r0w, r1w);
        axis_slots(col, n, bc, c0w, c1w);
        base = (br * n + bc) | ((br & 1) << 29) | ((bc & 1) << 28);   // cell and its row / column parity
        anyin = (r0w > 0.0 || r1w > 0.0) && (c0w > 0.0 || c1w > 0.0);
        if (!PAIR || r < n) aw = abs_weight(g, t, (double)(r - g.c));
        if constexpr (PAIR) {
          if (r >= r2_lo) aw2 = abs_weight(g, -t, (double)(g.c - r));
        }
        valid = anyin && (CLIP || aw != 0.0f || aw2 != 0.0f);
        // unpaired without the clip, the absorption folds into the weights;
        // paired, the sample is interpolated once and weighted per ray
        const float f = (CLIP || PAIR) ? 1.0f : aw;
        const float k00 = f * (float)(r0w * c0w), k01 = f * (float)(r0w * c1w);
        const float k10 = f * (float)(r1w * c0w), k11 = f * (float)(r1w * c1w);
        // parity (a, b) holds corner (a ^ row parity, b ^ column parity)
        const bool pr = br & 1, pc = bc & 1;
        const float e0 = pc ? k01 : k00, e1 = pc ? k00 : k01;   // row 0 by column parity
        const float f0 = pc ? k11 : k10, f1 = pc ? k10 : k11;   // row 1
        w00 = pr ? f0 : e0;
        w01 = pr ? f1 : e1;
        w10 = pr ? e0 : f0;
        w11 = pr ? e1 : f1;
      }
    }
    unsigned long long m = __ballot(valid);
    // the wave's previous reads of the buffers precede these writes (LDS
    // operations of a wavefront complete in order)
    __builtin_amdgcn_wave_barrier();
    sl.w[lane] = make_float4(w00, w01, w10, w11);
    sl.a[lane] = make_float2(aw, aw2);
    __builtin_amdgcn_wave_barrier();
    while (m) {
      const int j = __builtin_ctzll(m);
      m &= m - 1;
      const int bp = rl_i(base, j);
      const int b = bp & 0x0fffffff, par = (bp >> 28) & 3;   // par: row parity << 1 | column parity
      const float4 Q = sl.w[j];
      const float q00 = Q.x, q01 = Q.y, q10 = Q.z, q11 = Q.w;
      const __amdgpu_buffer_rsrc_t rs = buf_rsrc(vol + (size_t)b * nzs, span);
      // C00..C11 name the registers of the cell's corners (0,0) .. (1,1) for
      // this parity; each arm loads the corners the previous cell lacked
#define VAMX_TAP_ARMS(C00, C01, C10, C11, P)                                       \
  if (d == 1) {             /* one column right: the right column is new */      \
    asm volatile("; vamx fp " #P " d1");                                        \
    C01 = tload<ZV, SPLIT>(rs, zb, zb2, oc);                                                 \
    C11 = tload<ZV, SPLIT>(rs, zb, zb2, orc);                                                \
  } else if (d == n) {      /* one row down: the bottom row is new */           \
    asm volatile("; vamx fp " #P " dn");                                        \
    C10 = tload<ZV, SPLIT>(rs, zb, zb2, orr);                                                \
    C11 = tload<ZV, SPLIT>(rs, zb, zb2, orc);                                                \
  } else if (d == -1) {                                                         \
    asm volatile("; vamx fp " #P " dm1");                                       \
    C00 = tload<ZV, SPLIT>(rs, zb, zb2, 0);                                                  \
    C10 = tload<ZV, SPLIT>(rs, zb, zb2, orr);                                                \
  } else if (d == -n) {                                                         \
    asm volatile("; vamx fp " #P " dmn");                                       \
    C00 = tload<ZV, SPLIT>(rs, zb, zb2, 0);                                                  \
    C01 = tload<ZV, SPLIT>(rs, zb, zb2, oc);                                                 \
  } else if (d == n + 1) {  /* diagonal steps keep one tap */                   \
    asm volatile("; vamx fp " #P " dnp1");                                      \
    C01 = tload<ZV, SPLIT>(rs, zb, zb2, oc);                                                 \
    C10 = tload<ZV, SPLIT>(rs, zb, zb2, orr);                                                \
    C11 = tload<ZV, SPLIT>(rs, zb, zb2, orc);                                                \
  } else if (d == n - 1) {                                                      \
    asm volatile("; vamx fp " #P " dnm1");                                      \
    C00 = tload<ZV, SPLIT>(rs, zb, zb2, 0);                                                  \
    C10 = tload<ZV, SPLIT>(rs, zb, zb2, orr);                                                \
    C11 = tload<ZV, SPLIT>(rs, zb, zb2, orc);                                                \
  } else if (d == -n + 1) {                                                     \
    asm volatile("; vamx fp " #P " dmnp1");                                     \
    C00 = tload<ZV, SPLIT>(rs, zb, zb2, 0);                                                  \
    C01 = tload<ZV, SPLIT>(rs, zb, zb2, oc);                                                 \
    C11 = tload<ZV, SPLIT>(rs, zb, zb2, orc);                                                \
  } else if (d == -n - 1) {                                                     \
    asm volatile("; vamx fp " #P " dmnm1");                                     \
    C00 = tload<ZV, SPLIT>(rs, zb, zb2, 0);                                                  \
    C01 = tload<ZV, SPLIT>(rs, zb, zb2, oc);                                                 \
    C10 = tload<ZV, SPLIT>(rs, zb, zb2, orr);                                                \
  } else if (d != 0) {                                                          \
    asm volatile("; vamx fp " #P " full");                                      \
    C00 = tload<ZV, SPLIT>(rs, zb, zb2, 0);                                                  \
    C01 = tload<ZV, SPLIT>(rs, zb, zb2, oc);                                                 \
    C10 = tload<ZV, SPLIT>(rs, zb, zb2, orr);                                                \
    C11 = tload<ZV, SPLIT>(rs, zb, zb2, orc);                                                \
  }
      const int d = REUSE ? b - prev : 2 * n + 2;   // without reuse every sample reloads
      prev = b;
      switch (par) {
        case 0: { VAMX_TAP_ARMS(T00, T01, T10, T11, 0) break; }
        case 1: { VAMX_TAP_ARMS(T01, T00, T11, T10, 1) break; }
        case 2: { VAMX_TAP_ARMS(T10, T11, T00, T01, 2) break; }
        default: { VAMX_TAP_ARMS(T11, T10, T01, T00, 3) break; }
      }
#undef VAMX_TAP_ARMS
      if constexpr (CLIP) {
        V v = vfma(q11, T11, vfma(q10, T10, vfma(q01, T01, vmul(q00, T00))));
        v = vclamp(v, lo, hi);  // every enumerated sample has an in-range tap of positive weight
        acc = vfma(sl.a[j].x, v, acc);
      } else if constexpr (PAIR) {
        const V v = vfma(q11, T11, vfma(q10, T10, vfma(q01, T01, vmul(q00, T00))));
        const float2 A = sl.a[j];
        acc = vfma(A.x, v, acc);
        accp = vfma(A.y, v, accp);
      } else {
        acc = vfma(q11, T11, vfma(q10, T10, vfma(q01, T01, vfma(q00, T00, acc))));
      }
    }
  }
  if constexpr (PAIR) *acc2 = accp;
  return acc;
}

// Back projection of a 2x2 quad of voxel columns, lanes = z.  Pixel p
// (0..3) is at row r0 + (p >> 1), column c0 + (p & 1); bit p of `active`
// says whether it is computed (in range and not circle-clipped).  Within one
// angle the four pixels' detector coordinates t differ by at most
// |cos| + |sin| <= sqrt(2), so their lower np.interp nodes span at most two
// detector rows and the quad needs at most four distinct rows (usually
// three) where four separate columns would load eight.  Lanes first evaluate
// the float64 geometry of 16 angles x 4 pixels; the per-angle loop then
// loads the rows [bmin, bmax + 1] once and every pixel accumulates
// fma(w1, b[j+1], fma(w0, b[j], acc)) in angle order -- the same operations,
// in the same order, as one column at a time.  The 16 angles' weights pass
// through a per-wavefront LDS buffer and come back as one broadcast 16-byte
// read per row weight, and each angle's row offsets, span and first row are
// packed into one value: the per-angle loop moves one value from a lane.
// Requires n >= 2, all 64 lanes of the wavefront (cross-lane swizzles) and
// blocks of at most 4 wavefronts (the LDS buffer's size).
template <int ZV, bool SPLIT = false>
__device__ __forceinline__ void bp_quad_wave(const ParDev& g, const float* __restrict__ sino, int r0, int c0,
                                             unsigned active, int z, int lane, typename VecT<ZV>::T acc[4],
                                             int z2 = 0) {
  using V = typename VecT<ZV>::T;
  const int p = lane & 3, a = lane >> 2;
  const bool pin = (active >> p) & 1u;
  const double x = (double)(c0 + (p & 1) - g.c), y = (double)(r0 + (p >> 1) - g.c);
  const size_t nzs = (size_t)g.nz;
  const size_t step_j = (size_t)g.na * nzs;
  const uint32_t zb = 4u * (uint32_t)z, zb2 = 4u * (uint32_t)z2, row_b = (uint32_t)(4 * step_j);
  const uint32_t rows_b = 3 * row_b + 4u * (uint32_t)g.nz;
  float4* const bw0 = bp_lds_slot();   // [16] lower-row weights, then [16] upper-row weights
  float4* const bw1 = bw0 + 16;
#pragma unroll
  for (int q = 0; q < 4; ++q) acc[q] = vzero<ZV>();
  // One loop over the angles (a nested per-batch loop makes the compiler
  // copy the accumulators between two register sets on every angle); every
  // 16th angle first evaluates the geometry of the next 16.
  unsigned long long m = 0ull;   // valid pixels of the batch's angles (bit 4a + p)
  int pk = 0;
  for (int ia = 0; ia < g.na; ++ia) {
    if ((ia & 15) == 0) {
      const int i = ia + a;
      int jr = 0;
      float w0 = 0.f, w1 = 0.f;
      bool valid = false;
      if (pin && i < g.na) {
        const double co = g.cs[2 * i], sn = g.cs[2 * i + 1];
        const double t = x * co - y * sn;
        int j, j1;
        float w;
        if (interp_node(g, t, j, j1, w)) {
          const float aw = (g.mu != 0.0) ? abs_weight(g, t, x * sn + y * co) : 1.0f;
          valid = aw != 0.0f;
          if (j == g.n - 1) {        // exactly at the last node: fp[n-1]
            jr = j - 1;
            w1 = aw;
          } else {
            jr = j;
            w0 = aw * (1.0f - w);
            w1 = aw * w;
          }
        }
      }
      m = __ballot(valid);
      // the quad's row range over its valid pixels (the four lanes of one angle
      // are lanes 4a .. 4a+3), each pixel's row offset in it and the element
      // offset of its first row, (bmin * na + i) * nz -- lane-parallel for 16
      // angles at once, so the per-angle loop below only broadcasts them
      int lo = valid ? jr : 0x7fffffff, hi = valid ? jr : -1;
      lo = min(lo, __shfl_xor(lo, 1));
      lo = min(lo, __shfl_xor(lo, 2));
      hi = max(hi, __shfl_xor(hi, 1));
      hi = max(hi, __shfl_xor(hi, 2));
      // per angle: bits 0-7 the pixels' row offsets from bmin, 8-9 the span
      // bmax - bmin, 10.. bmin (n < 16384)
      int bits = valid ? (jr - lo) << (2 * p) : 0;
      bits |= __shfl_xor(bits, 1);
      bits |= __shfl_xor(bits, 2);
      pk = hi >= 0 ? bits | ((hi - lo) << 8) | (lo << 10) : 0;
      __builtin_amdgcn_wave_barrier();
      reinterpret_cast<float*>(&bw0[a])[p] = w0;
      reinterpret_cast<float*>(&bw1[a])[p] = w1;
      __builtin_amdgcn_wave_barrier();
    }
    const int l0 = 4 * (ia & 15);   // lane of pixel 0 of this angle
    const unsigned pm = (unsigned)(m >> l0) & 0xFu;
    if (!pm) continue;              // no valid pixel at this angle
    const int pkv = rl_i(pk, l0);
    const float4 U0 = bw0[l0 >> 2], U1 = bw1[l0 >> 2];
    const float u0[4] = {U0.x, U0.y, U0.z, U0.w}, u1[4] = {U1.x, U1.y, U1.z, U1.w};
    // rows bmin .. bmax + 1 <= n - 1 (jr <= n - 2); span <= 2 by the bound above
    // buffer loads from a descriptor at row bmin: lane byte offset 4z,
    // scalar row offsets
    const unsigned rowl = ((unsigned)(pkv >> 10) * (unsigned)g.na + (unsigned)ia) * (unsigned)g.nz;
    const __amdgpu_buffer_rsrc_t rs = buf_rsrc(sino + (size_t)rowl, rows_b);
    const int span = (pkv >> 8) & 3;
    const V T0 = tload<ZV, SPLIT>(rs, zb, zb2, 0), T1 = tload<ZV, SPLIT>(rs, zb, zb2, row_b);
    V T2 = vzero<ZV>(), T3 = vzero<ZV>();
    if (span >= 1) T2 = tload<ZV, SPLIT>(rs, zb, zb2, 2 * row_b);
    if (span >= 2) T3 = tload<ZV, SPLIT>(rs, zb, zb2, 3 * row_b);
    // one branch per pixel and row offset (the offsets are wave-uniform);
    // the empty asm markers keep the compiler from turning the branches
    // into register selects
#define VAMX_BP_PIXEL(Q)                                                        \
  if ((pm >> (Q)) & 1u) {                                                      \
    const int off = (pkv >> (2 * (Q))) & 3;                                    \
    if (off == 0) {                                                            \
      asm volatile("; vamx bp " #Q " 0");                                      \
      acc[Q] = vfma(u1[Q], T1, vfma(u0[Q], T0, acc[Q]));                       \
    } else if (off == 1) {                                                     \
      asm volatile("; vamx bp " #Q " 1");                                      \
      acc[Q] = vfma(u1[Q], T2, vfma(u0[Q], T1, acc[Q]));                       \
    } else {                                                                   \
      asm volatile("; vamx bp " #Q " 2");                                      \
      acc[Q] = vfma(u1[Q], T3, vfma(u0[Q], T2, acc[Q]));                       \
    }                                                                          \
  }
    VAMX_BP_PIXEL(0)
    VAMX_BP_PIXEL(1)
    VAMX_BP_PIXEL(2)
    VAMX_BP_PIXEL(3)
#undef VAMX_BP_PIXEL
  }
}

// L2-aware work order of the wavefront-uniform kernels.  About 1024 waves are
// resident per XCD at once (32 CUs x 32 waves) and consecutive logical blocks
// run on the same XCD (xcd_remap), so consecutive waves should form a compact
// tile whose data fits the XCD's 4 MB L2: forward = fp_ta consecutive angles x
// fp_tk consecutive rays (their samples at one along-ray step lie within a few
// tens of pixels), back projection = a bp_tile x bp_tile patch of 2x2 voxel
// quads (detector footprint ~2.8 bp_tile bins per angle).  When a ray (or
// column) spans several z chunks, the chunk is the slowest index, so that all
// resident waves share one z slab.  Padding waves exit.  Tiles never exceed
// the problem: a tile edge beyond the angle / ray / quad count would only add
// padding waves.
__host__ __device__ inline int fp_ta_eff(const ParDev& g) { return g.fp_ta < g.na ? g.fp_ta : g.na; }
__host__ __device__ inline int fp_tk_eff(const ParDev& g) { return g.fp_tk < g.n ? g.fp_tk : g.n; }
__host__ __device__ inline long long fp_tiled_rays(const ParDev& g) {
  const int ta = fp_ta_eff(g), tk = fp_tk_eff(g);
  return (long long)((g.na + ta - 1) / ta) * ((g.n + tk - 1) / tk) * ta * tk;
}
__device__ __forceinline__ bool fp_tile_ray(long long q, const ParDev& g, int& k, int& i) {
  const int ta = fp_ta_eff(g), tk = fp_tk_eff(g);
  const int ntk = (g.n + tk - 1) / tk;
  const int tsz = ta * tk;
  const long long tile = q / tsz;
  const int within = (int)(q % tsz);
  i = (int)(tile / ntk) * ta + within / tk;
  k = (int)(tile % ntk) * tk + within % tk;
  return i < g.na && k < g.n;
}
// back projection: quads of 2x2 voxel columns, tiles of bp_tile x bp_tile quads
__host__ __device__ inline int bp_nquad(const ParDev& g) { return (g.n + 1) / 2; }
__host__ __device__ inline int bp_tile_eff(const ParDev& g) {
  return g.bp_tile < bp_nquad(g) ? g.bp_tile : bp_nquad(g);
}
__host__ __device__ inline long long bp_tiled_quads(const ParDev& g) {
  const int bt = bp_tile_eff(g);
  const long long t = (bp_nquad(g) + bt - 1) / bt;
  return t * t * bt * bt;
}
// upper bound of bp_tiled_quads over every tile setting (tile <= nquad)
__host__ __device__ inline long long bp_tiled_quads_max(const ParDev& g) {
  const long long e = 2ll * bp_nquad(g);
  return e * e;
}
__device__ __forceinline__ bool bp_tile_quad(long long q, const ParDev& g, int& r0, int& c0) {
  const int bt = bp_tile_eff(g);
  const int nt = (bp_nquad(g) + bt - 1) / bt;
  const int tsz = bt * bt;
  const long long tile = q / tsz;
  const int within = (int)(q % tsz);
  const int qr = (int)(tile / nt) * bt + within / bt;
  const int qc = (int)(tile % nt) * bt + within % bt;
  r0 = 2 * qr;
  c0 = 2 * qc;
  return qr < bp_nquad(g) && qc < bp_nquad(g);
}

// Global index of the calling wavefront for logical block lb.  The wave's
// slot in its block is made wave-uniform explicitly (readfirstlane), so that
// everything derived from it -- ray, angle, voxel column, geometry constants
// -- lives in scalar registers.
__device__ __forceinline__ long long wave_index(unsigned lb) {
  const int slot = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  return (long long)lb * (blockDim.x >> 6) + slot;
}

// wave -> (ray, z) of the forward kernels; false for padding waves
// z chunks of a wavefront: 64 lanes x ZV values; the last chunk may be
// partial (lanes past nz load from a clamped z and store nothing)
__host__ __device__ inline int wave_zchunks(int nz, int zv) { return (nz + 64 * zv - 1) / (64 * zv); }

template <int ZV, bool SPLIT = false>
__device__ __forceinline__ bool fp_wave_item(long long wave, const ParDev& g, int lane, int& k, int& i, int& z) {
  const long long rays = fp_tiled_rays(g);
  const int nch = wave_zchunks(g.nz, ZV);
  if (wave >= rays * nch) return false;
  const int zc = (int)(wave / rays);
  if (!fp_tile_ray(wave % rays, g, k, i)) return false;
  z = SPLIT ? zc * 64 * ZV + lane : zc * 64 * ZV + lane * ZV;
  return true;
}

// Paired forward work order: unit u (an angle pair, or a single angle) has
// n + pair_u extended rays e: e < n is ray e of the primary angle (with its
// partner ray when that exists), e >= n a partner-angle ray without a
// primary.  Tiles are fp_ta units x fp_tk extended rays, z chunk slowest.
__host__ __device__ inline int fp_pair_next(const ParDev& g) { return g.n + g.pair_u; }
__host__ __device__ inline int fp_pair_ta(const ParDev& g) { return g.fp_ta < g.n_units ? g.fp_ta : g.n_units; }
__host__ __device__ inline int fp_pair_tk(const ParDev& g) {
  return g.fp_tk < fp_pair_next(g) ? g.fp_tk : fp_pair_next(g);
}
__host__ __device__ inline long long fp_pair_tiled(const ParDev& g) {
  const int ta = fp_pair_ta(g), tk = fp_pair_tk(g), ne = fp_pair_next(g);
  return (long long)((g.n_units + ta - 1) / ta) * ((ne + tk - 1) / tk) * ta * tk;
}
// wave -> ray (k, i), partner ray (k2, i2) or k2 = -1, and z; ok = false for
// padding waves
struct PairItem {
  int k, i, k2, i2, z;
  bool ok;
};
template <int ZV, bool SPLIT = false>
__host__ __device__ __forceinline__ PairItem fp_pair_item(long long wave, const ParDev& g, int lane) {
  PairItem it{0, 0, -1, -1, 0, false};
  const long long tiled = fp_pair_tiled(g);
  const int nch = wave_zchunks(g.nz, ZV);
  if (wave >= tiled * nch) return it;
  const int zc = (int)(wave / tiled);
  const long long q = wave % tiled;
  const int ta = fp_pair_ta(g), tk = fp_pair_tk(g), ne = fp_pair_next(g);
  const int ntk = (ne + tk - 1) / tk;
  const int tsz = ta * tk;
  const long long tile = q / tsz;
  const int within = (int)(q % tsz);
  const int u = (int)(tile / ntk) * ta + within / tk;
  const int e = (int)(tile % ntk) * tk + within % tk;
  if (u >= g.n_units || e >= ne) return it;
  const int u1 = g.units[2 * u], u2 = g.units[2 * u + 1];
  const int kk = 2 * g.c + g.pair_D - e;
  const bool has2 = e < g.n && u2 >= 0 && kk >= 0 && kk < g.n;
  if (e >= g.n && u2 < 0) return it;
  it.k = e < g.n ? e : g.pair_first2 + (e - g.n);
  it.i = e < g.n ? u1 : u2;
  it.k2 = has2 ? kk : -1;
  it.i2 = has2 ? u2 : -1;
  it.z = SPLIT ? zc * 64 * ZV + lane : zc * 64 * ZV + lane * ZV;   // split: z and z + 64
  it.ok = true;
  return it;
}

// wave -> (voxel quad, z) of the back-projection kernels; false for padding
template <int ZV, bool SPLIT = false>
__device__ __forceinline__ bool bp_wave_quad(long long wave, const ParDev& g, int lane, int& r0, int& c0, int& z) {
  const long long quads = bp_tiled_quads(g);
  const int nch = wave_zchunks(g.nz, ZV);
  if (wave >= quads * nch) return false;
  const int zc = (int)(wave / quads);
  if (!bp_tile_quad(wave % quads, g, r0, c0)) return false;
  z = SPLIT ? zc * 64 * ZV + lane : zc * 64 * ZV + lane * ZV;
  return true;
}

// pixels of quad (r0, c0) inside the grid (bit p = pixel p) and, of those,
// the ones computed (not zeroed by the circle clip)
__device__ __forceinline__ void bp_quad_masks(const ParDev& g, int r0, int c0, int clipc, unsigned& inrange,
                                              unsigned& active) {
  inrange = 0u;
  active = 0u;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int row = r0 + (q >> 1), col = c0 + (q & 1);
    if (row < g.n && col < g.n) {
      inrange |= 1u << q;
      if (!(clipc && g.outside[row * g.n + col])) active |= 1u << q;
    }
  }
}

// ---------------------------------------------------------------------------
// A^T on voxel quads.  Within one angle the detector coordinates of a 2x2
// quad's voxels lie within |cos| + |sin| <= sqrt(2) of each other, so the
// candidate bins of the four voxels (fpT_voxel: floor(k*) - 1 .. + 2 each)
// span at most 6 consecutive bins.  Lane l evaluates angle i0 + l for all
// four voxels with fpT_voxel's arithmetic and writes a 6-bin x 4-voxel weight
// table to LDS; the wavefront then loads each of the angle's sinogram rows
// with a nonzero weight once for the four voxels.  Each voxel sums the same
// nonzero terms in the same order as fpT_voxel (a zero weight adds an exact
// zero), so the results are bitwise equal.
// ---------------------------------------------------------------------------
struct FpTQuadSlot {
  float4* w;   // [64][6]: bin kbase + j -> weights of the quad's voxels 0..3
  int2* kb;    // [64] first candidate bin of the angle, bit j = bin kbase + j has a nonzero weight
};
__device__ __forceinline__ FpTQuadSlot fpT_quad_slot() {
  __shared__ float4 w[4][64 * 6];
  __shared__ int2 kb[4][64];
  const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)) & 3;
  return {w[wv], kb[wv]};
}

template <int ZV>
__device__ __forceinline__ void fpT_quad_wave(const ParDev& g, const float* __restrict__ sino, int r0, int c0,
                                              unsigned inrange, int zl, int lane, typename VecT<ZV>::T acc[4]) {
  using V = typename VecT<ZV>::T;
  const int n = g.n;
  const size_t zstride = (size_t)g.nz;
  const FpTQuadSlot sl = fpT_quad_slot();
#pragma unroll
  for (int q = 0; q < 4; ++q) acc[q] = vzero<ZV>();
  for (int i0 = 0; i0 < g.na; i0 += 64) {
    const int i = i0 + lane;
    // the lane's own table row: cleared, then filled with the nonzero weights
    __builtin_amdgcn_wave_barrier();
    float4* const wrow = sl.w + lane * 6;
#pragma unroll
    for (int j = 0; j < 6; ++j) wrow[j] = make_float4(0.f, 0.f, 0.f, 0.f);
    float* const wf = reinterpret_cast<float*>(wrow);
    int kbase = 0, mask = 0;
    if (i < g.na) {
      const double* M = g.mats + 6 * i;
      const double co = M[0], sn = M[1];
      int kbq[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const double X = (double)(c0 + (q & 1) - g.c), Y = (double)(r0 + (q >> 1) - g.c);
        const double kstar = co * X - sn * Y + (double)g.c + g.det_offset;
        kbq[q] = (int)floor(kstar) - 1;
      }
      kbase = min(min(kbq[0], kbq[1]), min(kbq[2], kbq[3]));
#pragma unroll 1
      for (int q = 0; q < 4; ++q) {
        if (!((inrange >> q) & 1u)) continue;
        const int row = r0 + (q >> 1), col = c0 + (q & 1);
        const double X = (double)(col - g.c), Y = (double)(row - g.c);
        const double rstar = sn * X + co * Y + (double)g.c;
        const int rb = (int)floor(rstar) - 1;
        for (int d = 0; d < 4; ++d) {
          const int kk = kbq[q] + d;
          if (kk < 0 || kk >= n) continue;
          const double kp = (double)kk - g.det_offset;
          const double a0 = M[0] * kp, a1 = M[3] * kp;
          const double t = (double)(kk - g.c) - g.det_offset;
          float wk = 0.0f;
          for (int r = rb; r < rb + 4; ++r) {
            if (r < 0 || r >= n) continue;
            const double rr = (double)r;
            const double colS = (a0 + M[1] * rr) + M[2];
            const double rowS = (a1 + M[4] * rr) + M[5];
            if (!(fabs(rowS - (double)row) < 1.0 && fabs(colS - (double)col) < 1.0)) continue;
            const float wr = tap_weight_axis(rowS, row);
            const float wc = tap_weight_axis(colS, col);
            wk = __builtin_fmaf(wr * wc, abs_weight(g, t, (double)(r - g.c)), wk);
          }
          // slot kk - kbase is in 0..5: the four voxels' first bins differ by
          // <= 2 (their k* lie within |cos| + |sin| <= sqrt(2))
          const int j = kk - kbase;
          if (wk != 0.0f && j < 6) {
            wf[j * 4 + q] = wk;
            mask |= 1 << j;
          }
        }
      }
    }
    sl.kb[lane] = make_int2(kbase, mask);
    __builtin_amdgcn_wave_barrier();
    const int rows = g.na - i0 < 64 ? g.na - i0 : 64;
    for (int a = 0; a < rows; ++a) {
      const int2 km = sl.kb[a];
      const int kb0 = __builtin_amdgcn_readfirstlane(km.x);
      const int msk = __builtin_amdgcn_readfirstlane(km.y);
      const int ia = i0 + a;
#pragma unroll
      for (int j = 0; j < 6; ++j) {
        // rows without a nonzero weight are skipped, among them every bin
        // outside 0..n-1 (its weights stay 0)
        if (!((msk >> j) & 1)) continue;
        const float4 w = sl.w[a * 6 + j];
        const V v = vload<ZV>(sino + ((size_t)(kb0 + j) * g.na + ia) * zstride + zl);
        acc[0] = vfma(w.x, v, acc[0]);
        acc[1] = vfma(w.y, v, acc[1]);
        acc[2] = vfma(w.z, v, acc[2]);
        acc[3] = vfma(w.w, v, acc[3]);
      }
    }
  }
}

// host helpers implemented in parallel.hip
int par_slice_minmax(vamx_plan* p, const float* vol, hipStream_t s);
int par_forward_launch(vamx_plan* p, const float* vol, float* sino, hipStream_t s, bool warp_clip);
bool par_use_vec4(const vamx_plan* p, const void* a, const void* b);
bool par_use_vec2(const vamx_plan* p, const void* a, const void* b);
int par_wave_zv(const vamx_plan* p, const void* a, const void* b, int requested);
int par_fp_wave_zv(const vamx_plan* p, const void* vol, const void* sino);
bool par_wave_split(const vamx_plan* p, int zv);
int wave_zv_choice(int nz, uintptr_t align_bits, int requested);
void par_pair_units(const std::vector<double>& cs, int na, int n, int c, double D, std::vector<int>& units,
                    int& npaired_rays, int& first2);

}  // namespace vamx
