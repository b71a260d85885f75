// Parallel-beam projector kernels for gfx950 (MI355X).
//
// Discretisation: the reference's CPU projector "D1"
// (vamtoolbox/projector/Projector3DParallel.py:113-207) plus the detector
// offset / absorption / inclination extensions.  Layouts: volume (n, n, nz)
// and sinogram (n, n_angles, nz), both z-contiguous, so the work item is a
// (ray or voxel column) x (z vector) and the bilinear / linear weights of a
// ray sample are identical for every z: a wavefront covers 64 z-vectors of
// one ray or voxel column, every tap is one coalesced load, and the weights
// are wavefront-uniform.
//
// All geometry (sample coordinates, floor/ceil decisions, weights) is
// computed in float64 with the reference's operation order (this file is
// compiled with -ffp-contract=off) so that every discontinuous decision
// matches the reference exactly; data arithmetic is float32 with explicit
// fused multiply-adds.  No kernel uses atomics: each output element is owned
// by exactly one lane and accumulated in a fixed order.
#include "vamx_internal.h"

namespace vamx {

// ---------------------------------------------------------------------------
// small vector helpers (ZV = z values per lane: 4 -> float4, 1 -> float)
// ---------------------------------------------------------------------------
template <int ZV> struct VecT;
template <> struct VecT<1> { using T = float; };
template <> struct VecT<4> { using T = float4; };

__device__ __forceinline__ float vzero1() { return 0.0f; }
__device__ __forceinline__ float4 vzero4() { return make_float4(0.f, 0.f, 0.f, 0.f); }
template <int ZV> __device__ __forceinline__ typename VecT<ZV>::T vzero();
template <> __device__ __forceinline__ float vzero<1>() { return 0.0f; }
template <> __device__ __forceinline__ float4 vzero<4>() { return vzero4(); }

template <int ZV> __device__ __forceinline__ typename VecT<ZV>::T vload(const float* p);
template <> __device__ __forceinline__ float vload<1>(const float* p) { return *p; }
template <> __device__ __forceinline__ float4 vload<4>(const float* p) { return *reinterpret_cast<const float4*>(p); }

template <int ZV> __device__ __forceinline__ void vstore(float* p, typename VecT<ZV>::T v);
template <> __device__ __forceinline__ void vstore<1>(float* p, float v) { *p = v; }
template <> __device__ __forceinline__ void vstore<4>(float* p, float4 v) { *reinterpret_cast<float4*>(p) = v; }

// a*x + y
__device__ __forceinline__ float vfma(float a, float x, float y) { return __builtin_fmaf(a, x, y); }
__device__ __forceinline__ float4 vfma(float a, float4 x, float4 y) {
  return make_float4(__builtin_fmaf(a, x.x, y.x), __builtin_fmaf(a, x.y, y.y),
                     __builtin_fmaf(a, x.z, y.z), __builtin_fmaf(a, x.w, y.w));
}
__device__ __forceinline__ float vmul(float a, float x) { return a * x; }
__device__ __forceinline__ float4 vmul(float a, float4 x) { return make_float4(a * x.x, a * x.y, a * x.z, a * x.w); }
__device__ __forceinline__ float vsub(float a, float b) { return a - b; }
__device__ __forceinline__ float4 vsub(float4 a, float4 b) { return make_float4(a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w); }
__device__ __forceinline__ float vadd(float a, float b) { return a + b; }
__device__ __forceinline__ float4 vadd(float4 a, float4 b) { return make_float4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w); }
__device__ __forceinline__ float clampf(float v, float lo, float hi) { return fminf(fmaxf(v, lo), hi); }
__device__ __forceinline__ float vclamp(float v, float lo, float hi) { return clampf(v, lo, hi); }
__device__ __forceinline__ float4 vclamp(float4 v, float4 lo, float4 hi) {
  return make_float4(clampf(v.x, lo.x, hi.x), clampf(v.y, lo.y, hi.y), clampf(v.z, lo.z, hi.z),
                     clampf(v.w, lo.w, hi.w));
}

// ---------------------------------------------------------------------------
// absorption weight of a sample at lateral coordinate t, along-ray s
// exp(-mu * max(0, sqrt(R^2 - t^2) - s) / cos(phi)) inside |t| < R, else 0
// (per-ray form of astra3Dabsorption.getExpabs, astra3Dabsorption.py:129-145)
// ---------------------------------------------------------------------------
__device__ __forceinline__ float abs_weight(const ParDev& g, double t, double s) {
  if (g.mu == 0.0) return 1.0f;
  if (!(fabs(t) < g.radius)) return 0.0f;
  double half = sqrt(g.radius * g.radius - t * t);
  double path = fmax(0.0, half - s) * g.inv_cos;
  return expf((float)(-g.mu * path));
}

// ---------------------------------------------------------------------------
// per-slice min / max (skimage warp(..., clip=True) output clip)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void slice_minmax_part(const float* __restrict__ vol, int npix, int nz,
                                                         int parts, float* __restrict__ part) {
  int pb = blockIdx.x;  // pixel chunk
  int chunk = (npix + parts - 1) / parts;
  int p0 = pb * chunk, p1 = min(npix, p0 + chunk);
  for (int z = threadIdx.x; z < nz; z += blockDim.x) {
    float lo = INFINITY, hi = -INFINITY;
    for (int p = p0; p < p1; ++p) {
      float v = vol[(size_t)p * nz + z];
      lo = fminf(lo, v);
      hi = fmaxf(hi, v);
    }
    part[(size_t)pb * 2 * nz + z] = lo;
    part[(size_t)pb * 2 * nz + nz + z] = hi;
  }
}

__global__ __launch_bounds__(256) void slice_minmax_final(const float* __restrict__ part, int nz, int parts,
                                                          float* __restrict__ out) {
  int z = blockIdx.x * blockDim.x + threadIdx.x;
  if (z >= nz) return;
  float lo = INFINITY, hi = -INFINITY;
  for (int b = 0; b < parts; ++b) {
    lo = fminf(lo, part[(size_t)b * 2 * nz + z]);
    hi = fmaxf(hi, part[(size_t)b * 2 * nz + nz + z]);
  }
  out[z] = lo;
  out[nz + z] = hi;
}

// ---------------------------------------------------------------------------
// forward projection, zero inclination (D1: Projector3DParallel.py:129-156)
// one lane = one ray (k, i) x ZV consecutive z
// ---------------------------------------------------------------------------
template <int ZV, bool CLIP>
__global__ __launch_bounds__(256) void par_fp_kernel(ParDev g, const float* __restrict__ vol,
                                                     float* __restrict__ sino,
                                                     const float* __restrict__ minmax) {
  using V = typename VecT<ZV>::T;
  const int nzv = g.nz / ZV;
  const long long total = (long long)g.na * g.n * nzv;
  const unsigned lb = xcd_remap(blockIdx.x, gridDim.x);
  const long long tid = (long long)lb * blockDim.x + threadIdx.x;
  if (tid >= total) return;
  const int zv = (int)(tid % nzv);
  const long long ray = tid / nzv;
  const int k = (int)(ray % g.n);
  const int i = (int)(ray / g.n);
  const int z = zv * ZV;
  const int n = g.n;
  const double* M = g.mats + 6 * i;
  const double M00 = M[0], M01 = M[1], M02 = M[2], M10 = M[3], M11 = M[4], M12 = M[5];
  const double kp = (double)k - g.det_offset;
  const double a0 = M00 * kp, a1 = M10 * kp;
  const double t = (double)(k - g.c) - g.det_offset;

  V lo, hi;
  if constexpr (CLIP) {
    lo = vload<ZV>(minmax + z);
    hi = vload<ZV>(minmax + g.nz + z);
  }
  V acc = vzero<ZV>();
  const size_t zstride = (size_t)g.nz;
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
    const float w = abs_weight(g, t, (double)(r - g.c));
    acc = vfma(w, v, acc);
  }
  vstore<ZV>(sino + ((size_t)k * g.na + i) * zstride + z, acc);
}

// forward projection with inclination: one lane = one ray x one detector row
__global__ __launch_bounds__(256) void par_fp_incl_kernel(ParDev g, const float* __restrict__ vol,
                                                          float* __restrict__ sino) {
  const long long total = (long long)g.na * g.n * g.nz;
  const unsigned lb = xcd_remap(blockIdx.x, gridDim.x);
  const long long tid = (long long)lb * blockDim.x + threadIdx.x;
  if (tid >= total) return;
  const int j = (int)(tid % g.nz);
  const long long ray = tid / g.nz;
  const int k = (int)(ray % g.n);
  const int i = (int)(ray / g.n);
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
      const float wz = (float)(q == 0 ? 1.0 - wz1 : wz1);
      if (zz < 0 || zz >= nz) continue;
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
  sino[((size_t)k * g.na + i) * nz + j] = acc;
}

// ---------------------------------------------------------------------------
// back projection (D1: Projector3DParallel.py:175-207): voxel-driven gather
// one lane = one voxel column (row, col) x ZV consecutive z
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool interp_node(const ParDev& g, double t, int& j, int& j1, float& w) {
  // np.interp(t, xp, fp, left=0, right=0), xp[k] = (k - c) - det_offset
  if (!(t >= g.xp0 && t <= g.xp_last)) return false;
  const double u = t - g.xp0;
  int jj = (int)floor(u);
  jj = jj < 0 ? 0 : (jj > g.n - 1 ? g.n - 1 : jj);
  j = jj;
  j1 = jj + 1 < g.n ? jj + 1 : g.n - 1;
  w = (jj == g.n - 1) ? 0.0f : (float)(t - ((double)(jj - g.c) - g.det_offset));
  return true;
}

template <int ZV, bool CLIPC>
__global__ __launch_bounds__(256) void par_bp_kernel(ParDev g, const float* __restrict__ sino,
                                                     float* __restrict__ vol) {
  using V = typename VecT<ZV>::T;
  const int nzv = g.nz / ZV;
  const long long total = (long long)g.n * g.n * nzv;
  const unsigned lb = xcd_remap(blockIdx.x, gridDim.x);
  const long long tid = (long long)lb * blockDim.x + threadIdx.x;
  if (tid >= total) return;
  const int zv = (int)(tid % nzv);
  const int pix = (int)(tid / nzv);
  const int z = zv * ZV;
  const size_t zstride = (size_t)g.nz;
  float* outp = vol + (size_t)pix * zstride + z;
  if (CLIPC && g.outside[pix]) {
    vstore<ZV>(outp, vzero<ZV>());
    return;
  }
  const int row = pix / g.n, col = pix % g.n;
  const double x = (double)(col - g.c), y = (double)(row - g.c);
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
    if (g.mu != 0.0) {
      const double s = x * sn + y * co;
      val = vmul(abs_weight(g, t, s), val);
    }
    acc = vadd(acc, val);
  }
  vstore<ZV>(outp, acc);
}

__global__ __launch_bounds__(256) void par_bp_incl_kernel(ParDev g, const float* __restrict__ sino,
                                                          float* __restrict__ vol, int clipc) {
  const long long total = (long long)g.n * g.n * g.nz;
  const unsigned lb = xcd_remap(blockIdx.x, gridDim.x);
  const long long tid = (long long)lb * blockDim.x + threadIdx.x;
  if (tid >= total) return;
  const int z = (int)(tid % g.nz);
  const int pix = (int)(tid / g.nz);
  if (clipc && g.outside[pix]) {
    vol[(size_t)pix * g.nz + z] = 0.0f;
    return;
  }
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
    acc += abs_weight(g, t, s) * val;
  }
  vol[(size_t)pix * g.nz + z] = acc;
}

// ---------------------------------------------------------------------------
// exact transpose of the forward projection (linear part, no output clip):
// voxel-driven gather over the <= 4x4 candidate samples (k, r) per angle
// whose bilinear footprint covers the voxel.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float tap_weight_axis(double coord, int idx) {
  // weight that a sample at `coord` gives to integer position idx
  // (skimage: floor tap weight 1-d, ceil tap weight d; floor==ceil -> 1 + 0)
  const double f = floor(coord), cl = ceil(coord);
  const double d = coord - f;
  double w = 0.0;
  if ((int)f == idx) w += 1.0 - d;
  if ((int)cl == idx) w += d;
  return (float)w;
}

template <int ZV>
__global__ __launch_bounds__(256) void par_fpT_kernel(ParDev g, const float* __restrict__ sino,
                                                      float* __restrict__ vol) {
  using V = typename VecT<ZV>::T;
  const int nzv = g.nz / ZV;
  const long long total = (long long)g.n * g.n * nzv;
  const unsigned lb = xcd_remap(blockIdx.x, gridDim.x);
  const long long tid = (long long)lb * blockDim.x + threadIdx.x;
  if (tid >= total) return;
  const int zv = (int)(tid % nzv);
  const int pix = (int)(tid / nzv);
  const int z = zv * ZV;
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
  vstore<ZV>(vol + (size_t)pix * zstride + z, acc);
}

__global__ __launch_bounds__(256) void par_fpT_incl_kernel(ParDev g, const float* __restrict__ sino,
                                                           float* __restrict__ vol) {
  const long long total = (long long)g.n * g.n * g.nz;
  const unsigned lb = xcd_remap(blockIdx.x, gridDim.x);
  const long long tid = (long long)lb * blockDim.x + threadIdx.x;
  if (tid >= total) return;
  const int z = (int)(tid % g.nz);
  const int pix = (int)(tid / g.nz);
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
  vol[(size_t)pix * nz + z] = acc;
}

// ---------------------------------------------------------------------------
// exact transpose of the back projection: sinogram-driven gather over the
// voxels whose detector coordinate falls within one bin of bin k.
// ---------------------------------------------------------------------------
template <int ZV>
__global__ __launch_bounds__(256) void par_bpT_kernel(ParDev g, const float* __restrict__ vol,
                                                      float* __restrict__ sino, int clipc) {
  using V = typename VecT<ZV>::T;
  const int nzv = g.nz / ZV;
  const long long total = (long long)g.n * g.na * nzv;
  const unsigned lb = xcd_remap(blockIdx.x, gridDim.x);
  const long long tid = (long long)lb * blockDim.x + threadIdx.x;
  if (tid >= total) return;
  const int zv = (int)(tid % nzv);
  const long long bin = tid / nzv;
  const int i = (int)(bin % g.na);
  const int k = (int)(bin / g.na);
  const int z = zv * ZV;
  const int n = g.n;
  const size_t zstride = (size_t)g.nz;
  const double co = g.cs[2 * i], sn = g.cs[2 * i + 1];
  const double xk = (double)(k - g.c) - g.det_offset;
  const bool rows_major = fabs(co) >= fabs(sn);
  V acc = vzero<ZV>();
  for (int a = 0; a < n; ++a) {
    // candidate range along the minor axis for this line of the major axis
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
    const int b0 = (int)floor(lo) - 1 + g.c, b1 = (int)ceil(hi) + 1 + g.c;
    for (int bb = b0 < 0 ? 0 : b0; bb <= b1 && bb < n; ++bb) {
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
  vstore<ZV>(sino + ((size_t)k * g.na + i) * zstride + z, acc);
}

__global__ __launch_bounds__(256) void par_bpT_incl_kernel(ParDev g, const float* __restrict__ vol,
                                                           float* __restrict__ sino, int clipc) {
  const long long total = (long long)g.n * g.na * g.nz;
  const unsigned lb = xcd_remap(blockIdx.x, gridDim.x);
  const long long tid = (long long)lb * blockDim.x + threadIdx.x;
  if (tid >= total) return;
  const int jz = (int)(tid % g.nz);
  const long long bin = tid / g.nz;
  const int i = (int)(bin % g.na);
  const int k = (int)(bin / g.na);
  const int n = g.n, nz = g.nz;
  const double co = g.cs[2 * i], sn = g.cs[2 * i + 1];
  const double xk = (double)(k - g.c) - g.det_offset;
  const bool rows_major = fabs(co) >= fabs(sn);
  float acc = 0.0f;
  for (int a = 0; a < n; ++a) {
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
    const int b0 = (int)floor(lo) - 1 + g.c, b1 = (int)ceil(hi) + 1 + g.c;
    for (int bb = b0 < 0 ? 0 : b0; bb <= b1 && bb < n; ++bb) {
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
  sino[((size_t)k * g.na + i) * nz + jz] = acc;
}

}  // namespace vamx

// ===========================================================================
// C ABI
// ===========================================================================
using namespace vamx;

static bool use_vec4(const vamx_plan* p, const void* a, const void* b) {
  return (p->par.nz % 4 == 0) && !p->par.incl && ((uintptr_t)a % 16 == 0) && ((uintptr_t)b % 16 == 0);
}

extern "C" int vamx_par_create(const vamx_par_geom* geom, vamx_plan** out_plan) {
  if (!geom || !out_plan) return set_error(VAMX_ERR_ARG, "vamx_par_create: null argument");
  *out_plan = nullptr;
  if (geom->n < 1 || geom->nz < 1 || geom->n_angles < 1)
    return set_error(VAMX_ERR_ARG, "vamx_par_create: n, nz and n_angles must be >= 1");
  if (!geom->angles_deg) return set_error(VAMX_ERR_ARG, "vamx_par_create: angles_deg is null");
  if (!geom->outside_mask) return set_error(VAMX_ERR_ARG, "vamx_par_create: outside_mask is null");
  if (!(std::fabs(geom->incl_deg) < 89.0))
    return set_error(VAMX_ERR_ARG, "vamx_par_create: |inclination| must be < 89 degrees");
  if (geom->mu < 0.0) return set_error(VAMX_ERR_ARG, "vamx_par_create: mu must be >= 0");
  vamx_plan* p = new vamx_plan();
  p->kind = PLAN_PARALLEL;
  if (hipGetDevice(&p->device) != hipSuccess) {
    delete p;
    return set_error(VAMX_ERR_HIP, "vamx_par_create: hipGetDevice failed");
  }
  const int n = geom->n, na = geom->n_angles, nz = geom->nz;
  const int c = n / 2;
  p->flags = geom->flags;
  p->h_angles.assign(geom->angles_deg, geom->angles_deg + na);
  std::vector<double> mats(6 * (size_t)na), cs(2 * (size_t)na);
  for (int i = 0; i < na; ++i) {
    // np.deg2rad(a) = a * (pi / 180); R of Projector3DParallel.py:137-143
    const double ang = geom->angles_deg[i] * (M_PI / 180.0);
    const double ca = std::cos(ang), sa = std::sin(ang);
    mats[6 * i + 0] = ca;
    mats[6 * i + 1] = sa;
    mats[6 * i + 2] = -c * (ca + sa - 1);
    mats[6 * i + 3] = -sa;
    mats[6 * i + 4] = ca;
    mats[6 * i + 5] = -c * (ca - sa - 1);
    cs[2 * i] = ca;
    cs[2 * i + 1] = sa;
  }
  ParDev& g = p->par;
  g.n = n;
  g.nz = nz;
  g.na = na;
  g.c = c;
  g.det_offset = geom->det_offset;
  g.xp0 = (double)(0 - c) - geom->det_offset;
  g.xp_last = (double)(n - 1 - c) - geom->det_offset;
  const double phi = geom->incl_deg * (M_PI / 180.0);
  g.tan_phi = geom->incl_deg != 0.0 ? std::tan(phi) : 0.0;
  g.inv_cos = 1.0 / std::cos(phi);
  g.incl = geom->incl_deg != 0.0;
  g.mu = geom->mu;
  g.radius = geom->radius > 0.0 ? geom->radius : (n - 1) / 2.0;
  p->minmax_parts = 256;
  hipError_t e = hipSuccess;
  e = hipMalloc((void**)&p->d_mats, mats.size() * sizeof(double));
  if (e == hipSuccess) e = hipMalloc((void**)&p->d_cs, cs.size() * sizeof(double));
  if (e == hipSuccess) e = hipMalloc((void**)&p->d_outside, (size_t)n * n);
  if (e == hipSuccess) e = hipMalloc((void**)&p->d_minmax_part, (size_t)p->minmax_parts * 2 * nz * sizeof(float));
  if (e == hipSuccess) e = hipMalloc((void**)&p->d_minmax, (size_t)2 * nz * sizeof(float));
  if (e == hipSuccess) e = hipMemcpy(p->d_mats, mats.data(), mats.size() * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(p->d_cs, cs.data(), cs.size() * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(p->d_outside, geom->outside_mask, (size_t)n * n, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    vamx_destroy(p);
    return hip_check(e, "vamx_par_create: allocation / upload");
  }
  g.mats = p->d_mats;
  g.cs = p->d_cs;
  g.outside = p->d_outside;
  *out_plan = p;
  return VAMX_OK;
}

namespace vamx {
int par_slice_minmax(vamx_plan* p, const float* vol, hipStream_t s) {
  const ParDev& g = p->par;
  hipLaunchKernelGGL(slice_minmax_part, dim3(p->minmax_parts), dim3(256), 0, s, vol, g.n * g.n, g.nz,
                     p->minmax_parts, p->d_minmax_part);
  hipLaunchKernelGGL(slice_minmax_final, dim3(blocks_for(g.nz, 256)), dim3(256), 0, s, p->d_minmax_part, g.nz,
                     p->minmax_parts, p->d_minmax);
  VAMX_HIP(hipGetLastError());
  return VAMX_OK;
}

int par_forward_launch(vamx_plan* p, const float* vol, float* sino, hipStream_t s, bool warp_clip) {
  const ParDev& g = p->par;
  if (g.incl) {
    const long long total = (long long)g.na * g.n * g.nz;
    hipLaunchKernelGGL(par_fp_incl_kernel, dim3(blocks_for(total, 256)), dim3(256), 0, s, g, vol, sino);
    VAMX_HIP(hipGetLastError());
    return VAMX_OK;
  }
  if (warp_clip) {
    int rc = par_slice_minmax(p, vol, s);
    if (rc) return rc;
  }
  const bool v4 = use_vec4(p, vol, sino);
  const int zv = v4 ? 4 : 1;
  const long long total = (long long)g.na * g.n * (g.nz / zv);
  const dim3 grid(blocks_for(total, 256)), block(256);
  if (v4) {
    if (warp_clip)
      hipLaunchKernelGGL((par_fp_kernel<4, true>), grid, block, 0, s, g, vol, sino, p->d_minmax);
    else
      hipLaunchKernelGGL((par_fp_kernel<4, false>), grid, block, 0, s, g, vol, sino, p->d_minmax);
  } else {
    if (warp_clip)
      hipLaunchKernelGGL((par_fp_kernel<1, true>), grid, block, 0, s, g, vol, sino, p->d_minmax);
    else
      hipLaunchKernelGGL((par_fp_kernel<1, false>), grid, block, 0, s, g, vol, sino, p->d_minmax);
  }
  VAMX_HIP(hipGetLastError());
  return VAMX_OK;
}

int par_backward_launch(vamx_plan* p, const float* sino, float* vol, hipStream_t s, bool clipc) {
  const ParDev& g = p->par;
  if (g.incl) {
    const long long total = (long long)g.n * g.n * g.nz;
    hipLaunchKernelGGL(par_bp_incl_kernel, dim3(blocks_for(total, 256)), dim3(256), 0, s, g, sino, vol,
                       clipc ? 1 : 0);
    VAMX_HIP(hipGetLastError());
    return VAMX_OK;
  }
  const bool v4 = use_vec4(p, sino, vol);
  const int zv = v4 ? 4 : 1;
  const long long total = (long long)g.n * g.n * (g.nz / zv);
  const dim3 grid(blocks_for(total, 256)), block(256);
  if (v4) {
    if (clipc)
      hipLaunchKernelGGL((par_bp_kernel<4, true>), grid, block, 0, s, g, sino, vol);
    else
      hipLaunchKernelGGL((par_bp_kernel<4, false>), grid, block, 0, s, g, sino, vol);
  } else {
    if (clipc)
      hipLaunchKernelGGL((par_bp_kernel<1, true>), grid, block, 0, s, g, sino, vol);
    else
      hipLaunchKernelGGL((par_bp_kernel<1, false>), grid, block, 0, s, g, sino, vol);
  }
  VAMX_HIP(hipGetLastError());
  return VAMX_OK;
}
}  // namespace vamx

extern "C" int vamx_par_forward(vamx_plan* plan, const float* vol, float* sino, void* stream) {
  if (!check_plan(plan, PLAN_PARALLEL)) return set_error(VAMX_ERR_STATE, "vamx_par_forward: not a parallel plan");
  if (!vol || !sino) return set_error(VAMX_ERR_ARG, "vamx_par_forward: null buffer");
  return par_forward_launch(plan, vol, sino, (hipStream_t)stream, !(plan->flags & VAMX_PAR_NO_WARP_CLIP));
}

extern "C" int vamx_par_backward(vamx_plan* plan, const float* sino, float* vol, void* stream) {
  if (!check_plan(plan, PLAN_PARALLEL)) return set_error(VAMX_ERR_STATE, "vamx_par_backward: not a parallel plan");
  if (!vol || !sino) return set_error(VAMX_ERR_ARG, "vamx_par_backward: null buffer");
  return par_backward_launch(plan, sino, vol, (hipStream_t)stream, !(plan->flags & VAMX_PAR_NO_CIRCLE_CLIP));
}

extern "C" int vamx_par_forward_adjoint(vamx_plan* plan, const float* sino, float* vol, void* stream) {
  if (!check_plan(plan, PLAN_PARALLEL))
    return set_error(VAMX_ERR_STATE, "vamx_par_forward_adjoint: not a parallel plan");
  if (!vol || !sino) return set_error(VAMX_ERR_ARG, "vamx_par_forward_adjoint: null buffer");
  const ParDev& g = plan->par;
  hipStream_t s = (hipStream_t)stream;
  if (g.incl) {
    const long long total = (long long)g.n * g.n * g.nz;
    hipLaunchKernelGGL(par_fpT_incl_kernel, dim3(blocks_for(total, 256)), dim3(256), 0, s, g, sino, vol);
  } else if (use_vec4(plan, sino, vol)) {
    const long long total = (long long)g.n * g.n * (g.nz / 4);
    hipLaunchKernelGGL((par_fpT_kernel<4>), dim3(blocks_for(total, 256)), dim3(256), 0, s, g, sino, vol);
  } else {
    const long long total = (long long)g.n * g.n * g.nz;
    hipLaunchKernelGGL((par_fpT_kernel<1>), dim3(blocks_for(total, 256)), dim3(256), 0, s, g, sino, vol);
  }
  VAMX_HIP(hipGetLastError());
  return VAMX_OK;
}

extern "C" int vamx_par_backward_adjoint(vamx_plan* plan, const float* vol, float* sino, void* stream) {
  if (!check_plan(plan, PLAN_PARALLEL))
    return set_error(VAMX_ERR_STATE, "vamx_par_backward_adjoint: not a parallel plan");
  if (!vol || !sino) return set_error(VAMX_ERR_ARG, "vamx_par_backward_adjoint: null buffer");
  const ParDev& g = plan->par;
  hipStream_t s = (hipStream_t)stream;
  const int clipc = (plan->flags & VAMX_PAR_NO_CIRCLE_CLIP) ? 0 : 1;
  if (g.incl) {
    const long long total = (long long)g.n * g.na * g.nz;
    hipLaunchKernelGGL(par_bpT_incl_kernel, dim3(blocks_for(total, 256)), dim3(256), 0, s, g, vol, sino, clipc);
  } else if (use_vec4(plan, vol, sino)) {
    const long long total = (long long)g.n * g.na * (g.nz / 4);
    hipLaunchKernelGGL((par_bpT_kernel<4>), dim3(blocks_for(total, 256)), dim3(256), 0, s, g, vol, sino, clipc);
  } else {
    const long long total = (long long)g.n * g.na * g.nz;
    hipLaunchKernelGGL((par_bpT_kernel<1>), dim3(blocks_for(total, 256)), dim3(256), 0, s, g, vol, sino, clipc);
  }
  VAMX_HIP(hipGetLastError());
  return VAMX_OK;
}
