// Parallel-beam projector kernels and their C ABI (gfx950).
//
// Each kernel is a thin wrapper around a device core of par_core.h: one lane
// owns one ray (forward, B^T) or one voxel column (backward, A^T) and ZV
// consecutive z values; every output element is written by exactly one lane
// after a fixed-order accumulation (no atomics, bitwise deterministic).
#include <algorithm>

#include "par_core.h"

namespace vamx {

// ---------------------------------------------------------------------------
// per-slice min / max of a volume (skimage warp(..., clip=True) output clip)
// ---------------------------------------------------------------------------
// Per-slice min / max of a volume (the warp clip range), pass 1: block b
// reduces pixels [b*chunk, (b+1)*chunk) for every z.  Threads form groups of
// min(nz, 1024) consecutive z values; the groups split the block's pixels and
// are combined through LDS.
__global__ __launch_bounds__(1024) void slice_minmax_part(const float* __restrict__ vol, int npix, int nz,
                                                          int parts, float* __restrict__ part) {
  __shared__ float slo[1024], shi[1024];
  const int t = threadIdx.x;
  const int zs = nz < 1024 ? nz : 1024;
  const int groups = 1024 / zs;
  const int grp = t / zs, zt = t % zs;
  const int chunk = (npix + parts - 1) / parts;
  const int p0 = blockIdx.x * chunk, p1 = min(npix, p0 + chunk);
  for (int z0 = 0; z0 < nz; z0 += zs) {
    const int z = z0 + zt;
    float lo = INFINITY, hi = -INFINITY;
    if (grp < groups && z < nz) {
#pragma unroll 8
      for (int p = p0 + grp; p < p1; p += groups) {
        const float v = vol[(size_t)p * nz + z];
        lo = fminf(lo, v);
        hi = fmaxf(hi, v);
      }
    }
    slo[t] = lo;
    shi[t] = hi;
    __syncthreads();
    if (grp == 0 && z < nz) {
      for (int q = 1; q < groups; ++q) {
        lo = fminf(lo, slo[q * zs + zt]);
        hi = fmaxf(hi, shi[q * zs + zt]);
      }
      part[(size_t)blockIdx.x * 2 * nz + z] = lo;
      part[(size_t)blockIdx.x * 2 * nz + nz + z] = hi;
    }
    __syncthreads();
  }
}

// pass 2: block b owns z in [64b, 64b+64); 16 groups of 64 threads split the
// parts.  out[0:nz] = min, out[nz:2nz] = max.
__global__ __launch_bounds__(1024) void slice_minmax_final(const float* __restrict__ part, int nz, int parts,
                                                           float* __restrict__ out) {
  __shared__ float slo[1024], shi[1024];
  const int t = threadIdx.x;
  const int zt = t & 63, grp = t >> 6;
  const int z = blockIdx.x * 64 + zt;
  float lo = INFINITY, hi = -INFINITY;
  if (z < nz) {
#pragma unroll 4
    for (int b = grp; b < parts; b += 16) {
      lo = fminf(lo, part[(size_t)b * 2 * nz + z]);
      hi = fmaxf(hi, part[(size_t)b * 2 * nz + nz + z]);
    }
  }
  slo[t] = lo;
  shi[t] = hi;
  __syncthreads();
  if (grp == 0 && z < nz) {
    for (int q = 1; q < 16; ++q) {
      lo = fminf(lo, slo[q * 64 + zt]);
      hi = fmaxf(hi, shi[q * 64 + zt]);
    }
    out[z] = lo;
    out[nz + z] = hi;
  }
}

// Occlusion sinogram (Projector3DParallel.py:77-111): for detector bin k,
// angle i and slice z, the row coordinate y = r - c of the first sample along
// the ray whose bilinear rotation of the insert is positive, i.e. that has a
// tap of positive weight on an occluding voxel (the warp clip keeps positive
// values positive); NaN when the ray meets no occluder.  One thread per
// (k, i, z), z fastest; runs once at plan creation.
__global__ __launch_bounds__(256) void occ_sino_kernel(ParDev g, const uint8_t* __restrict__ ins,
                                                       float* __restrict__ occ) {
  const long long total = (long long)g.n * g.na * g.nz;
  const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= total) return;
  const int z = (int)(tid % g.nz);
  const long long ki = tid / g.nz;
  const int i = (int)(ki % g.na), k = (int)(ki / g.na);
  const int n = g.n;
  const double* M = g.mats + 6 * i;
  const double kp = (double)k - g.det_offset;
  const double a0 = M[0] * kp, a1 = M[3] * kp;
  const size_t zs = (size_t)g.nz;
  float first = NAN;
  for (int r = 0; r < n; ++r) {
    const double rr = (double)r;
    const double col = (a0 + M[1] * rr) + M[2];
    const double row = (a1 + M[4] * rr) + M[5];
    if (!(row > -1.0 && row < (double)n && col > -1.0 && col < (double)n)) continue;
    const double fr = floor(row), fc = floor(col);
    const double dr = row - fr, dc = col - fc;
    const int r0 = (int)fr, c0 = (int)fc, r1 = (int)ceil(row), c1 = (int)ceil(col);
    const bool vr0 = r0 >= 0, vr1 = r1 < n, vc0 = c0 >= 0, vc1 = c1 < n;
    // tap weights: (1-dr)(1-dc) > 0 always; the others need dc > 0 / dr > 0
    const bool hit = (vr0 && vc0 && ins[((size_t)r0 * n + c0) * zs + z]) ||
                     (dc > 0.0 && vr0 && vc1 && ins[((size_t)r0 * n + c1) * zs + z]) ||
                     (dr > 0.0 && vr1 && vc0 && ins[((size_t)r1 * n + c0) * zs + z]) ||
                     (dr > 0.0 && dc > 0.0 && vr1 && vc1 && ins[((size_t)r1 * n + c1) * zs + z]);
    if (hit) {
      first = (float)(r - g.c);
      break;
    }
  }
  occ[((size_t)k * g.na + i) * zs + z] = first;
}

// whether the forward projection applies the shadow of angle i, slice z:
// the occlusion column holds a finite non-zero entry
// (count_nonzero - isnan != 0, Projector3DParallel.py:150)
__global__ __launch_bounds__(256) void occ_active_kernel(ParDev g, const float* __restrict__ occ,
                                                         uint8_t* __restrict__ active) {
  const long long total = (long long)g.na * g.nz;
  const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= total) return;
  const int z = (int)(tid % g.nz), i = (int)(tid / g.nz);
  uint8_t a = 0;
  for (int k = 0; k < g.n && !a; ++k) {
    const float o = occ[((size_t)k * g.na + i) * g.nz + z];
    if (!isnan(o) && o != 0.0f) a = 1;
  }
  active[(size_t)i * g.nz + z] = a;
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
template <int ZV, bool CLIP>
__global__ __launch_bounds__(256) void par_fp_kernel(ParDev g, const float* __restrict__ vol,
                                                     float* __restrict__ sino, const float* __restrict__ minmax) {
  const int nzv = g.nz / ZV;
  const long long total = (long long)g.na * g.n * nzv;
  const long long tid = (long long)xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (tid >= total) return;
  const int z = (int)(tid % nzv) * ZV;
  const long long ray = tid / nzv;
  const int k = (int)(ray % g.n), i = (int)(ray / g.n);
  vstore<ZV>(sino + ((size_t)k * g.na + i) * g.nz + z, fp_ray<ZV, CLIP>(g, vol, minmax, k, i, z));
}

// wavefront-uniform forward projection: one wavefront = one ray x 64*ZV z
// (SPLIT: ZV = 2 with lanes owning z and z + 64)
template <int ZV, bool REUSE, bool SPLIT>
__global__ __launch_bounds__(256) void par_fp_wave_kernel(ParDev g, const float* __restrict__ vol,
                                                          float* __restrict__ sino, const float* __restrict__ minmax,
                                                          int clip) {
  const int lane = threadIdx.x & 63;
  const long long wave = wave_index(xcd_remap(blockIdx.x, gridDim.x));
  int k, i, z, zl, zl2;
  if (!fp_wave_item<ZV, SPLIT>(wave, g, lane, k, i, z)) return;
  wave_zload<ZV, SPLIT>(z, g.nz, zl, zl2);
  const bool clip_on = clip && wave_clip_active<ZV, SPLIT>(minmax, g.nz, zl, zl2);
  typename VecT<ZV>::T acc =
      clip_on ? fp_ray_wave<ZV, true, REUSE, false, SPLIT>(g, vol, minmax, k, i, zl, lane, nullptr, zl2)
              : fp_ray_wave<ZV, false, REUSE, false, SPLIT>(g, vol, minmax, k, i, zl, lane, nullptr, zl2);
  wave_zstore<ZV, SPLIT>(sino + ((size_t)k * g.na + i) * g.nz, z, g.nz, acc);
}

// paired forward projection: one wavefront = ray (k, i) and, when it exists,
// the opposite ray (k2, i2) x 64*ZV z, from one pass over the taps.  Held to
// 64 VGPRs (8 waves per SIMD; a 12-byte spill) -- faster than 71 VGPRs at 7.
template <int ZV, bool REUSE, bool SPLIT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) void par_fp_pair_kernel(ParDev g, const float* __restrict__ vol,
                                                          float* __restrict__ sino, const float* __restrict__ minmax,
                                                          int clip) {
  using V = typename VecT<ZV>::T;
  const int lane = threadIdx.x & 63;
  const long long wave = wave_index(xcd_remap(blockIdx.x, gridDim.x));
  const PairItem it = fp_pair_item<ZV, SPLIT>(wave, g, lane);
  if (!it.ok) return;
  const int k = it.k, i = it.i, k2 = it.k2, i2 = it.i2, z = it.z;
  int zl, zl2;
  wave_zload<ZV, SPLIT>(z, g.nz, zl, zl2);
  const bool clip_on = clip && wave_clip_active<ZV, SPLIT>(minmax, g.nz, zl, zl2);
  if (clip_on) {
    // the clipped sample is not continuous in its position: the two rays are
    // projected one after the other (one call site keeps the register count
    // of the unpaired kernel)
    for (int h = 0; h < (k2 >= 0 ? 2 : 1); ++h) {
      const int kh = h ? k2 : k, ih = h ? i2 : i;
      const V acc = fp_ray_wave<ZV, true, REUSE, false, SPLIT>(g, vol, minmax, kh, ih, zl, lane, nullptr, zl2);
      wave_zstore<ZV, SPLIT>(sino + ((size_t)kh * g.na + ih) * g.nz, z, g.nz, acc);
    }
    return;
  }
  V acc, acc2 = vzero<ZV>();
  if (k2 >= 0)
    acc = fp_ray_wave<ZV, false, REUSE, true, SPLIT>(g, vol, minmax, k, i, zl, lane, &acc2, zl2);
  else
    acc = fp_ray_wave<ZV, false, REUSE, false, SPLIT>(g, vol, minmax, k, i, zl, lane, nullptr, zl2);
  wave_zstore<ZV, SPLIT>(sino + ((size_t)k * g.na + i) * g.nz, z, g.nz, acc);
  if (k2 >= 0) wave_zstore<ZV, SPLIT>(sino + ((size_t)k2 * g.na + i2) * g.nz, z, g.nz, acc2);
}

// wavefront-uniform back projection: one wavefront = one 2x2 quad of voxel
// columns x 64*ZV z
template <int ZV, bool SPLIT>
__global__ __launch_bounds__(256) void par_bp_wave_kernel(ParDev g, const float* __restrict__ sino,
                                                          float* __restrict__ vol, int clipc) {
  const int lane = threadIdx.x & 63;
  const long long wave = wave_index(xcd_remap(blockIdx.x, gridDim.x));
  int r0, c0, z, zl, zl2;
  if (!bp_wave_quad<ZV, SPLIT>(wave, g, lane, r0, c0, z)) return;
  wave_zload<ZV, SPLIT>(z, g.nz, zl, zl2);
  unsigned inrange, active;
  bp_quad_masks(g, r0, c0, clipc, inrange, active);
  typename VecT<ZV>::T acc[4];
  bp_quad_wave<ZV, SPLIT>(g, sino, r0, c0, active, zl, lane, acc, zl2);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    if (!((inrange >> q) & 1u)) continue;
    float* outp = vol + ((size_t)(r0 + (q >> 1)) * g.n + c0 + (q & 1)) * g.nz;
    wave_zstore<ZV, SPLIT>(outp, z, g.nz, ((active >> q) & 1u) ? acc[q] : vzero<ZV>());
  }
}

__global__ __launch_bounds__(256) void par_fp_incl_kernel(ParDev g, const float* __restrict__ vol,
                                                          float* __restrict__ sino) {
  const long long total = (long long)g.na * g.n * g.nz;
  const long long tid = (long long)xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (tid >= total) return;
  const int j = (int)(tid % g.nz);
  const long long ray = tid / g.nz;
  const int k = (int)(ray % g.n), i = (int)(ray / g.n);
  sino[((size_t)k * g.na + i) * g.nz + j] = fp_ray_incl(g, vol, k, i, j);
}

template <int ZV, bool CLIPC>
__global__ __launch_bounds__(256) void par_bp_kernel(ParDev g, const float* __restrict__ sino,
                                                     float* __restrict__ vol) {
  const int nzv = g.nz / ZV;
  const long long total = (long long)g.n * g.n * nzv;
  const long long tid = (long long)xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (tid >= total) return;
  const int z = (int)(tid % nzv) * ZV;
  const int pix = (int)(tid / nzv);
  float* outp = vol + (size_t)pix * g.nz + z;
  if (CLIPC && g.outside[pix]) {
    vstore<ZV>(outp, vzero<ZV>());
    return;
  }
  vstore<ZV>(outp, bp_voxel<ZV>(g, sino, pix, z));
}

__global__ __launch_bounds__(256) void par_bp_incl_kernel(ParDev g, const float* __restrict__ sino,
                                                          float* __restrict__ vol, int clipc) {
  const long long total = (long long)g.n * g.n * g.nz;
  const long long tid = (long long)xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (tid >= total) return;
  const int z = (int)(tid % g.nz);
  const int pix = (int)(tid / g.nz);
  vol[(size_t)pix * g.nz + z] = (clipc && g.outside[pix]) ? 0.0f : bp_voxel_incl(g, sino, pix, z);
}

template <int ZV>
__global__ __launch_bounds__(256) void par_fpT_kernel(ParDev g, const float* __restrict__ sino,
                                                      float* __restrict__ vol) {
  const int nzv = g.nz / ZV;
  const long long total = (long long)g.n * g.n * nzv;
  const long long tid = (long long)xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (tid >= total) return;
  const int z = (int)(tid % nzv) * ZV;
  const int pix = (int)(tid / nzv);
  vstore<ZV>(vol + (size_t)pix * g.nz + z, fpT_voxel<ZV>(g, sino, pix, z));
}

// wavefront-uniform exact transposes: one wavefront = one voxel column (A^T)
// or one detector bin (B^T) x 64*ZV z; bitwise equal to the per-lane kernels
template <int ZV>
__global__ __launch_bounds__(256) void par_fpT_wave_kernel(ParDev g, const float* __restrict__ sino,
                                                           float* __restrict__ vol) {
  const int lane = threadIdx.x & 63;
  const long long w = wave_index(xcd_remap(blockIdx.x, gridDim.x));
  const int nch = wave_zchunks(g.nz, ZV);
  if (w >= (long long)g.n * g.n * nch) return;   // whole wavefronts
  const int pix = (int)(w / nch);
  const int z = (int)(w % nch) * 64 * ZV + lane * ZV;
  int zl, zl2;
  wave_zload<ZV, false>(z, g.nz, zl, zl2);
  const typename VecT<ZV>::T acc = fpT_voxel_wave<ZV>(g, sino, pix, zl, lane);
  wave_zstore<ZV, false>(vol + (size_t)pix * g.nz, z, g.nz, acc);
}

template <int ZV>
__global__ __launch_bounds__(256) void par_bpT_wave_kernel(ParDev g, const float* __restrict__ vol,
                                                           float* __restrict__ sino, int clipc) {
  const int lane = threadIdx.x & 63;
  const long long w = wave_index(xcd_remap(blockIdx.x, gridDim.x));
  const int nch = wave_zchunks(g.nz, ZV);
  if (w >= (long long)g.n * g.na * nch) return;  // whole wavefronts
  // consecutive wavefronts: neighbouring bins of one angle (overlapping strips)
  const long long bin = w / nch;
  const int k = (int)(bin % g.n), i = (int)(bin / g.n);
  const int z = (int)(w % nch) * 64 * ZV + lane * ZV;
  int zl, zl2;
  wave_zload<ZV, false>(z, g.nz, zl, zl2);
  const typename VecT<ZV>::T acc = bpT_bin_wave<ZV>(g, vol, k, i, zl, lane, clipc);
  wave_zstore<ZV, false>(sino + ((size_t)k * g.na + i) * g.nz, z, g.nz, acc);
}

// A^T on voxel quads: one wavefront = one 2x2 quad of voxel columns x 64*ZV z
// (the back projection's quad tiles); bitwise equal to par_fpT_wave_kernel.
// Held to 6 waves per SIMD (a small spill in the geometry phase): the kernel
// is bound by latency, and 6 waves measured faster than 4 without the spill.
template <int ZV>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6))) void par_fpT_quad_kernel(ParDev g, const float* __restrict__ sino,
                                                           float* __restrict__ vol) {
  const int lane = threadIdx.x & 63;
  const long long wave = wave_index(xcd_remap(blockIdx.x, gridDim.x));
  int r0, c0, z, zl, zl2;
  if (!bp_wave_quad<ZV, false>(wave, g, lane, r0, c0, z)) return;
  wave_zload<ZV, false>(z, g.nz, zl, zl2);
  unsigned inrange, active;
  bp_quad_masks(g, r0, c0, 0, inrange, active);
  typename VecT<ZV>::T acc[4];
  fpT_quad_wave<ZV>(g, sino, r0, c0, inrange, zl, lane, acc);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    if (!((inrange >> q) & 1u)) continue;
    wave_zstore<ZV, false>(vol + ((size_t)(r0 + (q >> 1)) * g.n + c0 + (q & 1)) * g.nz, z, g.nz, acc[q]);
  }
}

// B^T paired like the forward projection: one wavefront = bin (k, i) and,
// when it exists, the opposite bin (k2, i2) x 64*ZV z, from one pass over the
// strip's voxel columns (the forward's work units and tile order)
template <int ZV>
__global__ __launch_bounds__(256) void par_bpT_pair_kernel(ParDev g, const float* __restrict__ vol,
                                                           float* __restrict__ sino, int clipc) {
  using V = typename VecT<ZV>::T;
  const int lane = threadIdx.x & 63;
  const long long wave = wave_index(xcd_remap(blockIdx.x, gridDim.x));
  const PairItem it = fp_pair_item<ZV, false>(wave, g, lane);
  if (!it.ok) return;
  int zl, zl2;
  wave_zload<ZV, false>(it.z, g.nz, zl, zl2);
  if (it.k2 < 0) {
    const V acc = bpT_bin_wave<ZV>(g, vol, it.k, it.i, zl, lane, clipc);
    wave_zstore<ZV, false>(sino + ((size_t)it.k * g.na + it.i) * g.nz, it.z, g.nz, acc);
    return;
  }
  V acc, acc2;
  bpT_pair_wave<ZV>(g, vol, it.k, it.i, it.k2, it.i2, zl, lane, clipc, acc, acc2);
  wave_zstore<ZV, false>(sino + ((size_t)it.k * g.na + it.i) * g.nz, it.z, g.nz, acc);
  wave_zstore<ZV, false>(sino + ((size_t)it.k2 * g.na + it.i2) * g.nz, it.z, g.nz, acc2);
}

__global__ __launch_bounds__(256) void par_fpT_incl_kernel(ParDev g, const float* __restrict__ sino,
                                                           float* __restrict__ vol) {
  const long long total = (long long)g.n * g.n * g.nz;
  const long long tid = (long long)xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (tid >= total) return;
  const int z = (int)(tid % g.nz);
  const int pix = (int)(tid / g.nz);
  vol[(size_t)pix * g.nz + z] = fpT_voxel_incl(g, sino, pix, z);
}

template <int ZV>
__global__ __launch_bounds__(256) void par_bpT_kernel(ParDev g, const float* __restrict__ vol,
                                                      float* __restrict__ sino, int clipc) {
  const int nzv = g.nz / ZV;
  const long long total = (long long)g.n * g.na * nzv;
  const long long tid = (long long)xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (tid >= total) return;
  const int z = (int)(tid % nzv) * ZV;
  const long long bin = tid / nzv;
  const int i = (int)(bin % g.na), k = (int)(bin / g.na);
  vstore<ZV>(sino + ((size_t)k * g.na + i) * g.nz + z, bpT_bin<ZV>(g, vol, k, i, z, clipc));
}

__global__ __launch_bounds__(256) void par_bpT_incl_kernel(ParDev g, const float* __restrict__ vol,
                                                           float* __restrict__ sino, int clipc) {
  const long long total = (long long)g.n * g.na * g.nz;
  const long long tid = (long long)xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (tid >= total) return;
  const int jz = (int)(tid % g.nz);
  const long long bin = tid / g.nz;
  const int i = (int)(bin % g.na), k = (int)(bin / g.na);
  sino[((size_t)k * g.na + i) * g.nz + jz] = bpT_bin_incl(g, vol, k, i, jz, clipc);
}

}  // namespace vamx

// ===========================================================================
// C ABI
// ===========================================================================
using namespace vamx;

int vamx::par_wave_zv(const vamx_plan* p, const void* a, const void* b, int requested) {
  // z-vector width <= requested for the wavefront kernels: nz a multiple of
  // zv, buffers aligned to it, and the best lane utilisation
  // nz / (chunks * 64 * zv) (ties -> the wider vector, fewer waves); 0 (use
  // the per-lane kernels) when even the best leaves half the lanes idle
  if ((p->flags & VAMX_PAR_PER_LANE_GEOMETRY) || p->par.incl || p->par.n < 2 || p->par.occ) return 0;
  return wave_zv_choice(p->par.nz, (uintptr_t)a | (uintptr_t)b, requested);
}

int vamx::wave_zv_choice(int nz, uintptr_t align_bits, int requested) {
  int best = 0;
  double best_util = 0.0;
  for (int zv = 1; zv <= requested && zv <= 4; zv <<= 1) {
    if (nz % zv != 0 || align_bits % (4 * zv) != 0) continue;
    const double util = (double)nz / ((double)wave_zchunks(nz, zv) * 64.0 * zv);
    if (util >= best_util) {
      best_util = util;
      best = zv;
    }
  }
  return best_util >= 0.5 ? best : 0;
}

int vamx::par_fp_wave_zv(const vamx_plan* p, const void* vol, const void* sino) {
  // the forward's wavefront z width, 0 for the per-lane kernel.  At one z
  // per lane a wavefront issues one 4-byte tap load per lane, and the forward
  // is bound by load issue, not bytes; the per-lane kernel with four z per
  // thread (several rays per wavefront, 16-byte tap loads) is then faster
  // up to nz = 56.  At 256^2 x 360 with absorption, nz 32 / 40 / 48 measured
  // 0.66 / 0.82 / 0.93 ms per-lane against 1.23 / 1.25 / 1.22 ms for the
  // wavefront kernel, and they tie at nz = 64.
  const int wz = par_wave_zv(p, vol, sino, p->par.fp_zv);
  if (wz == 1 && p->par.nz <= 56 && par_use_vec4(p, vol, sino)) return 0;
  return wz;
}

bool vamx::par_wave_split(const vamx_plan* p, int zv) {
  // the split z layout (ZV = 2, lanes own z and z + 64) where the contiguous
  // choice is one z per lane: half the wavefronts at the same lane
  // utilisation when the 64-z chunk count is even
  if (zv != 1 || !p->par.z_split) return false;
  const int ch = wave_zchunks(p->par.nz, 1);
  return ch % 2 == 0;
}

bool vamx::par_use_vec2(const vamx_plan* p, const void* a, const void* b) {
  return (p->par.nz % 2 == 0) && !p->par.incl && ((uintptr_t)a % 8 == 0) && ((uintptr_t)b % 8 == 0);
}

bool vamx::par_use_vec4(const vamx_plan* p, const void* a, const void* b) {
  return (p->par.nz % 4 == 0) && !p->par.incl && ((uintptr_t)a % 16 == 0) && ((uintptr_t)b % 16 == 0);
}

// Opposite-angle work units of the paired forward projection (see ParDev):
// angle b is angle a + 180 deg when its cos / sin are the negated ones to
// 1e-9; the detector shift D = 2 det_offset must be an integer.  Units are
// (primary, partner) with the lower index primary, or (angle, -1); empty when
// nothing pairs.  npaired_rays rays of a primary angle have a partner ray;
// the partner angle's rays without one start at first2.
void vamx::par_pair_units(const std::vector<double>& cs, int na, int n, int c, double D, std::vector<int>& units,
                          int& npaired_rays, int& first2) {
  units.clear();
  npaired_rays = 0;
  first2 = 0;
  if (n >= 2 && D == std::nearbyint(D) && std::fabs(D) < n) {
    const int Di = (int)D;
    const int lo2 = std::max(0, 2 * c + Di - (n - 1)), hi2 = std::min(n - 1, 2 * c + Di);
    npaired_rays = hi2 >= lo2 ? hi2 - lo2 + 1 : 0;
    first2 = lo2 > 0 ? 0 : hi2 + 1;
    std::vector<int> order(na);
    std::vector<double> phase(na);
    for (int i = 0; i < na; ++i) {
      order[i] = i;
      phase[i] = std::atan2(cs[2 * i + 1], cs[2 * i]);
    }
    std::sort(order.begin(), order.end(), [&](int x, int y) { return phase[x] < phase[y]; });
    std::vector<int> partner(na, -1);
    auto opposite = [&](int i, int j) {
      return std::fabs(cs[2 * i] + cs[2 * j]) < 1e-9 && std::fabs(cs[2 * i + 1] + cs[2 * j + 1]) < 1e-9;
    };
    // search the phases within 1e-7 of `want` for an unpaired opposite angle
    auto find = [&](int i, double want) {
      auto it = std::lower_bound(order.begin(), order.end(), want - 1e-7,
                                 [&](int x, double v) { return phase[x] < v; });
      for (; it != order.end() && phase[*it] <= want + 1e-7; ++it)
        if (*it != i && partner[*it] < 0 && opposite(i, *it)) return *it;
      return -1;
    };
    for (int i = 0; i < na; ++i) {
      if (partner[i] >= 0) continue;
      const double want = phase[i] > 0.0 ? phase[i] - M_PI : phase[i] + M_PI;
      int j = find(i, want);
      if (j < 0 && want + M_PI < 1e-6) j = find(i, want + 2 * M_PI);   // near -pi: also look near +pi
      if (j < 0 && M_PI - want < 1e-6) j = find(i, want - 2 * M_PI);
      if (j >= 0) {
        partner[i] = j;
        partner[j] = i;
      }
    }
    bool any = false;
    for (int i = 0; i < na; ++i) {
      if (partner[i] >= 0 && partner[i] < i) continue;   // the primary is the lower index
      units.push_back(i);
      units.push_back(partner[i]);
      any |= partner[i] >= 0;
    }
    if (!any || npaired_rays == 0) units.clear();
  }
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
  if (geom->occlusion && geom->incl_deg != 0.0)
    return set_error(VAMX_ERR_ARG, "vamx_par_create: occlusion shadowing needs zero inclination");
  // the wavefront kernels address a cell's two volume rows, and a voxel
  // quad's four sinogram rows, through 32-bit buffer offsets
  if (4.0 * geom->nz * ((double)geom->n + 2) >= 2147483648.0 ||
      4.0 * geom->nz * (3.0 * geom->n_angles + 1) >= 2147483648.0)
    return set_error(VAMX_ERR_ARG, "vamx_par_create: n * nz or n_angles * nz too large for 32-bit row offsets");
  // the forward kernels pack a cell index below 2^28 with its parity bits
  if (geom->n >= 16384) return set_error(VAMX_ERR_ARG, "vamx_par_create: n must be < 16384");
  // the back projection's sinogram row offsets are 32-bit element counts
  if ((double)geom->n * geom->n_angles * geom->nz >= 4294967296.0)
    return set_error(VAMX_ERR_ARG, "vamx_par_create: n * n_angles * nz too large for 32-bit sinogram offsets");
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
  std::vector<int> units;
  int npaired_rays = 0, first2 = 0;
  const double D = 2.0 * geom->det_offset;
  if (geom->incl_deg == 0.0 && !geom->occlusion)
    par_pair_units(cs, na, n, c, D, units, npaired_rays, first2);
  g.pair_ok = units.empty() ? 0 : 1;
  g.n_units = (int)(units.size() / 2);
  g.pair_D = (int)D;
  g.pair_u = n - npaired_rays;
  g.pair_first2 = first2;
  g.fp_pair = 1;
  g.z_split = 1;
  // defaults: the fastest measured work order at 256^3 x 360 angles on
  // MI355X (10 angle pairs x 4 rays, paired: 10 divides the 180 pairs)
  g.fp_ta = 10;
  g.fp_tk = 4;
  g.fp_zv = 4;
  g.fp_reuse = 1;
  g.bp_tile = 8;
  g.bp_zv = 4;
  p->minmax_parts = 256;
  hipError_t e = hipSuccess;
  e = hipMalloc((void**)&p->d_mats, mats.size() * sizeof(double));
  if (e == hipSuccess) e = hipMalloc((void**)&p->d_cs, cs.size() * sizeof(double));
  if (e == hipSuccess) e = hipMalloc((void**)&p->d_outside, (size_t)n * n);
  if (e == hipSuccess) e = hipMalloc((void**)&p->d_minmax_part, (size_t)p->minmax_parts * 2 * nz * sizeof(float));
  if (e == hipSuccess) e = hipMalloc((void**)&p->d_minmax, ((size_t)2 * nz + 4) * sizeof(float));
  if (e == hipSuccess) e = hipMemcpy(p->d_mats, mats.data(), mats.size() * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(p->d_cs, cs.data(), cs.size() * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(p->d_outside, geom->outside_mask, (size_t)n * n, hipMemcpyHostToDevice);
  if (e == hipSuccess && !units.empty()) {
    e = hipMalloc((void**)&p->d_units, units.size() * sizeof(int));
    if (e == hipSuccess) e = hipMemcpy(p->d_units, units.data(), units.size() * sizeof(int), hipMemcpyHostToDevice);
  }
  if (e != hipSuccess) {
    vamx_destroy(p);
    return hip_check(e, "vamx_par_create: allocation / upload");
  }
  g.mats = p->d_mats;
  g.cs = p->d_cs;
  g.units = p->d_units;
  g.outside = p->d_outside;
  g.occ = nullptr;
  g.occ_active = nullptr;
  if (geom->occlusion) {
    // occlusion sinogram of the insert, computed once like the reference's
    // constructor does (Projector3DParallel.py:72-75)
    const size_t nvox = (size_t)n * n * nz, nsino = (size_t)n * na * nz;
    uint8_t* d_ins = nullptr;
    e = hipMalloc((void**)&d_ins, nvox);
    if (e == hipSuccess) e = hipMalloc((void**)&p->d_occ, nsino * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void**)&p->d_occ_active, (size_t)na * nz);
    if (e == hipSuccess) e = hipMemcpy(d_ins, geom->occlusion, nvox, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
      hipLaunchKernelGGL(occ_sino_kernel, dim3(blocks_for((long long)nsino, 256)), dim3(256), 0, 0, g, d_ins,
                         p->d_occ);
      hipLaunchKernelGGL(occ_active_kernel, dim3(blocks_for((long long)na * nz, 256)), dim3(256), 0, 0, g,
                         p->d_occ, p->d_occ_active);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (d_ins) (void)hipFree(d_ins);
    if (e != hipSuccess) {
      vamx_destroy(p);
      return hip_check(e, "vamx_par_create: occlusion sinogram");
    }
    g.occ = p->d_occ;
    g.occ_active = p->d_occ_active;
  }
  *out_plan = p;
  return VAMX_OK;
}

namespace vamx {
int par_slice_minmax(vamx_plan* p, const float* vol, hipStream_t s) {
  const ParDev& g = p->par;
  hipLaunchKernelGGL(slice_minmax_part, dim3(p->minmax_parts), dim3(1024), 0, s, vol, g.n * g.n, g.nz,
                     p->minmax_parts, p->d_minmax_part);
  hipLaunchKernelGGL(slice_minmax_final, dim3((g.nz + 63) / 64), dim3(1024), 0, s, p->d_minmax_part, g.nz,
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
  int wz = par_fp_wave_zv(p, vol, sino);
  const bool split = par_wave_split(p, wz);
  if (split) wz = 2;
  if (wz && g.fp_pair && g.pair_ok) {
    const long long threads = fp_pair_tiled(g) * wave_zchunks(g.nz, wz) * 64;
    const dim3 grid(blocks_for(threads, 256)), block(256);
    const int clip = warp_clip ? 1 : 0;
#define VAMX_FPP(Z, R, S) \
  hipLaunchKernelGGL((par_fp_pair_kernel<Z, R, S>), grid, block, 0, s, g, vol, sino, p->d_minmax, clip)
    if (g.fp_reuse) {
      if (split) VAMX_FPP(2, true, true);
      else if (wz == 4) VAMX_FPP(4, true, false);
      else if (wz == 2) VAMX_FPP(2, true, false);
      else VAMX_FPP(1, true, false);
    } else {
      if (split) VAMX_FPP(2, false, true);
      else if (wz == 4) VAMX_FPP(4, false, false);
      else if (wz == 2) VAMX_FPP(2, false, false);
      else VAMX_FPP(1, false, false);
    }
#undef VAMX_FPP
    VAMX_HIP(hipGetLastError());
    return VAMX_OK;
  }
  if (wz) {
    const long long threads = fp_tiled_rays(g) * wave_zchunks(g.nz, wz) * 64;
    const dim3 grid(blocks_for(threads, 256)), block(256);
    const int clip = warp_clip ? 1 : 0;
#define VAMX_FPW(Z, R, S) \
  hipLaunchKernelGGL((par_fp_wave_kernel<Z, R, S>), grid, block, 0, s, g, vol, sino, p->d_minmax, clip)
    if (g.fp_reuse) {
      if (split) VAMX_FPW(2, true, true);
      else if (wz == 4) VAMX_FPW(4, true, false);
      else if (wz == 2) VAMX_FPW(2, true, false);
      else VAMX_FPW(1, true, false);
    } else {
      if (split) VAMX_FPW(2, false, true);
      else if (wz == 4) VAMX_FPW(4, false, false);
      else if (wz == 2) VAMX_FPW(2, false, false);
      else VAMX_FPW(1, false, false);
    }
#undef VAMX_FPW
    VAMX_HIP(hipGetLastError());
    return VAMX_OK;
  }
  const bool v4 = par_use_vec4(p, vol, sino);
  const bool v2 = !v4 && par_use_vec2(p, vol, sino);
  const int zv = v4 ? 4 : v2 ? 2 : 1;
  const long long total = (long long)g.na * g.n * (g.nz / zv);
  const dim3 grid(blocks_for(total, 256)), block(256);
  if (v4) {
    if (warp_clip)
      hipLaunchKernelGGL((par_fp_kernel<4, true>), grid, block, 0, s, g, vol, sino, p->d_minmax);
    else
      hipLaunchKernelGGL((par_fp_kernel<4, false>), grid, block, 0, s, g, vol, sino, p->d_minmax);
  } else if (v2) {
    if (warp_clip)
      hipLaunchKernelGGL((par_fp_kernel<2, true>), grid, block, 0, s, g, vol, sino, p->d_minmax);
    else
      hipLaunchKernelGGL((par_fp_kernel<2, false>), grid, block, 0, s, g, vol, sino, p->d_minmax);
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
  int wz = par_wave_zv(p, sino, vol, g.bp_zv);
  const bool split = par_wave_split(p, wz);
  if (split) wz = 2;
  if (wz) {
    const long long threads = bp_tiled_quads(g) * wave_zchunks(g.nz, wz) * 64;
    const dim3 grid(blocks_for(threads, 256)), block(256);
    const int cc = clipc ? 1 : 0;
    if (split)
      hipLaunchKernelGGL((par_bp_wave_kernel<2, true>), grid, block, 0, s, g, sino, vol, cc);
    else if (wz == 4)
      hipLaunchKernelGGL((par_bp_wave_kernel<4, false>), grid, block, 0, s, g, sino, vol, cc);
    else if (wz == 2)
      hipLaunchKernelGGL((par_bp_wave_kernel<2, false>), grid, block, 0, s, g, sino, vol, cc);
    else
      hipLaunchKernelGGL((par_bp_wave_kernel<1, false>), grid, block, 0, s, g, sino, vol, cc);
    VAMX_HIP(hipGetLastError());
    return VAMX_OK;
  }
  const bool v4 = par_use_vec4(p, sino, vol);
  const bool v2 = !v4 && par_use_vec2(p, sino, vol);
  const int zv = v4 ? 4 : v2 ? 2 : 1;
  const long long total = (long long)g.n * g.n * (g.nz / zv);
  const dim3 grid(blocks_for(total, 256)), block(256);
  if (v4) {
    if (clipc)
      hipLaunchKernelGGL((par_bp_kernel<4, true>), grid, block, 0, s, g, sino, vol);
    else
      hipLaunchKernelGGL((par_bp_kernel<4, false>), grid, block, 0, s, g, sino, vol);
  } else if (v2) {
    if (clipc)
      hipLaunchKernelGGL((par_bp_kernel<2, true>), grid, block, 0, s, g, sino, vol);
    else
      hipLaunchKernelGGL((par_bp_kernel<2, false>), grid, block, 0, s, g, sino, vol);
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
  if (plan->par.occ) return set_error(VAMX_ERR_STATE, "vamx_par_forward_adjoint: not available with occlusion");
  const ParDev& g = plan->par;
  hipStream_t s = (hipStream_t)stream;
  const int wz = (!g.incl && !(plan->flags & VAMX_PAR_PER_LANE_GEOMETRY)) ? par_wave_zv(plan, sino, vol, 4) : 0;
  if (wz && g.n >= 2) {
    const long long threads = bp_tiled_quads(g) * wave_zchunks(g.nz, wz) * 64;
    const dim3 grid(blocks_for(threads, 256)), block(256);
    if (wz == 4) hipLaunchKernelGGL((par_fpT_quad_kernel<4>), grid, block, 0, s, g, sino, vol);
    else if (wz == 2) hipLaunchKernelGGL((par_fpT_quad_kernel<2>), grid, block, 0, s, g, sino, vol);
    else hipLaunchKernelGGL((par_fpT_quad_kernel<1>), grid, block, 0, s, g, sino, vol);
  } else if (wz) {
    const long long threads = (long long)g.n * g.n * wave_zchunks(g.nz, wz) * 64;
    const dim3 grid(blocks_for(threads, 256)), block(256);
    if (wz == 4) hipLaunchKernelGGL((par_fpT_wave_kernel<4>), grid, block, 0, s, g, sino, vol);
    else if (wz == 2) hipLaunchKernelGGL((par_fpT_wave_kernel<2>), grid, block, 0, s, g, sino, vol);
    else hipLaunchKernelGGL((par_fpT_wave_kernel<1>), grid, block, 0, s, g, sino, vol);
  } else if (g.incl) {
    const long long total = (long long)g.n * g.n * g.nz;
    hipLaunchKernelGGL(par_fpT_incl_kernel, dim3(blocks_for(total, 256)), dim3(256), 0, s, g, sino, vol);
  } else if (par_use_vec4(plan, sino, vol)) {
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
  if (plan->par.occ) return set_error(VAMX_ERR_STATE, "vamx_par_backward_adjoint: not available with occlusion");
  const ParDev& g = plan->par;
  hipStream_t s = (hipStream_t)stream;
  const int clipc = (plan->flags & VAMX_PAR_NO_CIRCLE_CLIP) ? 0 : 1;
  const int wz = (!g.incl && !(plan->flags & VAMX_PAR_PER_LANE_GEOMETRY)) ? par_wave_zv(plan, vol, sino, 4) : 0;
  if (wz && g.fp_pair && g.pair_ok) {
    const long long threads = fp_pair_tiled(g) * wave_zchunks(g.nz, wz) * 64;
    const dim3 grid(blocks_for(threads, 256)), block(256);
    if (wz == 4) hipLaunchKernelGGL((par_bpT_pair_kernel<4>), grid, block, 0, s, g, vol, sino, clipc);
    else if (wz == 2) hipLaunchKernelGGL((par_bpT_pair_kernel<2>), grid, block, 0, s, g, vol, sino, clipc);
    else hipLaunchKernelGGL((par_bpT_pair_kernel<1>), grid, block, 0, s, g, vol, sino, clipc);
  } else if (wz) {
    const long long threads = (long long)g.n * g.na * wave_zchunks(g.nz, wz) * 64;
    const dim3 grid(blocks_for(threads, 256)), block(256);
    if (wz == 4) hipLaunchKernelGGL((par_bpT_wave_kernel<4>), grid, block, 0, s, g, vol, sino, clipc);
    else if (wz == 2) hipLaunchKernelGGL((par_bpT_wave_kernel<2>), grid, block, 0, s, g, vol, sino, clipc);
    else hipLaunchKernelGGL((par_bpT_wave_kernel<1>), grid, block, 0, s, g, vol, sino, clipc);
  } else if (g.incl) {
    const long long total = (long long)g.n * g.na * g.nz;
    hipLaunchKernelGGL(par_bpT_incl_kernel, dim3(blocks_for(total, 256)), dim3(256), 0, s, g, vol, sino, clipc);
  } else if (par_use_vec4(plan, vol, sino)) {
    const long long total = (long long)g.n * g.na * (g.nz / 4);
    hipLaunchKernelGGL((par_bpT_kernel<4>), dim3(blocks_for(total, 256)), dim3(256), 0, s, g, vol, sino, clipc);
  } else {
    const long long total = (long long)g.n * g.na * g.nz;
    hipLaunchKernelGGL((par_bpT_kernel<1>), dim3(blocks_for(total, 256)), dim3(256), 0, s, g, vol, sino, clipc);
  }
  VAMX_HIP(hipGetLastError());
  return VAMX_OK;
}

extern "C" int vamx_par_set_tuning(vamx_plan* plan, const vamx_par_tuning* t) {
  if (!check_plan(plan, PLAN_PARALLEL)) return set_error(VAMX_ERR_STATE, "vamx_par_set_tuning: not a parallel plan");
  if (!t) return set_error(VAMX_ERR_ARG, "vamx_par_set_tuning: null tuning");
  auto pow2 = [](int v) { return v == 1 || v == 2 || v == 4; };
  if (t->fp_tile_angles < 1 || t->fp_tile_rays < 1 || t->bp_tile < 1 || !pow2(t->fp_zv) || !pow2(t->bp_zv) ||
      t->fp_tile_angles * t->fp_tile_rays > (1 << 20) || t->bp_tile > 1024 || t->fp_reuse < 0 || t->fp_reuse > 1 ||
      t->fp_pair < 0 || t->fp_pair > 1 || t->z_split < 0 || t->z_split > 1)
    return set_error(VAMX_ERR_ARG, "vamx_par_set_tuning: tiles must be >= 1, z widths 1, 2 or 4, fp_reuse, "
                                   "fp_pair and z_split 0 or 1");
  plan->par.fp_ta = t->fp_tile_angles;
  plan->par.fp_tk = t->fp_tile_rays;
  plan->par.fp_zv = t->fp_zv;
  plan->par.fp_reuse = t->fp_reuse;
  plan->par.bp_tile = t->bp_tile;
  plan->par.bp_zv = t->bp_zv;
  plan->par.fp_pair = t->fp_pair;
  plan->par.z_split = t->z_split;
  return VAMX_OK;
}

extern "C" int vamx_par_get_tuning(vamx_plan* plan, vamx_par_tuning* t) {
  if (!check_plan(plan, PLAN_PARALLEL)) return set_error(VAMX_ERR_STATE, "vamx_par_get_tuning: not a parallel plan");
  if (!t) return set_error(VAMX_ERR_ARG, "vamx_par_get_tuning: null tuning");
  t->fp_tile_angles = plan->par.fp_ta;
  t->fp_tile_rays = plan->par.fp_tk;
  t->fp_zv = plan->par.fp_zv;
  t->fp_reuse = plan->par.fp_reuse;
  t->bp_tile = plan->par.bp_tile;
  t->bp_zv = plan->par.bp_zv;
  t->fp_pair = plan->par.fp_pair;
  t->z_split = plan->par.z_split;
  return VAMX_OK;
}
