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

// ---------------------------------------------------------------------------
// The same normalisation for k arrays packed one after another in one buffer
// (segment s is v[off[s], off[s+1]), each clipped to [0, n)), as the batched
// OSMO runs OSMO.py:60-69 on every problem: one grid pass for all the maxima
// and one for all the divisions.  The maximum is exact (atomicMax of ordered
// keys) and a NaN anywhere in a segment makes its maximum NaN, as np.max and
// torch.amax do.
// ---------------------------------------------------------------------------
__device__ __forceinline__ unsigned max_key(float v) {
  if (v != v) return 0xffffffffu;   // every NaN above +inf
  const unsigned b = __float_as_uint(v);
  return (b >> 31) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float max_value(unsigned k) {
  return __uint_as_float((k >> 31) ? (k & 0x7fffffffu) : ~k);
}

// One block per chunk of kSegChunk elements of the whole buffer, so every
// block has the same work whatever the segment sizes; a chunk that straddles
// segment ends walks them in order (block-uniform loop).
constexpr int kSegChunk = 8192;

// segment s's range, clipped to [0, n)
__device__ __forceinline__ void seg_range(const long long* off, int s, long long n, long long& b, long long& e) {
  b = min(max(off[s], 0ll), n);
  e = max(b, min(off[s + 1], n));
}

// the last segment whose start is <= pos (offsets non-decreasing)
__device__ __forceinline__ int seg_find(const long long* off, int k, long long pos) {
  int lo = 0, hi = k - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= pos) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__global__ void seg_max_init(unsigned* key, int k) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < k) key[s] = 0u;   // below every key of a float
}

__global__ __launch_bounds__(256) void seg_max(const float* __restrict__ v, long long n,
                                               const long long* __restrict__ off, int k,
                                               unsigned* __restrict__ key) {
  __shared__ unsigned s_m[4];
  long long pos = (long long)blockIdx.x * kSegChunk;
  const long long cend = min(pos + kSegChunk, n);
  while (pos < cend) {
    const int s = seg_find(off, k, pos);
    long long b, e;
    seg_range(off, s, n, b, e);
    if (e <= pos) break;   // pos in no segment: offsets not as documented
    e = min(e, cend);
    unsigned m = 0u;
#pragma unroll 4
    for (long long i = pos + threadIdx.x; i < e; i += 256) m = max(m, max_key(v[i]));
    for (int o = 32; o > 0; o >>= 1) m = max(m, (unsigned)__shfl_xor((int)m, o));
    if (__lane_id() == 0) s_m[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
      const unsigned bm = max(max(s_m[0], s_m[1]), max(s_m[2], s_m[3]));
      if (bm) atomicMax(&key[s], bm);
    }
    __syncthreads();   // s_m is reused by the next segment of this chunk
    pos = e;
  }
}

__global__ __launch_bounds__(256) void seg_scale_clamp(float* __restrict__ v, long long n,
                                                       const long long* __restrict__ off, int k,
                                                       const unsigned* __restrict__ key, float lo) {
  long long pos = (long long)blockIdx.x * kSegChunk;
  const long long cend = min(pos + kSegChunk, n);
  while (pos < cend) {
    const int s = seg_find(off, k, pos);
    long long b, e;
    seg_range(off, s, n, b, e);
    if (e <= pos) break;
    e = min(e, cend);
    const float mv = max_value(key[s]);
#pragma unroll 4
    for (long long i = pos + threadIdx.x; i < e; i += 256) {
      const float q = v[i] / mv;
      v[i] = q < lo ? lo : q;
    }
    pos = e;
  }
}

// ---------------------------------------------------------------------------
// The volumetric error rate (metrics.py:47-87): #(void voxels with dose >=
// the least gel dose) / (#gel + #void), as two grid passes over the dose with
// integer atomics (order-independent, so the value does not depend on the
// launch).  A NaN gel dose makes the least gel dose NaN and the count 0, as
// the reference's np.min and >= comparison do.
// ---------------------------------------------------------------------------
struct VerWork {
  unsigned min_key;         // least gel dose as an ordered key
  unsigned nan_gel;         // a gel dose is NaN
  unsigned long long count;
};

// float -> unsigned key with the float order (-0 just below +0)
__device__ __forceinline__ unsigned ver_key(float v) {
  const unsigned b = __float_as_uint(v);
  return (b >> 31) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float ver_value(unsigned k) {
  return __uint_as_float((k >> 31) ? (k & 0x7fffffffu) : ~k);
}

__global__ void ver_init(VerWork* w) {
  w->min_key = 0xffffffffu;
  w->nan_gel = 0u;
  w->count = 0ull;
}

// VEC: sixteen voxels per step (16 mask bytes, four float4 doses; n a multiple of 16)
template <bool VEC>
__global__ __launch_bounds__(256) void ver_min_gel(const float* __restrict__ d, const unsigned char* __restrict__ gel,
                                                   long long n, VerWork* w) {
  __shared__ unsigned s_min[4], s_nan[4];
  unsigned m = 0xffffffffu, nan = 0u;
  auto take = [&](bool g, float v) {
    if (!g) return;
    if (v != v) nan = 1u;
    else m = min(m, ver_key(v));
  };
  const long long stride = (long long)gridDim.x * blockDim.x;
  if (VEC) {
    const float4* d4 = reinterpret_cast<const float4*>(d);
    const uint4* g16 = reinterpret_cast<const uint4*>(gel);
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n / 16; i += stride) {
      const uint4 gm = g16[i];
      const unsigned gw[4] = {gm.x, gm.y, gm.z, gm.w};
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        if (!gw[q]) continue;
        const float4 v = d4[4 * i + q];
        take(gw[q] & 0xffu, v.x);
        take((gw[q] >> 8) & 0xffu, v.y);
        take((gw[q] >> 16) & 0xffu, v.z);
        take(gw[q] >> 24, v.w);
      }
    }
  } else {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) take(gel[i], d[i]);
  }
  for (int o = 32; o > 0; o >>= 1) {
    m = min(m, (unsigned)__shfl_xor((int)m, o));
    nan |= (unsigned)__shfl_xor((int)nan, o);
  }
  if (__lane_id() == 0) {
    s_min[threadIdx.x >> 6] = m;
    s_nan[threadIdx.x >> 6] = nan;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned bm = s_min[0], bn = s_nan[0];
    for (int i = 1; i < 4; ++i) {
      bm = min(bm, s_min[i]);
      bn |= s_nan[i];
    }
    if (bm != 0xffffffffu) atomicMin(&w->min_key, bm);
    if (bn) atomicOr(&w->nan_gel, 1u);
  }
}

template <bool VEC>
__global__ __launch_bounds__(256) void ver_count_void(const float* __restrict__ d,
                                                      const unsigned char* __restrict__ vd, long long n, VerWork* w) {
  __shared__ unsigned long long s_c[4];
  if (w->nan_gel || w->min_key == 0xffffffffu) return;   // grid-uniform: nothing counts
  const float mg = ver_value(w->min_key);
  unsigned long long c = 0ull;
  const long long stride = (long long)gridDim.x * blockDim.x;
  if (VEC) {
    const float4* d4 = reinterpret_cast<const float4*>(d);
    const uint4* v16 = reinterpret_cast<const uint4*>(vd);
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n / 16; i += stride) {
      const uint4 vm = v16[i];
      const unsigned vw[4] = {vm.x, vm.y, vm.z, vm.w};
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        if (!vw[q]) continue;
        const float4 v = d4[4 * i + q];
        c += ((vw[q] & 0xffu) && v.x >= mg) + (((vw[q] >> 8) & 0xffu) && v.y >= mg) +
             (((vw[q] >> 16) & 0xffu) && v.z >= mg) + ((vw[q] >> 24) && v.w >= mg);
      }
    }
  } else {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
      if (vd[i] && d[i] >= mg) ++c;
  }
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
  if (__lane_id() == 0) s_c[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned long long b = s_c[0] + s_c[1] + s_c[2] + s_c[3];
    if (b) atomicAdd(&w->count, b);
  }
}

__global__ void ver_finish(const VerWork* w, double num, double* out) { *out = (double)w->count / num; }

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

extern "C" int vamx_scale_clamp_segments(float* v, int64_t n, const int64_t* seg_off, int32_t k, float lo,
                                         void* workspace, void* stream) {
  if (!v || !seg_off || !workspace) return set_error(VAMX_ERR_ARG, "vamx_scale_clamp_segments: null argument");
  if (n < 0 || k < 1 || k > 65535)
    return set_error(VAMX_ERR_ARG, "vamx_scale_clamp_segments: need n >= 0 and 1 <= k <= 65535");
  if ((uintptr_t)workspace % 4 || (uintptr_t)seg_off % 8)
    return set_error(VAMX_ERR_ARG, "vamx_scale_clamp_segments: misaligned workspace or offsets");
  if (n == 0) return VAMX_OK;
  hipStream_t s = (hipStream_t)stream;
  unsigned* key = (unsigned*)workspace;
  const long long* off = (const long long*)seg_off;
  const long long chunks = (n + kSegChunk - 1) / kSegChunk;
  if (chunks > 0x7fffffffll) return set_error(VAMX_ERR_ARG, "vamx_scale_clamp_segments: array too large");
  const dim3 grid((unsigned)chunks);
  hipLaunchKernelGGL(seg_max_init, dim3((unsigned)((k + 255) / 256)), dim3(256), 0, s, key, (int)k);
  hipLaunchKernelGGL(seg_max, grid, dim3(256), 0, s, (const float*)v, (long long)n, off, (int)k, key);
  hipLaunchKernelGGL(seg_scale_clamp, grid, dim3(256), 0, s, v, (long long)n, off, (int)k, (const unsigned*)key, lo);
  VAMX_HIP(hipGetLastError());
  return VAMX_OK;
}

extern "C" int vamx_ver_workspace_bytes(int64_t* bytes) {
  if (!bytes) return set_error(VAMX_ERR_ARG, "vamx_ver_workspace_bytes: null argument");
  *bytes = (int64_t)sizeof(VerWork);
  return VAMX_OK;
}

extern "C" int vamx_ver(const float* dose, const uint8_t* gel, const uint8_t* void_mask, int64_t n, int64_t num,
                        double* ver, void* workspace, void* stream) {
  if (!dose || !gel || !void_mask || !ver || !workspace) return set_error(VAMX_ERR_ARG, "vamx_ver: null argument");
  if (n < 1 || num < 1) return set_error(VAMX_ERR_ARG, "vamx_ver: empty volume or no gel / void voxels");
  if ((uintptr_t)workspace % 8) return set_error(VAMX_ERR_ARG, "vamx_ver: workspace must be 8-byte aligned");
  hipStream_t s = (hipStream_t)stream;
  VerWork* w = (VerWork*)workspace;
  const bool vec = n % 16 == 0 && (uintptr_t)dose % 16 == 0 && (uintptr_t)gel % 16 == 0 &&
                   (uintptr_t)void_mask % 16 == 0;
  const dim3 grid((unsigned)ew_blocks(vec ? n / 16 : n));
  hipLaunchKernelGGL(ver_init, dim3(1), dim3(1), 0, s, w);
  const unsigned char* g8 = (const unsigned char*)gel;
  const unsigned char* v8 = (const unsigned char*)void_mask;
  if (vec) {
    hipLaunchKernelGGL(ver_min_gel<true>, grid, dim3(256), 0, s, dose, g8, (long long)n, w);
    hipLaunchKernelGGL(ver_count_void<true>, grid, dim3(256), 0, s, dose, v8, (long long)n, w);
  } else {
    hipLaunchKernelGGL(ver_min_gel<false>, grid, dim3(256), 0, s, dose, g8, (long long)n, w);
    hipLaunchKernelGGL(ver_count_void<false>, grid, dim3(256), 0, s, dose, v8, (long long)n, w);
  }
  hipLaunchKernelGGL(ver_finish, dim3(1), dim3(1), 0, s, (const VerWork*)w, (double)num, ver);
  VAMX_HIP(hipGetLastError());
  return VAMX_OK;
}
