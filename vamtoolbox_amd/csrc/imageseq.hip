// Image-sequence export for gfx950: the frames a DLP projector shows, one per
// angle, composed from the sinogram on the device.
//
// Reference: vamtoolbox/imagesequence.py, ImageSeq.__init__ (85-182) and its
// helpers (300-399):
//   flips of the u / v axes, rotation in the (u, v) plane
//   (scipy.ndimage.rotate, order 1, reshape=True, constant 0 outside),
//   size scaling (OpenCV resize, INTER_LINEAR), percentile normalisation,
//   intensity scale, truncation at 2^bit_depth - 1, cast to the bit depth's
//   unsigned type, and insertion of one or more copies of each angle's
//   (v, u) image into a zeroed uint8 frame (a later copy overwrites an earlier
//   one; wider types wrap modulo 256 on insertion, as NumPy's assignment does).
//
// Three kernels, all float64 like the reference's NumPy arrays:
//   imgseq_rotate   one thread per output sample (v fastest), flips fused into
//                   the read;
//   imgseq_resize   one thread per output sample, flips fused into the read;
//   imgseq_compose  one 256-thread block per 64 x 64 tile of one frame: the
//                   tile's samples are read along v (the sinogram's fastest
//                   axis, coalesced), converted to bytes, transposed through
//                   LDS and written along u (the frame's fastest axis) as
//                   dwords.
#include <algorithm>

#include "vamx_internal.h"

namespace vamx {

namespace {

// scipy's order-1 affine transform in 'constant' mode: a sample whose
// coordinate lies outside [0, n - 1] on either axis is 0; inside, the four
// neighbours are summed as ((v * wy) * wx) in (0,0), (0,1), (1,0), (1,1)
// order, a neighbour beyond the last row / column counting as 0.
__global__ __launch_bounds__(256) void imgseq_rotate(const double* __restrict__ in, long long nu, long long na,
                                                     long long nv, int flip_u, int flip_v, double m00, double m01,
                                                     double m10, double m11, double off0, double off1,
                                                     double* __restrict__ out, long long mu, long long mv) {
#pragma clang fp contract(off)
  const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= mu * na * mv) return;
  const long long k = tid % mv;
  const long long j = (tid / mv) % na;
  const long long i = tid / (mv * na);
  const double fi = (double)i, fk = (double)k;
  const double cy = m00 * fi + m01 * fk + off0;
  const double cx = m10 * fi + m11 * fk + off1;
  double v = 0.0;
  if (cy >= 0.0 && cy <= (double)(nu - 1) && cx >= 0.0 && cx <= (double)(nv - 1)) {
    const double y0f = floor(cy), x0f = floor(cx);
    const double ty = cy - y0f, tx = cx - x0f;
    const long long y0 = (long long)y0f, x0 = (long long)x0f;
    const double wy[2] = {1.0 - ty, ty}, wx[2] = {1.0 - tx, tx};
    double t = 0.0;
#pragma unroll
    for (int a = 0; a < 2; ++a) {
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        const long long yy = y0 + a, xx = x0 + b;
        double s = 0.0;
        if (yy < nu && xx < nv) {
          const long long ys = flip_u ? nu - 1 - yy : yy, xs = flip_v ? nv - 1 - xx : xx;
          s = in[(ys * na + j) * nv + xs];
        }
        t += (s * wy[a]) * wx[b];
      }
    }
    v = t;
  }
  out[tid] = v;
}

// OpenCV INTER_LINEAR: destination index d samples (d + 0.5) * (src / dst) -
// 0.5; the integer part is clamped to [0, src - 1] with a zero fraction at
// either edge.
__device__ __forceinline__ void linear_tap(long long d, long long n_src, long long n_dst, long long& i0,
                                           long long& i1, double& t) {
#pragma clang fp contract(off)
  const double f = ((double)d + 0.5) * ((double)n_src / (double)n_dst) - 0.5;
  const double f0 = floor(f);
  i0 = (long long)f0;
  t = f - f0;
  if (i0 < 0) {
    t = 0.0;
    i0 = 0;
  }
  if (i0 >= n_src - 1) {
    t = 0.0;
    i0 = n_src - 1;
  }
  i1 = i0 + 1 < n_src ? i0 + 1 : n_src - 1;
}

__global__ __launch_bounds__(256) void imgseq_resize(const double* __restrict__ in, long long nu, long long na,
                                                     long long nv, int flip_u, int flip_v, double* __restrict__ out,
                                                     long long mu, long long mv) {
#pragma clang fp contract(off)
  const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= mu * na * mv) return;
  const long long k = tid % mv;
  const long long j = (tid / mv) % na;
  const long long i = tid / (mv * na);
  long long y0, y1, x0, x1;
  double ty, tx;
  linear_tap(i, nu, mu, y0, y1, ty);
  linear_tap(k, nv, mv, x0, x1, tx);
  auto at = [&](long long y, long long x) {
    const long long ys = flip_u ? nu - 1 - y : y, xs = flip_v ? nv - 1 - x : x;
    return in[(ys * na + j) * nv + xs];
  };
  const double top = at(y0, x0) * (1.0 - tx) + at(y0, x1) * tx;
  const double bot = at(y1, x0) * (1.0 - tx) + at(y1, x1) * tx;
  out[tid] = top * (1.0 - ty) + bot * ty;
}

constexpr int kTile = 64;
constexpr int kPitch = 68;   // bytes per LDS row: 17 dwords, so 64 lanes writing one column hit 64 banks

struct ComposeDev {
  int n_u, n_v, u1, n_copies, bit_depth, normalize, flip_u, flip_v;
  int tile_u0, tile_v0;   // first tile of the launch (the copies' bounding box)
  double norm_value, max_value, intensity_scale;
  int v1[VAMX_IMGSEQ_MAX_COPIES];
};

// truncate at max_value (NaN passes, as np.minimum), cast toward zero to the
// unsigned type (NaN and negatives give 0), keep the low byte
__device__ __forceinline__ unsigned char to_byte(const ComposeDev& d, double x) {
#pragma clang fp contract(off)
  if (d.normalize) x = x / d.norm_value * d.max_value;
  if (d.intensity_scale != 1.0) x = x * d.intensity_scale;
  if (x > d.max_value) x = d.max_value;
  if (!(x >= 0.0)) return 0;
  return (unsigned char)((unsigned)x & 0xffu);
}

// zero frames: 16-byte stores (WIDE) or dword stores, grid-stride
template <bool WIDE>
__global__ __launch_bounds__(256) void imgseq_clear(unsigned char* __restrict__ frames, long long nbytes) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  if (WIDE) {
    uint4* f = reinterpret_cast<uint4*>(frames);
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nbytes / 16; i += stride)
      f[i] = make_uint4(0u, 0u, 0u, 0u);
  } else {
    unsigned* f = reinterpret_cast<unsigned*>(frames);
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nbytes / 4; i += stride) f[i] = 0u;
  }
}

__global__ __launch_bounds__(256) void imgseq_compose(ComposeDev d, const double* __restrict__ img, long long mu,
                                                      long long na, long long mv, unsigned char* __restrict__ frames) {
  __shared__ __align__(16) unsigned char tile[kTile * kPitch];
  const int u0 = (d.tile_u0 + (int)blockIdx.x) * kTile, v0 = (d.tile_v0 + (int)blockIdx.y) * kTile;
  const long long j = blockIdx.z;
  const int t = threadIdx.x;
  // a tile outside every copy of the image stays as imgseq_clear left it
  // (block-uniform test; the grid spans the copies' bounding box)
  bool empty = u0 + kTile <= d.u1 || u0 >= d.u1 + (int)mu;
  if (!empty) {
    empty = true;
    for (int c = 0; c < d.n_copies; ++c)
      if (v0 < d.v1[c] + (int)mv && v0 + kTile > d.v1[c]) empty = false;
  }
  if (empty) return;
  // load: lanes along v (the image's contiguous axis), 4 columns per pass
  {
    const int vl = t & 63;
    const int v = v0 + vl;
    int k = -1;
    if (v < d.n_v) {
      for (int c = 0; c < d.n_copies; ++c)
        if (v >= d.v1[c] && v < d.v1[c] + (int)mv) k = c;   // the last copy covering the row wins
    }
    const long long r = k >= 0 ? (long long)(v - d.v1[k]) : 0;
    const long long rs = d.flip_v ? mv - 1 - r : r;
    for (int ul = t >> 6; ul < kTile; ul += 4) {
      const int u = u0 + ul;
      unsigned char b = 0;
      if (k >= 0 && u >= d.u1 && u < d.u1 + (int)mu) {
        const long long c = u - d.u1;
        const long long cs = d.flip_u ? mu - 1 - c : c;
        b = to_byte(d, img[(cs * na + j) * mv + rs]);
      }
      tile[vl * kPitch + ul] = b;
    }
  }
  __syncthreads();
  // store: 4 threads per frame row, 16 bytes each as 4 dwords
  {
    const int vl = t >> 2, ch = t & 3;
    const int v = v0 + vl;
    if (v >= d.n_v) return;
    unsigned char* row = frames + ((size_t)j * d.n_v + v) * d.n_u;
    const unsigned* src = reinterpret_cast<const unsigned*>(tile + vl * kPitch + ch * 16);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int u = u0 + ch * 16 + q * 4;
      if (u + 4 <= d.n_u) *reinterpret_cast<unsigned*>(row + u) = src[q];
    }
  }
}

// ---------------------------------------------------------------------------
// Order statistics for np.percentile (imagesequence.py:128-131): the values
// of rank r and r + 1 of a float64 array, by an exact 8-pass radix selection
// on 64-bit keys whose unsigned order is the numeric order.  Pass p
// histograms digit p (8 bits, most significant first) of the elements whose
// higher digits equal the ones selected so far; a one-block launch then picks
// the bucket holding the rank and leaves the digits and the rank within them
// for the next pass.  The counts are integers, so the result is exact and
// does not depend on the launch.
// ---------------------------------------------------------------------------
constexpr int OS_PASSES = 8;
struct OsWork {
  unsigned long long counts[OS_PASSES][256];
  unsigned long long prefix[OS_PASSES + 1];   // digits selected before pass p
  unsigned long long rem[OS_PASSES + 1];      // rank within the elements matching them
  unsigned long long bucket[OS_PASSES + 1];   // elements matching them
  unsigned long long nans;     // NaN elements (counted in pass 0)
  unsigned long long min_gt;   // least key above the rank-r key (tail pass)
};

// negative values bit-inverted, others with the sign bit set: -0 sorts just
// below +0, NaN beyond -inf / +inf (a NaN anywhere makes the result NaN)
__device__ __forceinline__ unsigned long long os_key(double v) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
__device__ __forceinline__ double os_value(unsigned long long k) {
  const unsigned long long b = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
  return __longlong_as_double((long long)b);
}

__global__ __launch_bounds__(256) void os_init(OsWork* w, long long rank, long long n) {
  unsigned long long* c = &w->counts[0][0];
  for (int i = threadIdx.x; i < OS_PASSES * 256; i += blockDim.x) c[i] = 0ull;
  if (threadIdx.x == 0) {
    w->prefix[0] = 0ull;
    w->rem[0] = (unsigned long long)rank;
    w->bucket[0] = (unsigned long long)n;
    w->nans = 0ull;
    w->min_gt = ~0ull;
  }
}

// adds value v's digit to the block histogram when its higher digits match;
// lanes sharing the first active lane's digit add once, together
__device__ __forceinline__ void os_count(unsigned* h, double v, int pass, int shift, unsigned long long prefix,
                                         unsigned long long& nan_count) {
  if (pass == 0 && v != v) ++nan_count;
  const unsigned long long k = os_key(v);
  const bool match = pass == 0 || (k >> (shift + 8)) == (prefix >> (shift + 8));
  const unsigned d = (unsigned)(k >> shift) & 255u;
  const unsigned lead = __builtin_amdgcn_readfirstlane(match ? d : 256u);
  const unsigned long long same = __ballot(match && d == lead);
  if (match && d == lead) {
    if (__lane_id() == (unsigned)__ffsll((long long)same) - 1u) atomicAdd(&h[d], (unsigned)__popcll(same));
  } else if (match) {
    atomicAdd(&h[d], 1u);
  }
}

template <bool VEC>
__global__ __launch_bounds__(256) void os_hist(const double* __restrict__ x, long long n, OsWork* w, int pass) {
  __shared__ unsigned h[256];
  const int t = threadIdx.x;
  h[t] = 0u;
  __syncthreads();
  const unsigned long long prefix = w->prefix[pass];
  const int shift = 56 - 8 * pass;
  unsigned long long nan_count = 0ull;
  const long long gid = (long long)blockIdx.x * blockDim.x + t;
  const long long stride = (long long)gridDim.x * blockDim.x;
  if (VEC) {
    const double2* x2 = reinterpret_cast<const double2*>(x);
    for (long long i = gid; i < n / 2; i += stride) {
      const double2 v = x2[i];
      os_count(h, v.x, pass, shift, prefix, nan_count);
      os_count(h, v.y, pass, shift, prefix, nan_count);
    }
    if (gid == 0 && (n & 1)) os_count(h, x[n - 1], pass, shift, prefix, nan_count);
  } else {
    for (long long i = gid; i < n; i += stride) os_count(h, x[i], pass, shift, prefix, nan_count);
  }
  if (nan_count) atomicAdd(&w->nans, nan_count);
  __syncthreads();
  const unsigned c = h[t];
  if (c) atomicAdd(&w->counts[pass][t], (unsigned long long)c);
}

// one block: the bucket of pass `pass` holding the rank (an inclusive scan of
// its complete histogram), leaving the digits and the rank within them for the
// next pass
__global__ __launch_bounds__(256) void os_select(OsWork* w, int pass) {
  __shared__ unsigned long long sc[256];
  const int t = threadIdx.x;
  const int shift = 56 - 8 * pass;
  const unsigned long long prefix = w->prefix[pass];
  const unsigned long long cnt = w->counts[pass][t];
  sc[t] = cnt;
  __syncthreads();
  for (int o = 1; o < 256; o <<= 1) {
    const unsigned long long add = t >= o ? sc[t - o] : 0ull;
    __syncthreads();
    sc[t] += add;
    __syncthreads();
  }
  const unsigned long long incl = sc[t], excl = incl - cnt, rem = w->rem[pass];
  if (excl <= rem && rem < incl) {   // exactly one digit
    w->prefix[pass + 1] = prefix | ((unsigned long long)t << shift);
    w->rem[pass + 1] = rem - excl;
    w->bucket[pass + 1] = cnt;
  }
}

// rank r + 1 is an equal key when the rank-r key's bucket holds it
__device__ __forceinline__ bool os_need_next(const OsWork* w, long long n, long long rank) {
  return rank + 1 < n && w->rem[OS_PASSES] + 1 >= w->bucket[OS_PASSES];
}

// least key above the rank-r key, when rank r + 1 is not an equal key
template <bool VEC>
__global__ __launch_bounds__(256) void os_tail(const double* __restrict__ x, long long n, long long rank, OsWork* w) {
  __shared__ unsigned long long s_min[4];
  if (!os_need_next(w, n, rank)) return;   // grid-uniform
  const unsigned long long key = w->prefix[OS_PASSES];
  unsigned long long m = ~0ull;
  auto take = [&](double v) {
    const unsigned long long k = os_key(v);
    if (k > key && k < m) m = k;
  };
  const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long stride = (long long)gridDim.x * blockDim.x;
  if (VEC) {
    const double2* x2 = reinterpret_cast<const double2*>(x);
    for (long long i = gid; i < n / 2; i += stride) {
      const double2 v = x2[i];
      take(v.x);
      take(v.y);
    }
    if (gid == 0 && (n & 1)) take(x[n - 1]);
  } else {
    for (long long i = gid; i < n; i += stride) take(x[i]);
  }
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long q = __shfl_xor(m, o);
    m = q < m ? q : m;
  }
  if (__lane_id() == 0) s_min[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long b = s_min[0];
    for (int i = 1; i < 4; ++i) b = s_min[i] < b ? s_min[i] : b;
    if (b != ~0ull) atomicMin(&w->min_gt, b);
  }
}

__global__ void os_finish(long long n, long long rank, const OsWork* w, double* out) {
  const double a = os_value(w->prefix[OS_PASSES]);
  const double b = os_need_next(w, n, rank) ? os_value(w->min_gt) : a;
  const double nan = __longlong_as_double(0x7ff8000000000000ll);
  out[0] = w->nans ? nan : a;
  out[1] = w->nans ? nan : b;
}

}  // namespace

}  // namespace vamx

using namespace vamx;

extern "C" int vamx_order_stats_workspace_bytes(int64_t* bytes) {
  if (!bytes) return set_error(VAMX_ERR_ARG, "vamx_order_stats_workspace_bytes: null argument");
  *bytes = (int64_t)sizeof(OsWork);
  return VAMX_OK;
}

extern "C" int vamx_order_stats(const double* x, int64_t n, int64_t rank, double* out, void* workspace,
                                void* stream) {
  if (!x || !out || !workspace) return set_error(VAMX_ERR_ARG, "vamx_order_stats: null argument");
  if (n < 1) return set_error(VAMX_ERR_ARG, "vamx_order_stats: empty array");
  if (rank < 0 || rank >= n) return set_error(VAMX_ERR_ARG, "vamx_order_stats: rank outside 0 .. n-1");
  if ((uintptr_t)workspace % 8) return set_error(VAMX_ERR_ARG, "vamx_order_stats: workspace must be 8-byte aligned");
  hipStream_t s = (hipStream_t)stream;
  OsWork* w = (OsWork*)workspace;
  const long long blocks = std::min<long long>(2048, std::max<long long>(1, (n + 511) / 512));
  hipLaunchKernelGGL(os_init, dim3(1), dim3(256), 0, s, w, (long long)rank, (long long)n);
  const bool vec = (uintptr_t)x % 16 == 0;
  for (int p = 0; p < OS_PASSES; ++p) {
    if (vec) hipLaunchKernelGGL(os_hist<true>, dim3((unsigned)blocks), dim3(256), 0, s, x, (long long)n, w, p);
    else hipLaunchKernelGGL(os_hist<false>, dim3((unsigned)blocks), dim3(256), 0, s, x, (long long)n, w, p);
    hipLaunchKernelGGL(os_select, dim3(1), dim3(256), 0, s, w, p);
  }
  if (vec) hipLaunchKernelGGL(os_tail<true>, dim3((unsigned)blocks), dim3(256), 0, s, x, (long long)n, (long long)rank, w);
  else hipLaunchKernelGGL(os_tail<false>, dim3((unsigned)blocks), dim3(256), 0, s, x, (long long)n, (long long)rank, w);
  hipLaunchKernelGGL(os_finish, dim3(1), dim3(1), 0, s, (long long)n, (long long)rank, w, out);
  VAMX_HIP(hipGetLastError());
  return VAMX_OK;
}

extern "C" int vamx_imgseq_rotate(const double* in, int64_t nu, int64_t na, int64_t nv, int32_t flip_u,
                                  int32_t flip_v, const double* matrix, const double* offset, double* out,
                                  int64_t mu, int64_t mv, void* stream) {
  if (!in || !out || !matrix || !offset) return set_error(VAMX_ERR_ARG, "vamx_imgseq_rotate: null argument");
  if (nu < 1 || na < 1 || nv < 1 || mu < 1 || mv < 1)
    return set_error(VAMX_ERR_ARG, "vamx_imgseq_rotate: empty array");
  hipLaunchKernelGGL(imgseq_rotate, dim3(blocks_for(mu * na * mv, 256)), dim3(256), 0, (hipStream_t)stream, in,
                     (long long)nu, (long long)na, (long long)nv, flip_u ? 1 : 0, flip_v ? 1 : 0, matrix[0],
                     matrix[1], matrix[2], matrix[3], offset[0], offset[1], out, (long long)mu, (long long)mv);
  VAMX_HIP(hipGetLastError());
  return VAMX_OK;
}

extern "C" int vamx_imgseq_resize(const double* in, int64_t nu, int64_t na, int64_t nv, int32_t flip_u,
                                  int32_t flip_v, double* out, int64_t mu, int64_t mv, void* stream) {
  if (!in || !out) return set_error(VAMX_ERR_ARG, "vamx_imgseq_resize: null argument");
  if (nu < 1 || na < 1 || nv < 1 || mu < 1 || mv < 1)
    return set_error(VAMX_ERR_ARG, "vamx_imgseq_resize: empty array");
  hipLaunchKernelGGL(imgseq_resize, dim3(blocks_for(mu * na * mv, 256)), dim3(256), 0, (hipStream_t)stream, in,
                     (long long)nu, (long long)na, (long long)nv, flip_u ? 1 : 0, flip_v ? 1 : 0, out,
                     (long long)mu, (long long)mv);
  VAMX_HIP(hipGetLastError());
  return VAMX_OK;
}

extern "C" int vamx_imgseq_compose(const double* img, int64_t mu, int64_t na, int64_t mv,
                                   const vamx_imgseq_params* prm, uint8_t* frames, void* stream) {
  if (!img || !prm || !frames) return set_error(VAMX_ERR_ARG, "vamx_imgseq_compose: null argument");
  if (mu < 1 || na < 1 || mv < 1) return set_error(VAMX_ERR_ARG, "vamx_imgseq_compose: empty image");
  if (na > 65535) return set_error(VAMX_ERR_ARG, "vamx_imgseq_compose: at most 65535 frames per call");
  if (prm->n_u < 4 || prm->n_v < 4 || prm->n_u % 4 || prm->n_v % 4)
    return set_error(VAMX_ERR_ARG, "vamx_imgseq_compose: frame dimensions must be positive multiples of 4");
  if (prm->n_copies < 1 || prm->n_copies > VAMX_IMGSEQ_MAX_COPIES)
    return set_error(VAMX_ERR_ARG, "vamx_imgseq_compose: n_copies out of range");
  if (prm->bit_depth < 1 || prm->bit_depth > 32)
    return set_error(VAMX_ERR_ARG, "vamx_imgseq_compose: bit_depth must be in 1..32");
  if (prm->u1 < 0 || (int64_t)prm->u1 + mu > prm->n_u)
    return set_error(VAMX_ERR_ARG, "vamx_imgseq_compose: image exceeds the frame in u");
  for (int c = 0; c < prm->n_copies; ++c)
    if (prm->v1[c] < 0 || (int64_t)prm->v1[c] + mv > prm->n_v)
      return set_error(VAMX_ERR_ARG, "vamx_imgseq_compose: image copy exceeds the frame in v");
  if ((uintptr_t)frames % 4) return set_error(VAMX_ERR_ARG, "vamx_imgseq_compose: frames must be 4-byte aligned");
  // a NaN normalisation value (a percentile of data holding NaN) makes every
  // sample NaN and every byte 0, as the reference's cast of NaN does
  ComposeDev d;
  d.n_u = prm->n_u;
  d.n_v = prm->n_v;
  d.u1 = prm->u1;
  d.n_copies = prm->n_copies;
  d.bit_depth = prm->bit_depth;
  d.normalize = prm->normalize ? 1 : 0;
  d.flip_u = prm->flip_u ? 1 : 0;
  d.flip_v = prm->flip_v ? 1 : 0;
  d.norm_value = prm->norm_value;
  d.max_value = prm->bit_depth == 32 ? 4294967295.0 : (double)((1ull << prm->bit_depth) - 1ull);
  d.intensity_scale = prm->intensity_scale;
  for (int c = 0; c < VAMX_IMGSEQ_MAX_COPIES; ++c) d.v1[c] = c < prm->n_copies ? prm->v1[c] : 0;
  // every frame byte zeroed once, then the tiles of the copies' bounding box
  hipStream_t s = (hipStream_t)stream;
  const long long nbytes = (long long)na * prm->n_v * prm->n_u;   // a multiple of 16 (n_u, n_v multiples of 4)
  const long long cblocks = std::min<long long>(8192, (nbytes / 16 + 255) / 256);
  if ((uintptr_t)frames % 16 == 0)
    hipLaunchKernelGGL(imgseq_clear<true>, dim3((unsigned)cblocks), dim3(256), 0, s, frames, nbytes);
  else
    hipLaunchKernelGGL(imgseq_clear<false>, dim3((unsigned)cblocks), dim3(256), 0, s, frames, nbytes);
  int vlo = prm->v1[0], vhi = prm->v1[0] + (int)mv;
  for (int c = 1; c < prm->n_copies; ++c) {
    vlo = std::min(vlo, prm->v1[c]);
    vhi = std::max(vhi, prm->v1[c] + (int)mv);
  }
  d.tile_u0 = prm->u1 / kTile;
  d.tile_v0 = vlo / kTile;
  const int tu1 = (prm->u1 + (int)mu - 1) / kTile, tv1 = (vhi - 1) / kTile;
  const dim3 grid(tu1 - d.tile_u0 + 1, tv1 - d.tile_v0 + 1, (unsigned)na);
  hipLaunchKernelGGL(imgseq_compose, grid, dim3(256), 0, s, d, img, (long long)mu, (long long)na, (long long)mv,
                     frames);
  VAMX_HIP(hipGetLastError());
  return VAMX_OK;
}
