// Batched parallel-beam projection of many independent problems of uneven
// size in one persistent launch (gfx950); in the forward direction, problems
// large enough to fill the GPU on their own get single-problem launches
// instead (the same kernels and results as vamx_par_forward), spread over
// three streams.
//
// Each problem is a parallel plan (vamx_par_create) with its own volume and
// sinogram.  A problem's work is cut into tiles of one 256-thread block:
// four wavefronts of the wavefront-uniform kernels (one ray -- or a ray and
// its opposite ray, when the plan pairs opposite angles -- or one 2x2 voxel
// quad, and 64 x ZV z values per wavefront -- with ZV = 2 possibly split into
// z and z + 64 -- in the plan's L2-aware order) or,
// for shapes those do not take (inclination, very thin z), 256 items of the
// per-lane kernels.  Tiles are ordered largest problem first and handed out
// by a device-side counter, so the resident blocks of the single launch drain
// one queue until it is empty (every block exits when the counter passes the
// tile count).  The counter is reset by a one-thread kernel on the same
// stream, so the sequence is graph-capturable.
#include <algorithm>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>

#include "par_core.h"

namespace vamx {

struct BatchItem {
  ParDev g;
  const float* minmax;
  long long tile_begin;  // first tile of this problem
  long long items;       // wavefronts (wave mode) or threads (per-lane mode) of this problem
  int zv;                // z values per lane: 1, 2, 4 (wave mode) or 1, 4 (per-lane mode)
  int clip;              // forward: warp clip; backward: circle clip
  int wave;              // 1: wavefront-uniform kernels
  int pair;              // forward, wave mode: opposite-angle pairs (the plan's paired work order)
  int split;             // wave mode, zv 2: lanes hold z and z + 64
};

// Buffers of the problems of one launch, passed by value as a kernel
// argument (so a captured graph holds them and no device table changes from
// call to call): entry r belongs to the problem at table position r.
constexpr int kBatchChunk = 64;
struct BatchPtrs {
  const float* in[kBatchChunk];
  float* out[kBatchChunk];
};

__global__ void batch_reset(unsigned long long* counter) { *counter = 0ull; }

template <int ZV, bool SPLIT>
__device__ __forceinline__ void batch_fp_wave(const BatchItem& it, const float* in, float* out, long long w,
                                              int lane) {
  const ParDev& g = it.g;
  int k, i, z, zl, zl2;
  if (!fp_wave_item<ZV, SPLIT>(w, g, lane, k, i, z)) return;
  wave_zload<ZV, SPLIT>(z, g.nz, zl, zl2);
  const bool clip_on = it.clip && wave_clip_active<ZV, SPLIT>(it.minmax, g.nz, zl, zl2);
  const typename VecT<ZV>::T acc =
      clip_on ? fp_ray_wave<ZV, true, true, false, SPLIT>(g, in, it.minmax, k, i, zl, lane, nullptr, zl2)
              : fp_ray_wave<ZV, false, true, false, SPLIT>(g, in, it.minmax, k, i, zl, lane, nullptr, zl2);
  wave_zstore<ZV, SPLIT>(out + ((size_t)k * g.na + i) * g.nz, z, g.nz, acc);
}

// paired forward: ray (k, i) and its opposite ray from one pass over the
// taps, as the single-problem paired kernel (clipped waves project the two
// rays one after the other)
template <int ZV, bool SPLIT>
__device__ __forceinline__ void batch_fp_pair(const BatchItem& it, const float* in, float* out, long long w,
                                              int lane) {
  using V = typename VecT<ZV>::T;
  const ParDev& g = it.g;
  const PairItem pi = fp_pair_item<ZV, SPLIT>(w, g, lane);
  if (!pi.ok) return;
  const int z = pi.z;
  int zl, zl2;
  wave_zload<ZV, SPLIT>(z, g.nz, zl, zl2);
  const bool clip_on = it.clip && wave_clip_active<ZV, SPLIT>(it.minmax, g.nz, zl, zl2);
  if (clip_on) {
    for (int h = 0; h < (pi.k2 >= 0 ? 2 : 1); ++h) {
      const int kh = h ? pi.k2 : pi.k, ih = h ? pi.i2 : pi.i;
      const V acc = fp_ray_wave<ZV, true, true, false, SPLIT>(g, in, it.minmax, kh, ih, zl, lane, nullptr, zl2);
      wave_zstore<ZV, SPLIT>(out + ((size_t)kh * g.na + ih) * g.nz, z, g.nz, acc);
    }
    return;
  }
  V acc, acc2 = vzero<ZV>();
  if (pi.k2 >= 0)
    acc = fp_ray_wave<ZV, false, true, true, SPLIT>(g, in, it.minmax, pi.k, pi.i, zl, lane, &acc2, zl2);
  else
    acc = fp_ray_wave<ZV, false, true, false, SPLIT>(g, in, it.minmax, pi.k, pi.i, zl, lane, nullptr, zl2);
  wave_zstore<ZV, SPLIT>(out + ((size_t)pi.k * g.na + pi.i) * g.nz, z, g.nz, acc);
  if (pi.k2 >= 0) wave_zstore<ZV, SPLIT>(out + ((size_t)pi.k2 * g.na + pi.i2) * g.nz, z, g.nz, acc2);
}

template <int ZV, bool SPLIT>
__device__ __forceinline__ void batch_bp_wave(const BatchItem& it, const float* in, float* out, long long w,
                                              int lane) {
  const ParDev& g = it.g;
  int r0, c0, z, zl, zl2;
  if (!bp_wave_quad<ZV, SPLIT>(w, g, lane, r0, c0, z)) return;
  wave_zload<ZV, SPLIT>(z, g.nz, zl, zl2);
  unsigned inrange, active;
  bp_quad_masks(g, r0, c0, it.clip, inrange, active);
  typename VecT<ZV>::T acc[4];
  bp_quad_wave<ZV, SPLIT>(g, in, r0, c0, active, zl, lane, acc, zl2);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    if (!((inrange >> q) & 1u)) continue;
    float* outp = out + ((size_t)(r0 + (q >> 1)) * g.n + c0 + (q & 1)) * g.nz;
    wave_zstore<ZV, SPLIT>(outp, z, g.nz, ((active >> q) & 1u) ? acc[q] : vzero<ZV>());
  }
}

template <bool FORWARD>
__device__ __forceinline__ void batch_lane_item(const BatchItem& it, const float* in, float* out, long long w) {
  const ParDev& g = it.g;
  const int nzv = g.nz / it.zv;
  const int z = (int)(w % nzv) * it.zv;
  const long long outer = w / nzv;
  if (FORWARD) {
    const int k = (int)(outer % g.n), i = (int)(outer / g.n);
    float* dst = out + ((size_t)k * g.na + i) * g.nz + z;
    if (g.incl) {
      *dst = fp_ray_incl(g, in, k, i, z);
    } else if (it.zv == 4) {
      vstore<4>(dst, it.clip ? fp_ray<4, true>(g, in, it.minmax, k, i, z)
                             : fp_ray<4, false>(g, in, it.minmax, k, i, z));
    } else {
      *dst = it.clip ? fp_ray<1, true>(g, in, it.minmax, k, i, z)
                     : fp_ray<1, false>(g, in, it.minmax, k, i, z);
    }
  } else {
    const int pix = (int)outer;
    float* dst = out + (size_t)pix * g.nz + z;
    const bool zero = it.clip && g.outside[pix];
    if (g.incl) {
      *dst = zero ? 0.0f : bp_voxel_incl(g, in, pix, z);
    } else if (it.zv == 4) {
      vstore<4>(dst, zero ? vzero<4>() : bp_voxel<4>(g, in, pix, z));
    } else {
      *dst = zero ? 0.0f : bp_voxel<1>(g, in, pix, z);
    }
  }
}

template <bool FORWARD>
__global__ __launch_bounds__(256) void batch_kernel(const BatchItem* __restrict__ items, int nitems,
                                                    long long ntiles, unsigned long long* counter,
                                                    BatchPtrs ptrs) {
  __shared__ long long tile_sh;
  const int lane = threadIdx.x & 63;
  const int slot = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  for (;;) {
    if (threadIdx.x == 0) tile_sh = (long long)atomicAdd(counter, 1ull);
    __syncthreads();
    const long long tv = tile_sh;
    __syncthreads();
    // the tile is block-uniform: keep it (and the item search) scalar
    const long long tile = (long long)(((unsigned long long)(unsigned)__builtin_amdgcn_readfirstlane((int)(tv >> 32))
                                        << 32) |
                                       (unsigned)__builtin_amdgcn_readfirstlane((int)(tv & 0xffffffffll)));
    if (tile >= ntiles) return;
    // problem owning this tile (items are ordered by tile_begin)
    int lo = 0, hi = nitems - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (items[mid].tile_begin <= tile) lo = mid; else hi = mid - 1;
    }
    const BatchItem& it = items[lo];
    const float* in = ptrs.in[lo];
    float* out = ptrs.out[lo];
    // No `continue` below: a per-thread continue makes the loop header (and
    // its barriers) run for a subset of lanes, so waves of the block would
    // disagree on the barrier count.  Every thread reaches the loop end.
    if (it.wave) {
      const long long w = (tile - it.tile_begin) * 4 + slot;
      if (w < it.items) {
        if (FORWARD && it.pair) {
          if (it.split) batch_fp_pair<2, true>(it, in, out, w, lane);
          else if (it.zv == 4) batch_fp_pair<4, false>(it, in, out, w, lane);
          else if (it.zv == 2) batch_fp_pair<2, false>(it, in, out, w, lane);
          else batch_fp_pair<1, false>(it, in, out, w, lane);
        } else if (FORWARD) {
          if (it.split) batch_fp_wave<2, true>(it, in, out, w, lane);
          else if (it.zv == 4) batch_fp_wave<4, false>(it, in, out, w, lane);
          else if (it.zv == 2) batch_fp_wave<2, false>(it, in, out, w, lane);
          else batch_fp_wave<1, false>(it, in, out, w, lane);
        } else {
          if (it.split) batch_bp_wave<2, true>(it, in, out, w, lane);
          else if (it.zv == 4) batch_bp_wave<4, false>(it, in, out, w, lane);
          else if (it.zv == 2) batch_bp_wave<2, false>(it, in, out, w, lane);
          else batch_bp_wave<1, false>(it, in, out, w, lane);
        }
      }
    } else {
      const long long w = (tile - it.tile_begin) * 256 + threadIdx.x;
      if (w < it.items) batch_lane_item<FORWARD>(it, in, out, w);
    }
  }
}


}  // namespace vamx

using namespace vamx;

// A batch: the plans of its problems, and everything the batched launches
// need besides the caller's buffers -- the device tables of work items (one
// per distinct buffer alignment and work order seen, each kept until the
// batch is destroyed, so a graph that captured a launch stays valid), the
// tile counter of the persistent launches and the side streams of the
// forward.  Calls on one batch are ordered on one stream at a time.
struct vamx_batch {
  int magic = 0;
  int device = -1;
  std::vector<vamx_plan*> plans;
  std::vector<unsigned long long> serials;
  unsigned long long* d_counter = nullptr;
  struct Chunk {
    BatchItem* d_items = nullptr;  // table, largest problem first
    int m = 0;                     // problems in this launch (<= kBatchChunk)
    long long tiles = 0;
    std::vector<int> order;        // table position -> problem index
    std::vector<int> minmax;       // forward: problems whose per-slice min/max the warp clip needs
  };
  struct Layout {
    std::vector<int> sig;          // per problem: kernel path chosen for the buffers' alignment
    std::vector<char> devs;        // the plans' ParDev bytes (work order set by vamx_par_set_tuning)
    std::vector<int> single;       // forward: problems projected by their own single-problem launch
    std::vector<Chunk> chunks;
  };
  std::vector<Layout> layouts[2];  // [0] backward, [1] forward
  bool side_ok = false;
  hipStream_t side[2] = {};
  hipEvent_t fork = nullptr, join[2] = {};
};

namespace {

constexpr int kBatchMagic = 0x62617463;
constexpr int kSideStreams = 2;
constexpr size_t kMaxLayouts = 8;

bool check_batch(const vamx_batch* b) { return b && b->magic == kBatchMagic; }

void release_batch(vamx_batch* b) {
  int cur = 0;
  (void)hipGetDevice(&cur);
  if (cur != b->device) (void)hipSetDevice(b->device);
  (void)hipDeviceSynchronize();
  for (auto& ls : b->layouts)
    for (auto& l : ls)
      for (auto& c : l.chunks) (void)hipFree(c.d_items);
  if (b->d_counter) (void)hipFree(b->d_counter);
  for (int j = 0; j < kSideStreams; ++j) {
    if (b->join[j]) (void)hipEventDestroy(b->join[j]);
    if (b->side[j]) (void)hipStreamDestroy(b->side[j]);
  }
  if (b->fork) (void)hipEventDestroy(b->fork);
  if (cur != b->device) (void)hipSetDevice(cur);
  b->magic = 0;
  delete b;
}

int create_batch(vamx_plan* const* plans, int n, vamx_batch** out) {
  if (!out) return set_error(VAMX_ERR_ARG, "batch: null output");
  *out = nullptr;
  if (!plans || n < 1) return set_error(VAMX_ERR_ARG, "batch: empty or null batch");
  int dev = 0;
  VAMX_HIP(hipGetDevice(&dev));
  for (int i = 0; i < n; ++i) {
    if (!check_plan(plans[i], PLAN_PARALLEL)) return set_error(VAMX_ERR_STATE, "batch: not a parallel plan");
    if (plans[i]->device != dev) return set_error(VAMX_ERR_ARG, "batch: plans must live on the current device");
    for (int j = 0; j < i; ++j)
      if (plans[j] == plans[i])
        return set_error(VAMX_ERR_ARG, "batch: plan " + std::to_string(i) + " repeats plan " + std::to_string(j) +
                                           " (a plan's slice min/max scratch serves one problem)");
  }
  vamx_batch* b = new vamx_batch();
  b->magic = kBatchMagic;
  b->device = dev;
  b->plans.assign(plans, plans + n);
  for (int i = 0; i < n; ++i) b->serials.push_back(plans[i]->serial);
  hipError_t e = hipMalloc((void**)&b->d_counter, sizeof(unsigned long long));
  if (e != hipSuccess) {
    release_batch(b);
    return hip_check(e, "hipMalloc(batch counter)");
  }
  *out = b;
  return VAMX_OK;
}

// kernel path of problem i for these buffers: wave z width (1, 2, 4; 3 = z
// and z + 64), or -1 / -4 for the per-lane kernels with 1 / 4 z per thread
int problem_path(vamx_plan* p, const float* in, float* out, bool forward) {
  int wz = forward ? par_fp_wave_zv(p, in, out) : par_wave_zv(p, in, out, p->par.bp_zv);
  if (par_wave_split(p, wz)) return 3;
  if (wz) return wz;
  return par_use_vec4(p, in, out) ? -4 : -1;
}

int build_layout(vamx_batch* b, const std::vector<int>& sig, const std::vector<char>& devs, bool forward,
                 hipStream_t s, vamx_batch::Layout** out) {
  std::vector<vamx_batch::Layout>& ls = b->layouts[forward ? 1 : 0];
  if (ls.size() >= kMaxLayouts)
    return set_error(VAMX_ERR_ARG, "batch: more than 8 distinct buffer alignments or work orders on one batch");
  int cus = 0;
  VAMX_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->device));
  const int n = (int)b->plans.size();
  vamx_batch::Layout L;
  L.sig = sig;
  L.devs = devs;
  // Forward: a problem whose wavefront work alone fills the GPU (16 waves
  // per CU) runs as its own single-problem launch: the persistent kernel,
  // whose registers cover every z width and the per-lane fallback, reaches a
  // lower occupancy than the single-problem forward kernels.
  std::vector<int> shared;
  std::vector<BatchItem> items((size_t)n);
  for (int i = 0; i < n; ++i) {
    vamx_plan* p = b->plans[(size_t)i];
    BatchItem& it = items[(size_t)i];
    std::memset(&it, 0, sizeof(it));
    it.g = p->par;
    it.minmax = p->d_minmax;
    it.clip = forward ? !(p->flags & VAMX_PAR_NO_WARP_CLIP) : !(p->flags & VAMX_PAR_NO_CIRCLE_CLIP);
    const int path = sig[(size_t)i];
    if (path > 0) {
      const int wz = path == 3 ? 2 : path;
      it.wave = 1;
      it.zv = wz;
      it.split = path == 3 ? 1 : 0;
      it.pair = forward && it.g.fp_pair && it.g.pair_ok ? 1 : 0;
      const long long units = it.pair ? fp_pair_tiled(it.g) : forward ? fp_tiled_rays(it.g) : bp_tiled_quads(it.g);
      it.items = units * wave_zchunks(it.g.nz, wz);
      if (forward && !it.g.incl && it.items >= 16ll * cus) {
        L.single.push_back(i);
        continue;
      }
    } else {
      it.wave = 0;
      it.zv = -path;
      const long long outer = forward ? (long long)it.g.na * it.g.n : (long long)it.g.n * it.g.n;
      it.items = outer * (it.g.nz / it.zv);
    }
    shared.push_back(i);
  }
  // largest work first (cost ~ interpolated samples), then cut into launches
  // of at most kBatchChunk problems
  std::stable_sort(shared.begin(), shared.end(), [&](int x, int y) {
    const ParDev& a = items[(size_t)x].g;
    const ParDev& c = items[(size_t)y].g;
    return (double)a.n * a.n * a.na * a.nz > (double)c.n * c.n * c.na * c.nz;
  });
  if (!shared.empty()) {
    // a new table: blocking upload (not allowed under graph capture, where
    // hipStreamSynchronize fails and the error is returned)
    VAMX_HIP(hipStreamSynchronize(s));
  }
  int rc = VAMX_OK;
  for (size_t c0 = 0; c0 < shared.size() && rc == VAMX_OK; c0 += kBatchChunk) {
    vamx_batch::Chunk ch;
    const size_t m = std::min(shared.size() - c0, (size_t)kBatchChunk);
    std::vector<BatchItem> tab(m);
    for (size_t r = 0; r < m; ++r) {
      const int i = shared[c0 + r];
      tab[r] = items[(size_t)i];
      tab[r].tile_begin = ch.tiles;
      ch.tiles += (tab[r].items + (tab[r].wave ? 3 : 255)) / (tab[r].wave ? 4 : 256);
      ch.order.push_back(i);
      if (forward && tab[r].clip && !tab[r].g.incl) ch.minmax.push_back(i);
    }
    ch.m = (int)m;
    hipError_t e = hipMalloc((void**)&ch.d_items, m * sizeof(BatchItem));
    if (e == hipSuccess) e = hipMemcpy(ch.d_items, tab.data(), m * sizeof(BatchItem), hipMemcpyHostToDevice);
    if (e != hipSuccess) rc = hip_check(e, "batch table upload");
    L.chunks.push_back(std::move(ch));
  }
  if (rc != VAMX_OK) {
    for (auto& c : L.chunks) (void)hipFree(c.d_items);
    return rc;
  }
  ls.push_back(std::move(L));
  *out = &ls.back();
  return VAMX_OK;
}

int ensure_side_streams(vamx_batch* b) {
  if (b->side_ok) return VAMX_OK;
  if (!b->fork) VAMX_HIP(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
  for (int j = 0; j < kSideStreams; ++j) {
    if (!b->side[j]) VAMX_HIP(hipStreamCreateWithFlags(&b->side[j], hipStreamNonBlocking));
    if (!b->join[j]) VAMX_HIP(hipEventCreateWithFlags(&b->join[j], hipEventDisableTiming));
  }
  b->side_ok = true;
  return VAMX_OK;
}

int run_batch(vamx_batch* b, const float* const* ins, float* const* outs, hipStream_t s, bool forward) {
  if (!check_batch(b)) return set_error(VAMX_ERR_STATE, "batch: not a batch handle");
  if (!ins || !outs) return set_error(VAMX_ERR_ARG, "batch: null buffer table");
  int dev = 0;
  VAMX_HIP(hipGetDevice(&dev));
  if (dev != b->device) return set_error(VAMX_ERR_ARG, "batch: the batch's device must be the current device");
  const int n = (int)b->plans.size();
  std::vector<int> sig((size_t)n);
  std::vector<char> devs((size_t)n * sizeof(ParDev));
  for (int i = 0; i < n; ++i) {
    if (!ins[i] || !outs[i]) return set_error(VAMX_ERR_ARG, "batch: null buffer of problem " + std::to_string(i));
    vamx_plan* p = b->plans[(size_t)i];
    sig[(size_t)i] = problem_path(p, ins[i], outs[i], forward);
    std::memcpy(devs.data() + (size_t)i * sizeof(ParDev), &p->par, sizeof(ParDev));
  }
  vamx_batch::Layout* L = nullptr;
  for (auto& l : b->layouts[forward ? 1 : 0])
    if (l.sig == sig && l.devs == devs) {
      L = &l;
      break;
    }
  if (!L) {
    int rc = build_layout(b, sig, devs, forward, s, &L);
    if (rc) return rc;
  }
  // single-problem forward launches round-robin on the caller's stream and
  // two side streams forked from and joined back to it with events (so the
  // call stays graph-capturable): one launch's tail overlaps the next one's
  // start
  if (!L->single.empty()) {
    int ns = 1;
    if (L->single.size() > 1 && ensure_side_streams(b) == VAMX_OK) ns = 1 + kSideStreams;
    if (ns > 1) {
      VAMX_HIP(hipEventRecord(b->fork, s));
      for (int j = 0; j < kSideStreams; ++j) VAMX_HIP(hipStreamWaitEvent(b->side[j], b->fork, 0));
    }
    int rc = VAMX_OK;
    for (size_t k = 0; k < L->single.size() && rc == VAMX_OK; ++k) {
      const int i = L->single[k];
      vamx_plan* p = b->plans[(size_t)i];
      hipStream_t st = (k % ns == 0) ? s : b->side[k % ns - 1];
      rc = par_forward_launch(p, ins[i], outs[i], st, !(p->flags & VAMX_PAR_NO_WARP_CLIP));
    }
    if (ns > 1) {
      // joined also after a failed launch, so a capture on s is never left
      // with unjoined side streams
      for (int j = 0; j < kSideStreams; ++j) {
        const hipError_t e1 = hipEventRecord(b->join[j], b->side[j]);
        const hipError_t e2 = hipStreamWaitEvent(s, b->join[j], 0);
        if (rc == VAMX_OK && e1 != hipSuccess) rc = hip_check(e1, "hipEventRecord(join)");
        if (rc == VAMX_OK && e2 != hipSuccess) rc = hip_check(e2, "hipStreamWaitEvent(join)");
      }
    }
    if (rc) return rc;
  }
  int cus = 0;
  VAMX_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  for (const auto& ch : L->chunks) {
    for (int i : ch.minmax) {
      int rc = par_slice_minmax(b->plans[(size_t)i], ins[i], s);
      if (rc) return rc;
    }
    BatchPtrs ptrs;
    std::memset(&ptrs, 0, sizeof(ptrs));
    for (int r = 0; r < ch.m; ++r) {
      ptrs.in[r] = ins[ch.order[(size_t)r]];
      ptrs.out[r] = outs[ch.order[(size_t)r]];
    }
    hipLaunchKernelGGL(batch_reset, dim3(1), dim3(1), 0, s, b->d_counter);
    long long grid = (long long)std::max(cus, 1) * 8;
    if (grid > ch.tiles) grid = ch.tiles;
    if (forward)
      hipLaunchKernelGGL(batch_kernel<true>, dim3((unsigned)grid), dim3(256), 0, s, ch.d_items, ch.m, ch.tiles,
                         b->d_counter, ptrs);
    else
      hipLaunchKernelGGL(batch_kernel<false>, dim3((unsigned)grid), dim3(256), 0, s, ch.d_items, ch.m, ch.tiles,
                         b->d_counter, ptrs);
    VAMX_HIP(hipGetLastError());
  }
  return VAMX_OK;
}

// Batches behind vamx_par_{forward,backward}_batched, one per distinct list
// of plans (by plan serial number), released when one of their plans is
// destroyed.
std::mutex g_cache_mu;
std::vector<vamx_batch*> g_cache;

int cached_batch(vamx_plan* const* plans, int n, vamx_batch** out) {
  if (!plans || n < 1) return set_error(VAMX_ERR_ARG, "batched: empty or null batch");
  for (int i = 0; i < n; ++i)
    if (!check_plan(plans[i], PLAN_PARALLEL)) return set_error(VAMX_ERR_STATE, "batched: not a parallel plan");
  std::lock_guard<std::mutex> lock(g_cache_mu);
  for (vamx_batch* b : g_cache) {
    if ((int)b->serials.size() != n) continue;
    bool same = true;
    for (int i = 0; i < n && same; ++i) same = b->serials[(size_t)i] == plans[i]->serial;
    if (same) {
      *out = b;
      return VAMX_OK;
    }
  }
  int rc = create_batch(plans, n, out);
  if (rc) return rc;
  g_cache.push_back(*out);
  return VAMX_OK;
}

}  // namespace

namespace vamx {
void batch_forget_plan(const vamx_plan* p) {
  std::lock_guard<std::mutex> lock(g_cache_mu);
  for (size_t k = 0; k < g_cache.size();) {
    vamx_batch* b = g_cache[k];
    if (std::find(b->serials.begin(), b->serials.end(), p->serial) != b->serials.end()) {
      release_batch(b);
      g_cache.erase(g_cache.begin() + (long)k);
    } else {
      ++k;
    }
  }
}
}  // namespace vamx

extern "C" int vamx_batch_create(vamx_plan* const* plans, int32_t n_plans, vamx_batch** out_batch) {
  return create_batch(plans, n_plans, out_batch);
}

extern "C" int vamx_batch_forward(vamx_batch* batch, const float* const* vols, float* const* sinos, void* stream) {
  return run_batch(batch, vols, sinos, (hipStream_t)stream, true);
}

extern "C" int vamx_batch_backward(vamx_batch* batch, const float* const* sinos, float* const* vols, void* stream) {
  return run_batch(batch, sinos, vols, (hipStream_t)stream, false);
}

extern "C" void vamx_batch_destroy(vamx_batch* batch) {
  if (check_batch(batch)) release_batch(batch);
}

extern "C" int vamx_par_forward_batched(vamx_plan* const* plans, int32_t n_plans, const float* const* vols,
                                        float* const* sinos, void* stream) {
  vamx_batch* b = nullptr;
  int rc = cached_batch(plans, n_plans, &b);
  if (rc) return rc;
  return run_batch(b, vols, sinos, (hipStream_t)stream, true);
}

extern "C" int vamx_par_backward_batched(vamx_plan* const* plans, int32_t n_plans, const float* const* sinos,
                                         float* const* vols, void* stream) {
  vamx_batch* b = nullptr;
  int rc = cached_batch(plans, n_plans, &b);
  if (rc) return rc;
  return run_batch(b, sinos, vols, (hipStream_t)stream, false);
}
