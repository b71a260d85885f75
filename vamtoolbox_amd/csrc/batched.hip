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
#include <numeric>

#include "par_core.h"

namespace vamx {

struct BatchItem {
  ParDev g;
  const float* in;
  float* out;
  const float* minmax;
  long long tile_begin;  // first tile of this problem
  long long items;       // wavefronts (wave mode) or threads (per-lane mode) of this problem
  int zv;                // z values per lane: 1, 2, 4 (wave mode) or 1, 4 (per-lane mode)
  int clip;              // forward: warp clip; backward: circle clip
  int wave;              // 1: wavefront-uniform kernels
  int pair;              // forward, wave mode: opposite-angle pairs (the plan's paired work order)
  int split;             // wave mode, zv 2: lanes hold z and z + 64
};

__global__ void batch_reset(unsigned long long* counter) { *counter = 0ull; }

template <int ZV, bool SPLIT>
__device__ __forceinline__ void batch_fp_wave(const BatchItem& it, long long w, int lane) {
  const ParDev& g = it.g;
  int k, i, z, zl, zl2;
  if (!fp_wave_item<ZV, SPLIT>(w, g, lane, k, i, z)) return;
  wave_zload<ZV, SPLIT>(z, g.nz, zl, zl2);
  const bool clip_on = it.clip && wave_clip_active<ZV, SPLIT>(it.minmax, g.nz, zl, zl2);
  const typename VecT<ZV>::T acc =
      clip_on ? fp_ray_wave<ZV, true, true, false, SPLIT>(g, it.in, it.minmax, k, i, zl, lane, nullptr, zl2)
              : fp_ray_wave<ZV, false, true, false, SPLIT>(g, it.in, it.minmax, k, i, zl, lane, nullptr, zl2);
  wave_zstore<ZV, SPLIT>(it.out + ((size_t)k * g.na + i) * g.nz, z, g.nz, acc);
}

// paired forward: ray (k, i) and its opposite ray from one pass over the
// taps, as the single-problem paired kernel (clipped waves project the two
// rays one after the other)
template <int ZV, bool SPLIT>
__device__ __forceinline__ void batch_fp_pair(const BatchItem& it, long long w, int lane) {
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
      const V acc = fp_ray_wave<ZV, true, true, false, SPLIT>(g, it.in, it.minmax, kh, ih, zl, lane, nullptr, zl2);
      wave_zstore<ZV, SPLIT>(it.out + ((size_t)kh * g.na + ih) * g.nz, z, g.nz, acc);
    }
    return;
  }
  V acc, acc2 = vzero<ZV>();
  if (pi.k2 >= 0)
    acc = fp_ray_wave<ZV, false, true, true, SPLIT>(g, it.in, it.minmax, pi.k, pi.i, zl, lane, &acc2, zl2);
  else
    acc = fp_ray_wave<ZV, false, true, false, SPLIT>(g, it.in, it.minmax, pi.k, pi.i, zl, lane, nullptr, zl2);
  wave_zstore<ZV, SPLIT>(it.out + ((size_t)pi.k * g.na + pi.i) * g.nz, z, g.nz, acc);
  if (pi.k2 >= 0) wave_zstore<ZV, SPLIT>(it.out + ((size_t)pi.k2 * g.na + pi.i2) * g.nz, z, g.nz, acc2);
}

template <int ZV, bool SPLIT>
__device__ __forceinline__ void batch_bp_wave(const BatchItem& it, long long w, int lane) {
  const ParDev& g = it.g;
  int r0, c0, z, zl, zl2;
  if (!bp_wave_quad<ZV, SPLIT>(w, g, lane, r0, c0, z)) return;
  wave_zload<ZV, SPLIT>(z, g.nz, zl, zl2);
  unsigned inrange, active;
  bp_quad_masks(g, r0, c0, it.clip, inrange, active);
  typename VecT<ZV>::T acc[4];
  bp_quad_wave<ZV, SPLIT>(g, it.in, r0, c0, active, zl, lane, acc, zl2);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    if (!((inrange >> q) & 1u)) continue;
    float* outp = it.out + ((size_t)(r0 + (q >> 1)) * g.n + c0 + (q & 1)) * g.nz;
    wave_zstore<ZV, SPLIT>(outp, z, g.nz, ((active >> q) & 1u) ? acc[q] : vzero<ZV>());
  }
}

template <bool FORWARD>
__device__ __forceinline__ void batch_lane_item(const BatchItem& it, long long w) {
  const ParDev& g = it.g;
  const int nzv = g.nz / it.zv;
  const int z = (int)(w % nzv) * it.zv;
  const long long outer = w / nzv;
  if (FORWARD) {
    const int k = (int)(outer % g.n), i = (int)(outer / g.n);
    float* dst = it.out + ((size_t)k * g.na + i) * g.nz + z;
    if (g.incl) {
      *dst = fp_ray_incl(g, it.in, k, i, z);
    } else if (it.zv == 4) {
      vstore<4>(dst, it.clip ? fp_ray<4, true>(g, it.in, it.minmax, k, i, z)
                             : fp_ray<4, false>(g, it.in, it.minmax, k, i, z));
    } else {
      *dst = it.clip ? fp_ray<1, true>(g, it.in, it.minmax, k, i, z)
                     : fp_ray<1, false>(g, it.in, it.minmax, k, i, z);
    }
  } else {
    const int pix = (int)outer;
    float* dst = it.out + (size_t)pix * g.nz + z;
    const bool zero = it.clip && g.outside[pix];
    if (g.incl) {
      *dst = zero ? 0.0f : bp_voxel_incl(g, it.in, pix, z);
    } else if (it.zv == 4) {
      vstore<4>(dst, zero ? vzero<4>() : bp_voxel<4>(g, it.in, pix, z));
    } else {
      *dst = zero ? 0.0f : bp_voxel<1>(g, it.in, pix, z);
    }
  }
}

template <bool FORWARD>
__global__ __launch_bounds__(256) void batch_kernel(const BatchItem* __restrict__ items, int nitems,
                                                    long long ntiles, unsigned long long* counter) {
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
    // No `continue` below: a per-thread continue makes the loop header (and
    // its barriers) run for a subset of lanes, so waves of the block would
    // disagree on the barrier count.  Every thread reaches the loop end.
    if (it.wave) {
      const long long w = (tile - it.tile_begin) * 4 + slot;
      if (w < it.items) {
        if (FORWARD && it.pair) {
          if (it.split) batch_fp_pair<2, true>(it, w, lane);
          else if (it.zv == 4) batch_fp_pair<4, false>(it, w, lane);
          else if (it.zv == 2) batch_fp_pair<2, false>(it, w, lane);
          else batch_fp_pair<1, false>(it, w, lane);
        } else if (FORWARD) {
          if (it.split) batch_fp_wave<2, true>(it, w, lane);
          else if (it.zv == 4) batch_fp_wave<4, false>(it, w, lane);
          else if (it.zv == 2) batch_fp_wave<2, false>(it, w, lane);
          else batch_fp_wave<1, false>(it, w, lane);
        } else {
          if (it.split) batch_bp_wave<2, true>(it, w, lane);
          else if (it.zv == 4) batch_bp_wave<4, false>(it, w, lane);
          else if (it.zv == 2) batch_bp_wave<2, false>(it, w, lane);
          else batch_bp_wave<1, false>(it, w, lane);
        }
      }
    } else {
      const long long w = (tile - it.tile_begin) * 256 + threadIdx.x;
      if (w < it.items) batch_lane_item<FORWARD>(it, w);
    }
  }
}

namespace {

// Device copies of recently used item tables.  A call whose table matches a
// cached one enqueues only its two kernels (no copy, no synchronisation), so
// a fixed sequence of batched calls -- e.g. forward then backward of an
// iteration -- can be captured in a HIP graph once each table was uploaded
// by an earlier call.
struct BatchTable {
  std::vector<char> bytes;
  BatchItem* d_items = nullptr;
};

struct BatchScratch {
  int device = -1;
  unsigned long long* d_counter = nullptr;
  std::vector<BatchTable> tables;   // most recently used last
};

constexpr size_t kMaxTables = 16;

// side streams of the batched forward (default priority), created once per
// host thread and device and kept for the life of the process
constexpr int kSideStreams = 2;
struct SideStreams {
  int device = -1;
  bool ok = false;
  hipStream_t st[kSideStreams] = {};
  hipEvent_t fork = nullptr, join[kSideStreams] = {};
};
thread_local SideStreams g_side;

thread_local BatchScratch g_scratch;

int run_batched(vamx_plan* const* plans, int n, const float* const* ins, float* const* outs, hipStream_t s,
                bool forward) {
  if (!plans || n < 1 || !ins || !outs) return set_error(VAMX_ERR_ARG, "batched: empty or null batch");
  int dev = 0;
  VAMX_HIP(hipGetDevice(&dev));
  int cus = 0;
  VAMX_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  for (int i = 0; i < n; ++i) {
    vamx_plan* p = plans[i];
    if (!check_plan(p, PLAN_PARALLEL)) return set_error(VAMX_ERR_STATE, "batched: not a parallel plan");
    if (p->device != dev) return set_error(VAMX_ERR_ARG, "batched: plans must live on the current device");
    if (!ins[i] || !outs[i]) return set_error(VAMX_ERR_ARG, "batched: null buffer");
  }
  // Forward: a problem whose wavefront work alone fills the GPU (16 waves per
  // CU) runs as its own single-problem launch; the persistent kernel, whose
  // registers cover every z width and the per-lane fallback, reaches a lower
  // occupancy than the single-problem forward kernels.  The rest share the
  // persistent launch.  The single-problem launches go round-robin to the
  // caller's stream and two side streams (forked from and joined back to the
  // caller's stream with events, so the call stays graph-capturable): one
  // launch's tail overlaps the next one's start.
  std::vector<int> shared, single;
  for (int i = 0; i < n; ++i) {
    vamx_plan* p = plans[i];
    if (forward && !p->par.incl) {
      int wz = par_wave_zv(p, ins[i], outs[i], p->par.fp_zv);
      if (par_wave_split(p, wz)) wz = 2;
      const long long units =
          p->par.fp_pair && p->par.pair_ok ? fp_pair_tiled(p->par) : fp_tiled_rays(p->par);
      if (wz && units * wave_zchunks(p->par.nz, wz) >= 16ll * cus) {
        single.push_back(i);
        continue;
      }
    }
    shared.push_back(i);
  }
  if (!single.empty()) {
    SideStreams& ss = g_side;
    if (ss.device != dev) {
      ss.ok = hipEventCreateWithFlags(&ss.fork, hipEventDisableTiming) == hipSuccess;
      for (int j = 0; j < kSideStreams && ss.ok; ++j)
        ss.ok = hipStreamCreateWithFlags(&ss.st[j], hipStreamNonBlocking) == hipSuccess &&
                hipEventCreateWithFlags(&ss.join[j], hipEventDisableTiming) == hipSuccess;
      ss.device = dev;
    }
    const int ns = (ss.ok && single.size() > 1) ? 1 + kSideStreams : 1;
    if (ns > 1) {
      VAMX_HIP(hipEventRecord(ss.fork, s));
      for (int j = 0; j < kSideStreams; ++j) VAMX_HIP(hipStreamWaitEvent(ss.st[j], ss.fork, 0));
    }
    for (size_t k = 0; k < single.size(); ++k) {
      const int i = single[k];
      vamx_plan* p = plans[i];
      hipStream_t st = (k % ns == 0) ? s : ss.st[k % ns - 1];
      int rc = par_forward_launch(p, ins[i], outs[i], st, !(p->flags & VAMX_PAR_NO_WARP_CLIP));
      if (rc) return rc;
    }
    if (ns > 1) {
      for (int j = 0; j < kSideStreams; ++j) {
        VAMX_HIP(hipEventRecord(ss.join[j], ss.st[j]));
        VAMX_HIP(hipStreamWaitEvent(s, ss.join[j], 0));
      }
    }
  }
  if (shared.empty()) return VAMX_OK;
  const int m = (int)shared.size();
  std::vector<BatchItem> items((size_t)m);
  std::vector<int> order((size_t)m);
  for (int si = 0; si < m; ++si) {
    const int i = shared[(size_t)si];
    vamx_plan* p = plans[i];
    BatchItem& it = items[(size_t)si];
    it.g = p->par;
    it.in = ins[i];
    it.out = outs[i];
    it.minmax = p->d_minmax;
    it.clip = forward ? !(p->flags & VAMX_PAR_NO_WARP_CLIP) : !(p->flags & VAMX_PAR_NO_CIRCLE_CLIP);
    int wz = par_wave_zv(p, ins[i], outs[i], forward ? it.g.fp_zv : it.g.bp_zv);
    it.split = par_wave_split(p, wz) ? 1 : 0;
    if (it.split) wz = 2;
    it.pair = 0;
    if (wz) {
      it.wave = 1;
      it.zv = wz;
      it.pair = forward && it.g.fp_pair && it.g.pair_ok ? 1 : 0;
      const long long units = it.pair ? fp_pair_tiled(it.g) : forward ? fp_tiled_rays(it.g) : bp_tiled_quads(it.g);
      it.items = units * wave_zchunks(it.g.nz, wz);
    } else {
      it.wave = 0;
      it.zv = par_use_vec4(p, ins[i], outs[i]) ? 4 : 1;
      const long long outer = forward ? (long long)it.g.na * it.g.n : (long long)it.g.n * it.g.n;
      it.items = outer * (it.g.nz / it.zv);
    }
    if (forward && it.clip && !it.g.incl) {
      int rc = par_slice_minmax(p, ins[i], s);
      if (rc) return rc;
    }
  }
  // largest work first (cost ~ items * inner loop length)
  std::iota(order.begin(), order.end(), 0);
  auto cost = [&](int i) {   // interpolated samples of the problem
    const BatchItem& it = items[(size_t)i];
    return (double)it.g.n * it.g.n * it.g.na * it.g.nz;
  };
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cost(a) > cost(b); });
  std::vector<BatchItem> sorted((size_t)m);
  long long tiles = 0;
  for (int r = 0; r < m; ++r) {
    sorted[(size_t)r] = items[(size_t)order[(size_t)r]];
    sorted[(size_t)r].tile_begin = tiles;
    tiles += (sorted[(size_t)r].items + (sorted[(size_t)r].wave ? 3 : 255)) / (sorted[(size_t)r].wave ? 4 : 256);
  }
  BatchScratch& sc = g_scratch;
  if (sc.device != dev) {
    for (auto& t : sc.tables) (void)hipFree(t.d_items);
    sc.tables.clear();
    if (sc.d_counter) (void)hipFree(sc.d_counter);
    sc.d_counter = nullptr;
    VAMX_HIP(hipMalloc((void**)&sc.d_counter, sizeof(unsigned long long)));
    sc.device = dev;
  }
  const size_t bytes = (size_t)m * sizeof(BatchItem);
  const char* raw = reinterpret_cast<const char*>(sorted.data());
  BatchItem* d_items = nullptr;
  for (size_t t = 0; t < sc.tables.size(); ++t) {
    if (sc.tables[t].bytes.size() == bytes && std::memcmp(sc.tables[t].bytes.data(), raw, bytes) == 0) {
      BatchTable hit = std::move(sc.tables[t]);
      sc.tables.erase(sc.tables.begin() + (long)t);
      sc.tables.push_back(std::move(hit));
      d_items = sc.tables.back().d_items;
      break;
    }
  }
  if (!d_items) {
    // a new table: blocking upload (not allowed under graph capture, where
    // hipStreamSynchronize fails and the error is returned)
    VAMX_HIP(hipStreamSynchronize(s));
    if (sc.tables.size() >= kMaxTables) {
      (void)hipFree(sc.tables.front().d_items);
      sc.tables.erase(sc.tables.begin());
    }
    BatchTable nt;
    VAMX_HIP(hipMalloc((void**)&nt.d_items, bytes));
    hipError_t e = hipMemcpy(nt.d_items, raw, bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      (void)hipFree(nt.d_items);
      return hip_check(e, "hipMemcpy(batch table)");
    }
    nt.bytes.assign(raw, raw + bytes);
    sc.tables.push_back(std::move(nt));
    d_items = sc.tables.back().d_items;
  }
  hipLaunchKernelGGL(batch_reset, dim3(1), dim3(1), 0, s, sc.d_counter);
  long long grid = (long long)std::max(cus, 1) * 8;
  if (grid > tiles) grid = tiles;
  if (forward)
    hipLaunchKernelGGL(batch_kernel<true>, dim3((unsigned)grid), dim3(256), 0, s, d_items, m, tiles,
                       sc.d_counter);
  else
    hipLaunchKernelGGL(batch_kernel<false>, dim3((unsigned)grid), dim3(256), 0, s, d_items, m, tiles,
                       sc.d_counter);
  VAMX_HIP(hipGetLastError());
  return VAMX_OK;
}

}  // namespace
}  // namespace vamx

extern "C" int vamx_par_forward_batched(vamx_plan* const* plans, int32_t n_plans, const float* const* vols,
                                        float* const* sinos, void* stream) {
  return vamx::run_batched(plans, n_plans, vols, sinos, (hipStream_t)stream, true);
}

extern "C" int vamx_par_backward_batched(vamx_plan* const* plans, int32_t n_plans, const float* const* sinos,
                                         float* const* vols, void* stream) {
  return vamx::run_batched(plans, n_plans, sinos, vols, (hipStream_t)stream, false);
}
