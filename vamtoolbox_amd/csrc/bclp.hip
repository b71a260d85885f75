// Fused BCLP iteration (band-constraint Lp-norm sinogram optimisation) for gfx950.
//
// Reference: vamtoolbox/optimizer/BCLP.py, BCLPNorm.updateVariables /
// computeLoss / computeLossGradient (277-386) and gradientDescent (477-534);
// material response vamtoolbox/response.py:191-291.
//
// One iteration is three kernels on one stream and no host synchronisation:
//   A  back projection of g fused with the response map, band error, Lp
//      weighting and a per-block partial of the loss (fixed-order tree sum);
//   B  one block reduces the partials in a fixed order, logs the loss, and
//      derives the gradient scale q L^((q-p)/q);
//   C  forward projection of the operand (or the exact transpose of the back
//      projection) fused with g <- clip(g - lr * scale * grad, glb, gub).
// The per-voxel epilogue runs in float64, as the reference does in NumPy.
#include <mutex>

#include "bclp_core.h"
#include "par_core.h"

namespace vamx {

// A: back projection + response + band/Lp epilogue + loss partials
template <int ZV, bool INCL>
__global__ __launch_bounds__(256) void bclp_bp_epilogue(ParDev g, int clipc, RespDev r, BclpDev b,
                                                        const float* __restrict__ gsino,
                                                        const float* __restrict__ target,
                                                        float* __restrict__ dose, float* __restrict__ operand,
                                                        double* __restrict__ partials) {
  __shared__ double sh[256];
  const int nzv = g.nz / ZV;
  const long long total = (long long)g.n * g.n * nzv;
  const unsigned lb = xcd_remap(blockIdx.x, gridDim.x);
  const long long tid = (long long)lb * blockDim.x + threadIdx.x;
  double part = 0.0;
  if (tid < total) {
    const int z = (int)(tid % nzv) * ZV;
    const int pix = (int)(tid / nzv);
    const size_t base = (size_t)pix * g.nz + z;
    typename VecT<ZV>::T d = vzero<ZV>();
    if (!(clipc && g.outside[pix])) {
      if constexpr (INCL) d = bp_voxel_incl(g, gsino, pix, z);
      else d = bp_voxel<ZV>(g, gsino, pix, z);
    }
    vstore<ZV>(dose + base, d);
    typename VecT<ZV>::T op = vzero<ZV>();
    const typename VecT<ZV>::T tg = vload<ZV>(target + base);
#pragma unroll
    for (int q = 0; q < ZV; ++q) {
      float o;
      part += bclp_voxel(r, b, vget(d, q), vget(tg, q), base + q, o);
      vset(op, q, o);
    }
    vstore<ZV>(operand + base, op);
  }
  const double s = block_sum_256(part, sh);
  if (threadIdx.x == 0) partials[lb] = s;
}

// A (wavefront-uniform variant): one wavefront = one 2x2 quad of voxel
// columns x 64*ZV z
template <int ZV>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6))) void bclp_bp_epilogue_wave(ParDev g, int clipc, RespDev r, BclpDev b,
                                                             const float* __restrict__ gsino,
                                                             const float* __restrict__ target,
                                                             float* __restrict__ dose, float* __restrict__ operand,
                                                             double* __restrict__ partials) {
  __shared__ double sh[256];
  const int lane = threadIdx.x & 63;
  const unsigned lb = xcd_remap(blockIdx.x, gridDim.x);
  const long long wave = wave_index(lb);
  double part = 0.0;
  int r0 = 0, c0 = 0, z = 0;
  if (bp_wave_quad<ZV>(wave, g, lane, r0, c0, z)) {
    const int zl = z < g.nz - ZV ? z : g.nz - ZV;
    unsigned inrange, active;
    bp_quad_masks(g, r0, c0, clipc, inrange, active);
    typename VecT<ZV>::T acc[4];
    bp_quad_wave<ZV>(g, gsino, r0, c0, active, zl, lane, acc);
    if (z >= g.nz) inrange = 0u;   // lanes past nz: no store, no loss term
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (!((inrange >> q) & 1u)) continue;
      const size_t base = ((size_t)(r0 + (q >> 1)) * g.n + c0 + (q & 1)) * g.nz + z;
      const typename VecT<ZV>::T d = ((active >> q) & 1u) ? acc[q] : vzero<ZV>();
      vstore<ZV>(dose + base, d);
      typename VecT<ZV>::T op = vzero<ZV>();
      const typename VecT<ZV>::T tg = vload<ZV>(target + base);
#pragma unroll
      for (int e = 0; e < ZV; ++e) {
        float o;
        part += bclp_voxel(r, b, vget(d, e), vget(tg, e), base + e, o);
        vset(op, e, o);
      }
      vstore<ZV>(operand + base, op);
    }
  }
  const double s = block_sum_256(part, sh);
  if (threadIdx.x == 0) partials[lb] = s;
}

// C (wavefront-uniform variant)
template <int ZV, bool REUSE>
__global__ __launch_bounds__(256) void bclp_fp_update_wave(ParDev g, const float* __restrict__ operand,
                                                           float* __restrict__ gsino,
                                                           const float* __restrict__ minmax, int clip,
                                                           const double* __restrict__ scale_p, float lr, float glb,
                                                           float gub) {
  const int lane = threadIdx.x & 63;
  const long long wave = wave_index(xcd_remap(blockIdx.x, gridDim.x));
  int k, i, z;
  if (!fp_wave_item<ZV>(wave, g, lane, k, i, z)) return;
  const int zl = z < g.nz - ZV ? z : g.nz - ZV;
  const bool clip_on = clip && wave_clip_active<ZV>(minmax, g.nz, zl);
  const typename VecT<ZV>::T fp = clip_on ? fp_ray_wave<ZV, true, REUSE>(g, operand, minmax, k, i, zl, lane)
                                          : fp_ray_wave<ZV, false, REUSE>(g, operand, minmax, k, i, zl, lane);
  if (z >= g.nz) return;
  const float scale = (float)(*scale_p);
  float* gp = gsino + ((size_t)k * g.na + i) * g.nz + z;
  typename VecT<ZV>::T gv = vload<ZV>(gp);
#pragma unroll
  for (int q = 0; q < ZV; ++q) vset(gv, q, clampf(vget(gv, q) - lr * (scale * vget(fp, q)), glb, gub));
  vstore<ZV>(gp, gv);
}

// C, paired: the opposite ray (k2, i2) of a pair is projected from the same
// taps and updated too
template <int ZV, bool REUSE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) void bclp_fp_update_pair(ParDev g, const float* __restrict__ operand,
                                                           float* __restrict__ gsino,
                                                           const float* __restrict__ minmax, int clip,
                                                           const double* __restrict__ scale_p, float lr, float glb,
                                                           float gub) {
  using V = typename VecT<ZV>::T;
  const int lane = threadIdx.x & 63;
  const long long wave = wave_index(xcd_remap(blockIdx.x, gridDim.x));
  const PairItem it = fp_pair_item<ZV>(wave, g, lane);
  if (!it.ok) return;
  const int k = it.k, i = it.i, k2 = it.k2, i2 = it.i2, z = it.z;
  const int zl = z < g.nz - ZV ? z : g.nz - ZV;
  const bool clip_on = clip && wave_clip_active<ZV>(minmax, g.nz, zl);
  const float scale = (float)(*scale_p);
  auto update = [&](int kh, int ih, const V& fp) {
    float* gp = gsino + ((size_t)kh * g.na + ih) * g.nz + z;
    V gv = vload<ZV>(gp);
#pragma unroll
    for (int q = 0; q < ZV; ++q) vset(gv, q, clampf(vget(gv, q) - lr * (scale * vget(fp, q)), glb, gub));
    vstore<ZV>(gp, gv);
  };
  if (clip_on) {
    // the clipped sample is not continuous in its position: the two rays are
    // projected one after the other
    for (int h = 0; h < (k2 >= 0 ? 2 : 1); ++h) {
      const int kh = h ? k2 : k, ih = h ? i2 : i;
      const V fp = fp_ray_wave<ZV, true, REUSE>(g, operand, minmax, kh, ih, zl, lane);
      if (z < g.nz) update(kh, ih, fp);
    }
    return;
  }
  V fp, fp2 = vzero<ZV>();
  if (k2 >= 0)
    fp = fp_ray_wave<ZV, false, REUSE, true>(g, operand, minmax, k, i, zl, lane, &fp2);
  else
    fp = fp_ray_wave<ZV, false, REUSE>(g, operand, minmax, k, i, zl, lane);
  if (z >= g.nz) return;
  update(k, i, fp);
  if (k2 >= 0) update(k2, i2, fp2);
}

// B: fixed-order reduction of the partials, loss log and gradient scale
__global__ __launch_bounds__(256) void bclp_reduce(const double* __restrict__ partials, int nparts, double dvol,
                                                   double p, double q, double* loss_dst, int32_t* counter,
                                                   double* loss_log, double* scale_out) {
  __shared__ double sh[256];
  double acc = 0.0;
  for (int i = threadIdx.x; i < nparts; i += 256) acc += partials[i];
  const double s = block_sum_256(acc, sh);
  if (threadIdx.x == 0) {
    const double L = pow(s * dvol, q / p);                     // BCLP.py:345-347
    if (counter) {
      loss_log[*counter] = L;
      *counter = *counter + 1;
    } else {
      *loss_dst = L;
    }
    *scale_out = (L != 0.0) ? q * pow(L, (q - p) / q) : 1.0;    // 370-375
  }
}

// C: forward projection of the operand fused with the sinogram update
template <int ZV, bool INCL, bool CLIP>
__global__ __launch_bounds__(256) void bclp_fp_update(ParDev g, const float* __restrict__ operand,
                                                      float* __restrict__ gsino, const float* __restrict__ minmax,
                                                      const double* __restrict__ scale_p, float lr, float glb,
                                                      float gub) {
  const int nzv = g.nz / ZV;
  const long long total = (long long)g.na * g.n * nzv;
  const long long tid = (long long)xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (tid >= total) return;
  const int z = (int)(tid % nzv) * ZV;
  const long long ray = tid / nzv;
  const int k = (int)(ray % g.n), i = (int)(ray / g.n);
  typename VecT<ZV>::T fp;
  if constexpr (INCL) fp = fp_ray_incl(g, operand, k, i, z);
  else fp = fp_ray<ZV, CLIP>(g, operand, minmax, k, i, z);
  if (z >= g.nz) return;
  const float scale = (float)(*scale_p);
  float* gp = gsino + ((size_t)k * g.na + i) * g.nz + z;
  typename VecT<ZV>::T gv = vload<ZV>(gp);
#pragma unroll
  for (int q = 0; q < ZV; ++q) {
    const float grad = scale * vget(fp, q);
    vset(gv, q, clampf(vget(gv, q) - lr * grad, glb, gub));  // BCLP.py:486-487, 532
  }
  vstore<ZV>(gp, gv);
}

// g <- clip(g - lr * scale * grad) on a lane's z values of one sinogram row
template <int ZV>
__device__ __forceinline__ void wave_update_store(float* row, int z, int nz, typename VecT<ZV>::T gr, float scale,
                                                  float lr, float glb, float gub) {
  if (z >= nz) return;
  typename VecT<ZV>::T gv = vload<ZV>(row + z);
#pragma unroll
  for (int q = 0; q < ZV; ++q) vset(gv, q, clampf(vget(gv, q) - lr * (scale * vget(gr, q)), glb, gub));
  vstore<ZV>(row + z, gv);
}

// C' on wavefronts: the exact transpose of the back projection (one bin, or a
// bin and its opposite bin, x 64*ZV z; the cores of par_bpT_wave_kernel /
// par_bpT_pair_kernel) fused with the sinogram update; the same values as
// the per-lane kernel
template <int ZV, bool PAIR>
__global__ __launch_bounds__(256) void bclp_bpT_update_wave(ParDev g, int clipc, const float* __restrict__ operand,
                                                            float* __restrict__ gsino,
                                                            const double* __restrict__ scale_p, float lr, float glb,
                                                            float gub) {
  using V = typename VecT<ZV>::T;
  const int lane = threadIdx.x & 63;
  const long long wave = wave_index(xcd_remap(blockIdx.x, gridDim.x));
  int k, i, k2 = -1, i2 = -1, z;
  if constexpr (PAIR) {
    const PairItem it = fp_pair_item<ZV, false>(wave, g, lane);
    if (!it.ok) return;
    k = it.k;
    i = it.i;
    k2 = it.k2;
    i2 = it.i2;
    z = it.z;
  } else {
    const int nch = wave_zchunks(g.nz, ZV);
    if (wave >= (long long)g.n * g.na * nch) return;   // whole wavefronts
    const long long bin = wave / nch;
    k = (int)(bin % g.n);
    i = (int)(bin / g.n);
    z = (int)(wave % nch) * 64 * ZV + lane * ZV;
  }
  int zl, zl2;
  wave_zload<ZV, false>(z, g.nz, zl, zl2);
  const float scale = (float)(*scale_p);
  if (k2 < 0) {
    const V gr = bpT_bin_wave<ZV>(g, operand, k, i, zl, lane, clipc);
    wave_update_store<ZV>(gsino + ((size_t)k * g.na + i) * g.nz, z, g.nz, gr, scale, lr, glb, gub);
    return;
  }
  V gr, gr2;
  bpT_pair_wave<ZV>(g, operand, k, i, k2, i2, zl, lane, clipc, gr, gr2);
  wave_update_store<ZV>(gsino + ((size_t)k * g.na + i) * g.nz, z, g.nz, gr, scale, lr, glb, gub);
  wave_update_store<ZV>(gsino + ((size_t)k2 * g.na + i2) * g.nz, z, g.nz, gr2, scale, lr, glb, gub);
}

// C': exact transpose of the back projection fused with the sinogram update
template <int ZV, bool INCL>
__global__ __launch_bounds__(256) void bclp_bpT_update(ParDev g, int clipc, const float* __restrict__ operand,
                                                       float* __restrict__ gsino, const double* __restrict__ scale_p,
                                                       float lr, float glb, float gub) {
  const int nzv = g.nz / ZV;
  const long long total = (long long)g.n * g.na * nzv;
  const long long tid = (long long)xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (tid >= total) return;
  const int z = (int)(tid % nzv) * ZV;
  const long long bin = tid / nzv;
  const int i = (int)(bin % g.na), k = (int)(bin / g.na);
  typename VecT<ZV>::T gr;
  if constexpr (INCL) gr = bpT_bin_incl(g, operand, k, i, z, clipc);
  else gr = bpT_bin<ZV>(g, operand, k, i, z, clipc);
  if (z >= g.nz) return;
  const float scale = (float)(*scale_p);
  float* gp = gsino + ((size_t)k * g.na + i) * g.nz + z;
  typename VecT<ZV>::T gv = vload<ZV>(gp);
#pragma unroll
  for (int q = 0; q < ZV; ++q) vset(gv, q, clampf(vget(gv, q) - lr * (scale * vget(gr, q)), glb, gub));
  vstore<ZV>(gp, gv);
}

int resp_fill_dev(const vamx_response* resp, RespDev& r) {
  if (!resp) return set_error(VAMX_ERR_ARG, "response: null model");
  if (resp->kind < VAMX_RESP_GLF || resp->kind > VAMX_RESP_TABLE)
    return set_error(VAMX_ERR_ARG, "response: unknown response kind");
  if (resp->kind == VAMX_RESP_TABLE && (resp->n_pts < 2 || !resp->table_f || !resp->table_map || !resp->table_dmap))
    return set_error(VAMX_ERR_ARG, "response: interpolation response needs n_pts >= 2 and three device tables");
  r.kind = resp->kind;
  r.A = resp->A; r.K = resp->K; r.B = resp->B; r.M = resp->M; r.nu = resp->nu;
  r.lin_M = resp->lin_M; r.lin_C = resp->lin_C;
  r.n_pts = resp->n_pts;
  r.tf = resp->table_f; r.tm = resp->table_map; r.td = resp->table_dmap;
  return VAMX_OK;
}

int bclp_fill_devs(const vamx_response* resp, const vamx_bclp_params* prm, const float* eps_arr, const float* w_arr,
                   RespDev& r, BclpDev& b) {
  if (!resp || !prm) return set_error(VAMX_ERR_ARG, "bclp: null response or params");
  const int rc = resp_fill_dev(resp, r);
  if (rc) return rc;
  if (!(prm->p > 0.0) || !(prm->q > 0.0)) return set_error(VAMX_ERR_ARG, "bclp: p and q must be > 0");
  b.p = prm->p; b.q = prm->q; b.eps = prm->eps; b.weight = prm->weight; b.dvol = prm->dvol;
  b.eps_arr = eps_arr;
  b.w_arr = w_arr;
  return VAMX_OK;
}

int bclp_reduce_launch(const double* partials, int nparts, const BclpDev& b, double* loss_dst, int32_t* counter,
                       double* loss_log, double* scale_out, hipStream_t s) {
  hipLaunchKernelGGL(bclp_reduce, dim3(1), dim3(256), 0, s, partials, nparts, b.dvol, b.p, b.q, loss_dst, counter,
                     loss_log, scale_out);
  VAMX_HIP(hipGetLastError());
  return VAMX_OK;
}

}  // namespace vamx

using namespace vamx;

namespace {

struct WsLayout {
  size_t partials_off, scale_off, operand_off, total;
  int nparts;
};

WsLayout ws_layout(const vamx_plan* p) {
  const ParDev& g = p->par;
  WsLayout w;
  const long long items = (long long)g.n * g.n * g.nz;  // worst case ZV = 1
  // wavefront path: one wave per quad and z chunk (widest grid: ZV = 1), any tile setting
  const long long tiled = bp_tiled_quads_max(g) * wave_zchunks(g.nz, 1) * 64;
  w.nparts = (int)blocks_for(items > tiled ? items : tiled, 256);
  w.partials_off = 0;
  w.scale_off = ((size_t)w.nparts * sizeof(double) + 255) / 256 * 256;
  w.operand_off = w.scale_off + 256;
  w.total = w.operand_off + (size_t)items * sizeof(float);
  return w;
}

// Launches kernel A; its per-block loss partials go to partials[0, nparts).
// Fails before launching when nparts would exceed max_parts.
int launch_epilogue(vamx_plan* p, const RespDev& r, const BclpDev& b, const float* g, const float* target,
                    float* dose, float* operand, double* partials, int max_parts, int& nparts, hipStream_t s) {
  const ParDev& gd = p->par;
  const int clipc = (p->flags & VAMX_PAR_NO_CIRCLE_CLIP) ? 0 : 1;
  const bool v4 = !gd.incl && gd.nz % 4 == 0 && ((uintptr_t)g | (uintptr_t)target | (uintptr_t)dose |
                                                 (uintptr_t)operand) % 16 == 0 &&
                  (!b.eps_arr || (uintptr_t)b.eps_arr % 16 == 0) && (!b.w_arr || (uintptr_t)b.w_arr % 16 == 0);
  int wz = 0;
  if (!(p->flags & VAMX_PAR_PER_LANE_GEOMETRY) && !gd.incl && gd.n >= 2 && !gd.occ) {
    const uintptr_t al = (uintptr_t)g | (uintptr_t)dose | (uintptr_t)target | (uintptr_t)operand |
                         (uintptr_t)b.eps_arr | (uintptr_t)b.w_arr;
    wz = wave_zv_choice(gd.nz, al, gd.bp_zv);
  }
  if (wz) {
    const long long threads = bp_tiled_quads(gd) * wave_zchunks(gd.nz, wz) * 64;
    nparts = (int)blocks_for(threads, 256);
    if (nparts > max_parts)
      return set_error(VAMX_ERR_ARG, "bclp: " + std::to_string(nparts) + " loss partials exceed the " +
                                          std::to_string(max_parts) + " slots of the scratch");
    const dim3 grid(nparts), block(256);
    if (wz == 4)
      hipLaunchKernelGGL(bclp_bp_epilogue_wave<4>, grid, block, 0, s, gd, clipc, r, b, g, target, dose, operand,
                         partials);
    else if (wz == 2)
      hipLaunchKernelGGL(bclp_bp_epilogue_wave<2>, grid, block, 0, s, gd, clipc, r, b, g, target, dose, operand,
                         partials);
    else
      hipLaunchKernelGGL(bclp_bp_epilogue_wave<1>, grid, block, 0, s, gd, clipc, r, b, g, target, dose, operand,
                         partials);
    VAMX_HIP(hipGetLastError());
    return VAMX_OK;
  }
  const long long total = (long long)gd.n * gd.n * (gd.nz / (v4 ? 4 : 1));
  nparts = (int)blocks_for(total, 256);
  if (nparts > max_parts)
    return set_error(VAMX_ERR_ARG, "bclp: " + std::to_string(nparts) + " loss partials exceed the " +
                                        std::to_string(max_parts) + " slots of the scratch");
  const dim3 grid(nparts), block(256);
  if (gd.incl)
    hipLaunchKernelGGL((bclp_bp_epilogue<1, true>), grid, block, 0, s, gd, clipc, r, b, g, target, dose, operand,
                       partials);
  else if (v4)
    hipLaunchKernelGGL((bclp_bp_epilogue<4, false>), grid, block, 0, s, gd, clipc, r, b, g, target, dose, operand,
                       partials);
  else
    hipLaunchKernelGGL((bclp_bp_epilogue<1, false>), grid, block, 0, s, gd, clipc, r, b, g, target, dose, operand,
                       partials);
  VAMX_HIP(hipGetLastError());
  return VAMX_OK;
}

}  // namespace

extern "C" int vamx_bclp_workspace_bytes(vamx_plan* plan, int64_t* bytes) {
  if (!check_plan(plan, PLAN_PARALLEL)) return set_error(VAMX_ERR_STATE, "vamx_bclp_workspace_bytes: not a parallel plan");
  if (!bytes) return set_error(VAMX_ERR_ARG, "vamx_bclp_workspace_bytes: null output");
  *bytes = (int64_t)ws_layout(plan).total;
  return VAMX_OK;
}

namespace {

// grad_op 0 (forward projector) or 1 (exact transpose of the back projection,
// which does not model occlusion: rejected for plans with an occlusion map,
// like vamx_par_backward_adjoint)
int check_grad_op(const vamx_plan* plan, const vamx_bclp_params* prm) {
  if (prm->grad_op != 0 && prm->grad_op != 1) return set_error(VAMX_ERR_ARG, "bclp: grad_op must be 0 or 1");
  if (prm->grad_op == 1 && plan->par.occ)
    return set_error(VAMX_ERR_STATE, "bclp: the exact gradient (grad_op 1) is not available with occlusion");
  return VAMX_OK;
}

// Kernel C (with the slice min/max pre-pass of the warp clip): the update
// g <- clip(g - lr * scale * A operand) with the device scale
int bclp_update_impl(vamx_plan* plan, const vamx_bclp_params* prm, float* g, const float* operand,
                     const double* scale, hipStream_t s) {
  int rc = check_grad_op(plan, prm);
  if (rc) return rc;
  const ParDev& gd = plan->par;
  const float lr = (float)prm->learning_rate, glb = (float)prm->glb, gub = (float)prm->gub;
  const bool v4 = !gd.incl && gd.nz % 4 == 0 && ((uintptr_t)g | (uintptr_t)operand) % 16 == 0;
  const int zv = v4 ? 4 : 1;
  if (prm->grad_op == 0) {
    const bool clip = !(plan->flags & VAMX_PAR_NO_WARP_CLIP);
    if (clip && !gd.incl) {
      rc = par_slice_minmax(plan, operand, s);
      if (rc) return rc;
    }
    const int wz = par_fp_wave_zv(plan, operand, g);
    if (wz && gd.fp_pair && gd.pair_ok) {
      const long long threads = fp_pair_tiled(gd) * wave_zchunks(gd.nz, wz) * 64;
      const dim3 grid(blocks_for(threads, 256)), block(256);
      const int cl = clip ? 1 : 0;
#define VAMX_FPP(Z, R)                                                                                 \
  hipLaunchKernelGGL((bclp_fp_update_pair<Z, R>), grid, block, 0, s, gd, operand, g, plan->d_minmax, cl, scale, \
                     lr, glb, gub)
      if (gd.fp_reuse) {
        if (wz == 4) VAMX_FPP(4, true);
        else if (wz == 2) VAMX_FPP(2, true);
        else VAMX_FPP(1, true);
      } else {
        if (wz == 4) VAMX_FPP(4, false);
        else if (wz == 2) VAMX_FPP(2, false);
        else VAMX_FPP(1, false);
      }
#undef VAMX_FPP
      VAMX_HIP(hipGetLastError());
      return VAMX_OK;
    }
    if (wz) {
      const long long threads = fp_tiled_rays(gd) * wave_zchunks(gd.nz, wz) * 64;
      const dim3 grid(blocks_for(threads, 256)), block(256);
      const int cl = clip ? 1 : 0;
#define VAMX_FPU(Z, R)                                                                                 \
  hipLaunchKernelGGL((bclp_fp_update_wave<Z, R>), grid, block, 0, s, gd, operand, g, plan->d_minmax, cl, scale, \
                     lr, glb, gub)
      if (gd.fp_reuse) {
        if (wz == 4) VAMX_FPU(4, true);
        else if (wz == 2) VAMX_FPU(2, true);
        else VAMX_FPU(1, true);
      } else {
        if (wz == 4) VAMX_FPU(4, false);
        else if (wz == 2) VAMX_FPU(2, false);
        else VAMX_FPU(1, false);
      }
#undef VAMX_FPU
      VAMX_HIP(hipGetLastError());
      return VAMX_OK;
    }
    const long long total = (long long)gd.na * gd.n * (gd.nz / zv);
    const dim3 grid(blocks_for(total, 256)), block(256);
    if (gd.incl)
      hipLaunchKernelGGL((bclp_fp_update<1, true, false>), grid, block, 0, s, gd, operand, g, plan->d_minmax, scale,
                         lr, glb, gub);
    else if (v4 && clip)
      hipLaunchKernelGGL((bclp_fp_update<4, false, true>), grid, block, 0, s, gd, operand, g, plan->d_minmax, scale,
                         lr, glb, gub);
    else if (v4)
      hipLaunchKernelGGL((bclp_fp_update<4, false, false>), grid, block, 0, s, gd, operand, g, plan->d_minmax, scale,
                         lr, glb, gub);
    else if (clip)
      hipLaunchKernelGGL((bclp_fp_update<1, false, true>), grid, block, 0, s, gd, operand, g, plan->d_minmax, scale,
                         lr, glb, gub);
    else
      hipLaunchKernelGGL((bclp_fp_update<1, false, false>), grid, block, 0, s, gd, operand, g, plan->d_minmax, scale,
                         lr, glb, gub);
  } else {
    const int clipc = (plan->flags & VAMX_PAR_NO_CIRCLE_CLIP) ? 0 : 1;
    const int wz = (!gd.incl && !(plan->flags & VAMX_PAR_PER_LANE_GEOMETRY)) ? par_wave_zv(plan, g, operand, 4) : 0;
    if (wz) {
      const bool pair = gd.fp_pair && gd.pair_ok;
      const long long waves = (pair ? fp_pair_tiled(gd) : (long long)gd.n * gd.na) * wave_zchunks(gd.nz, wz);
      const dim3 grid(blocks_for(waves * 64, 256)), block(256);
#define VAMX_BTU(Z, P) \
  hipLaunchKernelGGL((bclp_bpT_update_wave<Z, P>), grid, block, 0, s, gd, clipc, operand, g, scale, lr, glb, gub)
      if (pair) {
        if (wz == 4) VAMX_BTU(4, true);
        else if (wz == 2) VAMX_BTU(2, true);
        else VAMX_BTU(1, true);
      } else {
        if (wz == 4) VAMX_BTU(4, false);
        else if (wz == 2) VAMX_BTU(2, false);
        else VAMX_BTU(1, false);
      }
#undef VAMX_BTU
      VAMX_HIP(hipGetLastError());
      return VAMX_OK;
    }
    const long long total = (long long)gd.n * gd.na * (gd.nz / zv);
    const dim3 grid(blocks_for(total, 256)), block(256);
    if (gd.incl)
      hipLaunchKernelGGL((bclp_bpT_update<1, true>), grid, block, 0, s, gd, clipc, operand, g, scale, lr, glb, gub);
    else if (v4)
      hipLaunchKernelGGL((bclp_bpT_update<4, false>), grid, block, 0, s, gd, clipc, operand, g, scale, lr, glb, gub);
    else
      hipLaunchKernelGGL((bclp_bpT_update<1, false>), grid, block, 0, s, gd, clipc, operand, g, scale, lr, glb, gub);
  }
  VAMX_HIP(hipGetLastError());
  return VAMX_OK;
}

// Kernels A, B, C of one iteration; partials has max_parts slots, scale one.
int bclp_step_impl(vamx_plan* plan, const RespDev& r, const BclpDev& b, const vamx_bclp_params* prm, float* g,
                   const float* target, float* dose, float* operand, double* partials, int max_parts,
                   double* scale, double* loss_log, int32_t* iter_counter, hipStream_t s) {
  int nparts = 0;
  int rc = check_grad_op(plan, prm);
  if (rc) return rc;
  rc = launch_epilogue(plan, r, b, g, target, dose, operand, partials, max_parts, nparts, s);
  if (rc) return rc;
  hipLaunchKernelGGL(bclp_reduce, dim3(1), dim3(256), 0, s, partials, nparts, b.dvol, b.p, b.q, (double*)nullptr,
                     iter_counter, loss_log, scale);
  VAMX_HIP(hipGetLastError());
  return bclp_update_impl(plan, prm, g, operand, scale, s);
}


}  // namespace

extern "C" int vamx_bclp_step(vamx_plan* plan, const vamx_response* resp, const vamx_bclp_params* prm, float* g,
                              const float* target, const float* eps_arr, const float* weight_arr, float* dose,
                              double* loss_log, int32_t* iter_counter, void* workspace, void* stream) {
  if (!check_plan(plan, PLAN_PARALLEL)) return set_error(VAMX_ERR_STATE, "vamx_bclp_step: not a parallel plan");
  if (!g || !target || !dose || !loss_log || !iter_counter || !workspace)
    return set_error(VAMX_ERR_ARG, "vamx_bclp_step: null buffer");
  RespDev r;
  BclpDev b;
  int rc = bclp_fill_devs(resp, prm, eps_arr, weight_arr, r, b);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  const WsLayout w = ws_layout(plan);
  char* ws = (char*)workspace;
  double* partials = (double*)(ws + w.partials_off);
  double* scale = (double*)(ws + w.scale_off);
  float* operand = (float*)(ws + w.operand_off);
  return bclp_step_impl(plan, r, b, prm, g, target, dose, operand, partials, w.nparts, scale, loss_log,
                        iter_counter, s);
}

extern "C" int vamx_bclp_update(vamx_plan* plan, const vamx_bclp_params* prm, float* g, void* workspace,
                                void* stream) {
  if (!check_plan(plan, PLAN_PARALLEL)) return set_error(VAMX_ERR_STATE, "vamx_bclp_update: not a parallel plan");
  if (!prm || !g || !workspace) return set_error(VAMX_ERR_ARG, "vamx_bclp_update: null argument");
  const WsLayout w = ws_layout(plan);
  char* ws = (char*)workspace;
  return bclp_update_impl(plan, prm, g, (const float*)(ws + w.operand_off), (const double*)(ws + w.scale_off),
                          (hipStream_t)stream);
}

extern "C" int vamx_bclp_eval(vamx_plan* plan, const vamx_response* resp, const vamx_bclp_params* prm,
                              const float* g, const float* target, const float* eps_arr, const float* weight_arr,
                              float* dose, float* operand, double* loss_out, void* workspace, void* stream) {
  if (!check_plan(plan, PLAN_PARALLEL)) return set_error(VAMX_ERR_STATE, "vamx_bclp_eval: not a parallel plan");
  if (!g || !target || !dose || !loss_out || !workspace)
    return set_error(VAMX_ERR_ARG, "vamx_bclp_eval: null buffer");
  RespDev r;
  BclpDev b;
  int rc = bclp_fill_devs(resp, prm, eps_arr, weight_arr, r, b);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  const WsLayout w = ws_layout(plan);
  char* ws = (char*)workspace;
  double* partials = (double*)(ws + w.partials_off);
  double* scale = (double*)(ws + w.scale_off);
  float* op = operand ? operand : (float*)(ws + w.operand_off);
  int nparts = 0;
  rc = launch_epilogue(plan, r, b, g, target, dose, op, partials, w.nparts, nparts, s);
  if (rc) return rc;
  hipLaunchKernelGGL(bclp_reduce, dim3(1), dim3(256), 0, s, partials, nparts, b.dvol, b.p, b.q, loss_out,
                     (int32_t*)nullptr, (double*)nullptr, scale);
  VAMX_HIP(hipGetLastError());
  return VAMX_OK;
}

// ---------------------------------------------------------------------------
// Element-wise material response: mapped = map(x), dmap = map'(x), in float64
// from float32 input (response.py:191-291), the forward and the saved
// derivative of a response autograd node.
// ---------------------------------------------------------------------------
namespace {

__global__ __launch_bounds__(256) void response_map_kernel(RespDev r, const float* __restrict__ x,
                                                           float* __restrict__ mapped, float* __restrict__ dmap,
                                                           long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double m, dm;
  response_eval(r, (double)x[i], m, dm);
  if (mapped) mapped[i] = (float)m;
  if (dmap) dmap[i] = (float)dm;
}

}  // namespace

extern "C" int vamx_response_map(const vamx_response* resp, const float* x, float* mapped, float* dmap, int64_t n,
                                 void* stream) {
  if (!x || (!mapped && !dmap)) return set_error(VAMX_ERR_ARG, "vamx_response_map: null buffer");
  if (n < 0) return set_error(VAMX_ERR_ARG, "vamx_response_map: negative length");
  RespDev r;
  const int rc = resp_fill_dev(resp, r);
  if (rc) return rc;
  if (n == 0) return VAMX_OK;
  hipLaunchKernelGGL(response_map_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, (hipStream_t)stream, r, x, mapped,
                     dmap, (long long)n);
  VAMX_HIP(hipGetLastError());
  return VAMX_OK;
}

// ---------------------------------------------------------------------------
// Allocation-free variant of vamx_bclp_step.  The loss partials and the
// gradient scale live in `scratch`, a file-scope device array that every GPU
// the library runs on holds until the process ends (1 MiB per GPU).  Host
// code never uses the array's name as an address: its address on each device
// comes from hipGetSymbolAddress, called once per device with that device
// current, and only that pointer is passed to kernels.  Every call on one GPU
// shares the array, so calls that overlap on one GPU race (one stream at a
// time per GPU).  Each call writes every partial slot it reads and the scale
// slot before reading them, so nothing has to be cleared between calls.
// ---------------------------------------------------------------------------
namespace {

constexpr int kScratchDoubles = 131072;              // 1 MiB
constexpr int kScratchParts = kScratchDoubles - 32;  // loss partial slots; the scale sits after them
static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxScratchDevices] = {};

// device address of `scratch` on the current device (looked up once per device)
int scratch_on_device(int dev, double** out) {
  if (dev < 0 || dev >= kMaxScratchDevices) return set_error(VAMX_ERR_ARG, "bclp noalloc: device id out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_check(e, "hipGetSymbolAddress(scratch)");
    if (!ptr) return set_error(VAMX_ERR_HIP, "bclp noalloc: scratch has no device address");
    g_scratch_ptr[dev] = (double*)ptr;
  }
  *out = g_scratch_ptr[dev];
  return VAMX_OK;
}

}  // namespace

extern "C" int vamx_bclp_step_noalloc(vamx_plan* plan, const vamx_response* resp, const vamx_bclp_params* prm,
                                      float* g, const float* target, const float* eps_arr, const float* weight_arr,
                                      float* dose, float* operand, double* loss_log, int32_t* iter_counter,
                                      void* stream) {
  if (!check_plan(plan, PLAN_PARALLEL)) return set_error(VAMX_ERR_STATE, "vamx_bclp_step_noalloc: not a parallel plan");
  if (!g || !target || !dose || !operand || !loss_log || !iter_counter)
    return set_error(VAMX_ERR_ARG, "vamx_bclp_step_noalloc: null buffer");
  RespDev r;
  BclpDev b;
  int rc = bclp_fill_devs(resp, prm, eps_arr, weight_arr, r, b);
  if (rc) return rc;
  int dev = -1;
  VAMX_HIP(hipGetDevice(&dev));
  if (dev != plan->device)
    return set_error(VAMX_ERR_ARG, "vamx_bclp_step_noalloc: the plan's device must be the current device");
  double* sp = nullptr;
  rc = scratch_on_device(dev, &sp);
  if (rc) return rc;
  return bclp_step_impl(plan, r, b, prm, g, target, dose, operand, sp, kScratchParts, sp + kScratchParts, loss_log,
                        iter_counter, (hipStream_t)stream);
}
