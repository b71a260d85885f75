// Algebraic (sparse-matrix) propagation: y = P x for a CSR matrix P with
// z-tiling (gfx950).
//
// Reference: vamtoolbox/projector/pyTorchAlgebraicPropagation.py:13-182 --
// P encodes one z layer (n_rows x n_cols); a volume of n_cols x Z voxels,
// flattened with z fastest, is multiplied as the (n_cols, Z) matrix X and the
// sinogram is the (n_rows, Z) product, again z fastest (torch.sparse.mm of
// lines 152-182).  The adjoint uses the explicit CSR transpose.
//
// Z >= 2: one wavefront per (row, 64*ZV consecutive z): the row's column
// indices and values are wave-uniform (scalar loads) and every lane reads
// ZV consecutive floats of X's row -- coalesced like the projector kernels.
// Z == 1 (plain SpMV): one wavefront per row, lanes stride the row's entries
// and a fixed butterfly reduction sums them.  Both orders are fixed, so the
// result is deterministic.
#include "par_core.h"

namespace vamx {

template <int ZV>
__global__ __launch_bounds__(256) void csr_rowz_kernel(const int64_t* __restrict__ rp, const int32_t* __restrict__ ci,
                                                       const float* __restrict__ val, long long n_rows, int Z,
                                                       const float* __restrict__ x, float* __restrict__ y) {
  const int lane = threadIdx.x & 63;
  const long long wave = wave_index(blockIdx.x);
  const int nch = wave_zchunks(Z, ZV);
  if (wave >= n_rows * nch) return;
  const long long r = wave / nch;
  const int z = (int)(wave % nch) * 64 * ZV + lane * ZV;
  const int zl = z < Z - ZV ? z : Z - ZV;   // lanes past Z read a valid copy and store nothing
  using V = typename VecT<ZV>::T;
  V acc = vzero<ZV>();
  const long long k1 = rp[r + 1];
  for (long long k = rp[r]; k < k1; ++k)
    acc = vfma(val[k], vload<ZV>(x + (size_t)ci[k] * Z + zl), acc);
  if (z < Z) vstore<ZV>(y + (size_t)r * Z + z, acc);
}

__global__ __launch_bounds__(256) void csr_vec_kernel(const int64_t* __restrict__ rp, const int32_t* __restrict__ ci,
                                                      const float* __restrict__ val, long long n_rows,
                                                      const float* __restrict__ x, float* __restrict__ y) {
  const int lane = threadIdx.x & 63;
  const long long r = wave_index(blockIdx.x);
  if (r >= n_rows) return;
  float acc = 0.0f;
  const long long k1 = rp[r + 1];
  for (long long k = rp[r] + lane; k < k1; k += 64) acc = fmaf(val[k], x[ci[k]], acc);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
  if (lane == 0) y[r] = acc;
}

}  // namespace vamx

using namespace vamx;

extern "C" int vamx_csr_spmm(const int64_t* row_ptr, const int32_t* col_idx, const float* values, int64_t n_rows,
                             int64_t n_cols, int32_t z_tiling, const float* x, float* y, void* stream) {
  if (!row_ptr || !col_idx || !values || !x || !y) return set_error(VAMX_ERR_ARG, "vamx_csr_spmm: null buffer");
  if (n_rows < 0 || n_cols < 1 || z_tiling < 1) return set_error(VAMX_ERR_ARG, "vamx_csr_spmm: invalid shape");
  if (n_rows == 0) return VAMX_OK;
  hipStream_t s = (hipStream_t)stream;
  if (z_tiling == 1) {
    hipLaunchKernelGGL(csr_vec_kernel, dim3(blocks_for(n_rows * 64, 256)), dim3(256), 0, s, row_ptr, col_idx, values,
                       (long long)n_rows, x, y);
  } else {
    int zv = wave_zv_choice(z_tiling, (uintptr_t)x | (uintptr_t)y, 4);
    if (!zv) zv = 1;   // thin z: one value per lane, part of the lanes idle
    const long long waves = (long long)n_rows * wave_zchunks(z_tiling, zv);
    const dim3 grid(blocks_for(waves * 64, 256)), block(256);
    if (zv == 4)
      hipLaunchKernelGGL(csr_rowz_kernel<4>, grid, block, 0, s, row_ptr, col_idx, values, (long long)n_rows,
                         (int)z_tiling, x, y);
    else if (zv == 2)
      hipLaunchKernelGGL(csr_rowz_kernel<2>, grid, block, 0, s, row_ptr, col_idx, values, (long long)n_rows,
                         (int)z_tiling, x, y);
    else
      hipLaunchKernelGGL(csr_rowz_kernel<1>, grid, block, 0, s, row_ptr, col_idx, values, (long long)n_rows,
                         (int)z_tiling, x, y);
  }
  VAMX_HIP(hipGetLastError());
  return VAMX_OK;
}
