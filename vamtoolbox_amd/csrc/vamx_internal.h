// Internal declarations shared by the translation units of libvamx_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/vamx.h"

namespace vamx {

enum PlanKind { PLAN_PARALLEL = 1, PLAN_RAYTRACE = 2 };

// Device-side description of a parallel-beam problem, passed by value.
struct ParDev {
  int n, nz, na, c;          // c = n / 2 (Projector3DParallel.py:64-67)
  double det_offset;         // detector-u offset
  double xp0, xp_last;       // first / last np.interp node: (0 - c) - d, (n-1-c) - d
  double tan_phi, inv_cos;   // inclination
  double mu, radius;         // absorption
  int incl;                  // tan_phi != 0
  const double* mats;        // [na][6]: R00 R01 R02 R10 R11 R12 (Projector3DParallel.py:137-143)
  const double* cs;          // [na][2]: cos, sin of deg2rad(angle)
  const uint8_t* outside;    // [n*n] clipToCircle mask
  // occlusion shadowing (null = none): first occluding row coordinate per
  // (detector bin, angle, slice), NaN where none, and per (angle, slice)
  // whether the forward projection applies it
  const float* occ;          // [n][na][nz]
  const uint8_t* occ_active; // [na][nz]
  // work order of the wavefront-uniform kernels (set by vamx_par_set_tuning)
  int fp_ta, fp_tk, fp_zv;   // forward: tile of fp_ta angles x fp_tk rays, z values per lane
  int fp_reuse;              // forward: keep the previous sample's taps in registers
  int bp_tile, bp_zv;        // back projection: bp_tile x bp_tile voxel columns, z values per lane
  int z_split;               // allow lanes holding z and z + 64 where one z per lane is the choice
  // Opposite-angle pairing of the forward projection.  Ray k at angle a and
  // ray 2c + D - k at angle a + 180 deg (D = 2 det_offset, an integer) sample
  // the same points in reverse order, so one pass over the taps serves both.
  // Work units are angles: (primary, partner) or (angle, -1).  Of the partner
  // angle's rays, pair_u have no primary ray; they start at pair_first2.
  int fp_pair;               // tuning: pair opposite angles when pair_ok
  int pair_ok;               // the geometry has opposite angles and D is an integer
  int n_units, pair_D, pair_u, pair_first2;
  const int* units;          // [n_units][2]
};

// Geometry of the curved-ray tracer, passed by value.
struct RtDev {
  int nx, ny, nz;
  double xmin, ymin, zmin, dx, dy, dz;
  double inv_dx, inv_dy, inv_dz;
  double span_x, span_y, span_z;
  double xmax, ymax, zmax;
  double ixmax, iymax, izmax;  // largest index coordinate: size - 1
  int nt, na, nzr;
  double incl;               // radians
  double step;
  // derived on the host once (kernel arguments in scalar registers, not
  // per-lane copies): step / dx, the exit bounds min - d and max + d per axis
  // (+-inf on z in 2-D), att_radius^2
  double hw;
  double exit_lo_x, exit_lo_y, exit_lo_z, exit_hi_x, exit_hi_y, exit_hi_z;
  double att_r2;
  int max_steps, exit_every, solver, param;
  float energy0;
  const float4* field;       // (n, gx, gy, gz)
  int index_kind;            // VAMX_RT_INDEX_*
  double lune_R, lune_p, lune_nsur, lune_c;   // analytical Luneburg lens; lune_c = (n_sur / R)^p
  int att_kind;
  double att_alpha, att_radius;
  const float* att_grid;
  const float* abs_grid;
  const double* ray_t;       // [nt] lateral coordinates   (sino_n0)
  const double* ray_z;       // [nzr] vertical coordinates (sino_n2)
  const double* ray_cs;      // [na][2] cos, sin of azimuth
  double radial_offset;      // proj_plane_radial_offset
};

// serial number of a new plan (unique for the process: a batch names its
// plans by serial, so a new plan at a freed plan's address is another plan)
unsigned long long next_plan_serial();

}  // namespace vamx

struct vamx_plan {
  int kind = 0;
  int device = 0;
  unsigned long long serial = vamx::next_plan_serial();
  // parallel
  vamx::ParDev par{};
  int flags = 0;
  double* d_mats = nullptr;
  double* d_cs = nullptr;
  uint8_t* d_outside = nullptr;
  float* d_minmax_part = nullptr;   // [parts][2*nz]
  float* d_minmax = nullptr;        // [2*nz]
  float* d_occ = nullptr;           // [n*na*nz] occlusion sinogram
  uint8_t* d_occ_active = nullptr;  // [na*nz]
  int* d_units = nullptr;           // [n_units][2] forward work units (opposite-angle pairs)
  int minmax_parts = 0;
  std::vector<double> h_angles;
  // ray tracer
  vamx::RtDev rt{};
  double* d_rt_tables = nullptr;
  unsigned long long* d_accum = nullptr;  // fixed-point deposit accumulator
  float* d_gmax = nullptr;                 // device scalar: max |g| for the fixed-point scale
  unsigned long long* d_stats = nullptr;   // [2]
  hipStream_t last_stream = nullptr;
};

namespace vamx {
int set_error(int code, const std::string& msg);
int hip_check(hipError_t e, const char* what);
bool check_plan(vamx_plan* p, int kind);
unsigned blocks_for(long long total, int block);
// releases the batches of vamx_par_{forward,backward}_batched that hold p
void batch_forget_plan(const vamx_plan* p);
}  // namespace vamx

#define VAMX_HIP(call)                                         \
  do {                                                         \
    hipError_t _e = (call);                                    \
    if (_e != hipSuccess) return vamx::hip_check(_e, #call);   \
  } while (0)

// XCD-aware block remap: HIP dispatches consecutive workgroups round robin to
// the 8 XCDs of an MI355X; this maps the hardware block id to a logical id so
// that each XCD works on one contiguous range of logical blocks (shared L2).
__device__ __forceinline__ unsigned xcd_remap(unsigned b, unsigned nblocks) {
  const unsigned nx = 8u;
  unsigned per = nblocks / nx;
  unsigned rem = nblocks % nx;
  unsigned xcd = b % nx;
  unsigned idx = b / nx;
  // XCDs [0, rem) own per+1 logical blocks, the others own per.
  unsigned base = (xcd < rem) ? xcd * (per + 1) : rem * (per + 1) + (xcd - rem) * per;
  unsigned lb = base + idx;
  return lb < nblocks ? lb : b;
}
