/*
 * vamx.h -- C ABI of the MI355X (gfx950) light-propagation library for
 * volumetric additive manufacturing (libvamx_hip.so).
 *
 * Plain pointers and sizes only.  Every device pointer is caller-owned
 * (HIP device memory); every call enqueues work on the caller's HIP stream
 * (passed as void*; NULL = the null stream) and returns immediately.  Every
 * entry point returns VAMX_OK (0) or a negative status; the message of the
 * last failure on the calling thread is returned by vamx_last_error().
 *
 * Layouts (float32, C order, the reference's NumPy layouts):
 *   volume    (ny, nx, nz)        z contiguous     vamtoolbox/geometry.py:223-225
 *   sinogram  (nu, n_angles, nz)  z contiguous     vamtoolbox/geometry.py:232-234
 *   ray-trace medium (nx, ny, nz) z contiguous (axis 0 = x) and its sinogram
 *             (nt, n_angles, nz)  flat C order    vamtoolbox/projector/pyTorchRayTrace.py:1522-1540
 */
#ifndef VAMX_H
#define VAMX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VAMX_OK 0
#define VAMX_ERR_ARG (-1)      /* invalid argument / shape                      */
#define VAMX_ERR_HIP (-2)      /* a HIP runtime call failed                     */
#define VAMX_ERR_STATE (-3)    /* plan of the wrong kind or not initialised     */

typedef struct vamx_plan vamx_plan;

/* ---------------------------------------------------------------------------
 * Parallel-beam projector (reference discretisation "D1" plus extensions)
 * ------------------------------------------------------------------------- */
#define VAMX_PAR_NO_WARP_CLIP 1   /* forward: skip the per-slice output clip of
                                     skimage.transform.warp(..., clip=True)    */
#define VAMX_PAR_NO_CIRCLE_CLIP 2 /* backward: skip util.data.clipToCircle     */
#define VAMX_PAR_PER_LANE_GEOMETRY 4 /* use the per-lane-geometry kernels instead of
                                        the wavefront-uniform ones (same results
                                        to float32 rounding; for cross-checks)  */

typedef struct vamx_par_geom {
  int32_t n;                /* nx == ny == nu (square slices)                   */
  int32_t nz;               /* slices == detector rows                          */
  int32_t n_angles;
  const double* angles_deg; /* host array [n_angles], copied into the plan      */
  double det_offset;        /* detector-u offset in voxels (0: reference)       */
  double incl_deg;          /* inclination phi (0: reference)                   */
  double mu;                /* absorption per voxel inside the cylinder (0: none) */
  double radius;            /* cylinder radius in voxels; <= 0 -> (n-1)/2        */
  const uint8_t* outside_mask; /* host array [n*n], 1 where clipToCircle zeroes,
                                  computed by the caller with the reference's
                                  formula; copied into the plan                 */
  int32_t flags;            /* VAMX_PAR_* bits                                   */
  const uint8_t* occlusion; /* host array [n*n*nz], non-zero where the attenuation
                               field is > 0 (an insert); NULL = no occlusion.
                               Binary shadowing of Projector3DParallel.py:72-111,
                               147-153, 192-199, 209-216; needs incl_deg == 0;
                               the exact adjoints are then unavailable         */
} vamx_par_geom;

/* Replaces the construction of Projector3DParallelPython / Projector3DParallelCUDAAstra
 * (vamtoolbox/projector/Projector3DParallel.py:57-75,
 *  vamtoolbox/projector/Projector3DParallelCUDA.py:13-38). */
int vamx_par_create(const vamx_par_geom* geom, vamx_plan** out_plan);

/* b = A x.  Replaces Projector3DParallelPython.forward
 * (vamtoolbox/projector/Projector3DParallel.py:113-156) and the ASTRA call
 * astra.create_sino3d_gpu (vamtoolbox/projector/Projector3DParallelCUDA.py:40-52). */
int vamx_par_forward(vamx_plan* plan, const float* vol, float* sino, void* stream);

/* x = B b (voxel-driven, np.interp gather, circle clip).  Replaces
 * Projector3DParallelPython.backward (vamtoolbox/projector/Projector3DParallel.py:158-207)
 * and astra.creators.create_backprojection3d_gpu
 * (vamtoolbox/projector/Projector3DParallelCUDA.py:54-72). */
int vamx_par_backward(vamx_plan* plan, const float* sino, float* vol, void* stream);

/* x = A^T b, the exact transpose of the linear part of vamx_par_forward
 * (autograd adjoint).  The reference has no such operator; it stands in for
 * the transpose that vamtoolbox/optimizer/BCLP.py:372 approximates with
 * P.forward. */
int vamx_par_forward_adjoint(vamx_plan* plan, const float* sino, float* vol, void* stream);

/* b = B^T x, the exact transpose of vamx_par_backward (autograd adjoint of
 * the dose operator; BCLP.py:353-375 uses P.forward in its place). */
int vamx_par_backward_adjoint(vamx_plan* plan, const float* vol, float* sino, void* stream);

/* Work order of the parallel-beam kernels (chosen by an autotuner on the host,
 * see vamtoolbox_amd/parallel.py ParallelPlan.autotune).  No reference
 * counterpart: ASTRA picks its launch shapes internally.  Results are
 * identical for every setting up to float32 rounding: tiles and tap reuse
 * give bitwise equal sinograms; opposite-angle pairing sums the second ray
 * of a pair in reverse sample order. */
typedef struct vamx_par_tuning {
  int32_t fp_tile_angles;   /* forward: angles per tile of consecutive waves (default 10) */
  int32_t fp_tile_rays;     /* forward: rays per tile (default 4)                         */
  int32_t fp_zv;            /* forward: z values per lane, 1, 2 or 4 (default 4)         */
  int32_t fp_reuse;         /* forward: 1 = keep the previous sample's taps in registers (default 1) */
  int32_t bp_tile;          /* back projection: tile edge in 2x2 quads of voxel columns (default 8) */
  int32_t bp_zv;            /* back projection: z values per lane (default 4)            */
  int32_t fp_pair;          /* forward: 1 = project angles a and a + 180 deg in one pass over the
                               taps when 2 * det_offset is an integer and there is no inclination
                               or occlusion (default 1); fp_tile_angles then counts angle pairs */
  int32_t z_split;          /* both: 1 = where a lane would hold one z value (nz odd or buffers
                               not 8-byte aligned) and the 64-z chunk count is even, lanes hold
                               z and z + 64 instead: half the wavefronts (default 1) */
} vamx_par_tuning;
int vamx_par_set_tuning(vamx_plan* plan, const vamx_par_tuning* tuning);
int vamx_par_get_tuning(vamx_plan* plan, vamx_par_tuning* tuning);

/* ---------------------------------------------------------------------------
 * Fused BCLP iteration (vamtoolbox/optimizer/BCLP.py:277-386, 477-534)
 * ------------------------------------------------------------------------- */
#define VAMX_RESP_GLF 0       /* analytical generalised logistic, response.py:191-218 */
#define VAMX_RESP_LINEAR 1    /* response.py:233-241 */
#define VAMX_RESP_IDENTITY 2  /* response.py:245-253 */
#define VAMX_RESP_TABLE 3     /* interpolation tables, response.py:256-279 */

typedef struct vamx_response {
  int32_t kind;               /* VAMX_RESP_*                                      */
  double A, K, B, M, nu;      /* GLF parameters                                   */
  double lin_M, lin_C;        /* linear: map = lin_M * f + lin_C                  */
  int32_t n_pts;              /* table: np.interp nodes (increasing)              */
  const double* table_f;      /* device [n_pts] node positions (interp_f_0)       */
  const double* table_map;    /* device [n_pts] (interp_map_0)                    */
  const double* table_dmap;   /* device [n_pts] (interp_dmapdf_0)                 */
} vamx_response;

typedef struct vamx_bclp_params {
  double p, q;                /* Lp norm and outer power                          */
  double eps;                 /* band half width (used when eps_arr == NULL)      */
  double weight;              /* used when weight_arr == NULL                     */
  double learning_rate;
  double glb, gub;            /* sinogram bounds; use -inf / +inf for "none"      */
  double dvol;                /* differential volume (BCLP.py:184)                */
  int32_t grad_op;            /* 0: gradient through the forward projector, as
                                    BCLP.py:372 does; 1: through the exact
                                    transpose of the back projector            */
} vamx_bclp_params;

/* Workspace bytes needed by vamx_bclp_step for this plan. */
int vamx_bclp_workspace_bytes(vamx_plan* plan, int64_t* bytes);

/* One BCLP gradient-descent iteration on sinogram g (in place):
 *   dose = B g;  m = map(dose);  e = m - target;  b = |e| - eps;  v = b > 0
 *   L = (sum v w |b|^p * dvol)^(q/p)      -> loss_log[*iter_counter]
 *   operand = v w |b|^(p-1) sign(e) map'(dose)
 *   g <- clip(g - lr * s * A operand, glb, gub),  s = q L^((q-p)/q) (1 if L == 0)
 * (A = forward projector, or B^T when prm->grad_op == 1; grad_op must be 0
 * or 1, and 1 returns VAMX_ERR_STATE for a plan with an occlusion map, whose
 * exact transpose is not available) and ++*iter_counter.  dose (device [V]) receives B g of the input g.
 * eps_arr / weight_arr are optional device arrays of the volume's size.
 * iter_counter is a device int32; workspace has vamx_bclp_workspace_bytes.
 * No host synchronisation: the step can be captured in a HIP graph. */
int vamx_bclp_step(vamx_plan* plan, const vamx_response* resp, const vamx_bclp_params* prm,
                   float* g, const float* target, const float* eps_arr, const float* weight_arr,
                   float* dose, double* loss_log, int32_t* iter_counter,
                   void* workspace, void* stream);

/* Loss, dose and operand of the current g without updating it
 * (BCLP.py:277-351 for the final evaluation after the loop, 502-505). */
int vamx_bclp_eval(vamx_plan* plan, const vamx_response* resp, const vamx_bclp_params* prm,
                   const float* g, const float* target, const float* eps_arr,
                   const float* weight_arr, float* dose, float* operand, double* loss_out,
                   void* workspace, void* stream);

/* The update half of vamx_bclp_step: g <- clip(g - lr * s * A operand, glb, gub)
 * with the operand and the gradient scale s that the last vamx_bclp_step, or
 * vamx_bclp_eval called with operand == NULL, left in this workspace
 * (BCLP.py:353-375, 486-487, 532).  vamx_bclp_eval (operand NULL) followed by
 * vamx_bclp_update on one stream is one vamx_bclp_step, bit for bit, without
 * the iteration counter and loss log. */
int vamx_bclp_update(vamx_plan* plan, const vamx_bclp_params* prm, float* g, void* workspace, void* stream);

/* Element-wise material response (ResponseModel.map and .dmapdf,
 * vamtoolbox/response.py:191-291): mapped[i] = map(x[i]) and
 * dmap[i] = map'(x[i]) for device float arrays of length n, evaluated in
 * float64; either output may be NULL.  No host synchronisation. */
int vamx_response_map(const vamx_response* resp, const float* x, float* mapped, float* dmap, int64_t n,
                      void* stream);

/* vamx_bclp_step without a workspace argument and without any allocation:
 * no counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Same kernels and results as vamx_bclp_step; the BCLP
 * operand, an intermediate inside vamx_bclp_step's workspace, is written to
 * the caller's device buffer `operand` [V], and the loss partials and the
 * gradient scale use a 1 MiB device array of the library (one per GPU,
 * held until the process ends): at most 131040 partial slots, i.e. one per
 * 256-thread block of the back-projection epilogue (n*n*nz / 256 on the
 * per-lane path, e.g. 384 x 384 x 227 at most); larger problems fail with
 * VAMX_ERR_ARG before anything is launched.  Calls that overlap on one GPU
 * race on that array: use one stream at a time per GPU.  The plan's device
 * must be the current device. */
int vamx_bclp_step_noalloc(vamx_plan* plan, const vamx_response* resp, const vamx_bclp_params* prm,
                           float* g, const float* target, const float* eps_arr, const float* weight_arr,
                           float* dose, float* operand, double* loss_log, int32_t* iter_counter,
                           void* stream);

/* ---------------------------------------------------------------------------
 * Curved-ray tracer (vamtoolbox/projector/pyTorchRayTrace.py)
 * ------------------------------------------------------------------------- */
#define VAMX_RT_SYMPLECTIC_EULER 0   /* pyTorchRayTrace.py:738-775 */
#define VAMX_RT_FORWARD_EULER 1      /* pyTorchRayTrace.py:777-814 */
#define VAMX_RT_LEAPFROG 2           /* pyTorchRayTrace.py:816-828 + symplectic Euler */
#define VAMX_RT_RK4 3                /* classical RK4 (placeholder in the reference, 289-292) */

#define VAMX_RT_CANONICAL 0          /* dx = v,     dv = n grad n   (pyTorchRayTrace.py:695-705) */
#define VAMX_RT_PHYSICAL 1           /* dx = v/n,   dv = grad n     (708-718) */
#define VAMX_RT_OPTICAL 2            /* dx = v/n^2, dv = grad n / n (721-735) */

#define VAMX_RT_ATT_NONE 0
#define VAMX_RT_ATT_CYLINDER 1       /* medium.py:895-903 */
#define VAMX_RT_ATT_BALL 2           /* medium.py:907-915 */
#define VAMX_RT_ATT_GRID 3           /* grid_sample, zeros padding, medium.py:161-185 */

#define VAMX_RT_INDEX_GRID 0         /* n and grad n sampled from n_grad (medium.py:161-211) */
#define VAMX_RT_INDEX_LUNEBURG 1     /* analytical Luneburg lens (medium.py:486-549) */

typedef struct vamx_rt_geom {
  int32_t nx, ny, nz;            /* medium grid (axis 0 = x)                           */
  double xmin, ymin, zmin;       /* physical coordinates of voxel (0,0,0)              */
  double dx, dy, dz;             /* voxel sizes (dz ignored when nz == 1)               */
  int32_t nt, n_angles, nzr;     /* rays: (nt, n_angles, nzr), C order                  */
  const double* angles_deg;      /* host [n_angles]                                     */
  double incl_deg;
  double step;                   /* parametrisation step (reference: dx / 2)           */
  int32_t max_steps;             /* reference: 3 * nx                                   */
  int32_t exit_every;            /* reference: 10                                       */
  int32_t solver;                /* VAMX_RT_* solver                                    */
  int32_t param;                 /* VAMX_RT_* parametrisation                           */
  double ray_energy0;            /* 1 / ray_density^2 (3-D) or 1 / ray_density (2-D)    */
  /* medium: float4 (n, dn/dx, dn/dy, dn/dz) per voxel, device, caller-owned,
   * sampled trilinearly with border padding (medium.py:161-211)          */
  const float* n_grad;
  int32_t att_kind;              /* attenuation model (VAMX_RT_ATT_*)                   */
  double att_alpha, att_radius;  /* analytic attenuation                                 */
  const float* att_grid;         /* device [nx*ny*nz] for VAMX_RT_ATT_GRID               */
  const float* abs_grid;         /* device [nx*ny*nz] absorption at grid points (alpha_abs) */
  int32_t index_kind;            /* VAMX_RT_INDEX_*; with VAMX_RT_INDEX_LUNEBURG, n_grad may be NULL
                                    and n(x) = n_sur (2 - (|x|/R)^p)^(1/p) for |x| < R, n_sur outside */
  double index_R, index_p, index_n_sur;   /* Luneburg lens: physical radius, exponent, surrounding index */
} vamx_rt_geom;

/* Replaces PyTorchRayTracingPropagator.__init__ / RayState.setupRaysParallel
 * (vamtoolbox/projector/pyTorchRayTrace.py:33-85, 1515-1666). */
int vamx_rt_create(const vamx_rt_geom* geom, vamx_plan** out_plan);

/* g = W (alpha_abs * f): replaces PyTorchRayTracingPropagator.forward
 * (pyTorchRayTrace.py:87-107 -> integrateEnergyUntilExit 471-571). */
int vamx_rt_integrate(vamx_plan* plan, const float* f, float* g, void* stream);

/* dose = alpha_abs * (W^T g): replaces PyTorchRayTracingPropagator.backward
 * (pyTorchRayTrace.py:109-133 -> depositEnergyUntilExit 378-469).  The
 * deposit is deterministic (64-bit fixed-point accumulation). */
int vamx_rt_deposit(vamx_plan* plan, const float* g, float* dose, void* stream);

/* Ray-step statistics of the last integrate/deposit call on this plan:
 * steps[0] = total ray-steps executed, steps[1] = rays still active at
 * max_steps (the reference logs an error for those, pyTorchRayTrace.py:566-569).
 * Synchronises the plan's stream. */
int vamx_rt_stats(vamx_plan* plan, int64_t* steps);

/* Workspace bytes of vamx_rt_bclp_step for this plan. */
int vamx_rt_bclp_workspace_bytes(vamx_plan* plan, int64_t* bytes);

/* One BCLP iteration on the ray tracer (BCLPNorm with
 * PyTorchRayTracingPropagator, vamtoolbox/optimizer/BCLP.py:277-386, 477-534;
 * pyTorchRayTrace.py:87-133): dose = alpha_abs W^T g (the deposit), the
 * response, band and Lp weighting fused into the deposit's finalisation,
 * the loss logged at loss_log[*iter_counter], and the gradient
 * W (alpha_abs operand) fused with g <- clip(g - lr * s * gradient) in the
 * integrate pass; ++*iter_counter.  g is the flat ray sinogram (nt, n_angles,
 * nzr), target and dose the medium grid (nx, ny, nz).  prm->grad_op must be
 * 0 (the forward is the exact transpose of the deposit).  No host
 * synchronisation: the step can be captured in a HIP graph. */
int vamx_rt_bclp_step(vamx_plan* plan, const vamx_response* resp, const vamx_bclp_params* prm,
                      float* g, const float* target, const float* eps_arr, const float* weight_arr,
                      float* dose, double* loss_log, int32_t* iter_counter, void* workspace, void* stream);

/* Loss, dose and operand of the current ray sinogram g on the ray tracer,
 * without updating g (BCLPNorm.updateVariables / computeLoss with
 * PyTorchRayTracingPropagator, BCLP.py:277-351): the first half of
 * vamx_rt_bclp_step.  operand (device [V], or NULL to leave it in the
 * workspace) holds v w |e - eps|^(p-1) sign(e) map'(dose) without alpha_abs,
 * so that q L^((q-p)/q) vamx_rt_integrate(operand) is the gradient of L(g);
 * loss_out is a device double.  No host synchronisation. */
int vamx_rt_bclp_eval(vamx_plan* plan, const vamx_response* resp, const vamx_bclp_params* prm,
                      const float* g, const float* target, const float* eps_arr, const float* weight_arr,
                      float* dose, float* operand, double* loss_out, void* workspace, void* stream);

/* Propagation matrix of the ray tracer (PyTorchRayTracingPropagator.buildPropagationMatrix,
 * pyTorchRayTrace.py:168-197 -> recordEnergyUntilExit 573-690), in two passes
 * over the same traces:
 *   vamx_rt_record_counts: counts[ray] = number of interaction taps of the ray
 *                          (device int64 [n_rays]);
 *   vamx_rt_record:        with offsets = exclusive prefix sum of counts, writes the
 *                          ray's taps (voxel index, weight = step * E * trilinear weight)
 *                          to voxel/weight[offsets[ray] ...] in step order.
 * Summing equal (ray, voxel) pairs and scaling column v by alpha_abs[v] gives the
 * matrix P with P f = vamx_rt_integrate(f) and P^T g = vamx_rt_deposit(g). */
int vamx_rt_record_counts(vamx_plan* plan, int64_t* counts, void* stream);
int vamx_rt_record(vamx_plan* plan, const int64_t* offsets, int32_t* voxel, float* weight, void* stream);

/* ---------------------------------------------------------------------------
 * Image-sequence export (vamtoolbox/imagesequence.py:85-399)
 * ------------------------------------------------------------------------- */
#define VAMX_IMGSEQ_MAX_COPIES 64

/* A sinogram-shaped float64 device array (nu, na, nv), v fastest, rotated in
 * the (u, v) plane like imagesequence.py:397-399 (scipy.ndimage.rotate,
 * axes (0, 2), order 1, constant 0 outside the input): out (mu, na, mv) at
 * (i, j, k) samples the input at (matrix[0] i + matrix[1] k + offset[0],
 * matrix[2] i + matrix[3] k + offset[1]); the caller computes mu, mv, matrix
 * and offset as scipy does for reshape=True.  flip_u / flip_v reverse the
 * input's u / v axes first (_invertU / _invertV, 363-370). */
int vamx_imgseq_rotate(const double* in, int64_t nu, int64_t na, int64_t nv, int32_t flip_u, int32_t flip_v,
                       const double* matrix, const double* offset, double* out, int64_t mu, int64_t mv,
                       void* stream);

/* Size scaling of each angle's (u, v) image (_scaleSize, imagesequence.py:
 * 373-384: OpenCV resize with INTER_LINEAR) from (nu, na, nv) to (mu, na, mv),
 * float64; flips as in vamx_imgseq_rotate. */
int vamx_imgseq_resize(const double* in, int64_t nu, int64_t na, int64_t nv, int32_t flip_u, int32_t flip_v,
                       double* out, int64_t mu, int64_t mv, void* stream);

typedef struct vamx_imgseq_params {
  int32_t n_u, n_v;           /* frame width and height (multiples of 4)             */
  int32_t u1;                 /* first frame column of every copy                    */
  int32_t n_copies;           /* 1 .. VAMX_IMGSEQ_MAX_COPIES                          */
  int32_t v1[VAMX_IMGSEQ_MAX_COPIES];  /* first frame row of each copy, insertion order */
  int32_t bit_depth;          /* 1 .. 32                                             */
  int32_t normalize;          /* 1: x / norm_value * (2^bit_depth - 1)               */
  double norm_value;          /* the normalisation percentile of the image           */
  double intensity_scale;     /* applied when != 1                                   */
  int32_t flip_u, flip_v;     /* reverse the image's u / v axes on read              */
} vamx_imgseq_params;

/* ImageSeq.__init__ (imagesequence.py:85-182) from the (possibly rotated and
 * scaled) float64 image array (mu, na, mv): frames (na, n_v, n_u) uint8,
 * frame j holding copies of the transposed (v, u) image of angle j at rows
 * [v1[c], v1[c] + mv) and columns [u1, u1 + mu), a later copy overwriting an
 * earlier one, zero elsewhere.  Each sample is normalised, scaled, truncated at
 * 2^bit_depth - 1, cast toward zero to the bit depth's unsigned type and kept
 * modulo 256 (NumPy's assignment into the reference's uint8 frames); NaN and
 * negative samples give 0.  frames must be 4-byte aligned. */
int vamx_imgseq_compose(const double* img, int64_t mu, int64_t na, int64_t mv, const vamx_imgseq_params* prm,
                        uint8_t* frames, void* stream);

/* The order statistics np.percentile (linear method) interpolates between,
 * for ImageSeq's percentile normalisation (np.percentile(mod_sinogram, q),
 * imagesequence.py:128-131): out[0] = sorted(x)[rank], out[1] =
 * sorted(x)[rank + 1] (sorted(x)[rank] when rank = n - 1), both NaN when x
 * holds a NaN.  x and out (2 doubles) are device memory; workspace has
 * vamx_order_stats_workspace_bytes and is 8-byte aligned.  Exact (a radix
 * selection on the values' bits), no host synchronisation. */
int vamx_order_stats_workspace_bytes(int64_t* bytes);
int vamx_order_stats(const double* x, int64_t n, int64_t rank, double* out, void* workspace, void* stream);

/* ---------------------------------------------------------------------------
 * OSMO element-wise steps (vamtoolbox/optimizer/OSMO.py:36-110)
 * ------------------------------------------------------------------------- */
/* The model update of minimizeOSMO's stepVoid / stepGel on n voxels (device
 * float32 x and x_model, uint8 mask): gel = 0: x_model -= max(x - d, 0), gel =
 * 1: x_model += max(d - x, 0), on the voxels where mask is nonzero
 * (OSMO.py:54-56, 92-94; NaN passes the bound as the reference's masked
 * assignment lets it). */
int vamx_osmo_model_update(const float* x, float* x_model, const uint8_t* mask, int64_t n, float d, int32_t gel,
                           void* stream);
/* v = v / *m, then v[v < lo] = lo, in place (the max normalisations of
 * OSMO.py:60-65 and 67-69 with the device scalar m = max(v)). */
int vamx_scale_clamp(float* v, int64_t n, const float* m, float lo, void* stream);
/* vamx_scale_clamp on k arrays packed in one buffer, each with its own
 * maximum: segment s is v[seg_off[s], seg_off[s+1]) (device int64 offsets,
 * k + 1 of them, non-decreasing, off[k] = n; each range is clipped to
 * [0, n)); v[i] = v[i] / max(segment), then v[i] < lo -> lo.  A NaN in a
 * segment makes its maximum NaN.  workspace: k uint32, 4-byte aligned.  The batched OSMO's per-problem normalisations
 * (OSMO.py:60-69 for every problem); no reference counterpart as one call. */
int vamx_scale_clamp_segments(float* v, int64_t n, const int64_t* seg_off, int32_t k, float lo, void* workspace,
                              void* stream);

/* calcVER (vamtoolbox/metrics.py:47-87, not layerwise) of a device float32
 * dose of n voxels with device uint8 gel / void masks: *ver (device double)
 * = #(void voxels with dose >= the least gel dose) / num, num = #gel + #void;
 * 0 when a gel dose is NaN.  workspace has vamx_ver_workspace_bytes, 8-byte
 * aligned.  No host synchronisation. */
int vamx_ver_workspace_bytes(int64_t* bytes);
int vamx_ver(const float* dose, const uint8_t* gel, const uint8_t* void_mask, int64_t n, int64_t num, double* ver,
             void* workspace, void* stream);

/* ---------------------------------------------------------------------------
 * Algebraic (sparse-matrix) propagation
 * ------------------------------------------------------------------------- */
/* y = P x for a CSR matrix P (n_rows x n_cols; device row_ptr [n_rows+1] int64,
 * col_idx / values [nnz] int32 / float32) applied to each of z_tiling z layers:
 * x is (n_cols, z_tiling) and y (n_rows, z_tiling), z fastest.  Replaces
 * torch.sparse.mm in PyTorchAlgebraicPropagator.forward / backward
 * (vamtoolbox/projector/pyTorchAlgebraicPropagation.py:152-182; the backward
 * passes the CSR transpose).  Column indices must be < n_cols (not checked on
 * the device).  Deterministic (fixed summation order). */
int vamx_csr_spmm(const int64_t* row_ptr, const int32_t* col_idx, const float* values, int64_t n_rows,
                  int64_t n_cols, int32_t z_tiling, const float* x, float* y, void* stream);

/* ---------------------------------------------------------------------------
 * Sinogram resampling (fan-beam rebinning)
 * ------------------------------------------------------------------------- */
/* out[p, z] = factors[2p+1] * (bilinear(sino[., ., z]) * factors[2p]) at the
 * grid cell (cell[2p], cell[2p+1]) with weights (weights[2p] along the
 * detector, weights[2p+1] along the angles); sino is (n_r, n_a, nz) float32,
 * z fastest, out is (m, nz).  A cell with cell[2p] == -1 samples 0 (a point
 * outside the grid: interpn's fill value); any other cell outside
 * 0 <= cell[2p] <= n_r-2, 0 <= cell[2p+1] <= n_a-2 samples NaN.  The factors
 * are the differential-area factor and the inverse Fresnel transmission.
 * The resampling of rebinFanBeam (vamtoolbox/geometry.py:736-944; interpn at
 * line 805, the factors at 814-821). */
int vamx_sino_resample(const float* sino, int32_t n_r, int32_t n_a, int32_t nz, const int32_t* cell,
                       const float* weights, const float* factors, int64_t m, float* out, void* stream);

/* ---------------------------------------------------------------------------
 * Batched uneven parallel-beam problems in one launch
 * ------------------------------------------------------------------------- */
/* A batch of n_plans independent problems, each a parallel plan created with
 * vamx_par_create; no plan may appear twice (VAMX_ERR_ARG).  The batch owns
 * its device tables of work items, the tile counter of its persistent
 * launches and two side streams, all released by vamx_batch_destroy (destroy
 * the batch before its plans).  A batch is used from one stream at a time;
 * different batches may run concurrently on different streams.  No
 * reference counterpart: VAMToolbox solves one target per call
 * (vamtoolbox/optimize.py:174-217). */
typedef struct vamx_batch vamx_batch;
int vamx_batch_create(vamx_plan* const* plans, int32_t n_plans, vamx_batch** out_batch);

/* Forward-project every problem of the batch: vols[i] / sinos[i] are device
 * pointers of problem i.  Problems share persistent launches (at most 64
 * problems per launch) whose blocks drain a device-side queue of tiles,
 * largest problem first; work is balanced by n*n_angles*n*nz.  A problem
 * whose wavefront work alone fills the GPU (16 waves per CU) is projected by
 * its own single-problem launch instead (results as vamx_par_forward); those
 * launches are spread over the caller's stream and the batch's two side
 * streams, which fork from and join back to it through events, so the call is
 * ordered on `stream` as a whole.  The buffer pointers travel as kernel
 * arguments; the first call with a new buffer alignment or work order
 * uploads a table and synchronises the stream, after which calls can be
 * captured in a HIP graph (the tables live as long as the batch).  Replaces
 * Projector3DParallelPython.forward (Projector3DParallel.py:113-156) for every
 * problem at once. */
int vamx_batch_forward(vamx_batch* batch, const float* const* vols, float* const* sinos, void* stream);

/* Back-project every problem of the batch (one persistent launch per 64
 * problems; Projector3DParallelPython.backward, Projector3DParallel.py:158-207). */
int vamx_batch_backward(vamx_batch* batch, const float* const* sinos, float* const* vols, void* stream);
void vamx_batch_destroy(vamx_batch* batch);

/* The same for a list of plans without an explicit batch: the library keeps
 * one batch per distinct list of plans and releases it when one of those
 * plans is destroyed. */
int vamx_par_forward_batched(vamx_plan* const* plans, int32_t n_plans,
                             const float* const* vols, float* const* sinos, void* stream);
int vamx_par_backward_batched(vamx_plan* const* plans, int32_t n_plans,
                              const float* const* sinos, float* const* vols, void* stream);

/* ---------------------------------------------------------------------------
 * common
 * ------------------------------------------------------------------------- */
void vamx_destroy(vamx_plan* plan);
const char* vamx_last_error(void);
int vamx_version(void);

#ifdef __cplusplus
}
#endif

#endif /* VAMX_H */
