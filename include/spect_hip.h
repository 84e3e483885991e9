/*
 * spect_hip.h -- C ABI of libspect_hip.so, the MI355X (gfx950) kernels behind
 * the SPECT rotate+sum system matrix.
 *
 * The reference (PyTomography 3.3.2) is pure Python with no FFI; its
 * boundary is the duck-typed SystemMatrix / Transform interface.  Each entry
 * point below replaces one reference call chain, cited as
 * src/pytomography/<file>:<line>.  The Python drop-in (pytomography_amd)
 * binds them with ctypes.
 *
 * Conventions
 *  - Every function returns 0 on success or a negative SPECT_E* code; the
 *    message of the last failure on the calling thread is spect_last_error().
 *  - All array arguments are device pointers to C-contiguous buffers owned by
 *    the caller.  Object arrays are [nx][ny][nz] (z fastest), projections are
 *    [n][ny][nz] (the reference's [Ltheta, Lr, Lz]).
 *  - dtype selects float32 (SPECT_F32) or float64 (SPECT_F64) for every
 *    real-valued array of the call.
 *  - Work is enqueued on `stream` (a hipStream_t, NULL = default stream);
 *    nothing synchronises the device, so calls can be captured in a hipGraph.
 *  - A plan is immutable after creation, so calls with the same plan may run
 *    on different streams as long as they do not share a workspace.
 *  - Real-valued arrays must be 16-byte aligned (torch allocations are).
 */
#ifndef SPECT_HIP_H
#define SPECT_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  SPECT_OK = 0,
  SPECT_E_ARG = -1,      /* invalid argument / shape / unsupported combination */
  SPECT_E_HIP = -2,      /* HIP runtime error */
  SPECT_E_WORKSPACE = -3 /* workspace too small */
};

enum { SPECT_F32 = 0, SPECT_F64 = 1 };

/* Flag OR-ed into the dtype argument of spect_forward, spect_backward,
 * spect_backward_transpose and spect_osem_subiteration: att_table then holds
 * the attenuation field of spect_att_field_build instead of the suffix table
 * (mu must still be non-NULL; it is not read). */
enum { SPECT_ATT_FIELD = 0x100 };

/* Padding modes of spect_smooth_axis (torch Conv1d padding_mode). */
enum { SPECT_PAD_CIRCULAR = 0, SPECT_PAD_REPLICATE = 1, SPECT_PAD_ZEROS = 2 };

/* Rotation used by the back projection. */
enum {
  SPECT_BP_REFERENCE = 0, /* gather at 270-beta: spect_system_matrix.py:196 */
  SPECT_BP_ADJOINT = 1    /* exact transpose of the forward bilinear gather */
};

/* What a back projection does with its result. */
enum {
  SPECT_OUT_STORE = 0,   /* out = b */
  SPECT_OUT_ADD = 1,     /* out += b */
  SPECT_OUT_OSEM = 2     /* out = f + f*(1/(n+delta))*(b-n), then clamp <= floor -> floor */
};

typedef struct spect_plan spect_plan;

/* Geometry and transforms of one SPECT problem.
 * Replaces SPECTObjectMeta / SPECTProjMeta (metadata/SPECT/spect_metadata.py:8-64),
 * compute_pad_size (utils/spatial.py:36-45), the PSF layers built by
 * SPECTPSFTransform._configure_simple_model (transforms/SPECT/psf.py:123-142)
 * and the angle handling of SPECTSystemMatrix.forward/backward
 * (projectors/SPECT/spect_system_matrix.py:145-154,188-198). */
typedef struct spect_geom {
  int32_t nx, ny, nz;      /* object shape; nx == ny <= 16384, nz <= 256 */
  int32_t n_angles;        /* number of projections */
  double dx, dy, dz;       /* voxel size in cm (dx scales the attenuation integral) */
  const float *angles_deg; /* host [n_angles]: detector angle beta (float32, as proj_meta.angles) */
  const double *angles_deg64; /* host [n_angles]: beta for float64 runs (NULL: no float64 tables) */
  /* PSF (optional): psf_kr == 0 disables it, psf_kz == 0 disables the z blur ('1D'). */
  int32_t psf_kr, psf_kz;  /* taps per plane (odd) */
  int32_t n_radii;         /* distinct detector radii */
  const int32_t *radius_index; /* host [n_angles] -> [0, n_radii) */
  const float *taps_r;     /* host [n_radii][P][psf_kr], P = nx + 2*pad */
  const float *taps_z;     /* host [n_radii][P][psf_kz] */
  const double *taps_r64;  /* float64 copies (NULL: no float64 PSF) */
  const double *taps_z64;
  int32_t chunk_planes;    /* planes per uniform work chunk (0: default, 16 planes) */
  int32_t sched_rounds;    /* forward and field kernel without attenuation or with the
                              attenuation field: their visited planes are split evenly over
                              sched_rounds x the blocks the GPU holds at once; 0: automatic
                              -- one round when the call's pieces of at most 32 planes fit
                              the resident blocks (the forward: also at most 19 planes per
                              block on average), else uniform chunks; -1: always uniform
                              chunks of chunk_planes planes */
} spect_geom;

const char *spect_last_error(void);
int spect_version(void);

/* Build a plan: uploads trig, linspace and PSF tables to the current device. */
int spect_plan_create(const spect_geom *g, spect_plan **out);
int spect_plan_destroy(spect_plan *p);
/* Padded side P, pad, PSF half width H used by the kernels and chunk count. */
int spect_plan_dims(const spect_plan *p, int32_t *padded, int32_t *pad, int32_t *psf_half,
                    int32_t *chunks);

/* Bytes of workspace needed by forward / backward / osem_subiteration over n angles. */
size_t spect_workspace_bytes(const spect_plan *p, int dtype, int n);

/* The plane schedule an n-angle call runs with this plan, dtype and the
 * device current now (sched_rounds of spect_geom): *forward = 1 when the
 * forward projection (no attenuation, or the attenuation field) runs the
 * balanced schedule, 0 for uniform chunks; *field likewise for the back
 * projection's field kernel.  No counterpart in the reference: a diagnostic of
 * this library's launch choice, for callers that tune or report it. */
int spect_schedule_info(const spect_plan *p, int dtype, int n, int32_t *forward, int32_t *field);

/* Attenuation table: FP64 suffix sums of mu_rot*dx at every chunk boundary
 * for every angle -- the reverse cumsum of utils/misc.py:18 evaluated once
 * per attenuation map (attenuation.py:87-91 recomputes it on every call).
 * Pass it to forward/backward as att_table; NULL makes each call rebuild the
 * rows of its own angles in the workspace. */
size_t spect_att_table_bytes(const spect_plan *p);
int spect_att_table_build(const spect_plan *p, int dtype, const void *mu, void *table,
                          void *stream);

/* Attenuation field: the factor exp(-(S - mu_rot*dx/2)) itself (S = FP64
 * reverse cumsum, as above; attenuation.py:21 with utils/misc.py:18) at every
 * rotated-frame voxel the projectors visit, for every angle, in dtype:
 * [n_angles][P][ny + 2*psf_half][nz] (P = padded width).  With SPECT_ATT_FIELD
 * the forward and back projections read it instead of resampling mu (same
 * factors, bitwise); it trades HBM capacity (n_angles * P * (ny + 2*psf_half) *
 * nz elements) for the per-call gathers of mu. */
size_t spect_att_field_bytes(const spect_plan *p, int dtype);
int spect_att_field_build(const spect_plan *p, int dtype, const void *mu, void *field,
                          void *stream);

/* Forward projection H_m f.
 * Replaces SPECTSystemMatrix.forward (spect_system_matrix.py:120-157) with
 * SPECTAttenuationTransform.forward (transforms/SPECT/attenuation.py:72-92)
 * and SPECTPSFTransform.forward (transforms/SPECT/psf.py:199-214) applied in
 * that order.
 *   f     [nx][ny][nz]         mu [nx][ny][nz] or NULL (no attenuation)
 *   angle_idx  device int32 [n] global angle indices
 *   proj  [n][ny][nz] output (unpadded)                                   */
int spect_forward(const spect_plan *p, int dtype, const void *f, const void *mu,
                  const void *att_table, const int32_t *angle_idx, int n, void *proj,
                  void *workspace, size_t workspace_bytes, void *stream);

/* Largest scratch, in bytes, that one spect_forward_noalloc call may use. */
#define SPECT_NOALLOC_SCRATCH_BYTES ((size_t)128 << 20)

/* spect_forward without a workspace and without any device allocation: the
 * chunk partials (and, when mu comes without att_table, the attenuation rows)
 * live in a scratch array of SPECT_NOALLOC_SCRATCH_BYTES that the library
 * holds on every device it runs on until the process ends.  Same arguments
 * and result as spect_forward minus the workspace; nothing in the scratch has
 * to start at zero.  A call whose scratch would exceed the limit returns
 * SPECT_E_WORKSPACE before launching anything (project fewer angles per
 * call: at 128^3 with the 29-tap PSF, 15 angles need 56 MB).  All calls on
 * one device share that device's scratch, so calls that overlap in time on
 * one device (different streams or threads) race; calls on one stream do not.
 * The current device must be the plan's device.
 * No counterpart in the reference: it is for callers that cannot allocate
 * device memory -- code under hipGraph capture, tight serving loops and
 * harnesses that forbid hipMalloc. */
int spect_forward_noalloc(const spect_plan *p, int dtype, const void *f, const void *mu,
                          const void *att_table, const int32_t *angle_idx, int n, void *proj,
                          void *stream);

/* Back projection.
 * mode SPECT_BP_REFERENCE replaces SPECTSystemMatrix.backward
 * (spect_system_matrix.py:159-201: PSF (psf.py:216-231), attenuation
 * (attenuation.py:94-115), rotation by 270-beta); SPECT_BP_ADJOINT is the
 * exact transpose of spect_forward.
 * out_mode SPECT_OUT_OSEM fuses the OSEM update of
 * PreconditionedGradientAscentAlgorithm.__call__ (algorithms/
 * preconditioned_gradient_ascent.py:119-122) with OSEM's preconditioner
 * (:319-329): out = f + (f*(1/(norm+delta)))*(b - norm), then values <= floor
 * become floor.  f/norm are ignored for the other out modes.             */
int spect_backward(const spect_plan *p, int dtype, int mode, int out_mode, const void *proj,
                   const void *mu, const void *att_table, const int32_t *angle_idx, int n,
                   void *out, const void *f, const void *norm, double delta, double floor_value,
                   void *workspace, size_t workspace_bytes, void *stream);

/* Transpose of the reference back projection: proj = B_m^T v with
 * B_m = spect_backward(REFERENCE) (spect_system_matrix.py:159-201), i.e. the
 * forward projector with the rotation replaced by the transpose of the
 * gather at 270-beta (spect_system_matrix.py:196).  No counterpart in the
 * reference; needed to differentiate through OSEM/BSREM iterations.
 *   v [nx][ny][nz]  ->  proj [n][ny][nz]                               */
int spect_backward_transpose(const spect_plan *p, int dtype, const void *v, const void *mu,
                             const void *att_table, const int32_t *angle_idx, int n, void *proj,
                             void *workspace, size_t workspace_bytes, void *stream);

/* One OSEM sub-iteration in place on f (the body of the subset loop of
 * preconditioned_gradient_ascent.py:103-122 for OSEM with
 * PoissonLogLikelihood.compute_gradient, poisson_log_likelihood.py:32-36):
 *   pred = H_m f + s;  r = g / (pred + delta);  b = B_m r (reference back
 *   projection);  f = f + (f*(1/(norm+delta)))*(b - norm);  f[f<=floor] = floor.
 * g, s (NULL: no additive term) and pred are [n][ny][nz]; pred may be NULL. */
int spect_osem_subiteration(const spect_plan *p, int dtype, void *f, const void *mu,
                            const void *att_table, const int32_t *angle_idx, int n, const void *g,
                            const void *s, const void *norm, void *pred, double delta,
                            double floor_value, void *workspace, size_t workspace_bytes,
                            void *stream);

/* Vector-Jacobian product of the forward projection proj = H(mu) f for the
 * differentiable projector (no counterpart in the reference, whose transforms
 * all run under torch.no_grad: transforms/SPECT/attenuation.py:72,94,
 * transforms/SPECT/psf.py:199,216):
 *   grad_f  = H(mu)^T u       (exact transpose of spect_forward)
 *   grad_mu = (d proj / d mu)^T u   (through the reverse cumsum and exp of
 *             utils/misc.py:18, attenuation.py:21; NULL to skip).
 * u is [n][ny][nz]; grad_f, grad_mu are [nx][ny][nz]; the attenuation rows
 * are rebuilt from mu (mu may change between calls).                    */
size_t spect_grad_workspace_bytes(const spect_plan *p, int dtype, int n);
int spect_forward_vjp(const spect_plan *p, int dtype, const void *u, const void *f, const void *mu,
                      const int32_t *angle_idx, int n, void *grad_f, void *grad_mu,
                      void *workspace, size_t workspace_bytes, void *stream);

/* ---------------------------------------------------------------------------
 * Batched OSEM over many independent problems (no counterpart in the
 * reference, which runs one reconstruction at a time through the Python loop
 * of preconditioned_gradient_ascent.py:103-122).  Every sub-iteration of every
 * problem is the spect_osem_subiteration of that problem; the batch launches
 * each kernel once per sub-iteration for all problems that share the PSF
 * width and z extent, and can capture the whole schedule in a hipGraph. */
typedef struct spect_batch spect_batch;

typedef struct spect_batch_problem {
  const spect_plan *plan;
  void *f;                          /* [nx][ny][nz], updated in place */
  const void *mu;                   /* attenuation map or NULL */
  const void *att_table;            /* spect_att_table_build(mu) result; required with mu */
  int32_t att_field;                /* != 0: att_table is a spect_att_field_build(mu) result */
  int32_t n_subsets;                /* equal for all problems of a batch */
  const int32_t *const *angle_idx;  /* host [n_subsets] of device int32 [count] */
  const int32_t *counts;            /* host [n_subsets] */
  const void *const *g;             /* host [n_subsets] of device [count][ny][nz] */
  const void *const *s;             /* host [n_subsets] of device [count][ny][nz] or NULL */
  const void *const *norm;          /* host [n_subsets] of device [nx][ny][nz] */
  void *const *pred;                /* host [n_subsets] of device [count][ny][nz] or NULL */
  void *workspace;                  /* >= spect_workspace_bytes(plan, dtype, max count) */
  size_t workspace_bytes;
} spect_batch_problem;

/* Problems may share read-only inputs (plan, mu, attenuation data, g, s,
 * norm); each needs its own f and workspace: spect_batch_create returns
 * SPECT_E_ARG when two of them overlap. */
int spect_batch_create(const spect_batch_problem *probs, int n, int dtype, double delta,
                       double floor_value, spect_batch **out);
/* n_iters full iterations (all subsets, all problems).  use_graph != 0 captures
 * the schedule into a hipGraph on the first call (per n_iters) and replays it;
 * the groups (problems of one PSF width and z extent) are parallel branches of
 * that graph, so one group's launches overlap another's.  use_graph == 0 runs
 * the groups one after another on `stream`. */
int spect_batch_run(spect_batch *b, int n_iters, int use_graph, void *stream);
int spect_batch_destroy(spect_batch *b);

/* Poisson ratio of PoissonLogLikelihood.compute_gradient
 * (likelihoods/poisson_log_likelihood.py:34-36):
 *   pred = hf + s (s may be NULL);  ratio = g / (pred + delta).
 * pred may be NULL when the caller does not keep projections_predicted. */
int spect_poisson_ratio(int dtype, const void *g, const void *hf, const void *s,
                        void *pred, void *ratio, int64_t count, double delta, void *stream);

/* Gradient of the relative difference prior
 * (priors/nearest_neighbour.py:44-75,156-183 with EuclideanNeighbourWeight
 * :207-220 and replicate-clamped neighbours utils/misc.py:40-57):
 *   out = beta*scale * sum_s w_rs phi1(f_r, f_s).                        */
int spect_rdp_gradient(int dtype, const void *f, void *out, int32_t nx, int32_t ny, int32_t nz,
                       double dx, double dy, double dz, double beta, double gamma,
                       double delta, double scale, void *stream);

/* Transposed Jacobian of spect_rdp_gradient times u (the RDP Hessian-vector
 * product), built from RelativeDifferencePrior.phi2_1 / phi2_2
 * (priors/nearest_neighbour.py:179-183) over the replicate-clamped 26
 * neighbours:  out = beta*scale * J_r^T u.  No counterpart in the reference
 * (its order-2 prior operator, nearest_neighbour.py:104-107, is not transposed). */
int spect_rdp_hessian_vjp(int dtype, const void *f, const void *u, void *out, int32_t nx,
                          int32_t ny, int32_t nz, double dx, double dy, double dz, double beta,
                          double gamma, double delta, double scale, void *stream);

/* The prior's derivative of order 2 applied to u: NearestNeighbourPrior.__call__(2)
 * (priors/nearest_neighbour.py:105-108) for the RDP, i.e.
 * out = beta*scale * [u * sum_s w phi2_1(f, f_s) + sum_s w phi2_2(f, f_s) u_s]
 * (replicate-clamped neighbours; FOV scale one). */
int spect_rdp_hessian_jvp(int dtype, const void *f, const void *u, void *out, int32_t nx,
                          int32_t ny, int32_t nz, double dx, double dy, double dz, double beta,
                          double gamma, double delta, double scale, void *stream);

/* Linear preconditioned update of PreconditionedGradientAscentAlgorithm.__call__
 * (preconditioned_gradient_ascent.py:119-122) for BSREM's preconditioner
 * (:461-474):  f = f + (f*(relax/(w*norm_sum + delta)))*((b - norm) - prior);
 * values <= floor become floor.  prior may be NULL (treated as 0).      */
int spect_bsrem_update(int dtype, void *f, const void *b, const void *norm,
                       const void *norm_sum, const void *prior, int64_t count,
                       double relax, double weight, double delta, double floor_value,
                       void *stream);

/* Energy-window scatter estimate, compute_EW_scatter (utils/scatter.py:52-99)
 * before smoothing: estimate = (lower/wl*kl + upper/wu*ku)*wp, and, when
 * variance_diag is not NULL, the diagonal of its variance estimate
 * (wp/wl*kl)^2*lower + (wp/wu*ku)^2*upper (:86).  upper may be NULL (DEW:
 * its terms are 0).  Element-wise over count values. */
int spect_ew_scatter(int dtype, const void *lower, const void *upper, void *estimate,
                     void *variance_diag, int64_t count, double width_lower, double width_upper,
                     double width_peak, double weighting_lower, double weighting_upper,
                     void *stream);

/* One axis of get_smoothed_scatter (utils/scatter.py:10-49): the 'same'
 * correlation of an [n0, n1, n2] array along `axis` with K (odd) device-resident
 * taps, padding mode SPECT_PAD_*.  in and out must not overlap. */
int spect_smooth_axis(int dtype, const void *in, void *out, int32_t n0, int32_t n1, int32_t n2,
                      int32_t axis, const void *taps, int32_t K, int32_t mode, void *stream);

/* Kernel-EM basis K applied to a coefficient image, KEMTransform.forward /
 * backward (transforms/shared/kem.py:113-146; K is symmetric so both are this):
 * out(v) = sum_{i,j,k in [-m,m]} kernel[i][j][k](v) * f(clamp(v + (i,j,k))),
 * m = (size-1)/2, with kernel laid out [size^3][nx*ny*nz] and neighbours
 * replicate-clamped (utils/misc.py:40-57).  f and out must not overlap. */
int spect_kem_apply(int dtype, const void *kernel, const void *f, void *out, int32_t nx,
                    int32_t ny, int32_t nz, int32_t size, void *stream);

/* DVFMotionTransform forward / backward (transforms/shared/motion.py:54-110):
 * out(v) = trilinear sample (zeros outside, align_corners) of f at v + dvf(v),
 * dvf laid out [nx][ny][nz][3].  nx, ny, nz >= 2; f and out must not overlap. */
int spect_dvf_warp(int dtype, const void *f, const void *dvf, void *out, int32_t nx, int32_t ny,
                   int32_t nz, void *stream);

/* ---- whole rotated frames: the per-angle projector path ----------------
 * SPECTSystemMatrix.forward/backward (spect_system_matrix.py:145-156,
 * 187-198) hand each angle's padded, rotated object to the obj2obj
 * transforms.  With transforms other than attenuation followed by PSF (or a
 * proj2proj transform) the projector runs that loop on these volumes,
 * [n][P][P][nz] (P = padded width, axis 1 the ray direction towards the
 * detector at +x after rotation).  n is 1..65535; entries of angle_idx out
 * of range give zero volumes; arrays 16-byte aligned. */

/* out[a] = RotationTransform.backward(pad_object(f), 270 - beta_a)
 * (spect_system_matrix.py:149-151; transforms/shared/spatial.py:39-53) for
 * beta_a = angles[angle_idx[a]]: the bilinear samples spect_forward takes.
 * f [nx][ny][nz] (unpadded; zeros outside). */
int spect_rotate_object(const spect_plan *p, int dtype, const void *f, const int32_t *angle_idx,
                        int n, void *out, void *stream);

/* out = unpad_object(sum_a RotationTransform.forward(vol[a], 270 - beta_a))
 * (spect_system_matrix.py:194-200; transforms/shared/spatial.py:22-37), the
 * rotation and sum of the reference back projection; vol [n][P][P][nz], out [nx][ny][nz] with out_mode
 * SPECT_OUT_STORE or SPECT_OUT_ADD. */
int spect_rotate_back(const spect_plan *p, int dtype, const void *vol, const int32_t *angle_idx,
                      int n, int out_mode, void *out, void *stream);

/* out[a] = get_prob_of_detection_matrix(rotate_detector_z(pad_object(mu),
 * beta_a), dx) (attenuation.py:11-21, 88-91): exp(-(S - mu_rot*dx/2)) with
 * the reverse cumsum S in FP64, the factors spect_att_field_build stores,
 * over the whole frame.  mu [nx][ny][nz]. */
int spect_att_volume(const spect_plan *p, int dtype, const void *mu, const int32_t *angle_idx,
                     int n, void *out, void *stream);

/* out[a] = SPECTPSFTransform.forward(vol[a], angle_idx[a]) (psf.py:199-214;
 * Seperable1DBlurNet, psf.py:27-40): every plane x blurred along y with the
 * plan's taps_r of that plane for the angle's radius, then along z with its
 * taps_z, zeros outside the frame.  The plan must have a PSF; vol != out. */
int spect_psf_volume(const spect_plan *p, int dtype, const void *vol, const int32_t *angle_idx,
                     int n, void *out, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* SPECT_HIP_H */
